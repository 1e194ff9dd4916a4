/*
 * sheen_scenes.h — seeded synthetic scene generators (libsheen_host.so).
 *
 * No counterpart in the reference, which builds scenes through its Scene
 * class (reference scene.h:26-481) and the .sheen parser
 * (api/scene_io.cpp:216-968).  These generators produce the benchmark and
 * parity workloads as a ready SheenScene that any sheen_c_api entry point
 * (or the reference library) accepts.  All arrays are owned by the handle.
 *
 * Workloads (SheenSceneParams.config):
 *   1  spheres: 64 spheres, centres U[-10,10]^3, radii U[0.2,1.0], ground
 *      plane y=-10, one point light, Lambert+Phong albedo U[0.1,0.9],
 *      a seeded quarter of the spheres reflective; 256x256, 1 spp.
 *   2  mesh: icosahedral geodesic sphere of frequency `detail` (20*f^2
 *      triangles; f=229 gives 1,048,820), radially displaced by seeded
 *      6-octave value-noise fBm of amplitude 0.15, smooth FP32 normals,
 *      70% diffuse / 15% mirror / 15% glass (IOR 1.5) per triangle,
 *      2 point lights + 1 area light; 1920x1080, 4 spp.
 *   3  analytic primitives + CSG + caustics (see sheen_scenes.cpp).
 *   4  one scene of the uneven batch: seeded triangle count, resolution
 *      and spp (sheen_scene_view); the batch uses seeds 0..255 + base.
 *   5  shading: `detail` spheres (default 10,000) with per-sphere albedo
 *      U[0.05,0.95], a 1024x1024 value-noise texture on the ground plane,
 *      3 sun lights with intensity U[0.5,2.0]; 512x512, 8 spp.
 *   100+k  small parity scenes exercising one feature each (see source).
 */
#ifndef SHEEN_SCENES_H
#define SHEEN_SCENES_H

#include "sheen_c_api.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct SheenSceneParams {
    int config;       /* workload id, see above */
    uint64_t seed;    /* master seed; 0 selects 0x5EED0000 + config */
    int width;        /* image size and samples the workload is quoted at */
    int height;
    int aa;
    int detail;       /* size knob: cfg1 sphere count, cfg2 geodesic frequency,
                         cfg3 primitive count, cfg5 sphere count */
    float scale;      /* parameter scale for perturbed targets (cfg5); 0 = 1 */
    int reserved[7];
} SheenSceneParams;

/* Fills the defaults of a workload.  0, or -1 for an unknown config. */
int sheen_scene_defaults(int config, SheenSceneParams* out);

/* Generates a scene; NULL on bad parameters. */
void* sheen_scene_generate(const SheenSceneParams* params);

/* The generated scene; valid until sheen_scene_free. */
const SheenScene* sheen_scene_get(void* handle);
/* A generator that picks its own image size and spp (config 4): 1 and the
 * values, else 0. */
int sheen_scene_view(void* handle, int* width, int* height, int* aa);

/* Mutable material table of a generated scene (for parameter sweeps). */
Material* sheen_scene_materials(void* handle, int* count);
Light* sheen_scene_lights(void* handle, int* count);
/* Mutable base-level texture pixels (w*h*3 bytes) of texture `tex`. */
unsigned char* sheen_scene_texture(void* handle, int tex, int* w, int* h);

void sheen_scene_free(void* handle);

/* ---- .sheen scene files (reference api/scene_io.cpp) -----------------------
 * sheen_file_load parses a .sheen file (curves tessellated on the host) and
 * always returns a handle; sheen_file_error is "" on success, else the
 * message with its line number.  The scene and the output path stay valid
 * until sheen_file_free.  Relative file names inside the scene are tried
 * as given, then next to the scene file. */
typedef struct SheenFile SheenFile;
SheenFile* sheen_file_load(const char* path);
const char* sheen_file_error(const SheenFile* file);
const SheenScene* sheen_file_scene(const SheenFile* file);
int sheen_file_settings(const SheenFile* file, int* width, int* height, int* aa, const char** output);
void sheen_file_free(SheenFile* file);

#ifdef __cplusplus
}
#endif

#endif /* SHEEN_SCENES_H */
