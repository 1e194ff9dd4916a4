/*
 * sheen_c_api.h — C ABI of the MI355X render path (libsheen_c_api.so).
 *
 * The first block is the reference's binary contract, unchanged: the
 * SheenScene layout (reference sheen_c_api.h:18-87) and the four exported
 * functions (reference sheen_c_api.h:91-107, implemented in
 * api/sheen_c_api.cu:160-365).  A caller that builds SheenScene for the
 * reference library (Blender add-on engine.py:691-766, ctypes contract
 * c_api_contract.py:233-274, CPython module api/sheen_py_module.cpp) can load
 * this library instead without changes.
 *
 * The second block is additive: a resident-scene context for repeated
 * renders, float radiance and primary-hit output for parity checks, ray
 * counters and kernel timings, batched multi-scene rendering and the
 * gradient (adjoint) entry points.  None of these exist in the reference.
 *
 * All pointers are host pointers unless a name says "device".  No HIP, CUDA
 * or torch types appear in any signature; a stream is passed as void*.
 */
#ifndef SHEEN_C_API_H
#define SHEEN_C_API_H

#include "sheen_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference sheen_c_api.h:18-87 — 352 bytes, photons@316, media@328. */
typedef struct SheenScene {
    CameraData camera;

    const Object* objects;
    int num_objects;

    const SweepPoint* sweep_points;
    int num_sweep_points;
    const SweepStrand* sweep_strands;
    int num_sweep_strands;

    const Triangle* triangles;
    int num_triangles;

    const HeightField* heightfields;
    int num_heightfields;

    const Superellipsoid* superellipsoids;
    int num_superellipsoids;

    const LathePoint* lathe_points;
    int num_lathe_points;
    const Lathe* lathes;
    int num_lathes;

    const PrismPoint* prism_points;
    int num_prism_points;
    const Prism* prisms;
    int num_prisms;

    const CSGNode* csg_nodes;
    int csg_root;
    int num_csg_nodes;

    const Plane* planes;
    int num_planes;

    const Material* materials;
    int num_materials;

    const unsigned char* const* texture_pixels; /* [num_textures], w*h*3 RGB, top row first */
    const int* texture_widths;
    const int* texture_heights;
    int num_textures;

    const Light* lights;
    int num_lights;

    int photons;        /* photons emitted per light; 0 disables the caustic path */
    float photon_cell;  /* hash-grid cell size; <= 0 selects 0.25 */
    float caustics;     /* caustic gather scale; 0 disables the gather */

    const Medium* media;
    int num_media;
} SheenScene;

/* ------------------------------------------------------------------------- */
/* Reference entry points                                                     */
/* ------------------------------------------------------------------------- */

/* Replaces reference sheen_c_api.h:91 / api/sheen_c_api.cu:160-180.
 * Selects the HIP device; cuda_arch is accepted and ignored.  0 or -1. */
int sheen_init(int device_id, int cuda_arch);

/* Replaces reference sheen_c_api.h:94 / api/sheen_c_api.cu:182-185.
 * Synchronises the device.  The reference also resets the device
 * (cudaDeviceReset), which would free the allocations of every other HIP
 * user in the process; here sheen_render_scene keeps no device buffer after
 * it returns and contexts free their own in sheen_context_destroy, so there
 * is nothing left to release.  sheen_init may be called again but is not
 * required before the next render. */
void sheen_cleanup(void);

/* Replaces reference sheen_c_api.h:99-104 / api/sheen_c_api.cu:187-361.
 * Synchronous: builds the BVH, uploads, renders and writes an 8-bit sRGB RGB
 * PNG to output_path.  Validation as the reference: scene != NULL,
 * width,height > 0, 1 <= aa_samples <= 64, output_path != NULL,
 * num_materials > 0.  Returns 0, or -1 with sheen_get_last_error() set. */
int sheen_render_scene(const SheenScene* scene, int width, int height,
                       int aa_samples, const char* output_path);

/* Replaces reference sheen_c_api.h:107 / api/sheen_c_api.cu:363-365.
 * Static buffer, valid until the next call into the library. */
const char* sheen_get_last_error(void);

/* ------------------------------------------------------------------------- */
/* Extensions (no counterpart in the reference)                               */
/* ------------------------------------------------------------------------- */

/* Ray counters and timings of the last render on a context.  A "ray" is one
 * traversal call: primary, shadow pass (each glass pass counts), reflection,
 * refraction face, photon bounce. */
typedef struct SheenRayStats {
    uint64_t primary_rays;
    uint64_t shadow_rays;
    uint64_t reflection_rays;
    uint64_t refraction_rays;
    uint64_t photon_rays;
    uint64_t total_rays;
    uint64_t samples;
    /* Filled only with counting enabled (sheen_context_set_counting): wide
     * BVH nodes fetched and leaf primitives tested, over all rays. */
    uint64_t node_fetches;
    uint64_t prim_tests;
    uint64_t shadow_node_fetches;
    uint64_t shadow_prim_tests;
    uint64_t primary_node_fetches;   /* of the primary trace kernel only */
    uint64_t primary_prim_tests;
    uint64_t primary_stack_spills;   /* traversal-stack entries moved out of LDS */
    uint64_t shadow_stack_spills;
    /* GPU time from the first to the last kernel of the frame (hipEvents). */
    float frame_ms;
    /* With profiling enabled (hipEvents on the render stream, no added
     * synchronisation): summed durations of the closest-hit trace kernels
     * (primary + secondary), of the primary trace kernel alone, of the shadow
     * kernel, and of the shading/accumulation kernels. */
    float trace_ms;
    float primary_ms;
    float shadow_ms;
    float shade_ms;
    int trace_launches;
    int shadow_launches;
    float shade_primary_ms;          /* the primary-hit shading kernel alone */
    /* With counting enabled: 64 x the node-visit iterations of the shadow
     * kernel's waves; shadow_node_fetches / shadow_visit_lanes is the share
     * of wave lanes that visit a node in the traversal loop. */
    uint64_t shadow_visit_lanes;
    /* The shadow_* counts and shadow_ms cover the shadow kernel (every
     * request's first pass); the second and third passes of requests that
     * went through glass run in continuation kernels of their own (scenes of
     * spheres, triangles and planes), counted here: */
    uint64_t cont_node_fetches;
    uint64_t cont_prim_tests;
    float shadow_cont_ms;
    int cont_launches;
} SheenRayStats;

typedef struct SheenContext SheenContext;

/* Uploads the scene to the current device and builds its wide BVHs there
 * (binned SAH, one BVH per primitive kind); replaces the reference's host
 * build per render call (api/sheen_c_api.cu:138, scene.h:410-481).  The
 * scene arrays are copied; the caller keeps ownership.  Returns NULL with
 * sheen_get_last_error() set on failure. */
SheenContext* sheen_context_create(const SheenScene* scene);
void sheen_context_destroy(SheenContext* ctx);

/* sheen_context_create with options.  SHEEN_CTX_HOST_BVH: build the BVHs
 * with the host builder instead (the same trees, byte for byte). */
#define SHEEN_CTX_HOST_BVH 1
SheenContext* sheen_context_create_ex(const SheenScene* scene, int flags);

/* The wide BVH of one primitive kind (0 spheres, 1 strands, 2 triangles,
 * 3 height fields, 4 superellipsoids, 5 lathes, 6 prisms).  info[4]: nodes,
 * leaf slots, worst-case traversal stack entries, wide levels.  nodes
 * (128-byte records, node_cap of them) and prims (primitive index per leaf
 * slot) may be NULL to query the sizes. */
int sheen_context_bvh(const SheenContext* ctx, int category, void* nodes, int node_cap, unsigned int* prims,
                      int prim_cap, int* info);

/* Milliseconds sheen_context_create spent on validation, bounds and BVH
 * builds (device builds included, photon pass excluded). */
double sheen_context_build_ms(const SheenContext* ctx);

/* Renders one frame.  out_rgba: host float[width*height*4] or NULL (the
 * frame then stays in device memory, see sheen_context_frame_device).
 * stats may be NULL.  Synchronous; 0 or -1. */
int sheen_context_render(SheenContext* ctx, int width, int height, int aa_samples,
                         float* out_rgba, SheenRayStats* stats);

/* Device pointer to the last frame: float4 per pixel, row-major. */
const float* sheen_context_frame_device(const SheenContext* ctx);

/* Per-sample primary hits of a frame: kind[i], index[i], t[i] for
 * i = (y*width + x)*aa + s.  kind: -1 miss, 0 sphere, 1 strand, 2 triangle,
 * 3 heightfield, 4 superellipsoid, 5 lathe, 6 prism, 7 CSG, 8 plane
 * (the reference's category order, integrator/raycast.h:33-68).  Host arrays. */
int sheen_context_primary_hits(SheenContext* ctx, int width, int height, int aa_samples,
                               int* kind, int* index, float* t);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Renders one frame like sheen_context_render but calls no
 * hipMalloc, hipFree or allocating library routine: every work buffer (sample
 * state, event and shadow queues, counters) lives in one static device array
 * of SHEEN_NOALLOC_SCRATCH_BYTES that each device running the library holds
 * until the process ends, and the frame is rendered in as many chunks of
 * whole 8x8 tiles as that array requires.  out_rgba_device: device memory the
 * caller owns, float[width*height*4], row-major.  The context's device must
 * be the current device.  The work space of one 8x8 tile must fit the array
 * (it grows with aa_samples and the light count); larger needs are refused
 * with an error.  Calls that overlap on one device race on the array, which
 * keeps the last call's contents; profiling events are not recorded.
 * Synchronous; 0 or -1. */
#define SHEEN_NOALLOC_SCRATCH_BYTES 268435456u /* 256 MiB */
int sheen_context_render_noalloc(SheenContext* ctx, int width, int height, int aa_samples, float* out_rgba_device,
                                 SheenRayStats* stats);

/* 1 = record hipEvents around every kernel (SheenRayStats.*_ms). */
int sheen_context_set_profiling(SheenContext* ctx, int enable);
/* 1 = count node fetches / primitive tests (slower; for roofline bytes). */
int sheen_context_set_counting(SheenContext* ctx, int enable);
/* Samples per wavefront chunk (0 = automatic). */
int sheen_context_set_chunk(SheenContext* ctx, int samples_per_chunk);
/* Launch-shape parameters set by the autotuner: "primary_grid", "sec_grid",
 * "shadow_grid", "shade_grid", "shade_sec_grid" (blocks per CU of each
 * grid-stride kernel; 0 = one work item per thread) and "chunk" (samples per
 * chunk, 0 = automatic); "overlap" (1: the primary shading's shadow rays are
 * traced on a second stream while the secondary rays run; default 0: one
 * shadow launch after them; the image is the same either way); "light_runs"
 * (default 1: each light's shadow requests are one run of the queue, traced
 * light after light; 0: one run shared by the lights, in shading order; a
 * scene with primitives other than spheres, triangles and planes always uses
 * one run; the image is the same either way); "shadow_order" (default 1: a
 * shading block queues its shadow requests that leave the surface before
 * those that enter it) and "sec_order" (default 1: its reflection rays
 * before its refraction rays), both only reorder queues; "sec_sort" (0:
 * off; 1: before each trace round the secondary rays are sorted into 16
 * bins, reflection or refraction times the direction's octant; 2: 48 bins,
 * also by the direction's dominant axis; a stable device sort, so a wave
 * traces rays of one kind heading one way; the image is the same either
 * way; default 2 for scenes of spheres, triangles and planes, 0 otherwise);
 * "skip_dark" (default 0; 1: a shadow request whose unshadowed contribution
 * is exactly zero, a light behind the shading normal of a non-hair
 * material, is not traced; the image is the same, the reference traces
 * them and counts them as rays); "shade_block"
 * (workgroup size of the shading kernels, a multiple of 64 up to 1024; 0 =
 * the compiled bound, 1024 for scenes of spheres, triangles and planes, 512
 * otherwise).  Returns 0, or -1 for an unknown name or an invalid value. */
int sheen_context_set_param(SheenContext* ctx, const char* name, int value);

/* ---- Batches of scenes -------------------------------------------------------
 * Many scenes (each with its own camera, image size and spp) rendered as one
 * wavefront pipeline over their concatenated sample slots: per-scene BVHs
 * and records stay separate (one device table of scene descriptors), queues
 * and work buffers are shared, and the frame's launches are captured once
 * into a hipGraph and replayed.  The reference renders one scene per call
 * (sheen_c_api.cu:187-361); sheen_batch_render_serial renders the same
 * scenes that way, one render per scene, for comparison.  Any scene may be
 * batched: when one holds primitives beyond spheres, triangles and planes,
 * media or spot-light beams, the whole batch runs the pipeline built for
 * every kind (slower per ray than the spheres/triangles/planes one); each
 * scene keeps its own photon map.
 */
typedef struct SheenBatch SheenBatch;
SheenBatch* sheen_batch_create(const SheenScene* const* scenes, const int* widths, const int* heights,
                               const int* aa_samples, int count);
void sheen_batch_destroy(SheenBatch* batch);
/* Launch shapes as sheen_context_set_param, plus "graph" (0 = launch the
 * kernels directly instead of replaying the captured graph) and
 * "scene_table" (1..512, default 512: entries of the per-block LDS copy of
 * the scenes' first slots that the kernels search to find a slot's scene;
 * beyond that many scenes the search ends in global memory); "sec_sort" is
 * also set on every scene's own context (sheen_batch_render_serial). */
int sheen_batch_set_param(SheenBatch* batch, const char* name, int value);
/* One frame of every scene.  stats: ray counts of the whole batch, frame_ms
 * (GPU time, hipEvents) and trace_ms (wall clock of the call). */
int sheen_batch_render(SheenBatch* batch, SheenRayStats* stats);
int sheen_batch_render_serial(SheenBatch* batch, SheenRayStats* stats);
/* Float RGBA image of scene `index` from the last batched (serial = 0) or
 * serial (serial = 1) render, to host memory. */
int sheen_batch_image(SheenBatch* batch, int index, int serial, float* out_rgba);
int sheen_batch_info(SheenBatch* batch, int* scenes, long long* samples, int* chunks);

/* ---- Differentiable shading ------------------------------------------------
 * Forward/backward entries for gradient-based inverse rendering (the
 * reference has none; its ABI lists them as additions a replacement should
 * provide).  Parameters, all float32 device arrays:
 *   albedo     [materials * 3]  Material.diffuse_color
 *   emission   [materials * 3]  emitted radiance, emission_color * emission
 *   intensity  [lights]         Light.intensity
 *   texels     [texels]         level-0 texels of every texture in texture
 *                               order, RGB in [0, 1] (8-bit value / 255)
 * The forward renders the context's scene with those values (textures as
 * float chains whose mips are 2x2 means); the image is float RGBA
 * [height * width * 4] in device memory.  The backward replays the same frame
 * (same RNG streams) and returns d(sum grad_image * image)/d(parameter) with
 * visibility held fixed.  Scenes with reflection, transmission, hair, debug
 * materials, caustics, media, beams or primitives other than spheres,
 * triangles and planes are refused (-1, reason in sheen_get_last_error).
 * `stream` is a hipStream_t (NULL = the null stream); work is ordered after
 * earlier work on it and later work on it waits for the results (an event
 * recorded on it and waited on by the library's stream, and the reverse; if
 * either step fails the host synchronises the earlier stream instead, see
 * sheen_context_stream_fallbacks).  The context's own materials, lights and
 * textures are restored after each call.
 */
int sheen_adjoint_sizes(SheenContext* ctx, int* materials, int* lights, long long* texels);
/* Number of caller-stream orderings of this context that fell back to host
 * synchronisation (0 when every event record and wait succeeded). */
int sheen_context_stream_fallbacks(const SheenContext* ctx);
int sheen_adjoint_get_params(SheenContext* ctx, float* albedo, float* emission, float* intensity, float* texels,
                             void* stream);
int sheen_adjoint_forward(SheenContext* ctx, int width, int height, int aa_samples, const float* albedo,
                          const float* emission, const float* intensity, const float* texels, float* image,
                          void* stream);
int sheen_adjoint_backward(SheenContext* ctx, int width, int height, int aa_samples, const float* albedo,
                           const float* emission, const float* intensity, const float* texels,
                           const float* grad_image, float* d_albedo, float* d_emission, float* d_intensity,
                           float* d_texels, void* stream);

/* Photon pass results of a context (scene.photons > 0): photons emitted and
 * stored (reference render_kernel.cu:928-953). */
int sheen_context_photon_stats(SheenContext* ctx, int* emitted, int* stored);
/* Copies the cell-sorted photon map: photons[stored], cell_start[2^18],
 * cell_count[2^18] (any may be NULL).  Returns the photon count or -1. */
int sheen_context_photon_map(SheenContext* ctx, Photon* photons, int* cell_start, int* cell_count);
/* Gather counts around every stride-th stored photon, through the hash map
 * (brute = 0) or by brute force over all photons (brute = 1); reference
 * render_kernel.cu:981-990 and 1092-1107.  Returns the probe count or -1. */
int sheen_context_photon_probe(SheenContext* ctx, int stride, float radius, int max_samples, int brute,
                               int* counts);

/* One-shot float render: create, render, download, destroy. */
int sheen_render_float(const SheenScene* scene, int width, int height, int aa_samples,
                       float* out_rgba, SheenRayStats* stats);

/* Linear -> sRGB 8-bit encoding of a float RGBA frame, as written to PNG
 * (reference png_color.h:11-21).  out_rgb: width*height*3 bytes. */
void sheen_encode_srgb(const float* rgba, int width, int height, unsigned char* out_rgb);

/* Writes an 8-bit RGB PNG.  0 or -1. */
int sheen_write_png(const char* path, int width, int height, const unsigned char* rgb);

#ifdef __cplusplus
} /* extern "C" */
SHEEN_STATIC_ASSERT(sizeof(SheenScene) == 352, "SheenScene 352");
SHEEN_STATIC_ASSERT(offsetof(SheenScene, photons) == 316, "SheenScene.photons@316");
SHEEN_STATIC_ASSERT(offsetof(SheenScene, photon_cell) == 320, "SheenScene.photon_cell@320");
SHEEN_STATIC_ASSERT(offsetof(SheenScene, caustics) == 324, "SheenScene.caustics@324");
SHEEN_STATIC_ASSERT(offsetof(SheenScene, media) == 328, "SheenScene.media@328");
SHEEN_STATIC_ASSERT(offsetof(SheenScene, num_media) == 336, "SheenScene.num_media@336");
#else
SHEEN_STATIC_ASSERT(sizeof(SheenScene) == 352, "SheenScene 352");
SHEEN_STATIC_ASSERT(offsetof(SheenScene, photons) == 316, "SheenScene.photons@316");
SHEEN_STATIC_ASSERT(offsetof(SheenScene, media) == 328, "SheenScene.media@328");
#endif

#endif /* SHEEN_C_API_H */
