/*
 * sheen_types.h — plain-old-data scene records shared by the C ABI, the HIP
 * kernels and host tools.  Valid C11 and C++17; no HIP, CUDA or torch types.
 *
 * Every record keeps the byte layout of the reference's types.h (the layout is
 * the drop-in contract a caller builds SheenScene from): 12-byte float3 with
 * 4-byte alignment, 8-byte aligned float2, and 16-byte record alignment.
 * Sizes and offsets are pinned by the static asserts at the bottom, taken from
 * reference tests/sizes_main.cpp:5-21 and the ctypes contract
 * addon_sheen/sheen_render/c_api_contract.py:40-274.
 */
#ifndef SHEEN_TYPES_H
#define SHEEN_TYPES_H

#include <stddef.h>
#include <stdint.h>

#define SHEEN_ALIGN(n) __attribute__((aligned(n)))

#ifdef __cplusplus
extern "C" {
#endif

/* float3 / float2 of the reference (CUDA vector types): 12 B align 4, 8 B align 8. */
typedef struct sheen_f3 { float x, y, z; } sheen_f3;
typedef struct SHEEN_ALIGN(8) sheen_f2 { float x, y; } sheen_f2;

/* reference types.h:49-58 */
typedef struct SHEEN_ALIGN(16) CameraData {
    sheen_f3 position;
    sheen_f3 right;
    sheen_f3 up;
    sheen_f3 forward;
    float focal_length;
    float sensor_width;
    float shift_x;
    float shift_y;
} CameraData;

/* reference types.h:63-85 */
typedef struct SHEEN_ALIGN(16) Material {
    sheen_f3 diffuse_color;
    float specular;
    float shininess;
    float ambient;
    int texture_id;
    int _pad_tex;
    float anisotropy;
    float hair_shift;
    float hair_spec_r;
    float hair_exp_r;
    float hair_spec_t;
    float hair_exp_t;
    float hair_backlight;
    float reflect;
    float transmission;
    float ior;
    sheen_f3 emission_color;
    float emission;
    int mask_id;
    int normal_id;
    float normal_strength;
} Material;

/* reference types.h:90-96 */
typedef struct SHEEN_ALIGN(16) Triangle {
    sheen_f3 v0, v1, v2;
    sheen_f3 n0, n1, n2;
    sheen_f2 uv0, uv1, uv2;
    int material_id;
    int _pad0, _pad1;
} Triangle;

/* reference types.h:100-105 (device-side texture view) */
typedef struct SHEEN_ALIGN(16) TextureRef {
    const unsigned char* pixels;
    int width;
    int height;
    int mips;
} TextureRef;

/* reference types.h:110-116; type 0 = sphere (centre param1, radius param2) */
typedef struct SHEEN_ALIGN(16) Object {
    int type;
    int material_id;
    sheen_f3 param1;
    float param2;
    float _pad0, _pad1, _pad2;
} Object;

/* reference types.h:130-145; type 0 SUN, 1 POINT, 2 AREA, 3 SPOT */
typedef struct SHEEN_ALIGN(16) Light {
    sheen_f3 direction;
    sheen_f3 position;
    sheen_f3 u_axis;
    sheen_f3 v_axis;
    sheen_f3 color;
    float intensity;
    float radius;
    float angle;
    float blend;
    float beam;
    int type;
    int flags;
    int texture_id;
    int _pad1;
} Light;

/* reference types.h:152-159 */
typedef struct SHEEN_ALIGN(16) Medium {
    sheen_f3 center;
    sheen_f3 half;
    sheen_f3 sigma_s;
    float sigma_t;
    int samples;
    int _pad0;
} Medium;

/* reference types.h:164-169; surface n.p + d = 0 */
typedef struct SHEEN_ALIGN(16) Plane {
    sheen_f3 normal;
    float d;
    int material_id;
    int _pad0, _pad1, _pad2;
} Plane;

/* reference types.h:174-187 */
typedef struct SHEEN_ALIGN(16) SweepPoint {
    sheen_f3 position;
    float radius;
} SweepPoint;

typedef struct SHEEN_ALIGN(16) SweepStrand {
    int point_offset;
    int point_count;
    int material_id;
    int _pad0;
} SweepStrand;

/* reference types.h:201-217 */
typedef struct SHEEN_ALIGN(16) HeightField {
    sheen_f3 center;
    float size_x;
    float size_z;
    float height;
    float freq;
    int octaves;
    int seed;
    int material_id;
    const unsigned char* image;
    int img_w;
    int img_h;
    int _pad0;
} HeightField;

/* reference types.h:225-237 */
typedef struct SHEEN_ALIGN(16) LathePoint {
    float y;
    float radius;
    int _pad0, _pad1;
} LathePoint;

typedef struct SHEEN_ALIGN(16) Lathe {
    sheen_f3 center;
    int point_offset;
    int point_count;
    int material_id;
    int _pad0;
} Lathe;

/* reference types.h:245-260 */
typedef struct SHEEN_ALIGN(16) PrismPoint {
    float x;
    float z;
    int _pad0, _pad1;
} PrismPoint;

typedef struct SHEEN_ALIGN(16) Prism {
    sheen_f3 center;
    float y_min;
    float y_max;
    float sweep;
    int point_offset;
    int point_count;
    int material_id;
    int _pad0;
} Prism;

/* reference types.h:266-287 */
typedef struct SHEEN_ALIGN(16) CSGInterval {
    float t_enter;
    float t_exit;
    int material_id;
    int leaf_id;
    int kind;
    int _pad0;
} CSGInterval;

typedef struct SHEEN_ALIGN(16) CSGNode {
    int op;          /* 0 union, 1 intersection, 2 difference (left - right) */
    int leaf_kind;   /* -1 interior; 0 sphere (objects[]), 1 superellipsoid */
    int leaf_index;
    int left;
    int right;
} CSGNode;

/* reference types.h:293-301 */
typedef struct SHEEN_ALIGN(16) Superellipsoid {
    sheen_f3 center;
    float a;
    float b;
    float c;
    float n;
    int material_id;
    int _pad0, _pad1;
} Superellipsoid;

/* reference render_kernel.h:10-15 */
typedef struct SHEEN_ALIGN(16) Photon {
    sheen_f3 position;
    sheen_f3 direction;
    sheen_f3 normal;
    sheen_f3 flux;
} Photon;

#ifdef __cplusplus
} /* extern "C" */
#define SHEEN_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define SHEEN_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

/* Layout pins (reference tests/sizes_main.cpp:5-21, HANDOFF.md:149-151). */
SHEEN_STATIC_ASSERT(sizeof(CameraData) == 64, "CameraData 64");
SHEEN_STATIC_ASSERT(sizeof(Material) == 112, "Material 112");
SHEEN_STATIC_ASSERT(offsetof(Material, emission_color) == 72, "Material.emission_color@72");
SHEEN_STATIC_ASSERT(offsetof(Material, emission) == 84, "Material.emission@84");
SHEEN_STATIC_ASSERT(offsetof(Material, mask_id) == 88, "Material.mask_id@88");
SHEEN_STATIC_ASSERT(offsetof(Material, normal_id) == 92, "Material.normal_id@92");
SHEEN_STATIC_ASSERT(offsetof(Material, normal_strength) == 96, "Material.normal_strength@96");
SHEEN_STATIC_ASSERT(sizeof(Triangle) == 112, "Triangle 112");
SHEEN_STATIC_ASSERT(offsetof(Triangle, uv0) == 72, "Triangle.uv0@72");
SHEEN_STATIC_ASSERT(offsetof(Triangle, material_id) == 96, "Triangle.material_id@96");
SHEEN_STATIC_ASSERT(sizeof(TextureRef) == 32, "TextureRef 32");
SHEEN_STATIC_ASSERT(sizeof(Object) == 48, "Object 48");
SHEEN_STATIC_ASSERT(sizeof(Light) == 96, "Light 96");
SHEEN_STATIC_ASSERT(sizeof(Medium) == 48, "Medium 48");
SHEEN_STATIC_ASSERT(sizeof(Plane) == 32, "Plane 32");
SHEEN_STATIC_ASSERT(sizeof(SweepPoint) == 16, "SweepPoint 16");
SHEEN_STATIC_ASSERT(sizeof(SweepStrand) == 16, "SweepStrand 16");
SHEEN_STATIC_ASSERT(sizeof(HeightField) == 64, "HeightField 64");
SHEEN_STATIC_ASSERT(sizeof(LathePoint) == 16, "LathePoint 16");
SHEEN_STATIC_ASSERT(sizeof(Lathe) == 32, "Lathe 32");
SHEEN_STATIC_ASSERT(sizeof(PrismPoint) == 16, "PrismPoint 16");
SHEEN_STATIC_ASSERT(sizeof(Prism) == 48, "Prism 48");
SHEEN_STATIC_ASSERT(sizeof(CSGInterval) == 32, "CSGInterval 32");
SHEEN_STATIC_ASSERT(sizeof(CSGNode) == 32, "CSGNode 32");
SHEEN_STATIC_ASSERT(sizeof(Superellipsoid) == 48, "Superellipsoid 48");
SHEEN_STATIC_ASSERT(sizeof(Photon) == 48, "Photon 48");

#endif /* SHEEN_TYPES_H */
