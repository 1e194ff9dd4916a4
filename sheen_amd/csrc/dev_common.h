// dev_common.h — vector math, RNG and the device scene view of the HIP path.
//
// All device arithmetic is compiled with contraction off: every product and
// sum rounds separately, in the reference's operation order, so that hit
// distances are reproducible bit for bit on the CPU.  normalize(v) is
// v * (1/sqrt(v.v)) with correctly rounded sqrt and division.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sheen_c_api.h"
#include "sheen_math.h"
#include "wbvh.h"

#define SHD __device__ __forceinline__
#define SHHD __host__ __device__ __forceinline__

struct v3 { float x, y, z; };
SHHD v3 mk3(float x, float y, float z) { return {x, y, z}; }
SHHD v3 cv3(const sheen_f3& f) { return {f.x, f.y, f.z}; }
SHHD v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
SHHD v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
SHHD v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
SHHD v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
SHHD v3 operator*(float s, v3 a) { return {a.x * s, a.y * s, a.z * s}; }
SHHD v3 operator*(v3 a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
SHHD v3 operator/(v3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
SHHD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SHHD v3 cross(v3 a, v3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
SHHD v3 normalize(v3 v) {
    float inv = 1.0f / sqrtf(dot(v, v));
    return {v.x * inv, v.y * inv, v.z * inv};
}

// Transcendentals of the shading, light-sampling, media and photon paths
// (sheen_math.h): evaluated in double by series and rounded once, the same
// bits on the GPU and in the CPU oracle.  Superellipsoid tests and normals
// keep powf: they are part of the hit distance, which matches bit for bit
// with powf on both sides.
SHHD float sin_r(float x) { return shm::sinf_r(x); }
SHHD float cos_r(float x) { return shm::cosf_r(x); }
SHHD float tan_r(float x) { return shm::tanf_r(x); }
SHHD float exp_r(float x) { return shm::expf_r(x); }
SHHD float log2_r(float x) { return shm::log2f_r(x); }
SHHD float pow_r(float x, float y) { return shm::powf_r(x, y); }

// Integer hash and uniform draws (reference sampling/aa.h:32-41).
SHHD uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
SHHD float rand01(uint32_t seed) { return float(mix32(seed) & 0xffffffu) / 16777216.0f; }

// Primitive categories in the reference's closest-hit order
// (integrator/raycast.h:33-68).
enum : int { C_SPHERE = 0, C_STRAND = 1, C_TRI = 2, C_HF = 3, C_SE = 4, C_LATHE = 5, C_PRISM = 6, C_CSG = 7, C_PLANE = 8 };
#define N_BVH_CATS 7

struct PhotonMapView {
    const Photon* photons;  // sorted by cell
    const int* cell_start;
    const int* cell_count;
    int num_cells;
    float cell;
};

// Hierarchy over the subtrees hanging off the CSG root's union spine (the
// maximal set of union nodes reachable from the root through unions).
// Internal node: children a, b (>= 0); leaf: a = -(first + 1) into the
// subtree list, b = count.  Boxes as the CSG node boxes.
struct CsgSpineNode {
    float lo[3];
    int a;
    float hi[3];
    int b;
};

// A load through the scalar cache: the address must be the same in every
// active lane (a value the wave reads once), and the memory read-only while
// the kernel runs.
template <class T>
SHD T sload(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "whole dwords");
    typedef const __attribute__((address_space(4))) uint32_t* CW;
    const CW q = (CW)p;
    T v;
    uint32_t* w = reinterpret_cast<uint32_t*>(&v);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); i++) w[i] = q[i];
    return v;
}

struct LatheFrustum;  // dev_prims.h

// Device scene: everything a kernel reads, passed by value.
struct DScene {
    const WNode* bvh[N_BVH_CATS];
    const uint32_t* prim_ids[N_BVH_CATS];  // leaf slot -> primitive index
    const float* bounds[N_BVH_CATS];       // 6 floats per primitive (reference bounds rules); tri: unused
    int count[N_BVH_CATS];
    int cat_opaque[N_BVH_CATS];            // 1 = no material of the category can pass a shadow ray
    const float4* tri_leaf;                // 3 float4 per triangle leaf slot: v0|idx, v1, v2
    float box_k[N_BVH_CATS][3];            // per BVH and axis: max of 2|origin| + 2 Q scale over its nodes

    const Object* objects;
    const SweepStrand* strands;
    const SweepPoint* sweep_points;
    const Triangle* triangles;
    int num_triangles;
    const HeightField* hfs;
    const Superellipsoid* ses;
    const Lathe* lathes;
    const LathePoint* lathe_points;
    const LatheFrustum* lathe_frusta;  // per profile point: the frustum to the next point (lathe_frustum)
    const float2* lathe_yr;            // per lathe: lowest and highest profile y of its frusta (+-inf: any)
    // lathes of the BVH in the order lathe_scan tests them, widest profile
    // first (lathe_scan of them; 0: the tree is traversed)
    const int* lathe_order;
    int lathe_scan;
    const Prism* prisms;
    const PrismPoint* prism_points;
    const CSGNode* csg;
    const float* csg_box;  // 6 floats per CSG node (min xyz, max xyz): contains every span of its subtree
    int csg_root;
    // 1: lathes, prisms and CSG leaves all have materials that block shadow
    // rays (query_opaque may then ask them in any order)
    int opaque_tail;
    int num_csg;
    // union spine of the CSG root (csg_spine != 0): subtrees as {node, in-order
    // rank, frame depth, 0}; the first spine_always of them have unbounded
    // boxes and are always evaluated, the rest are reached through spine_nodes
    int csg_spine;
    const CsgSpineNode* spine_nodes;
    const int4* spine_sub;
    int spine_n_nodes;
    int spine_always;
    const Plane* planes;
    int num_planes;
    const Material* materials;
    int num_materials;
    const TextureRef* textures;
    int num_textures;
    // float texel chains (same element layout as the 8-bit chains), set by the
    // differentiable entry points; null for the reference's 8-bit textures
    const float* tex_f;
    const uint32_t* tex_f_off;  // per texture: element offset of level 0
    const Light* lights;
    int num_lights;
    const Medium* media;
    int num_media;
    PhotonMapView pm;
    float caustics;
    int any_strands;
    int any_beams;
    int wave_one_scene;  // 1: every lane of a kernel traces this scene (a single context); 0 in batch tables
    CameraData cam;
};

struct DRay {
    v3 o, d;
    float tmin, tmax;
};

// Hit record (reference types.h:34-44 semantics; kind/index added).
struct DHit {
    float t;
    int material_id;
    int object_id;
    int kind, index;
    v3 point, normal, tangent;
    float uvx, uvy;
};
