// wbvh.h — 8-wide compressed BVH node shared by the host builder and the HIP
// traversal kernels.
//
// One node is exactly one 128-byte HBM/L2 line:
//   [0,12)   quantisation origin (float x,y,z) = lower corner of the node box
//   [12,15)  per-axis scale exponent bytes: scale = 2^(e-127), i.e. the float
//            whose bit pattern is e<<23
//   [15]     slot counts: low nibble = used child slots nc, high nibble =
//            internal children ni; slots 0..ni-1 hold the internal children,
//            ni..nc-1 the leaves, nc..7 are empty
//   [16,20)  child_base: node index of the first internal child; internal
//            children of a node are stored contiguously, child_base + slot
//   [20,24)  prim_base: first entry in the primitive reference array of the
//            leaf children of this node (stored contiguously, slot order)
//   [24,32)  meta[8]: 0xFF empty slot; (cnt<<5)|off with cnt in 1..7 for a
//            leaf of cnt primitives at prim_base+off; (0<<5)|k for the
//            internal child in slot k, at child_base+k
//   [32,128) quantised child boxes, SoA: lo_x[8] hi_x[8] lo_y[8] hi_y[8]
//            lo_z[8] hi_z[8]; each an integer q in 0..WBVH_QMAX held as the
//            bit pattern of the fp16 number q (exact)
// A child bound decodes as fmaf(q, scale, origin).  The builder checks
// every decoded bound against the exact float bound (same fmaf on the host),
// so a decoded child box always contains the boxes below it.
#pragma once
#include <stdint.h>

#define WBVH_WIDTH 8
#define WBVH_EMPTY 0xFFu
#ifndef WBVH_MAX_LEAF
#define WBVH_MAX_LEAF 3
#endif
// a leaf slot's meta byte holds its count in 3 bits and its offset in the
// node's primitive run in 5 bits; a node holds up to 8 leaves
static_assert(WBVH_MAX_LEAF >= 1 && WBVH_MAX_LEAF <= 7 && WBVH_MAX_LEAF * WBVH_WIDTH <= 32,
              "leaf offsets must fit the meta byte's 5 bits");
#define WBVH_QMAX 2048  // largest child-plane integer (fp16-exact)
#ifndef WBVH_BINS
#define WBVH_BINS 16  // binned-SAH bins per axis (32: 0.3 % faster frames, 30 % slower builds)
#endif

struct __attribute__((aligned(128))) WNode {
    float ox, oy, oz;
    uint8_t ex, ey, ez, slots;
    uint32_t child_base;
    uint32_t prim_base;
    uint8_t meta[8];
    uint16_t q[6][8];
};
static_assert(sizeof(WNode) == 128, "WNode must be one 128-byte line");

#ifdef __cplusplus
#include <vector>
struct WBvhHost {
    std::vector<WNode> nodes;
    std::vector<uint32_t> prims;  // original primitive index per leaf slot
    int max_stack = 0;            // worst-case traversal stack entries
    int depth = 0;                // wide levels
    int leaf_count = 0;
};

// Builds a wide BVH over n primitive boxes.  bmin/bmax are float[3*n].
// Binned SAH with `bins` bins over the centroid extent of each axis, leaves of
// at most max_leaf primitives, then collapsed into 8-wide nodes by repeatedly
// opening the largest-area internal child.  Primitives listed in `skip`
// (nullable, n flags) are left out.
void wbvh_build(const float* bmin, const float* bmax, int n, const unsigned char* skip, int max_leaf, int bins,
                WBvhHost& out);
#endif
