// wbvh_gpu.h — device builder of the 8-wide compressed BVH.
//
// Builds on the GPU the same tree, byte for byte, as the host builder
// wbvh_build() (wbvh.h): binned SAH over 16 centroid bins, leaves of at most
// max_leaf primitives, stable partitions, the area-greedy collapse into
// 8-wide nodes in breadth-first order and the same quantisation (the shared
// rules are in wbvh_rules.h).  Replaces the reference's per-render host build
// (api/sheen_c_api.cu:138 calling scene.h:410-481).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "wbvh.h"

struct WBvhDevice {
    WNode* nodes = nullptr;     // device, n_nodes
    uint32_t* prims = nullptr;  // device, n_prims: primitive index per leaf slot
    int n_nodes = 0, n_prims = 0;
    int max_stack = 0, depth = 0;
};

// box: device float[6 * n] (min xyz, max xyz per primitive); skip: device
// flags or null.  Allocates out.nodes / out.prims with hipMalloc (the caller
// frees them).  Runs on `stream` and synchronises it.  false + err on error.
bool wbvh_build_device(const float* box, int n, const unsigned char* skip, int max_leaf, hipStream_t stream,
                       WBvhDevice& out, std::string& err);

// Triangle bounds by the reference's rule (bvh.h:235-242): float[6 * n].
void wbvh_triangle_bounds(const void* triangles, int n, float* box, hipStream_t stream);
