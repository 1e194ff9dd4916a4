// render.hip — host side of the HIP render path and the C ABI
// (include/sheen_c_api.h).  One translation unit holds the kernels
// (kernels.h) so the device code is compiled once for gfx950.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "dev_adjoint.h"
#include "dev_photon.h"
#include "kernels.h"
#include "sheen_c_api.h"
#include "host_util.h"
#include "wbvh_gpu.h"

// ---------------------------------------------------------------------------
// Error state (reference api/sheen_c_api.cu:21-26: static 512-byte buffer)
// ---------------------------------------------------------------------------
static char g_error[512] = "no error";
void sheen_set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

#define HIP_OR(expr, ret)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            sheen_set_error("%s failed: %s", #expr, hipGetErrorString(e_));                \
            return ret;                                                                    \
        }                                                                                  \
    } while (0)

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct WorkBufs {
    size_t cap_slots = 0;
    int cap_L = -1, cap_beam = -1;
    bool cap_caustic = false, cap_media = false;
    int4* hit_p = nullptr;
    float4* samp = nullptr;
    float4* ev_base = nullptr;
    float4* ev_caustic = nullptr;
    float4* contrib = nullptr;
    float* trans = nullptr;
    ShadowReq* shadow_q = nullptr;
    ShadowReq* cont[2] = {nullptr, nullptr};  // shadow continuations (second, third pass)
    uint32_t* shadow_nl = nullptr;  // per-light request counts and their snapshot
    SecRay* sec_q[2] = {nullptr, nullptr};
    int4* sec_hits = nullptr;
    uint32_t* sort_hist = nullptr;  // secondary-ray sort: per-bin counts of each queue tile
    float4* beam_add = nullptr;
    float4* media_p = nullptr;
    float4* media_r = nullptr;
    void release() {
        void* ps[] = {hit_p, samp, ev_base, ev_caustic, contrib, trans, shadow_q, cont[0], cont[1], shadow_nl, sec_q[0],
                      sec_q[1], sec_hits, sort_hist, beam_add, media_p, media_r};
        for (void* p : ps) if (p) (void)hipFree(p);
        *this = WorkBufs();
    }
};

struct SheenContext {
    int device = 0;
    int num_cu = 256;
    hipStream_t stream = nullptr;
    std::vector<void*> owned;
    DScene ds{};
    HostSceneInfo info;
    int n_beam = 0;
    bool full = true;  // scene uses a kind beyond spheres, triangles and planes
    // some material reflects or transmits (secondary rays can exist), some
    // transmits above 0.5 (a shadow ray can pass glass): the rounds and
    // kernels for them are launched only then
    bool any_secondary = true, any_glass = true;
    WorkBufs wb;
    WorkBufs* wbp = &wb;  // the work buffers in use (a batch shares one set among its scenes)
    float4* frame = nullptr;
    size_t frame_cap = 0;
    int frame_w = 0, frame_h = 0;
    Counters* ctr = nullptr;
    bool profiling = false, counting = false;
    long long chunk_override = 0;
    // launch shapes: blocks per CU of each grid-stride kernel (0 = one work
    // item per thread); queue-fed kernels are launched over the queue capacity
    int primary_grid = 64, sec_grid = 6, shadow_grid = 64, shade_grid = 1, shade_sec_grid = 1, adjoint_grid = 8;
    int shade_block = 0;  // workgroup size of the shading kernels (0: the compiled bound)
    // differentiable shading: why the scene is not differentiable ("" = it is),
    // float texel chains, their gradient, and the scene's own materials/lights
    std::string adj_reason;
    std::vector<int> tex_w, tex_h, tex_mips;
    std::vector<uint32_t> tex_off;  // element offset of each texture's level 0 in the chain
    size_t tex_elems = 0;           // chain elements of all textures
    float* tex_f = nullptr;
    uint32_t* tex_f_off = nullptr;
    float* gtex = nullptr;
    Material* mats_orig = nullptr;
    Light* lights_orig = nullptr;
    hipEvent_t ev_join[2] = {nullptr, nullptr};
    int join_fallbacks = 0;  // caller-stream joins that fell back to host synchronisation
    hipEvent_t ev_frame[2] = {nullptr, nullptr};
    // second stream: the primary shading's shadow requests are traced there
    // while the secondary rays run on `stream`
    hipStream_t side = nullptr;
    hipEvent_t ev_side[2] = {nullptr, nullptr};
    bool overlap = false;  // measured faster on the analytic 10k workload only
    // shadow queue in one run per light (faster on the 1M-triangle frame) or
    // one run shared by the lights; scenes shaded by the FULL kernels always
    // use one run
    bool light_runs = true;
    // queue order of the shading kernels (EventBufs::order bits)
    int order = ORDER_SHADOW | ORDER_SEC;
    // secondary rays sorted before each trace round: 0 off, 1 kind + octant,
    // 2 kind + octant + dominant axis (sec_sort_key); set in context_create
    int sec_sort = 0;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    struct Mark { int slot; hipEvent_t a, b; };
    std::vector<Mark> marks;
    hipEvent_t event() {
        if (ev_used == ev_pool.size()) {
            hipEvent_t e = nullptr;
            (void)hipEventCreate(&e);
            ev_pool.push_back(e);
        }
        return ev_pool[ev_used++];
    }
    SheenRayStats last{};
    bool bvh_on_device = true;
    int bvh_nodes[N_BVH_CATS] = {}, bvh_prims[N_BVH_CATS] = {}, bvh_stack[N_BVH_CATS] = {}, bvh_depth[N_BVH_CATS] = {};
    Photon* photons = nullptr;  // sorted by cell
    int* cell_start = nullptr;
    int* cell_count = nullptr;
    int n_photons = 0;

    ~SheenContext() {
        if (stream) (void)hipStreamSynchronize(stream);
        wb.release();
        if (frame) (void)hipFree(frame);
        if (ctr) (void)hipFree(ctr);
        if (photons) (void)hipFree(photons);
        if (cell_start) (void)hipFree(cell_start);
        if (cell_count) (void)hipFree(cell_count);
        for (void* p : owned) (void)hipFree(p);
        for (hipEvent_t e : ev_frame) if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : ev_join) if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : ev_side) if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
        if (side) (void)hipStreamSynchronize(side);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }
    template <class T>
    T* alloc_owned(size_t n) {
        void* p = nullptr;
        if (n == 0 || hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
        owned.push_back(p);
        return (T*)p;
    }
    template <class T>
    T* upload(const T* src, size_t n) {
        if (!src || n == 0) return nullptr;
        void* p = nullptr;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
        owned.push_back(p);
        if (hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        return (T*)p;
    }
};

static int grid_n(int num_cu, size_t n, int per_cu, int block = TRACE_BLOCK) {
    size_t blocks = (n + block - 1) / block;
    size_t cap = per_cu > 0 ? (size_t)num_cu * per_cu : blocks;
    if (blocks > cap) blocks = cap;
    if (blocks > 0x7FFFFFFF) blocks = 0x7FFFFFFF;
    if (blocks < 1) blocks = 1;
    return (int)blocks;
}
static int grid_for(SheenContext* c, size_t n, int per_cu, int block = TRACE_BLOCK) { return grid_n(c->num_cu, n, per_cu, block); }
// Workgroup size of the shading kernels: the knob's value (a multiple of 64)
// up to the bound they are compiled for, which is also the default.
template <bool FULL>
static int shade_block_of(int knob) { return knob > 0 ? std::min(knob, (int)SHADE_BLOCK_OF(FULL)) : (int)SHADE_BLOCK_OF(FULL); }

static bool alloc_dev(void** p, size_t bytes) {
    *p = nullptr;
    return hipMalloc(p, bytes ? bytes : 16) == hipSuccess;
}

// Photon pass and deterministic hash-grid map (reference
// api/sheen_c_api.cu:271-307, render_kernel.cu:928-1045).
bool photon_pass(SheenContext* c, const SheenScene* in) {
    const long long total_ll = (long long)in->photons * c->ds.num_lights;
    if (total_ll <= 0) return true;
    if (total_ll > 0x7FFFFFFF) { sheen_set_error("too many photons"); return false; }
    const uint32_t total = (uint32_t)total_ll;
    const uint32_t cap = (uint32_t)std::min<long long>(total_ll, 4000000);  // sheen_c_api.cu:276-277
    const float cell = in->photon_cell > 0.0f ? in->photon_cell : 0.25f;    // sheen_c_api.cu:306
    const int num_cells = 1 << 18;                                           // render_kernel.cu:997
    hipStream_t s = c->stream;
    Photon *all = nullptr, *stored = nullptr;
    uint32_t *flag = nullptr, *pos = nullptr, *keys = nullptr, *keys2 = nullptr, *vals = nullptr, *vals2 = nullptr;
    PhotonCounters* pc = nullptr;
    void* tmp = nullptr;
    bool ok = false;
    do {
        if (!alloc_dev((void**)&all, (size_t)total * sizeof(Photon)) || !alloc_dev((void**)&flag, (size_t)total * 4) ||
            !alloc_dev((void**)&pos, (size_t)total * 4) || !alloc_dev((void**)&pc, sizeof(PhotonCounters))) {
            sheen_set_error("photon alloc failed");
            break;
        }
        PhotonCounters zero{};
        if (hipMemcpyAsync(pc, &zero, sizeof(zero), hipMemcpyHostToDevice, s) != hipSuccess) break;
        k_photon_emit<<<grid_for(c, total, 16), TRACE_BLOCK, 0, s>>>(c->ds, in->photons, total, all, flag, pc);
        size_t tb = 0;
        if (hipcub::DeviceScan::ExclusiveSum(nullptr, tb, flag, pos, (int)total, s) != hipSuccess) break;
        if (!alloc_dev(&tmp, tb)) { sheen_set_error("photon scan alloc failed"); break; }
        if (hipcub::DeviceScan::ExclusiveSum(tmp, tb, flag, pos, (int)total, s) != hipSuccess) break;
        uint32_t last[2] = {0, 0};
        PhotonCounters hpc{};
        if (hipMemcpyAsync(&last[0], pos + total - 1, 4, hipMemcpyDeviceToHost, s) != hipSuccess) break;
        if (hipMemcpyAsync(&last[1], flag + total - 1, 4, hipMemcpyDeviceToHost, s) != hipSuccess) break;
        if (hipMemcpyAsync(&hpc, pc, sizeof(hpc), hipMemcpyDeviceToHost, s) != hipSuccess) break;
        if (hipStreamSynchronize(s) != hipSuccess) break;
        c->info.photons_emitted = (int)total;
        c->info.photon_rays = hpc.rays;
        const uint32_t n = std::min(last[0] + last[1], cap);
        c->info.photons_stored = (int)n;
        c->n_photons = (int)n;
        if (n == 0) { ok = true; break; }
        if (!alloc_dev((void**)&stored, (size_t)n * sizeof(Photon)) || !alloc_dev((void**)&keys, (size_t)n * 4) ||
            !alloc_dev((void**)&keys2, (size_t)n * 4) || !alloc_dev((void**)&vals, (size_t)n * 4) ||
            !alloc_dev((void**)&vals2, (size_t)n * 4) || !alloc_dev((void**)&c->photons, (size_t)n * sizeof(Photon)) ||
            !alloc_dev((void**)&c->cell_count, (size_t)num_cells * 4) ||
            !alloc_dev((void**)&c->cell_start, (size_t)num_cells * 4)) {
            sheen_set_error("photon map alloc failed");
            break;
        }
        k_photon_compact<<<grid_for(c, total, 8), 256, 0, s>>>(all, flag, pos, total, cap, stored);
        k_zero_i32<<<grid_for(c, num_cells, 8), 256, 0, s>>>(c->cell_count, num_cells);
        k_photon_keys<<<grid_for(c, n, 8), 256, 0, s>>>(stored, n, cell, num_cells, keys, vals, c->cell_count);
        (void)hipFree(tmp);
        tmp = nullptr;
        size_t sb = 0;
        if (hipcub::DeviceRadixSort::SortPairs(nullptr, sb, keys, keys2, vals, vals2, (int)n, 0, 18, s) != hipSuccess) break;
        size_t sb2 = 0;
        if (hipcub::DeviceScan::ExclusiveSum(nullptr, sb2, c->cell_count, c->cell_start, num_cells, s) != hipSuccess) break;
        if (!alloc_dev(&tmp, std::max(sb, sb2))) { sheen_set_error("photon sort alloc failed"); break; }
        if (hipcub::DeviceRadixSort::SortPairs(tmp, sb, keys, keys2, vals, vals2, (int)n, 0, 18, s) != hipSuccess) break;
        if (hipcub::DeviceScan::ExclusiveSum(tmp, sb2, c->cell_count, c->cell_start, num_cells, s) != hipSuccess) break;
        k_photon_gather_sorted<<<grid_for(c, n, 8), 256, 0, s>>>(stored, vals2, n, c->photons);
        if (hipStreamSynchronize(s) != hipSuccess) break;
        c->ds.pm.photons = c->photons;
        c->ds.pm.cell_start = c->cell_start;
        c->ds.pm.cell_count = c->cell_count;
        c->ds.pm.num_cells = num_cells;
        c->ds.pm.cell = cell;
        ok = true;
    } while (false);
    if (!ok && std::strcmp(g_error, "no error") == 0) sheen_set_error("photon pass failed: %s", hipGetErrorString(hipGetLastError()));
    void* fr[] = {all, stored, flag, pos, keys, keys2, vals, vals2, pc, tmp};
    for (void* p : fr) if (p) (void)hipFree(p);
    return ok;
}

static_assert(HOST_CSG_DEPTH == CSG_MAX_DEPTH, "host and device CSG stack depths");
static_assert(sizeof(HostSpineNode) == sizeof(CsgSpineNode), "spine node layout");

// Triangle leaf copies in leaf-slot order: v0 + index, v1, v2 (48 bytes).
// Frustum records of the lathe profiles (lathe_frustum), one thread per lathe;
// point ranges were validated on the host, and are clipped here as well.
__global__ void k_lathe_frusta(const Lathe* lathes, int n, const LathePoint* pts, int n_pts, LatheFrustum* out,
                               float2* yr) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const Lathe L = lathes[l];
    // the y range of the frusta that can be hit; a frustum taking the
    // division path (an overflowing or NaN square) opens it to everything
    float lo = __builtin_huge_valf(), hi = -__builtin_huge_valf();
    for (int i = 0; i + 1 < L.point_count; i++) {
        const int g = L.point_offset + i;
        if (g < 0 || g + 1 >= n_pts) break;
        const LatheFrustum f = lathe_frustum(pts[g].y, pts[g].radius, pts[g + 1].y, pts[g + 1].radius);
        out[g] = f;
        if (f.mode == 0) {
            lo = fminf(lo, fminf(pts[g].y, pts[g + 1].y));
            hi = fmaxf(hi, fmaxf(pts[g].y, pts[g + 1].y));
        } else if (f.mode == 2) {
            lo = -__builtin_huge_valf();
            hi = __builtin_huge_valf();
        }
    }
    yr[l] = make_float2(lo, hi);  // (+inf, -inf): no frustum can be hit
}
__global__ void k_tri_leaf(const uint32_t* prims, int n, const Triangle* tris, float4* out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = prims[k];
    const Triangle& t = tris[i];
    out[3 * (size_t)k] = make_float4(t.v0.x, t.v0.y, t.v0.z, __int_as_float((int)i));
    out[3 * (size_t)k + 1] = make_float4(t.v1.x, t.v1.y, t.v1.z, 0.0f);
    out[3 * (size_t)k + 2] = make_float4(t.v2.x, t.v2.y, t.v2.z, 0.0f);
}

// Per axis, the largest 2|origin| + 2 WBVH_QMAX scale over `n` nodes, as
// float bits (non-negative floats order as their bits; a NaN orders last
// and sends every ray to the exact box test).
__global__ void k_box_k(const WNode* nodes, int n, uint32_t* out) {
    uint32_t m[3] = {0u, 0u, 0u};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const WNode& w = nodes[i];
        const float o[3] = {w.ox, w.oy, w.oz};
        const uint8_t e[3] = {w.ex, w.ey, w.ez};
        for (int a = 0; a < 3; a++) {
            const float v = 2.0f * fabsf(o[a]) + 2.0f * (float)WBVH_QMAX * __uint_as_float((uint32_t)e[a] << 23);
            m[a] = max(m[a], __float_as_uint(v));
        }
    }
    for (int a = 0; a < 3; a++) atomicMax(&out[a], m[a]);
}

// DScene::box_k of every category's BVH (the margins of the
// distance-space box test, trav_begin).
static bool scene_box_k(SheenContext* c) {
    DScene& d = c->ds;
    uint32_t* dk = nullptr;
    uint32_t hk[N_BVH_CATS][3] = {};
    if (!alloc_dev((void**)&dk, sizeof(hk))) { sheen_set_error("BVH bound alloc failed"); return false; }
    hipError_t e = hipMemcpyAsync(dk, hk, sizeof(hk), hipMemcpyHostToDevice, c->stream);
    for (int k = 0; k < N_BVH_CATS && e == hipSuccess; k++) {
        if (!d.bvh[k] || c->bvh_nodes[k] <= 0) continue;
        k_box_k<<<std::min(1024, (c->bvh_nodes[k] + 255) / 256), 256, 0, c->stream>>>(d.bvh[k], c->bvh_nodes[k], dk + 3 * k);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(hk, dk, sizeof(hk), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dk);
    if (e != hipSuccess) { sheen_set_error("BVH bound failed: %s", hipGetErrorString(e)); return false; }
    std::memcpy(d.box_k, hk, sizeof(hk));
    return true;
}

// Builds every category BVH on the device from the primitive bounds (the
// host's for all kinds but triangles, whose bounds come from the uploaded
// records).  Same trees, byte for byte, as the host build.
static bool device_bvhs(SheenContext* c, const SheenScene* in, const HostScene& hs) {
    DScene& d = c->ds;
    for (int k = 0; k < N_BVH_CATS; k++) {
        const HostBvh& hb = hs.bvh[k];
        d.count[k] = 0;
        d.cat_opaque[k] = hb.opaque;
        if (hb.count == 0) continue;
        float* box = nullptr;
        bool own_box = false;
        if (k == C_TRI) {
            if (!alloc_dev((void**)&box, 6 * sizeof(float) * (size_t)hb.count)) { sheen_set_error("BVH bounds alloc failed"); return false; }
            own_box = true;
            wbvh_triangle_bounds(d.triangles, hb.count, box, c->stream);
        } else {
            box = c->upload(hb.bounds.data(), hb.bounds.size());
            d.bounds[k] = box;
            if (!box) { sheen_set_error("BVH bounds upload failed"); return false; }
        }
        unsigned char* skip = nullptr;
        bool any_skip = false;
        for (unsigned char f : hb.skip) any_skip = any_skip || f;
        if (any_skip && !alloc_dev((void**)&skip, hb.skip.size())) { sheen_set_error("BVH flags alloc failed"); return false; }
        if (skip) HIP_OR(hipMemcpy(skip, hb.skip.data(), hb.skip.size(), hipMemcpyHostToDevice), false);
        WBvhDevice wd;
        std::string err;
        const bool ok = wbvh_build_device(box, hb.count, skip, WBVH_MAX_LEAF, c->stream, wd, err);
        if (skip) (void)hipFree(skip);
        if (own_box) (void)hipFree(box);
        if (!ok) { sheen_set_error("%s", err.c_str()); return false; }
        if (wd.n_nodes == 0) continue;
        c->owned.push_back(wd.nodes);
        c->owned.push_back(wd.prims);
        if (wd.max_stack > 128) { sheen_set_error("BVH too deep for the traversal stack"); return false; }
        d.bvh[k] = wd.nodes;
        d.prim_ids[k] = wd.prims;
        d.count[k] = hb.count;
        c->bvh_nodes[k] = wd.n_nodes;
        c->bvh_prims[k] = wd.n_prims;
        c->bvh_stack[k] = wd.max_stack;
        c->bvh_depth[k] = wd.depth;
        if (k == C_TRI) {
            float4* leaf = c->alloc_owned<float4>(3 * (size_t)wd.n_prims);
            if (!leaf) { sheen_set_error("triangle leaf alloc failed"); return false; }
            k_tri_leaf<<<(wd.n_prims + 255) / 256, 256, 0, c->stream>>>(wd.prims, wd.n_prims, (const Triangle*)d.triangles, leaf);
            d.tri_leaf = leaf;
            HIP_OR(hipGetLastError(), false);
        }
    }
    HIP_OR(hipStreamSynchronize(c->stream), false);
    return true;
}

// Uploads a scene: BVHs, records, textures.  flags: SHEEN_CTX_HOST_BVH builds
// the BVHs on the host instead of the device.  Returns nullptr on error.
static SheenContext* context_create(const SheenScene* in, int flags = 0) {
    if (!in) { sheen_set_error("null scene"); return nullptr; }
    if (in->num_materials <= 0 || !in->materials) { sheen_set_error("no materials"); return nullptr; }
    const auto t0 = std::chrono::steady_clock::now();
    std::unique_ptr<SheenContext> c(new SheenContext());
    HIP_OR(hipGetDevice(&c->device), nullptr);
    hipDeviceProp_t prop;
    HIP_OR(hipGetDeviceProperties(&prop, c->device), nullptr);
    c->num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    HIP_OR(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), nullptr);
    HIP_OR(hipEventCreate(&c->ev_frame[0]), nullptr);
    HIP_OR(hipEventCreate(&c->ev_frame[1]), nullptr);
    HIP_OR(hipEventCreateWithFlags(&c->ev_join[0], hipEventDisableTiming), nullptr);
    HIP_OR(hipEventCreateWithFlags(&c->ev_join[1], hipEventDisableTiming), nullptr);
    HIP_OR(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking), nullptr);
    HIP_OR(hipEventCreateWithFlags(&c->ev_side[0], hipEventDisableTiming), nullptr);
    HIP_OR(hipEventCreateWithFlags(&c->ev_side[1], hipEventDisableTiming), nullptr);

    const bool host_bvh = (flags & SHEEN_CTX_HOST_BVH) != 0;
    c->bvh_on_device = !host_bvh;
    HostScene hs;
    std::string err;
    if (!host_prepare_scene(*in, hs, err, host_bvh)) { sheen_set_error("%s", err.c_str()); return nullptr; }
    c->info = hs.info;
    DScene& d = c->ds;
    std::memset(&d, 0, sizeof(d));
    d.triangles = c->upload(in->triangles, in->triangles ? (size_t)in->num_triangles : 0);
    d.num_triangles = in->triangles ? in->num_triangles : 0;
    if (host_bvh) {
        for (int k = 0; k < N_BVH_CATS; k++) {
            const HostBvh& hb = hs.bvh[k];
            d.count[k] = hb.nodes.empty() ? 0 : hb.count;
            d.cat_opaque[k] = hb.opaque;
            if (hb.nodes.empty()) continue;
            d.bvh[k] = c->upload(hb.nodes.data(), hb.nodes.size());
            d.prim_ids[k] = c->upload(hb.prims.data(), hb.prims.size());
            d.bounds[k] = c->upload(hb.bounds.data(), hb.bounds.size());
            if (!d.bvh[k] || !d.prim_ids[k]) { sheen_set_error("BVH upload failed"); return nullptr; }
            c->bvh_nodes[k] = (int)hb.nodes.size();
            c->bvh_prims[k] = (int)hb.prims.size();
            c->bvh_stack[k] = hb.max_stack;
            c->bvh_depth[k] = hb.depth;
        }
        if (!hs.tri_leaf.empty()) {
            d.tri_leaf = c->upload((const float4*)hs.tri_leaf.data(), hs.tri_leaf.size() / 4);
            if (!d.tri_leaf) { sheen_set_error("triangle upload failed"); return nullptr; }
        }
    } else if (!device_bvhs(c.get(), in, hs)) {
        return nullptr;
    }
    if (!scene_box_k(c.get())) return nullptr;
    c->info.bvh_build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    d.objects = c->upload(in->objects, in->objects ? (size_t)in->num_objects : 0);
    d.strands = c->upload(in->sweep_strands, in->sweep_strands ? (size_t)in->num_sweep_strands : 0);
    d.sweep_points = c->upload(in->sweep_points, in->sweep_points ? (size_t)in->num_sweep_points : 0);
    // height fields: image pointers become device pointers (reference api/sheen_c_api.cu:220-236)
    if (in->heightfields && in->num_heightfields > 0) {
        std::vector<HeightField> hf(in->heightfields, in->heightfields + in->num_heightfields);
        for (auto& h : hf) {
            if (h.image && h.img_w > 0 && h.img_h > 0) {
                h.image = c->upload(h.image, (size_t)h.img_w * h.img_h);
                if (!h.image) { sheen_set_error("hf image alloc failed"); return nullptr; }
            } else {
                h.image = nullptr;
            }
        }
        d.hfs = c->upload(hf.data(), hf.size());
    }
    d.ses = c->upload(in->superellipsoids, in->superellipsoids ? (size_t)in->num_superellipsoids : 0);
    d.lathes = c->upload(in->lathes, in->lathes ? (size_t)in->num_lathes : 0);
    d.lathe_points = c->upload(in->lathe_points, in->lathe_points ? (size_t)in->num_lathe_points : 0);
    // up to LATHE_SCAN_MAX lathes are tested without their tree, widest
    // profile first, from per-frustum records (lathe_scan in dev_trace.h)
    if (d.count[C_LATHE] > 0 && d.lathe_points && c->bvh_prims[C_LATHE] > 0 && c->bvh_prims[C_LATHE] <= LATHE_SCAN_MAX) {
        const int n = c->bvh_prims[C_LATHE];
        std::vector<uint32_t> ids((size_t)n);
        HIP_OR(hipStreamSynchronize(c->stream), nullptr);
        HIP_OR(hipMemcpy(ids.data(), d.prim_ids[C_LATHE], (size_t)n * 4, hipMemcpyDeviceToHost), nullptr);
        std::vector<float> width((size_t)n, 0.0f);
        for (int j = 0; j < n; j++) {
            if (ids[j] >= (uint32_t)in->num_lathes) { sheen_set_error("lathe BVH index out of range"); return nullptr; }
            const Lathe& l = in->lathes[ids[j]];
            for (int k = 0; k < l.point_count; k++) width[j] = std::max(width[j], in->lathe_points[l.point_offset + k].radius);
        }
        std::vector<int> order((size_t)n);
        for (int j = 0; j < n; j++) order[j] = j;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return width[a] > width[b]; });
        std::vector<int> sorted((size_t)n);
        for (int j = 0; j < n; j++) sorted[j] = (int)ids[order[j]];
        LatheFrustum* fr = c->alloc_owned<LatheFrustum>((size_t)in->num_lathe_points);
        float2* yr = c->alloc_owned<float2>((size_t)in->num_lathes);
        d.lathe_order = c->upload(sorted.data(), sorted.size());
        if (!fr || !yr || !d.lathe_order) { sheen_set_error("lathe record upload failed"); return nullptr; }
        k_lathe_frusta<<<(in->num_lathes + 255) / 256, 256, 0, c->stream>>>(d.lathes, in->num_lathes, d.lathe_points,
                                                                            in->num_lathe_points, fr, yr);
        HIP_OR(hipGetLastError(), nullptr);
        d.lathe_frusta = fr;
        d.lathe_yr = yr;
        d.lathe_scan = n;
    }
    d.prisms = c->upload(in->prisms, in->prisms ? (size_t)in->num_prisms : 0);
    d.prism_points = c->upload(in->prism_points, in->prism_points ? (size_t)in->num_prism_points : 0);
    d.csg = c->upload(in->csg_nodes, in->csg_nodes ? (size_t)in->num_csg_nodes : 0);
    if (d.csg && (in->num_objects >= (1 << 30) || in->num_superellipsoids >= (1 << 30))) {
        sheen_set_error("CSG leaves index at most 2^30 spheres or superellipsoids");  // hit records carry 30 bits
        return nullptr;
    }
    {
        // opacity of the CSG leaves' materials, as cat_opaque for the BVH categories
        float maxtr = 0.0f;
        for (int i = 0; d.csg && i < in->num_csg_nodes; i++) {
            const CSGNode& n = in->csg_nodes[i];
            if (n.leaf_kind < 0) continue;
            int m = -1;
            if (n.leaf_kind == 0 && n.leaf_index >= 0 && n.leaf_index < in->num_objects) m = in->objects[n.leaf_index].material_id;
            if (n.leaf_kind == 1 && n.leaf_index >= 0 && n.leaf_index < in->num_superellipsoids)
                m = in->superellipsoids[n.leaf_index].material_id;
            maxtr = (m >= 0 && m < in->num_materials) ? std::max(maxtr, in->materials[m].transmission) : 1.0f;
        }
        const bool lathe_ok = d.count[C_LATHE] == 0 || d.cat_opaque[C_LATHE];
        const bool prism_ok = d.count[C_PRISM] == 0 || d.cat_opaque[C_PRISM];
        d.opaque_tail = (lathe_ok && prism_ok && maxtr <= 0.5f) ? 1 : 0;
    }
    if (d.csg) {
        std::vector<float> cbox;
        host_csg_boxes(*in, cbox);
        d.csg_box = c->upload(cbox.data(), cbox.size());
        HostCsgSpine sp;
        host_csg_spine(*in, cbox, sp);
        d.csg_spine = sp.spine ? 1 : 0;
        d.spine_nodes = c->upload(reinterpret_cast<const CsgSpineNode*>(sp.nodes.data()), sp.nodes.size());
        d.spine_sub = c->upload(reinterpret_cast<const int4*>(sp.sub.data()), sp.sub.size() / 4);
        d.spine_n_nodes = (int)sp.nodes.size();
        d.spine_always = sp.always;
    }
    d.num_csg = in->csg_nodes ? in->num_csg_nodes : 0;
    d.csg_root = (d.csg && in->csg_root >= 0 && in->csg_root < d.num_csg) ? in->csg_root : -1;
    d.planes = c->upload(in->planes, in->planes ? (size_t)in->num_planes : 0);
    d.num_planes = in->planes ? in->num_planes : 0;
    d.materials = c->upload(in->materials, (size_t)in->num_materials);
    d.num_materials = in->num_materials;
    d.lights = c->upload(in->lights, in->lights ? (size_t)in->num_lights : 0);
    d.num_lights = in->lights ? in->num_lights : 0;
    d.media = c->upload(in->media, in->media ? (size_t)in->num_media : 0);
    d.num_media = in->media ? in->num_media : 0;
    // textures with their mip chains (reference scene.h:230-261)
    if (in->num_textures > 0) {
        std::vector<TextureRef> refs(in->num_textures);
        std::memset(refs.data(), 0, refs.size() * sizeof(TextureRef));
        for (int i = 0; i < in->num_textures; i++) {
            const HostTexture& t = hs.tex[i];
            if (t.px.empty()) continue;
            refs[i].pixels = c->upload(t.px.data(), t.px.size());
            if (!refs[i].pixels) { sheen_set_error("tex alloc failed"); return nullptr; }
            refs[i].width = t.w;
            refs[i].height = t.h;
            refs[i].mips = t.mips;
        }
        d.textures = c->upload(refs.data(), refs.size());
        d.num_textures = in->num_textures;
        for (int i = 0; i < in->num_textures; i++) {
            c->tex_w.push_back(refs[i].pixels ? refs[i].width : 0);
            c->tex_h.push_back(refs[i].pixels ? refs[i].height : 0);
            c->tex_mips.push_back(refs[i].pixels ? refs[i].mips : 0);
            c->tex_off.push_back((uint32_t)c->tex_elems);
            c->tex_elems += hs.tex[i].px.size();
        }
    }
    d.cam = in->camera;
    d.caustics = in->caustics;
    for (int i = 0; i < d.num_lights; i++)
        if (in->lights[i].type == 3 && in->lights[i].beam > 0.0f) c->n_beam++;
    d.any_beams = c->n_beam > 0;
    d.wave_one_scene = 1;  // a context's kernels trace only this scene; batch tables clear it
    d.any_strands = d.count[C_STRAND] > 0;
    c->full = d.count[C_STRAND] || d.count[C_HF] || d.count[C_SE] || d.count[C_LATHE] || d.count[C_PRISM] ||
              d.csg_root >= 0 || d.num_media > 0 || c->n_beam > 0;
    // sorted secondary rays: 0.8 % faster on the 1M-triangle frame, 3 %
    // slower on the analytic 10k frame (whose shading favours rays that
    // start close together over rays that head one way)
    c->sec_sort = c->full ? 0 : 2;
    // the differentiable path covers primary hits on spheres, triangles and
    // planes under direct light
    c->any_secondary = c->any_glass = false;
    for (int i = 0; i < in->num_materials; i++) {
        const Material& m = in->materials[i];
        c->any_secondary = c->any_secondary || m.reflect > 0.0f || m.transmission > 0.0f;
        c->any_glass = c->any_glass || m.transmission > 0.5f;
    }
    for (int i = 0; i < in->num_materials && c->adj_reason.empty(); i++) {
        const Material& m = in->materials[i];
        if (m.reflect > 0.0f) c->adj_reason = "material with reflection";
        else if (m.transmission > 0.0f) c->adj_reason = "material with transmission";
        else if (m.anisotropy > 0.0f) c->adj_reason = "hair material";
        else if (m.shininess < 0.0f) c->adj_reason = "debug material";
    }
    if (c->adj_reason.empty() && d.num_media > 0) c->adj_reason = "participating media";
    if (c->adj_reason.empty() && c->n_beam > 0) c->adj_reason = "spot-light beams";
    if (c->adj_reason.empty() && c->full) c->adj_reason = "primitives other than spheres, triangles and planes";
    if (c->adj_reason.empty() && in->photons > 0 && in->caustics > 0.0f) c->adj_reason = "caustics";
    if (!alloc_dev((void**)&c->ctr, sizeof(Counters))) { sheen_set_error("counter alloc failed"); return nullptr; }
    if (in->photons > 0 && d.num_lights > 0) {
        if (!photon_pass(c.get(), in)) return nullptr;
    }
    HIP_OR(hipDeviceSynchronize(), nullptr);
    return c.release();
}

// Per-slot bytes of the work buffers.
// Entries of each shadow continuation queue for `shadow_entries` requests: a
// quarter of them (requests beyond it finish their passes in the first
// shadow kernel).
static size_t cont_entries(size_t shadow_entries) { return std::max<size_t>(64, shadow_entries / 4); }
// Tiles of a secondary queue of 2 * slots entries, and the sort's count words.
static uint32_t sort_max_tiles(size_t slots) { return (uint32_t)((2 * slots + SORT_TILE - 1) / SORT_TILE); }
static size_t sort_hist_words(size_t slots) { return (size_t)SORT_BINS * sort_max_tiles(slots); }

static size_t slot_bytes(int L, int n_beam, bool caustic, bool media) {
    size_t b = 16 /*hit*/ + 16 /*samp*/ + 3 * 16 /*base*/ + (caustic ? 3 * 16 : 0) + 3 * (size_t)L * (16 + 4) +
               4 * 32 /*sec queues*/ + 2 * 16 /*sec hits*/ + 1 /*sort counts*/ + 3 * (size_t)L * 32 /*shadow queue*/ +
               2 * 3 * (size_t)L * 32 / 4 /*shadow continuations*/ + (size_t)n_beam * 16 + (media ? 32 : 0);
    return b;
}

// Work buffers for `slots` sample slots of scenes with at most L lights and
// n_beam beams; kept when the current ones are large enough.
static bool ensure_bufs(WorkBufs& w, size_t slots, int L, int n_beam, bool caustic, bool media) {
    if (w.cap_slots >= slots && w.cap_L >= L && w.cap_beam >= n_beam && (w.cap_caustic || !caustic) &&
        (w.cap_media || !media))
        return true;
    w.release();
    const size_t ev = slots * 3;
    const size_t Lq = (size_t)std::max(L, 1);
    if (ev * Lq >= ((size_t)1 << 31)) {  // per-light event entries are indexed by 31 bits
        sheen_set_error("work buffers of %zu slots x %d lights exceed the event index range", slots, L);
        return false;
    }
    bool ok = alloc_dev((void**)&w.hit_p, slots * sizeof(int4)) && alloc_dev((void**)&w.samp, slots * sizeof(float4)) &&
              alloc_dev((void**)&w.ev_base, ev * sizeof(float4)) &&
              alloc_dev((void**)&w.contrib, ev * Lq * sizeof(float4)) &&
              alloc_dev((void**)&w.trans, ev * Lq * sizeof(float)) &&
              alloc_dev((void**)&w.shadow_q, ev * Lq * sizeof(ShadowReq)) &&
              alloc_dev((void**)&w.cont[0], cont_entries(ev * Lq) * sizeof(ShadowReq)) &&
              alloc_dev((void**)&w.cont[1], cont_entries(ev * Lq) * sizeof(ShadowReq)) &&
              alloc_dev((void**)&w.shadow_nl, 2 * Lq * sizeof(uint32_t)) &&
              alloc_dev((void**)&w.sec_q[0], 2 * slots * sizeof(SecRay)) &&
              alloc_dev((void**)&w.sec_q[1], 2 * slots * sizeof(SecRay)) &&
              alloc_dev((void**)&w.sec_hits, 2 * slots * sizeof(int4)) &&
              alloc_dev((void**)&w.sort_hist, sort_hist_words(slots) * sizeof(uint32_t));
    if (ok && caustic) ok = alloc_dev((void**)&w.ev_caustic, ev * sizeof(float4));
    if (ok && n_beam > 0) ok = alloc_dev((void**)&w.beam_add, slots * n_beam * sizeof(float4));
    if (ok && media) ok = alloc_dev((void**)&w.media_p, slots * sizeof(float4)) && alloc_dev((void**)&w.media_r, slots * sizeof(float4));
    if (!ok) {
        w.release();
        sheen_set_error("work buffer allocation failed (%zu slots)", slots);
        return false;
    }
    w.cap_slots = slots; w.cap_L = L; w.cap_beam = n_beam; w.cap_caustic = caustic; w.cap_media = media;
    return true;
}

static bool ensure_work(SheenContext* c, size_t slots) {
    return ensure_bufs(*c->wbp, slots, c->ds.num_lights, c->n_beam, c->ds.pm.photons && c->ds.caustics > 0.0f,
                       c->ds.num_media > 0);
}

// Kernel timing with events recorded on the render stream; elapsed times are
// read once the frame has finished, so timing adds no synchronisation.
struct Timer {
    SheenContext* c;
    int slot;
    hipStream_t s;
    hipEvent_t b = nullptr;
    Timer(SheenContext* c_, int slot_, hipStream_t s_ = nullptr) : c(c_), slot(slot_), s(s_ ? s_ : c_->stream) {
        if (!c->profiling) return;
        hipEvent_t a = c->event();
        b = c->event();
        (void)hipEventRecord(a, s);
        c->marks.push_back({slot, a, b});
    }
    ~Timer() {
        if (c->profiling) (void)hipEventRecord(b, s);
    }
};

// The kernels of one frame chunk.  FULL = false selects the instantiation
// compiled for spheres, triangles and planes only.
// Backward pass of the differentiable entry points: image gradient in, and
// the gradient buffers the adjoint kernel accumulates into.
struct AdjCall {
    const float* grad;
    AdjGrads G;
};

// Sorts secondary queue `in` (count C->sec_n[which]) into `out` by
// sec_sort_key (mode 1 or 2); cap_slots sizes the grids and the counts.
static void launch_sec_sort(hipStream_t s, int num_cu, const SecRay* in, SecRay* out, Counters* ctr, int which,
                            int mode, uint32_t* hist, size_t cap_slots) {
    const uint32_t max_tiles = sort_max_tiles(cap_slots);
    const int g = (int)std::min<uint32_t>(max_tiles, (uint32_t)num_cu * 8u);
    k_sec_hist<<<g, SORT_BLOCK, 0, s>>>(in, ctr, which, mode, hist, max_tiles);
    k_sec_scan<<<SORT_BINS, 1024, 0, s>>>(ctr, which, hist, max_tiles, ctr->sort_tot);
    k_sec_scatter<<<g, SORT_BLOCK, 0, s>>>(in, ctr, which, mode, hist, max_tiles, ctr->sort_tot, out);
}

// The second and third shadow passes of the requests that went through
// glass (scenes of spheres, triangles and planes; the others keep every pass
// in k_shadow).
template <bool FULL, class SRC>
static void launch_shadow_cont(hipStream_t s, int grid, bool counting, const SRC& src, const EventBufs& EB,
                               Counters* ctr) {
    if (FULL || !EB.cont_cap) return;
    for (int stage = 0; stage < 2; stage++) {
        if (counting) k_shadow_cont<true><<<grid, TRACE_BLOCK, 0, s>>>(src, EB, ctr, stage);
        else k_shadow_cont<false><<<grid, TRACE_BLOCK, 0, s>>>(src, EB, ctr, stage);
    }
}

template <bool FULL>
static void run_chunk(SheenContext* c, WorkBufs& w, Counters* ctr, float4* frame, const FrameDesc& F, SampleBufs& SB,
                      EventBufs& EB, SheenRayStats& S, bool hits_only, int* dk, int* di, float* dt, const AdjCall* adj) {
    hipStream_t s = c->stream;
    const int ns = F.ns;
    OneSrc src;
    src.S = c->ds;
    src.F = F;
        k_reset<<<1, 64, 0, s>>>(ctr, EB.shadow_nl, EB.runs);
        {
            Timer t(c, 0);
            if (c->counting) k_primary<true, FULL><<<grid_for(c, F.n_slots, c->primary_grid), TRACE_BLOCK, 0, s>>>(src, w.hit_p, ctr);
            else k_primary<false, FULL><<<grid_for(c, F.n_slots, c->primary_grid), TRACE_BLOCK, 0, s>>>(src, w.hit_p, ctr);
            S.trace_launches++;
        }
        if (hits_only) {
            k_export_hits<<<grid_for(c, F.n_slots, 8), 256, 0, s>>>(F, w.hit_p, dk, di, dt);
            return;
        }
        {
            Timer t(c, 4);
            const int sb = shade_block_of<FULL>(c->shade_block);
            k_shade_primary<FULL><<<grid_for(c, F.n_slots, c->shade_grid, sb), sb, 0, s>>>(src, SB, EB, w.sec_q[0], ctr);
        }
        if (adj) {
            // replay: shadow transmittances of the primary hits, then the adjoint
            // (differentiable scenes have no secondary rays)
            const size_t cap = 3 * (size_t)F.n_slots * std::max(1, c->ds.num_lights);
            k_shadow<false, FULL><<<grid_for(c, cap, c->shadow_grid), TRACE_BLOCK, 0, s>>>(src, EB, ctr, 0);
            if (c->any_glass) launch_shadow_cont<FULL>(s, grid_for(c, EB.cont_cap, c->shadow_grid), false, src, EB, ctr);
            S.shadow_launches++;
            // one block per 8x8 tile of samples at a time
            const size_t tiles = F.n_slots / (64u * (uint32_t)ns);
            k_adjoint<<<grid_for(c, tiles * TRACE_BLOCK, c->adjoint_grid), TRACE_BLOCK, 0, s>>>(c->ds, F, w.hit_p, EB, adj->grad, adj->G);
            return;
        }
        // The primary shading's shadow requests are complete here: trace them
        // on the side stream while the secondary rounds run on this one (their
        // requests queue after each light's mark; the two write disjoint entries).
        const bool overlap = c->overlap;
        if (overlap) {
            k_snap_shadow<<<1, 64, 0, s>>>(EB.shadow_nl, EB.runs);
            (void)hipEventRecord(c->ev_side[0], s);
            (void)hipStreamWaitEvent(c->side, c->ev_side[0], 0);
            Timer t(c, 2, c->side);
            const size_t cap = 3 * (size_t)F.n_slots * std::max(1, c->ds.num_lights);
            const int g = grid_for(c, cap, c->shadow_grid);
            if (c->counting) k_shadow<true, FULL><<<g, TRACE_BLOCK, 0, c->side>>>(src, EB, ctr, 1);
            else k_shadow<false, FULL><<<g, TRACE_BLOCK, 0, c->side>>>(src, EB, ctr, 1);
            S.shadow_launches++;
        }
        if (overlap) (void)hipEventRecord(c->ev_side[1], c->side);
        // three rounds: reflection + first refraction face, then faces 2 and 3
        for (int round = 0; round < (c->any_secondary ? 3 : 0); round++) {
            // unsorted: queues 0 and 1 alternate; sorted: a round's queue is
            // sorted from 0 into 1, traced and shaded from 1, and the next
            // faces queue into 0 again
            const bool sorted = c->sec_sort > 0;
            const int in_q = sorted ? 1 : (round & 1), out_q = sorted ? 0 : ((round + 1) & 1);
            const int in_ctr = round, out_ctr = round + 1;
            const size_t cap = 2 * (size_t)F.n_slots;
            if (sorted) {
                Timer t(c, 1);
                launch_sec_sort(s, c->num_cu, w.sec_q[0], w.sec_q[1], ctr, in_ctr, c->sec_sort, w.sort_hist, w.cap_slots);
            }
            {
                Timer t(c, 1);
                if (c->counting)
                    k_trace_sec<true, FULL><<<grid_for(c, cap, c->sec_grid), TRACE_BLOCK, 0, s>>>(src, w.sec_q[in_q], in_ctr, w.sec_hits, ctr);
                else
                    k_trace_sec<false, FULL><<<grid_for(c, cap, c->sec_grid), TRACE_BLOCK, 0, s>>>(src, w.sec_q[in_q], in_ctr, w.sec_hits, ctr);
                S.trace_launches++;
            }
            {
                Timer t(c, 3);
                const int sb = shade_block_of<FULL>(c->shade_block);
                k_shade_sec<FULL><<<grid_for(c, cap, c->shade_sec_grid, sb), sb, 0, s>>>(src, SB, EB, w.sec_q[in_q], in_ctr, w.sec_hits,
                                                                       w.sec_q[out_q], out_ctr, ctr);
            }
        }
        {
            Timer t(c, 2);
            const size_t cap = 3 * (size_t)F.n_slots * std::max(1, c->ds.num_lights);
            const int g = grid_for(c, cap, c->shadow_grid);
            if (c->counting) k_shadow<true, FULL><<<g, TRACE_BLOCK, 0, s>>>(src, EB, ctr, overlap ? 2 : 0);
            else k_shadow<false, FULL><<<g, TRACE_BLOCK, 0, s>>>(src, EB, ctr, overlap ? 2 : 0);
            S.shadow_launches++;
        }
        if (overlap) (void)hipStreamWaitEvent(s, c->ev_side[1], 0);
        if (!FULL && EB.cont_cap && c->any_glass) {
            // the continuations of both shadow launches (one queue)
            Timer t(c, 5);
            launch_shadow_cont<FULL>(s, grid_for(c, EB.cont_cap, c->shadow_grid), c->counting, src, EB, ctr);
            S.cont_launches += 2;
        }
        {
            Timer t(c, 3);
            k_accumulate<<<grid_for(c, F.n_slots / ns, 8), 256, 0, s>>>(src, SB, EB, frame);
        }
}

// Caller-provided buffers of a frame (the no-allocation entry point): work
// buffers carved from the static scratch array, counters, the output image,
// and the chunk size that fits them.
struct FixedBufs {
    WorkBufs wb;
    Counters* ctr = nullptr;
    float4* frame = nullptr;
    size_t chunk = 0;
};

static int render_frame(SheenContext* c, int W, int H, int aa, SheenRayStats* st, bool hits_only, int* hk, int* hi,
                        float* ht, const AdjCall* adj = nullptr, FixedBufs* fx = nullptr) {
    if (W <= 0 || H <= 0 || aa < 1 || aa > 64) { sheen_set_error("bad render params"); return -1; }
    if (c->ds.num_materials <= 0) { sheen_set_error("no materials"); return -1; }
    const int ns = aa;
    const int tiles_x = (W + 7) / 8, tiles_y = (H + 7) / 8;
    const size_t total = (size_t)tiles_x * tiles_y * ns * 64;
    const size_t unit = (size_t)ns * 64;
    const bool caustic = c->ds.pm.photons && c->ds.caustics > 0.0f;
    size_t chunk;
    if (fx) chunk = fx->chunk;
    else if (c->chunk_override > 0) chunk = (size_t)c->chunk_override;
    else {
        // 32 GiB of work buffers (of 288 GB), at most 16 M slots: the analytic
        // 10k frame (14.7 M slots) in one chunk measured 93.5 ms against
        // 101 ms in two chunks of an 8 GiB budget
        const size_t budget = (size_t)32 << 30;
        chunk = budget / slot_bytes(c->ds.num_lights, c->n_beam, caustic, c->ds.num_media > 0);
        chunk = std::min(chunk, (size_t)16 << 20);
        // per-light event entries are indexed by 31 bits (ensure_bufs)
        chunk = std::min(chunk, ((size_t)1 << 31) / (3 * (size_t)std::max(c->ds.num_lights, 1)) - 1);
    }
    chunk = std::max(unit, chunk / unit * unit);
    chunk = std::min(chunk, (total + unit - 1) / unit * unit);
    if (chunk > 0x0FFFFFFF) chunk = 0x0FFFFFFF / unit * unit;
    const size_t npix = (size_t)W * H;
    if (!fx) {
        if (!ensure_work(c, chunk)) return -1;
        if (c->frame_cap < npix) {
            if (c->frame) (void)hipFree(c->frame);
            c->frame = nullptr;
            c->frame_cap = 0;
            if (!alloc_dev((void**)&c->frame, npix * sizeof(float4))) { sheen_set_error("frame alloc failed"); return -1; }
            c->frame_cap = npix;
        }
        c->frame_w = W; c->frame_h = H;
    }
    WorkBufs& w = fx ? fx->wb : *c->wbp;
    Counters* const ctr = fx ? fx->ctr : c->ctr;
    float4* const frame = fx ? fx->frame : c->frame;
    int *dk = nullptr, *di = nullptr;
    float* dt = nullptr;
    if (hits_only) {
        const size_t n = npix * ns;
        if (!alloc_dev((void**)&dk, n * 4) || !alloc_dev((void**)&di, n * 4) || !alloc_dev((void**)&dt, n * 4)) {
            sheen_set_error("hit buffer alloc failed");
            return -1;
        }
    }
    SheenRayStats S{};
    EventBufs EB;
    EB.base = w.ev_base; EB.caustic = caustic ? w.ev_caustic : nullptr; EB.contrib = w.contrib; EB.trans = w.trans;
    EB.shadow_q = w.shadow_q; EB.shadow_nl = w.shadow_nl; EB.seg = 3u * (uint32_t)w.cap_slots;
    EB.cont[0] = w.cont[0]; EB.cont[1] = w.cont[1];
    EB.cont_cap = c->full ? 0u : (uint32_t)cont_entries(3 * w.cap_slots * (size_t)std::max(w.cap_L, 1));
    EB.L = c->ds.num_lights;
    EB.runs = (c->light_runs && !c->full) ? std::max(EB.L, 1) : 1;  // k_shadow<*, true> reads one run
    EB.order = c->order;
    EB.cap = (uint32_t)w.cap_slots;
    SampleBufs SB;
    SB.hit_p = w.hit_p; SB.samp = w.samp; SB.beam_add = c->n_beam ? w.beam_add : nullptr;
    SB.media_p = c->ds.num_media ? w.media_p : nullptr; SB.media_r = c->ds.num_media ? w.media_r : nullptr;
    SB.n_beam = c->n_beam;
    hipStream_t s = c->stream;
    k_zero_counters<<<1, 256, 0, s>>>(ctr);
    c->ev_used = 0;
    c->marks.clear();
    HIP_OR(hipEventRecord(c->ev_frame[0], s), -1);
    for (size_t g0 = 0; g0 < total; g0 += chunk) {
        FrameDesc F;
        F.width = W; F.height = H; F.ns = ns; F.tiles_x = tiles_x;
        F.g0 = (uint32_t)g0;
        F.n_slots = (uint32_t)std::min(chunk, total - g0);
        F.out0 = 0;
        if (c->full) run_chunk<true>(c, w, ctr, frame, F, SB, EB, S, hits_only, dk, di, dt, adj);
        else run_chunk<false>(c, w, ctr, frame, F, SB, EB, S, hits_only, dk, di, dt, adj);
    }
    HIP_OR(hipEventRecord(c->ev_frame[1], s), -1);
    HIP_OR(hipGetLastError(), -1);
    Counters host{};
    HIP_OR(hipMemcpyAsync(&host, ctr, sizeof(Counters), hipMemcpyDeviceToHost, s), -1);
    HIP_OR(hipStreamSynchronize(s), -1);
    HIP_OR(hipEventElapsedTime(&S.frame_ms, c->ev_frame[0], c->ev_frame[1]), -1);
    for (const auto& m : c->marks) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, m.a, m.b) != hipSuccess) continue;
        if (m.slot == 0) { S.primary_ms += ms; S.trace_ms += ms; }
        else if (m.slot == 1) S.trace_ms += ms;
        else if (m.slot == 2) S.shadow_ms += ms;
        else if (m.slot == 5) S.shadow_cont_ms += ms;
        else if (m.slot == 4) { S.shade_ms += ms; S.shade_primary_ms += ms; }
        else S.shade_ms += ms;
    }
    if (hits_only) {
        const size_t n = npix * ns;
        HIP_OR(hipMemcpy(hk, dk, n * 4, hipMemcpyDeviceToHost), -1);
        HIP_OR(hipMemcpy(hi, di, n * 4, hipMemcpyDeviceToHost), -1);
        HIP_OR(hipMemcpy(ht, dt, n * 4, hipMemcpyDeviceToHost), -1);
        (void)hipFree(dk); (void)hipFree(di); (void)hipFree(dt);
    }
    S.samples = npix * ns;
    S.primary_rays = S.samples;
    S.shadow_rays = ctr_total(host.shadow_passes) + ctr_total(host.extra_shadow);
    S.reflection_rays = ctr_total(host.reflection);
    S.refraction_rays = ctr_total(host.refraction);
    S.photon_rays = c->info.photon_rays;
    S.total_rays = S.primary_rays + S.shadow_rays + S.reflection_rays + S.refraction_rays;
    S.node_fetches = ctr_total(host.nodes);
    S.prim_tests = ctr_total(host.prims);
    S.primary_node_fetches = ctr_total(host.pnodes);
    S.primary_prim_tests = ctr_total(host.pprims);
    S.primary_stack_spills = ctr_total(host.spills);
    S.shadow_stack_spills = ctr_total(host.sspills);
    S.shadow_node_fetches = ctr_total(host.snodes);
    S.shadow_prim_tests = ctr_total(host.sprims);
    S.shadow_visit_lanes = ctr_total(host.switers) * 64ull;
    S.cont_node_fetches = ctr_total(host.cnodes);
    S.cont_prim_tests = ctr_total(host.cprims);
    c->last = S;
    if (st) *st = S;
    return 0;
}

// ---------------------------------------------------------------------------
// Work space of sheen_context_render_noalloc: one static device array, held by
// every device the library runs on until the process ends.  In host code the
// name `scratch` is a host-side stand-in, not a device address; the device
// address of each device's copy comes from hipGetSymbolAddress, taken once
// per device (with that device current) and kept in g_scratch.  Only that
// address is ever passed to kernels.  Calls that overlap on one device race
// on the array, which keeps the last call's contents.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[SHEEN_NOALLOC_SCRATCH_BYTES];
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch[kMaxDevices] = {};

// Device address of the current device's scratch array (looked up on the
// first call per device).  nullptr with the error set on failure.
static unsigned char* scratch_on_current_device() {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { sheen_set_error("hipGetDevice failed: %s", hipGetErrorString(e)); return nullptr; }
    if (dev < 0 || dev >= kMaxDevices) { sheen_set_error("device %d beyond the scratch table", dev); return nullptr; }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            sheen_set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
            return nullptr;
        }
        g_scratch[dev] = p;
    }
    return static_cast<unsigned char*>(g_scratch[dev]);
}

// Lays the work buffers of `slots` sample slots and the counters out from
// `base` (256-byte aligned pieces) as ensure_bufs allocates them.  Returns
// the bytes used; with base == nullptr only measures.
static size_t carve_work(unsigned char* base, size_t slots, int L, int n_beam, bool caustic, bool media, WorkBufs& w,
                         Counters** ctr) {
    size_t off = 0;
    auto take = [&](size_t bytes) -> void* {
        void* p = base ? base + off : nullptr;
        off += (bytes + 255) / 256 * 256;
        return p;
    };
    const size_t ev = slots * 3;
    const size_t Lq = (size_t)std::max(L, 1);
    w = WorkBufs();
    *ctr = (Counters*)take(sizeof(Counters));
    w.hit_p = (int4*)take(slots * sizeof(int4));
    w.samp = (float4*)take(slots * sizeof(float4));
    w.ev_base = (float4*)take(ev * sizeof(float4));
    w.contrib = (float4*)take(ev * Lq * sizeof(float4));
    w.trans = (float*)take(ev * Lq * sizeof(float));
    w.shadow_q = (ShadowReq*)take(ev * Lq * sizeof(ShadowReq));
    w.cont[0] = (ShadowReq*)take(cont_entries(ev * Lq) * sizeof(ShadowReq));
    w.cont[1] = (ShadowReq*)take(cont_entries(ev * Lq) * sizeof(ShadowReq));
    w.shadow_nl = (uint32_t*)take(2 * Lq * sizeof(uint32_t));
    w.sec_q[0] = (SecRay*)take(2 * slots * sizeof(SecRay));
    w.sec_q[1] = (SecRay*)take(2 * slots * sizeof(SecRay));
    w.sec_hits = (int4*)take(2 * slots * sizeof(int4));
    w.sort_hist = (uint32_t*)take(sort_hist_words(slots) * sizeof(uint32_t));
    if (caustic) w.ev_caustic = (float4*)take(ev * sizeof(float4));
    if (n_beam > 0) w.beam_add = (float4*)take(slots * n_beam * sizeof(float4));
    if (media) {
        w.media_p = (float4*)take(slots * sizeof(float4));
        w.media_r = (float4*)take(slots * sizeof(float4));
    }
    w.cap_slots = slots; w.cap_L = L; w.cap_beam = n_beam; w.cap_caustic = caustic; w.cap_media = media;
    return off;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int sheen_init(int device_id, int cuda_arch) {
    (void)cuda_arch;
    HIP_OR(hipSetDevice(device_id), -1);
    HIP_OR(hipFree(nullptr), -1);
    // set-up of the no-allocation entry point's scratch address on this device
    if (!scratch_on_current_device()) return -1;
    return 0;
}

void sheen_cleanup(void) { (void)hipDeviceSynchronize(); }

const char* sheen_get_last_error(void) { return g_error; }

SheenContext* sheen_context_create(const SheenScene* scene) { return context_create(scene); }

SheenContext* sheen_context_create_ex(const SheenScene* scene, int flags) { return context_create(scene, flags); }

int sheen_context_bvh(const SheenContext* ctx, int category, void* nodes, int node_cap, unsigned int* prims, int prim_cap,
                      int* info) {
    if (!ctx || category < 0 || category >= N_BVH_CATS) { sheen_set_error("bad context or category"); return -1; }
    const int nn = ctx->bvh_nodes[category], np = ctx->bvh_prims[category];
    if (info) {
        info[0] = nn;
        info[1] = np;
        info[2] = ctx->bvh_stack[category];
        info[3] = ctx->bvh_depth[category];
    }
    if (nodes && nn > 0) {
        if (node_cap < nn) { sheen_set_error("node buffer too small"); return -1; }
        HIP_OR(hipMemcpy(nodes, ctx->ds.bvh[category], (size_t)nn * sizeof(WNode), hipMemcpyDeviceToHost), -1);
    }
    if (prims && np > 0) {
        if (prim_cap < np) { sheen_set_error("primitive buffer too small"); return -1; }
        HIP_OR(hipMemcpy(prims, ctx->ds.prim_ids[category], (size_t)np * sizeof(uint32_t), hipMemcpyDeviceToHost), -1);
    }
    return 0;
}

double sheen_context_build_ms(const SheenContext* ctx) { return ctx ? ctx->info.bvh_build_ms : -1.0; }

void sheen_context_destroy(SheenContext* ctx) { delete ctx; }

int sheen_context_render(SheenContext* ctx, int width, int height, int aa_samples, float* out_rgba,
                         SheenRayStats* stats) {
    if (!ctx) { sheen_set_error("null context"); return -1; }
    if (render_frame(ctx, width, height, aa_samples, stats, false, nullptr, nullptr, nullptr) != 0) return -1;
    if (out_rgba) HIP_OR(hipMemcpy(out_rgba, ctx->frame, (size_t)width * height * sizeof(float4), hipMemcpyDeviceToHost), -1);
    return 0;
}

const float* sheen_context_frame_device(const SheenContext* ctx) { return ctx ? (const float*)ctx->frame : nullptr; }

int sheen_context_render_noalloc(SheenContext* ctx, int width, int height, int aa_samples, float* out_rgba_device,
                                 SheenRayStats* stats) {
    if (!ctx || !out_rgba_device) { sheen_set_error("null argument"); return -1; }
    if (width <= 0 || height <= 0 || aa_samples < 1 || aa_samples > 64) { sheen_set_error("bad render params"); return -1; }
    int dev = -1;
    HIP_OR(hipGetDevice(&dev), -1);
    if (dev != ctx->device) {
        sheen_set_error("the context's device (%d) is not the current device (%d)", ctx->device, dev);
        return -1;
    }
    unsigned char* base = scratch_on_current_device();
    if (!base) return -1;
    SheenContext* c = ctx;
    const int L = c->ds.num_lights;
    const bool caustic = c->ds.pm.photons && c->ds.caustics > 0.0f;
    const bool media = c->ds.num_media > 0;
    const size_t unit = (size_t)aa_samples * 64;  // one 8x8 tile of a pixel's samples
    FixedBufs fx;
    // bytes grow with the slot count (plus alignment): the largest whole
    // number of tiles that fits the array
    const size_t fixed = carve_work(nullptr, 0, L, c->n_beam, caustic, media, fx.wb, &fx.ctr);
    const size_t per_unit = carve_work(nullptr, unit, L, c->n_beam, caustic, media, fx.wb, &fx.ctr) - fixed;
    size_t chunk = 0;
    if (fixed < SHEEN_NOALLOC_SCRATCH_BYTES && per_unit > 0)
        chunk = (SHEEN_NOALLOC_SCRATCH_BYTES - fixed) / per_unit * unit;
    chunk = std::min(chunk, (size_t)0x0FFFFFFF / unit * unit);
    while (chunk >= unit &&
           carve_work(nullptr, chunk, L, c->n_beam, caustic, media, fx.wb, &fx.ctr) > SHEEN_NOALLOC_SCRATCH_BYTES)
        chunk -= unit;
    if (chunk < unit) {
        sheen_set_error("one tile of %d samples with %d lights needs %zu bytes of work space, beyond the %u-byte scratch",
                        aa_samples, L, fixed + per_unit, (unsigned)SHEEN_NOALLOC_SCRATCH_BYTES);
        return -1;
    }
    carve_work(base, chunk, L, c->n_beam, caustic, media, fx.wb, &fx.ctr);
    fx.frame = reinterpret_cast<float4*>(out_rgba_device);
    fx.chunk = chunk;
    const bool prof = c->profiling;
    c->profiling = false;  // profiling events would be created on demand
    const int rc = render_frame(c, width, height, aa_samples, stats, false, nullptr, nullptr, nullptr, nullptr, &fx);
    c->profiling = prof;
    return rc;
}

int sheen_context_primary_hits(SheenContext* ctx, int width, int height, int aa_samples, int* kind, int* index,
                               float* t) {
    if (!ctx || !kind || !index || !t) { sheen_set_error("null argument"); return -1; }
    return render_frame(ctx, width, height, aa_samples, nullptr, true, kind, index, t);
}

int sheen_context_set_profiling(SheenContext* ctx, int enable) {
    if (!ctx) return -1;
    ctx->profiling = enable != 0;
    return 0;
}
int sheen_context_set_counting(SheenContext* ctx, int enable) {
    if (!ctx) return -1;
    ctx->counting = enable != 0;
    return 0;
}
int sheen_context_set_chunk(SheenContext* ctx, int samples_per_chunk) {
    if (!ctx) return -1;
    ctx->chunk_override = samples_per_chunk > 0 ? samples_per_chunk : 0;
    return 0;
}

int sheen_context_photon_stats(SheenContext* ctx, int* emitted, int* stored) {
    if (!ctx) return -1;
    if (emitted) *emitted = ctx->info.photons_emitted;
    if (stored) *stored = ctx->info.photons_stored;
    return 0;
}

int sheen_context_photon_map(SheenContext* ctx, Photon* photons, int* cell_start, int* cell_count) {
    if (!ctx) return -1;
    if (!ctx->photons) return 0;
    const int nc = ctx->ds.pm.num_cells;
    if (photons) HIP_OR(hipMemcpy(photons, ctx->photons, (size_t)ctx->n_photons * sizeof(Photon), hipMemcpyDeviceToHost), -1);
    if (cell_start) HIP_OR(hipMemcpy(cell_start, ctx->cell_start, (size_t)nc * 4, hipMemcpyDeviceToHost), -1);
    if (cell_count) HIP_OR(hipMemcpy(cell_count, ctx->cell_count, (size_t)nc * 4, hipMemcpyDeviceToHost), -1);
    return ctx->n_photons;
}

int sheen_context_photon_probe(SheenContext* ctx, int stride, float radius, int max_samples, int brute, int* counts) {
    if (!ctx || stride <= 0 || !counts) { sheen_set_error("bad probe arguments"); return -1; }
    if (!ctx->photons || ctx->n_photons <= 0) return 0;
    const uint32_t n = (uint32_t)ctx->n_photons;
    const uint32_t probes = (n + stride - 1) / stride;
    int* d = nullptr;
    if (!alloc_dev((void**)&d, (size_t)probes * 4)) { sheen_set_error("probe alloc failed"); return -1; }
    k_photon_probe<<<grid_for(ctx, probes, 8), 256, 0, ctx->stream>>>(ctx->ds.pm, ctx->photons, n, stride, radius,
                                                                        max_samples, d, brute);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(counts, d, (size_t)probes * 4, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) { sheen_set_error("probe failed: %s", hipGetErrorString(e)); return -1; }
    return (int)probes;
}

int sheen_context_set_param(SheenContext* ctx, const char* name, int value) {
    if (!ctx || !name) { sheen_set_error("null argument"); return -1; }
    const std::string n(name);
    if (value < 0) { sheen_set_error("negative parameter value"); return -1; }
    if (n == "overlap") ctx->overlap = value != 0;
    else if (n == "light_runs") ctx->light_runs = value != 0;
    else if (n == "sec_sort") {
        if (value < 0 || value > 2) { sheen_set_error("sec_sort must be 0, 1 or 2"); return -1; }
        ctx->sec_sort = (int)value;
    }
    else if (n == "shadow_order") ctx->order = value ? (ctx->order | ORDER_SHADOW) : (ctx->order & ~ORDER_SHADOW);
    else if (n == "sec_order") ctx->order = value ? (ctx->order | ORDER_SEC) : (ctx->order & ~ORDER_SEC);
    else if (n == "skip_dark") ctx->order = value ? (ctx->order | ORDER_SKIP_DARK) : (ctx->order & ~ORDER_SKIP_DARK);
    else if (n == "primary_grid") ctx->primary_grid = value;
    else if (n == "sec_grid") ctx->sec_grid = value;
    else if (n == "shadow_grid") ctx->shadow_grid = value;
    else if (n == "shade_grid") ctx->shade_grid = value;
    else if (n == "shade_sec_grid") ctx->shade_sec_grid = value;
    else if (n == "shade_block") {
        if (value % 64 != 0 || value > 1024) { sheen_set_error("shade_block must be a multiple of 64 up to 1024"); return -1; }
        ctx->shade_block = value;
    }
    else if (n == "adjoint_grid") ctx->adjoint_grid = value;
    else if (n == "chunk") ctx->chunk_override = value > 0 ? value : 0;
    else { sheen_set_error("unknown parameter: %s", name); return -1; }
    return 0;
}

// ---- differentiable shading ------------------------------------------------

static int adj_check(SheenContext* c) {
    if (!c->adj_reason.empty()) {
        sheen_set_error("scene is not differentiable: %s", c->adj_reason.c_str());
        return -1;
    }
    return 0;
}

// Float chains, gradient chain and the scene's own materials/lights (once).
static int adj_alloc(SheenContext* c) {
    if (!c->mats_orig) {
        c->mats_orig = c->alloc_owned<Material>((size_t)c->ds.num_materials);
        if (!c->mats_orig) { sheen_set_error("adjoint alloc failed"); return -1; }
        HIP_OR(hipMemcpyAsync(c->mats_orig, c->ds.materials, sizeof(Material) * c->ds.num_materials, hipMemcpyDeviceToDevice,
                              c->stream), -1);
        if (c->ds.num_lights > 0) {
            c->lights_orig = c->alloc_owned<Light>((size_t)c->ds.num_lights);
            if (!c->lights_orig) { sheen_set_error("adjoint alloc failed"); return -1; }
            HIP_OR(hipMemcpyAsync(c->lights_orig, c->ds.lights, sizeof(Light) * c->ds.num_lights, hipMemcpyDeviceToDevice,
                                  c->stream), -1);
        }
    }
    if (c->tex_elems > 0 && !c->tex_f) {
        c->tex_f = c->alloc_owned<float>(c->tex_elems);
        c->gtex = c->alloc_owned<float>(c->tex_elems);
        c->tex_f_off = c->upload(c->tex_off.data(), c->tex_off.size());
        if (!c->tex_f || !c->gtex || !c->tex_f_off) { sheen_set_error("adjoint alloc failed"); return -1; }
    }
    return 0;
}

// Orders the render stream after the caller's stream (into_render) or the
// caller's stream after the render stream.  The caller's stream may be the
// null stream (user == NULL): the render stream is non-blocking, so nothing
// orders it against the null stream unless an event does.  When the event
// cannot be recorded or waited on, the host waits for the earlier stream
// instead, and the context counts the fallback.  0, or -1 with the error set.
static int adj_join(SheenContext* c, hipStream_t user, bool into_render) {
    hipEvent_t e = c->ev_join[into_render ? 0 : 1];
    hipStream_t from = into_render ? user : c->stream;
    hipStream_t to = into_render ? c->stream : user;
    hipError_t r = hipEventRecord(e, from);
    if (r == hipSuccess) r = hipStreamWaitEvent(to, e, 0);
    if (r == hipSuccess) return 0;
    (void)hipGetLastError();
    c->join_fallbacks++;
    r = hipStreamSynchronize(from);
    if (r != hipSuccess) {
        sheen_set_error("stream ordering failed: %s", hipGetErrorString(r));
        return -1;
    }
    return 0;
}

// Parameter set -> device materials, lights and float texel chains.
static int adj_set(SheenContext* c, const float* albedo, const float* emission, const float* intensity,
                   const float* texels) {
    hipStream_t s = c->stream;
    const int M = c->ds.num_materials, L = c->ds.num_lights;
    k_adj_set<<<grid_for(c, std::max(M, L), 8), 256, 0, s>>>((Material*)c->ds.materials, M, (Light*)c->ds.lights, L,
                                                             albedo, emission, intensity);
    size_t p = 0;
    for (size_t t = 0; t < c->tex_w.size(); t++) {
        if (!c->tex_w[t]) continue;
        int sw = c->tex_w[t], sh = c->tex_h[t];
        uint32_t src = c->tex_off[t];
        const size_t n0 = (size_t)sw * sh * 3;
        HIP_OR(hipMemcpyAsync(c->tex_f + src, texels + p, n0 * sizeof(float), hipMemcpyDeviceToDevice, s), -1);
        p += n0;
        for (int lv = 1; lv < c->tex_mips[t]; lv++) {
            const int nw = sw > 1 ? sw / 2 : 1, nh = sh > 1 ? sh / 2 : 1;
            const uint32_t dst = src + (uint32_t)sw * sh * 3u;
            k_mip_f<<<grid_for(c, (size_t)nw * nh, 8), 256, 0, s>>>(c->tex_f, src, sw, sh, dst, nw, nh);
            src = dst; sw = nw; sh = nh;
        }
    }
    c->ds.tex_f = c->tex_f;
    c->ds.tex_f_off = c->tex_f_off;
    return 0;
}

// Back to the scene's own materials, lights and 8-bit textures.
static void adj_restore(SheenContext* c) {
    hipStream_t s = c->stream;
    (void)hipMemcpyAsync((void*)c->ds.materials, c->mats_orig, sizeof(Material) * c->ds.num_materials,
                         hipMemcpyDeviceToDevice, s);
    if (c->lights_orig)
        (void)hipMemcpyAsync((void*)c->ds.lights, c->lights_orig, sizeof(Light) * c->ds.num_lights, hipMemcpyDeviceToDevice, s);
    c->ds.tex_f = nullptr;
    c->ds.tex_f_off = nullptr;
}

int sheen_context_stream_fallbacks(const SheenContext* ctx) { return ctx ? ctx->join_fallbacks : -1; }

int sheen_adjoint_sizes(SheenContext* ctx, int* materials, int* lights, long long* texels) {
    if (!ctx) { sheen_set_error("null context"); return -1; }
    long long n = 0;
    for (size_t t = 0; t < ctx->tex_w.size(); t++) n += (long long)ctx->tex_w[t] * ctx->tex_h[t] * 3;
    if (materials) *materials = ctx->ds.num_materials;
    if (lights) *lights = ctx->ds.num_lights;
    if (texels) *texels = n;
    return 0;
}

int sheen_adjoint_get_params(SheenContext* ctx, float* albedo, float* emission, float* intensity, float* texels,
                             void* stream) {
    if (!ctx || !albedo || !emission || (!intensity && ctx->ds.num_lights > 0)) { sheen_set_error("null argument"); return -1; }
    SheenContext* c = ctx;
    if (adj_alloc(c) != 0) return -1;
    if (adj_join(c, (hipStream_t)stream, true) != 0) return -1;
    const int M = c->ds.num_materials, L = c->ds.num_lights;
    k_adj_get<<<grid_for(c, std::max(M, L), 8), 256, 0, c->stream>>>(c->mats_orig, M, c->lights_orig, L, albedo, emission,
                                                                      intensity);
    size_t p = 0;
    std::vector<TextureRef> refs(c->ds.num_textures);
    if (c->ds.num_textures > 0)
        HIP_OR(hipMemcpy(refs.data(), c->ds.textures, sizeof(TextureRef) * refs.size(), hipMemcpyDeviceToHost), -1);
    for (size_t t = 0; t < c->tex_w.size(); t++) {
        if (!c->tex_w[t]) continue;
        const uint32_t n0 = (uint32_t)c->tex_w[t] * c->tex_h[t] * 3u;
        if (!texels) { sheen_set_error("null texel buffer"); return -1; }
        k_bytes_to_float<<<grid_for(c, n0, 8), 256, 0, c->stream>>>(refs[t].pixels, n0, texels + p);
        p += n0;
    }
    HIP_OR(hipGetLastError(), -1);
    return adj_join(c, (hipStream_t)stream, false);
}

int sheen_adjoint_forward(SheenContext* ctx, int width, int height, int aa_samples, const float* albedo,
                          const float* emission, const float* intensity, const float* texels, float* image,
                          void* stream) {
    if (!ctx || !albedo || !emission || !image) { sheen_set_error("null argument"); return -1; }
    SheenContext* c = ctx;
    if (adj_check(c) != 0 || adj_alloc(c) != 0) return -1;
    if (adj_join(c, (hipStream_t)stream, true) != 0) return -1;
    int rc = adj_set(c, albedo, emission, intensity, texels);
    if (rc == 0) rc = render_frame(c, width, height, aa_samples, nullptr, false, nullptr, nullptr, nullptr);
    if (rc == 0 && hipMemcpyAsync(image, c->frame, (size_t)width * height * sizeof(float4), hipMemcpyDeviceToDevice,
                                  c->stream) != hipSuccess) {
        sheen_set_error("image copy failed");
        rc = -1;
    }
    adj_restore(c);
    if (adj_join(c, (hipStream_t)stream, false) != 0) rc = -1;
    return rc;
}

int sheen_adjoint_backward(SheenContext* ctx, int width, int height, int aa_samples, const float* albedo,
                           const float* emission, const float* intensity, const float* texels, const float* grad_image,
                           float* d_albedo, float* d_emission, float* d_intensity, float* d_texels, void* stream) {
    if (!ctx || !albedo || !emission || !grad_image || !d_albedo || !d_emission) { sheen_set_error("null argument"); return -1; }
    SheenContext* c = ctx;
    if (adj_check(c) != 0 || adj_alloc(c) != 0) return -1;
    if (c->ds.num_lights > 0 && !d_intensity) { sheen_set_error("null intensity gradient"); return -1; }
    if (c->tex_elems > 0 && !d_texels) { sheen_set_error("null texel gradient"); return -1; }
    if (adj_join(c, (hipStream_t)stream, true) != 0) return -1;
    hipStream_t s = c->stream;
    const int M = c->ds.num_materials, L = c->ds.num_lights;
    k_fill_f<<<grid_for(c, (size_t)M * 3, 8), 256, 0, s>>>(d_albedo, (size_t)M * 3, 0.0f);
    k_fill_f<<<grid_for(c, (size_t)M * 3, 8), 256, 0, s>>>(d_emission, (size_t)M * 3, 0.0f);
    if (L > 0) k_fill_f<<<1, 256, 0, s>>>(d_intensity, (size_t)L, 0.0f);
    if (c->tex_elems > 0) k_fill_f<<<grid_for(c, c->tex_elems, 8), 256, 0, s>>>(c->gtex, c->tex_elems, 0.0f);
    int rc = adj_set(c, albedo, emission, intensity, texels);
    AdjCall call;
    call.grad = grad_image;
    call.G.albedo = d_albedo;
    call.G.emission = d_emission;
    call.G.intensity = d_intensity;
    call.G.tex = c->gtex;
    if (rc == 0) rc = render_frame(c, width, height, aa_samples, nullptr, false, nullptr, nullptr, nullptr, &call);
    if (rc == 0) {
        // mip chain adjoint, top level down, then the level-0 gradients out
        size_t p = 0;
        for (size_t t = 0; t < c->tex_w.size(); t++) {
            if (!c->tex_w[t]) continue;
            std::vector<uint32_t> off;
            std::vector<int> lw, lh;
            int sw = c->tex_w[t], sh = c->tex_h[t];
            uint32_t o = c->tex_off[t];
            for (int lv = 0; lv < c->tex_mips[t]; lv++) {
                off.push_back(o); lw.push_back(sw); lh.push_back(sh);
                o += (uint32_t)sw * sh * 3u;
                sw = sw > 1 ? sw / 2 : 1;
                sh = sh > 1 ? sh / 2 : 1;
            }
            for (int lv = c->tex_mips[t] - 1; lv >= 1; lv--)
                k_mip_adj<<<grid_for(c, (size_t)lw[lv] * lh[lv], 8), 256, 0, s>>>(c->gtex, off[lv - 1], lw[lv - 1], lh[lv - 1],
                                                                                  off[lv], lw[lv], lh[lv]);
            const size_t n0 = (size_t)c->tex_w[t] * c->tex_h[t] * 3;
            if (hipMemcpyAsync(d_texels + p, c->gtex + c->tex_off[t], n0 * sizeof(float), hipMemcpyDeviceToDevice, s) != hipSuccess) {
                sheen_set_error("gradient copy failed");
                rc = -1;
                break;
            }
            p += n0;
        }
    }
    adj_restore(c);
    if (rc == 0 && hipGetLastError() != hipSuccess) { sheen_set_error("adjoint launch failed"); rc = -1; }
    if (adj_join(c, (hipStream_t)stream, false) != 0) rc = -1;
    return rc;
}

// ---- batches of scenes ----------------------------------------------------
// Every scene keeps its own context (BVHs, records, camera); the batch adds a
// device table of their DScene records and frames and renders all of them as
// one wavefront pipeline over the concatenated sample slots, captured once
// into a hipGraph and replayed.  The scenes' own contexts still render one
// by one (the serial path), sharing one set of work buffers.

}  // extern "C"

struct SheenBatch {
    int n = 0;
    int num_cu = 256;
    std::vector<SheenContext*> ctx;
    std::vector<FrameDesc> frames;
    std::vector<uint32_t> base;  // n + 1 first slots
    std::vector<std::pair<uint32_t, uint32_t>> chunks;
    uint64_t samples = 0;
    size_t pixels = 0;
    int L = 1;
    DScene* d_scenes = nullptr;
    DScene* d_scenes1 = nullptr;  // copies with wave_one_scene set
    FrameDesc* d_frames = nullptr;
    uint32_t* d_base = nullptr;
    float4* frame = nullptr;
    Counters* ctr = nullptr;
    WorkBufs wb, serial_wb;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int primary_grid = 64, sec_grid = 4, shadow_grid = 128, shade_grid = 1, shade_sec_grid = 1;  // tuned on the 256-scene batch
    // some scene holds primitives beyond spheres, triangles and planes, media
    // or beams (the pipeline then runs the instantiations for every kind);
    // caustics, media and the largest beam count size the shared buffers
    bool full = false, caustic = false, media = false;
    bool any_secondary = false, any_glass = false;  // of any scene (SheenContext)
    int n_beam = 0;
    int order = ORDER_SHADOW | ORDER_SEC;  // as SheenContext::order
    int shade_block = 0;                   // as SheenContext::shade_block
    int sec_sort = 0;                      // as SheenContext::sec_sort
    size_t chunk_slots = 0;
    bool use_graph = true, dirty = true;
    int scene_table = SCENE_LDS;  // LDS entries of the slot -> scene search
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~SheenBatch() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (SheenContext* c : ctx) delete c;
        wb.release();
        serial_wb.release();
        void* ps[] = {d_scenes, d_scenes1, d_frames, d_base, frame, ctr};
        for (void* p : ps) if (p) (void)hipFree(p);
        for (hipEvent_t e : ev) if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// Chunks of at most `cap` slots, cut at tile boundaries (a pixel's samples
// stay in one chunk).
static void batch_chunks(SheenBatch* b, size_t cap) {
    b->chunks.clear();
    uint32_t start = 0, len = 0;
    for (int k = 0; k < b->n; k++) {
        const uint32_t unit = 64u * (uint32_t)b->frames[k].ns;
        uint32_t left = b->frames[k].n_slots;
        while (left > 0) {
            uint32_t room = (uint32_t)cap - len;
            if (room < unit) {
                b->chunks.push_back({start, len});
                start += len;
                len = 0;
                room = (uint32_t)cap;
            }
            const uint32_t take = std::min(left, room / unit * unit);
            len += take;
            left -= take;
        }
    }
    if (len > 0) b->chunks.push_back({start, len});
}

// The batch frame's kernels on the batch stream (captured into the graph).
// Default chunk of a batch: 64 M slots within 48 GiB of work buffers (an
// MI355X holds 288 GB; on the 256-scene batch, 64 M-slot chunks measured 2 %
// faster than 32 M-slot ones, fewer chunk tails).
static size_t batch_default_chunk(const SheenBatch* b) {
    return std::min((size_t)64 << 20, ((size_t)48 << 30) / slot_bytes(b->L, b->n_beam, b->caustic, b->media));
}
static bool batch_bufs(SheenBatch* b, size_t slots) {
    return ensure_bufs(b->wb, slots, b->L, b->n_beam, b->caustic, b->media);
}

template <bool FULL>
static void batch_enqueue_t(SheenBatch* b) {
    hipStream_t s = b->stream;
    WorkBufs& w = b->wb;
    const int cu = b->num_cu;
    const int SB_ = shade_block_of<FULL>(b->shade_block);
    k_zero_counters<<<1, 256, 0, s>>>(b->ctr);
    EventBufs EB;
    EB.base = w.ev_base; EB.caustic = b->caustic ? w.ev_caustic : nullptr; EB.contrib = w.contrib; EB.trans = w.trans;
    EB.shadow_q = w.shadow_q; EB.shadow_nl = w.shadow_nl; EB.seg = 3u * (uint32_t)w.cap_slots; EB.L = b->L;
    EB.cont[0] = w.cont[0]; EB.cont[1] = w.cont[1];
    EB.cont_cap = FULL ? 0u : (uint32_t)cont_entries(3 * w.cap_slots * (size_t)std::max(w.cap_L, 1));
    EB.runs = FULL ? 1 : std::max(b->L, 1);  // k_shadow<*, true> reads one run
    EB.order = b->order;
    EB.cap = (uint32_t)w.cap_slots;
    SampleBufs SB;
    SB.hit_p = w.hit_p; SB.samp = w.samp;
    SB.beam_add = b->n_beam ? w.beam_add : nullptr;
    SB.media_p = b->media ? w.media_p : nullptr;
    SB.media_r = b->media ? w.media_r : nullptr;
    SB.n_beam = b->n_beam;
    for (const auto& ch : b->chunks) {
        ManySrc src;
        src.S = b->d_scenes; src.S1 = b->d_scenes1; src.F = b->d_frames; src.base = b->d_base; src.n = b->n;
        src.cap = b->scene_table; src.tab = nullptr; src.step = 1; src.m = 1;
        src.g0 = ch.first; src.count = ch.second;
        const size_t n = ch.second;
        k_reset<<<1, 64, 0, s>>>(b->ctr, EB.shadow_nl, EB.runs);
        k_primary<false, FULL><<<grid_n(cu, n, b->primary_grid), TRACE_BLOCK, 0, s>>>(src, w.hit_p, b->ctr);
        k_shade_primary<FULL><<<grid_n(cu, n, b->shade_grid, SB_), SB_, 0, s>>>(src, SB, EB, w.sec_q[0], b->ctr);
        for (int round = 0; round < (b->any_secondary ? 3 : 0); round++) {
            const bool sorted = b->sec_sort > 0;  // as in run_chunk
            const int in_q = sorted ? 1 : (round & 1), out_q = sorted ? 0 : ((round + 1) & 1);
            if (sorted) launch_sec_sort(s, cu, w.sec_q[0], w.sec_q[1], b->ctr, round, b->sec_sort, w.sort_hist, w.cap_slots);
            k_trace_sec<false, FULL><<<grid_n(cu, 2 * n, b->sec_grid), TRACE_BLOCK, 0, s>>>(src, w.sec_q[in_q], round,
                                                                                           w.sec_hits, b->ctr);
            k_shade_sec<FULL><<<grid_n(cu, 2 * n, b->shade_sec_grid, SB_), SB_, 0, s>>>(
                src, SB, EB, w.sec_q[in_q], round, w.sec_hits, w.sec_q[out_q], round + 1, b->ctr);
        }
        k_shadow<false, FULL><<<grid_n(cu, 3 * n * (size_t)b->L, b->shadow_grid), TRACE_BLOCK, 0, s>>>(src, EB, b->ctr, 0);
        if (b->any_glass) launch_shadow_cont<FULL>(s, grid_n(cu, EB.cont_cap, b->shadow_grid), false, src, EB, b->ctr);
        k_accumulate<<<grid_n(cu, n, 8), 256, 0, s>>>(src, SB, EB, b->frame);
    }
}

static void batch_enqueue(SheenBatch* b) {
    if (b->full) batch_enqueue_t<true>(b);
    else batch_enqueue_t<false>(b);
}

static void batch_stats(SheenRayStats* S, const Counters& h, uint64_t samples) {
    S->samples = samples;
    S->primary_rays = samples;
    S->shadow_rays = ctr_total(h.shadow_passes) + ctr_total(h.extra_shadow);
    S->reflection_rays = ctr_total(h.reflection);
    S->refraction_rays = ctr_total(h.refraction);
    S->total_rays = S->primary_rays + S->shadow_rays + S->reflection_rays + S->refraction_rays;
}

extern "C" {

SheenBatch* sheen_batch_create(const SheenScene* const* scenes, const int* widths, const int* heights,
                               const int* aa_samples, int count) {
    if (!scenes || !widths || !heights || !aa_samples || count <= 0) { sheen_set_error("bad batch arguments"); return nullptr; }
    std::unique_ptr<SheenBatch> b(new SheenBatch());
    b->n = count;
    HIP_OR(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), nullptr);
    HIP_OR(hipEventCreate(&b->ev[0]), nullptr);
    HIP_OR(hipEventCreate(&b->ev[1]), nullptr);
    uint64_t slot0 = 0;
    size_t px0 = 0;
    for (int k = 0; k < count; k++) {
        const int W = widths[k], H = heights[k], A = aa_samples[k];
        if (W <= 0 || H <= 0 || A < 1 || A > 64) { sheen_set_error("bad render params for batch scene %d", k); return nullptr; }
        SheenContext* c = context_create(scenes[k]);
        if (!c) return nullptr;
        b->ctx.push_back(c);
        b->full = b->full || c->full;
        b->any_secondary = b->any_secondary || c->any_secondary;
        b->any_glass = b->any_glass || c->any_glass;
        b->caustic = b->caustic || (c->ds.pm.photons && c->ds.caustics > 0.0f);
        b->media = b->media || c->ds.num_media > 0;
        b->n_beam = std::max(b->n_beam, c->n_beam);
        c->wbp = &b->serial_wb;
        b->num_cu = c->num_cu;
        b->L = std::max(b->L, c->ds.num_lights);
        FrameDesc F;
        F.width = W; F.height = H; F.ns = A; F.tiles_x = (W + 7) / 8;
        const uint64_t slots = (uint64_t)F.tiles_x * ((H + 7) / 8) * A * 64;
        if (slot0 + slots > 0xFFFFFFFFull) { sheen_set_error("batch exceeds 2^32 sample slots"); return nullptr; }
        F.g0 = (uint32_t)slot0;
        F.n_slots = (uint32_t)slots;
        F.out0 = (uint32_t)px0;
        b->frames.push_back(F);
        b->base.push_back((uint32_t)slot0);
        slot0 += slots;
        px0 += (size_t)W * H;
        b->samples += (uint64_t)W * H * A;
    }
    b->base.push_back((uint32_t)slot0);
    b->pixels = px0;
    std::vector<DScene> table;
    for (SheenContext* c : b->ctx) {
        table.push_back(c->ds);
        table.back().wave_one_scene = 0;  // a wave of the batch pipeline may hold rays of several scenes
    }
    std::vector<DScene> table1(table);
    for (DScene& d : table1) d.wave_one_scene = 1;  // for waves whose lanes all trace one scene
    if (!alloc_dev((void**)&b->d_scenes, sizeof(DScene) * count) || !alloc_dev((void**)&b->d_frames, sizeof(FrameDesc) * count) ||
        !alloc_dev((void**)&b->d_base, sizeof(uint32_t) * (count + 1)) || !alloc_dev((void**)&b->frame, sizeof(float4) * px0) ||
        !alloc_dev((void**)&b->ctr, sizeof(Counters))) {
        sheen_set_error("batch allocation failed");
        return nullptr;
    }
    HIP_OR(hipMemcpy(b->d_scenes, table.data(), sizeof(DScene) * count, hipMemcpyHostToDevice), nullptr);
    if (!alloc_dev((void**)&b->d_scenes1, sizeof(DScene) * count)) { sheen_set_error("batch allocation failed"); return nullptr; }
    HIP_OR(hipMemcpy(b->d_scenes1, table1.data(), sizeof(DScene) * count, hipMemcpyHostToDevice), nullptr);
    HIP_OR(hipMemcpy(b->d_frames, b->frames.data(), sizeof(FrameDesc) * count, hipMemcpyHostToDevice), nullptr);
    HIP_OR(hipMemcpy(b->d_base, b->base.data(), sizeof(uint32_t) * (count + 1), hipMemcpyHostToDevice), nullptr);
    // as a context's default; the scenes' own renders sort alike
    b->sec_sort = b->full ? 0 : 2;
    for (SheenContext* c : b->ctx) c->sec_sort = b->sec_sort;
    b->chunk_slots = batch_default_chunk(b.get());
    b->chunk_slots = std::min<size_t>(b->chunk_slots, ((size_t)slot0 + 63) / 64 * 64 + 64 * 64);
    batch_chunks(b.get(), b->chunk_slots);
    if (!batch_bufs(b.get(), b->chunk_slots)) return nullptr;
    return b.release();
}

void sheen_batch_destroy(SheenBatch* b) { delete b; }

int sheen_batch_set_param(SheenBatch* b, const char* name, int value) {
    if (!b || !name) { sheen_set_error("null argument"); return -1; }
    if (value < 0) { sheen_set_error("negative parameter value"); return -1; }
    const std::string n(name);
    if (n == "primary_grid") b->primary_grid = value;
    else if (n == "sec_grid") b->sec_grid = value;
    else if (n == "shadow_grid") b->shadow_grid = value;
    else if (n == "shade_grid") b->shade_grid = value;
    else if (n == "shade_sec_grid") b->shade_sec_grid = value;
    else if (n == "shade_block") {
        if (value % 64 != 0 || value > 1024) { sheen_set_error("shade_block must be a multiple of 64 up to 1024"); return -1; }
        b->shade_block = value;
    }
    else if (n == "graph") b->use_graph = value != 0;
    else if (n == "shadow_order") b->order = value ? (b->order | ORDER_SHADOW) : (b->order & ~ORDER_SHADOW);
    else if (n == "sec_order") b->order = value ? (b->order | ORDER_SEC) : (b->order & ~ORDER_SEC);
    else if (n == "skip_dark") b->order = value ? (b->order | ORDER_SKIP_DARK) : (b->order & ~ORDER_SKIP_DARK);
    else if (n == "sec_sort") {
        if (value < 0 || value > 2) { sheen_set_error("sec_sort must be 0, 1 or 2"); return -1; }
        b->sec_sort = (int)value;
        for (SheenContext* c : b->ctx) c->sec_sort = (int)value;  // the serial renders sort alike
    }
    else if (n == "scene_table") {
        if (value < 1 || value > SCENE_LDS) { sheen_set_error("scene_table must be 1..%d", SCENE_LDS); return -1; }
        b->scene_table = value;
    } else if (n == "chunk") {
        size_t cap = value > 0 ? (size_t)value : batch_default_chunk(b);
        if (value == 0) cap = std::min<size_t>(cap, ((size_t)b->base[b->n] + 63) / 64 * 64 + 64 * 64);  // no larger than the batch
        size_t unit = 64;
        for (const FrameDesc& F : b->frames) unit = std::max(unit, (size_t)64 * F.ns);
        if (cap < unit) { sheen_set_error("chunk smaller than a tile of the batch"); return -1; }
        // a chunk-local slot index must fit the 28 bits SecRay.meta holds below its flags
        if (cap > 0x0FFFFFFF) { sheen_set_error("chunk above 2^28 - 1 slots"); return -1; }
        if (!batch_bufs(b, cap)) return -1;
        b->chunk_slots = cap;
        batch_chunks(b, cap);
    } else { sheen_set_error("unknown parameter: %s", name); return -1; }
    b->dirty = true;
    return 0;
}

int sheen_batch_render(SheenBatch* b, SheenRayStats* stats) {
    if (!b) { sheen_set_error("null batch"); return -1; }
    hipStream_t s = b->stream;
    if (b->use_graph && (b->dirty || !b->exec)) {
        if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
        if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
        HIP_OR(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), -1);
        batch_enqueue(b);
        HIP_OR(hipStreamEndCapture(s, &b->graph), -1);
        HIP_OR(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), -1);
        b->dirty = false;
    }
    const auto t0 = std::chrono::steady_clock::now();
    HIP_OR(hipEventRecord(b->ev[0], s), -1);
    if (b->use_graph) HIP_OR(hipGraphLaunch(b->exec, s), -1);
    else batch_enqueue(b);
    HIP_OR(hipEventRecord(b->ev[1], s), -1);
    HIP_OR(hipGetLastError(), -1);
    Counters host{};
    HIP_OR(hipMemcpyAsync(&host, b->ctr, sizeof(Counters), hipMemcpyDeviceToHost, s), -1);
    HIP_OR(hipStreamSynchronize(s), -1);
    const auto t1 = std::chrono::steady_clock::now();
    if (stats) {
        SheenRayStats S{};
        batch_stats(&S, host, b->samples);
        HIP_OR(hipEventElapsedTime(&S.frame_ms, b->ev[0], b->ev[1]), -1);
        S.trace_ms = (float)std::chrono::duration<double, std::milli>(t1 - t0).count();  // wall clock of the call
        S.shadow_launches = (int)b->chunks.size();
        *stats = S;
    }
    return 0;
}

int sheen_batch_render_serial(SheenBatch* b, SheenRayStats* stats) {
    if (!b) { sheen_set_error("null batch"); return -1; }
    SheenRayStats T{};
    const auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < b->n; k++) {
        SheenRayStats S{};
        const FrameDesc& F = b->frames[k];
        if (render_frame(b->ctx[k], F.width, F.height, F.ns, &S, false, nullptr, nullptr, nullptr) != 0) return -1;
        T.samples += S.samples; T.primary_rays += S.primary_rays; T.shadow_rays += S.shadow_rays;
        T.reflection_rays += S.reflection_rays; T.refraction_rays += S.refraction_rays; T.total_rays += S.total_rays;
        T.frame_ms += S.frame_ms;
        T.shadow_launches += S.shadow_launches;
    }
    const auto t1 = std::chrono::steady_clock::now();
    T.trace_ms = (float)std::chrono::duration<double, std::milli>(t1 - t0).count();
    if (stats) *stats = T;
    return 0;
}

int sheen_batch_image(SheenBatch* b, int index, int serial, float* out_rgba) {
    if (!b || !out_rgba || index < 0 || index >= b->n) { sheen_set_error("bad batch image request"); return -1; }
    const FrameDesc& F = b->frames[index];
    const size_t bytes = (size_t)F.width * F.height * sizeof(float4);
    if (serial) {
        SheenContext* c = b->ctx[index];
        if (!c->frame || c->frame_w != F.width || c->frame_h != F.height) { sheen_set_error("scene not rendered serially"); return -1; }
        HIP_OR(hipMemcpy(out_rgba, c->frame, bytes, hipMemcpyDeviceToHost), -1);
    } else {
        HIP_OR(hipStreamSynchronize(b->stream), -1);
        HIP_OR(hipMemcpy(out_rgba, b->frame + F.out0, bytes, hipMemcpyDeviceToHost), -1);
    }
    return 0;
}

int sheen_batch_info(SheenBatch* b, int* scenes, long long* samples, int* chunks) {
    if (!b) { sheen_set_error("null batch"); return -1; }
    if (scenes) *scenes = b->n;
    if (samples) *samples = (long long)b->samples;
    if (chunks) *chunks = (int)b->chunks.size();
    return 0;
}

int sheen_render_float(const SheenScene* scene, int width, int height, int aa_samples, float* out_rgba,
                       SheenRayStats* stats) {
    if (width <= 0 || height <= 0 || aa_samples < 1 || aa_samples > 64 || !out_rgba) {
        sheen_set_error("bad render params");
        return -1;
    }
    SheenContext* c = context_create(scene);
    if (!c) return -1;
    int rc = sheen_context_render(c, width, height, aa_samples, out_rgba, stats);
    delete c;
    return rc;
}

int sheen_render_scene(const SheenScene* scene, int width, int height, int aa_samples, const char* output_path) {
    if (!scene) { sheen_set_error("null scene"); return -1; }
    if (width <= 0 || height <= 0 || aa_samples < 1 || aa_samples > 64 || !output_path) {
        sheen_set_error("bad render params");
        return -1;
    }
    if (!scene->num_materials) { sheen_set_error("no materials"); return -1; }
    std::vector<float> rgba((size_t)width * height * 4);
    if (sheen_render_float(scene, width, height, aa_samples, rgba.data(), nullptr) != 0) return -1;
    std::vector<unsigned char> rgb((size_t)width * height * 3);
    sheen_encode_srgb(rgba.data(), width, height, rgb.data());
    if (sheen_write_png(output_path, width, height, rgb.data()) != 0) {
        sheen_set_error("stbi_write_png failed");  // the reference's message (api/sheen_c_api.cu:356)
        return -1;
    }
    return 0;
}

}  // extern "C"
