// wbvh_gpu.hip — device builder of the 8-wide BVH (see wbvh_gpu.h).
//
// Nodes over more than SMALL_MAX primitives are split one tree level at a
// time: each node's index range is cut into CHUNK-element tasks, one
// workgroup each; node bounds and the 16-bin histograms of the three axes
// reduce in LDS and meet in global memory through atomic min/max on
// order-preserving integer images of the floats (exact, so the order of
// arrival does not matter).  One thread per node takes the SAH choice; a
// stable partition places each primitive by prefix counts.  A node of at
// most SMALL_MAX primitives is built whole, subtree and all, by one
// wavefront working from LDS.  The binary nodes are then collapsed into wide
// nodes one wide level at a time; prefix sums over each level give the
// host builder's breadth-first numbering of nodes and primitive slots.
// Binary node ids are handed out by an atomic counter: they differ from run
// to run, the wide tree does not.
#include "wbvh_gpu.h"

#include <algorithm>
#include <cstring>
#include <vector>

#include "sheen_types.h"
#include "wbvh_rules.h"

namespace {

constexpr int SMALL_MAX = 512;  // primitives of a node built by one wavefront
constexpr int CHUNK = 4096;     // primitives per workgroup task of a large node
constexpr int NB = WBVH_BINS;
constexpr int BIN_WORDS = 3 * NB * 7;      // per axis and bin: min xyz, max xyz, count
constexpr int ACC_WORDS = 12 + BIN_WORDS;  // node box, centroid box, bins
constexpr int SMALL_STACK = 16;            // log2(SMALL_MAX) + 1 entries suffice
constexpr uint32_t ORD_PINF = 0xFF800000u;
constexpr uint32_t ORD_NINF = 0x007FFFFFu;

struct BNode {
    WbBox box;
    int left, right;   // binary children; -1 for a leaf
    int start, count;  // leaf: range in the partitioned index array
};
struct Seg { int node, s, e; };
struct Task { int seg, cs, ce, first; };  // first: task index of the segment's first task
struct SegInfo { int axis, bin, mid, nleft; float lo, k; };

// order-preserving image of a float as an unsigned integer
__device__ __forceinline__ uint32_t ord(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}
// initial accumulator word: boxes start empty (+inf, -inf), counts at 0
__device__ __forceinline__ uint32_t init_word(int w) {
    if (w < 12) return ((w / 3) & 1) ? ORD_NINF : ORD_PINF;
    const int r = (w - 12) % 7;
    return r < 3 ? ORD_PINF : (r < 6 ? ORD_NINF : 0u);
}
__device__ __forceinline__ float comp(const float4& c, int a) { return a == 0 ? c.x : (a == 1 ? c.y : c.z); }

// ---------------------------------------------------------------- scans

constexpr int SCAN_T = 256, SCAN_PER = 4, SCAN_BLK = SCAN_T * SCAN_PER;

__device__ uint32_t block_exclusive(uint32_t v, uint32_t* sh, uint32_t& total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
        const uint32_t x = t >= off ? sh[t - off] : 0u;
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    total = sh[blockDim.x - 1];
    const uint32_t r = sh[t] - v;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_local(const uint32_t* in, uint32_t* out, uint32_t* sums, int n) {
    __shared__ uint32_t sh[SCAN_T];
    const long base = (long)blockIdx.x * SCAN_BLK + threadIdx.x * SCAN_PER;
    uint32_t v[SCAN_PER], s = 0;
    for (int j = 0; j < SCAN_PER; j++) {
        v[j] = base + j < n ? in[base + j] : 0u;
        s += v[j];
    }
    uint32_t tot;
    uint32_t run = block_exclusive(s, sh, tot);
    for (int j = 0; j < SCAN_PER; j++) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_sums(uint32_t* sums, int nb, uint32_t* total) {
    __shared__ uint32_t sh[SCAN_T];
    uint32_t carry = 0;
    for (int c = 0; c < nb; c += SCAN_BLK) {
        const int base = c + threadIdx.x * SCAN_PER;
        uint32_t v[SCAN_PER], s = 0;
        for (int j = 0; j < SCAN_PER; j++) {
            v[j] = base + j < nb ? sums[base + j] : 0u;
            s += v[j];
        }
        uint32_t tot;
        uint32_t run = carry + block_exclusive(s, sh, tot);
        for (int j = 0; j < SCAN_PER; j++) {
            if (base + j < nb) sums[base + j] = run;
            run += v[j];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void k_scan_add(uint32_t* out, const uint32_t* sums, int n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += sums[i / SCAN_BLK];
}

// ---------------------------------------------------------------- inputs

__global__ void k_prep(const float* box, int n, const unsigned char* skip, WbBox* pb, float4* cen, uint32_t* flag) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    WbBox b;
    for (int a = 0; a < 3; a++) {
        b.mn[a] = wb_canon(box[6 * i + a]);
        b.mx[a] = wb_canon(box[6 * i + 3 + a]);
    }
    pb[i] = b;
    cen[i] = make_float4(wb_centroid(b.mn[0], b.mx[0]), wb_centroid(b.mn[1], b.mx[1]), wb_centroid(b.mn[2], b.mx[2]), 0.0f);
    flag[i] = (skip && skip[i]) ? 0u : 1u;
}

__global__ void k_compact(const uint32_t* flag, const uint32_t* pos, int n, uint32_t* idx) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) idx[pos[i]] = (uint32_t)i;
}

__global__ void k_tri_bounds(const Triangle* t, int n, float* box) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Triangle& r = t[i];
    float* o = box + 6 * i;
    o[0] = fminf(r.v0.x, fminf(r.v1.x, r.v2.x));
    o[1] = fminf(r.v0.y, fminf(r.v1.y, r.v2.y));
    o[2] = fminf(r.v0.z, fminf(r.v1.z, r.v2.z));
    o[3] = fmaxf(r.v0.x, fmaxf(r.v1.x, r.v2.x));
    o[4] = fmaxf(r.v0.y, fmaxf(r.v1.y, r.v2.y));
    o[5] = fmaxf(r.v0.z, fmaxf(r.v1.z, r.v2.z));
}

// ---------------------------------------------------------------- large nodes

__global__ void k_acc_init(uint32_t* acc, int nseg) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (long)nseg * ACC_WORDS) acc[i] = init_word((int)(i % ACC_WORDS));
}

__device__ __forceinline__ void reduce12(float* v) {
    for (int off = 32; off >= 1; off >>= 1)
        for (int k = 0; k < 12; k++) {
            const float o = __shfl_xor(v[k], off);
            v[k] = ((k / 3) & 1) ? fmaxf(v[k], o) : fminf(v[k], o);
        }
}

__device__ __forceinline__ void init12(float* v) {
    for (int k = 0; k < 12; k++) v[k] = ((k / 3) & 1) ? -INFINITY : INFINITY;
}

__device__ __forceinline__ void grow12(float* v, const WbBox& b, const float4& c) {
    for (int a = 0; a < 3; a++) {
        v[a] = fminf(v[a], b.mn[a]);
        v[3 + a] = fmaxf(v[3 + a], b.mx[a]);
        v[6 + a] = fminf(v[6 + a], comp(c, a));
        v[9 + a] = fmaxf(v[9 + a], comp(c, a));
    }
}

// node box and centroid box of each task's chunk, merged per node
__global__ __launch_bounds__(256) void k_lbox(const Task* tasks, const Seg* segs, const uint32_t* idx, const WbBox* pb,
                                              const float4* cen, uint32_t* acc) {
    __shared__ float sh[4][12];
    const Task t = tasks[blockIdx.x];
    float v[12];
    init12(v);
    for (int i = t.cs + threadIdx.x; i < t.ce; i += 256) {
        const uint32_t p = idx[i];
        grow12(v, pb[p], cen[p]);
    }
    reduce12(v);
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 12; k++) sh[wv][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 12) {
        const int k = threadIdx.x;
        float x = sh[0][k];
        for (int w = 1; w < 4; w++) x = ((k / 3) & 1) ? fmaxf(x, sh[w][k]) : fminf(x, sh[w][k]);
        uint32_t* a = acc + (size_t)t.seg * ACC_WORDS + k;
        if (x == x) {
            if ((k / 3) & 1) atomicMax(a, ord(x));
            else atomicMin(a, ord(x));
        }
    }
}

// adds primitive p to the LDS bins of the axes whose centroid extent is positive
__device__ __forceinline__ void bin_add(uint32_t* bins, const WbBox& b, const float4& c, const bool* ok, const float* lo,
                                        const float* k) {
    for (int a = 0; a < 3; a++) {
        if (!ok[a]) continue;
        uint32_t* w = bins + (a * NB + wb_bin_of(comp(c, a), lo[a], k[a])) * 7;
        for (int d = 0; d < 3; d++) {
            if (b.mn[d] == b.mn[d]) atomicMin(w + d, ord(b.mn[d]));
            if (b.mx[d] == b.mx[d]) atomicMax(w + 3 + d, ord(b.mx[d]));
        }
        atomicAdd(w + 6, 1u);
    }
}

__device__ __forceinline__ void axis_scales(const float* cmn, const float* cmx, bool* ok, float* lo, float* k) {
    for (int a = 0; a < 3; a++) {
        const float ext = cmx[a] - cmn[a];
        ok[a] = ext > 0.0f;
        lo[a] = cmn[a];
        k[a] = ok[a] ? wb_bin_scale(ext) : 0.0f;
    }
}

__global__ __launch_bounds__(256) void k_lbin(const Task* tasks, const uint32_t* idx, const WbBox* pb, const float4* cen,
                                              uint32_t* acc) {
    __shared__ uint32_t bins[BIN_WORDS];
    const Task t = tasks[blockIdx.x];
    uint32_t* A = acc + (size_t)t.seg * ACC_WORDS;
    for (int w = threadIdx.x; w < BIN_WORDS; w += 256) bins[w] = init_word(12 + w);
    float cmn[3], cmx[3], lo[3], k[3];
    bool ok[3];
    for (int a = 0; a < 3; a++) { cmn[a] = unord(A[6 + a]); cmx[a] = unord(A[9 + a]); }
    axis_scales(cmn, cmx, ok, lo, k);
    __syncthreads();
    for (int i = t.cs + threadIdx.x; i < t.ce; i += 256) {
        const uint32_t p = idx[i];
        bin_add(bins, pb[p], cen[p], ok, lo, k);
    }
    __syncthreads();
    for (int w = threadIdx.x; w < BIN_WORDS; w += 256) {
        const uint32_t x = bins[w];
        const int r = w % 7;
        uint32_t* d = A + 12 + w;
        if (r == 6) { if (x) atomicAdd(d, x); }
        else if (r < 3) { if (x != ORD_PINF) atomicMin(d, x); }
        else if (x != ORD_NINF) atomicMax(d, x);
    }
}

// SAH choice over bins held as accumulator words (LDS or global)
__device__ void pick_split(const uint32_t* bins, const float* cmn, const float* cmx, int& axis, int& bin, int& nleft) {
    wb_sah(cmn, cmx,
           [&](int a, int b) {
               const uint32_t* w = bins + (a * NB + b) * 7;
               WbBox r;
               for (int d = 0; d < 3; d++) { r.mn[d] = unord(w[d]); r.mx[d] = unord(w[3 + d]); }
               return r;
           },
           [&](int a, int b) { return (int)bins[(a * NB + b) * 7 + 6]; }, axis, bin);
    nleft = 0;
    if (axis >= 0)
        for (int b = 0; b < bin; b++) nleft += (int)bins[(axis * NB + b) * 7 + 6];
}

__global__ void k_lsah(const Seg* segs, int nseg, const uint32_t* acc, SegInfo* info, BNode* nodes) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nseg) return;
    const uint32_t* A = acc + (size_t)j * ACC_WORDS;
    const Seg sg = segs[j];
    WbBox box;
    float cmn[3], cmx[3];
    for (int a = 0; a < 3; a++) {
        box.mn[a] = unord(A[a]);
        box.mx[a] = unord(A[3 + a]);
        cmn[a] = unord(A[6 + a]);
        cmx[a] = unord(A[9 + a]);
    }
    nodes[sg.node].box = box;
    SegInfo in;
    pick_split(A + 12, cmn, cmx, in.axis, in.bin, in.nleft);
    if (in.axis < 0) {
        in.mid = sg.s + (sg.e - sg.s) / 2;  // all centroids coincide: split by index
        in.lo = in.k = 0.0f;
    } else {
        in.mid = sg.s + in.nleft;
        in.lo = cmn[in.axis];
        in.k = wb_bin_scale(cmx[in.axis] - cmn[in.axis]);
    }
    info[j] = in;
}

__global__ __launch_bounds__(256) void k_lcount(const Task* tasks, const SegInfo* info, const uint32_t* idx,
                                                const float4* cen, uint32_t* task_left) {
    __shared__ uint32_t sh[4];
    const Task t = tasks[blockIdx.x];
    const SegInfo in = info[t.seg];
    uint32_t c = 0;
    if (in.axis >= 0)
        for (int i = t.cs + threadIdx.x; i < t.ce; i += 256)
            c += wb_bin_of(comp(cen[idx[i]], in.axis), in.lo, in.k) < in.bin ? 1u : 0u;
    for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) task_left[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// stable partition: lefts keep their order at the front, rights after them
__global__ __launch_bounds__(256) void k_lscatter(const Task* tasks, const Seg* segs, const SegInfo* info,
                                                  const uint32_t* task_off, const uint32_t* idx, const float4* cen,
                                                  uint32_t* idx2) {
    __shared__ uint32_t sh[256];
    const Task t = tasks[blockIdx.x];
    const SegInfo in = info[t.seg];
    if (in.axis < 0) return;
    const Seg sg = segs[t.seg];
    const int before = (int)(task_off[blockIdx.x] - task_off[t.first]);  // lefts in earlier tasks
    int run = 0;
    for (int b0 = t.cs; b0 < t.ce; b0 += 256) {
        const int i = b0 + threadIdx.x;
        const bool valid = i < t.ce;
        const uint32_t p = valid ? idx[i] : 0u;
        const bool left = valid && wb_bin_of(comp(cen[p], in.axis), in.lo, in.k) < in.bin;
        uint32_t tot;
        const int ex = (int)block_exclusive(left ? 1u : 0u, sh, tot);
        if (valid) {
            const int lefts_before = before + run + ex;
            const int pos = left ? sg.s + lefts_before : sg.s + in.nleft + (i - sg.s) - lefts_before;
            idx2[pos] = p;
        }
        run += (int)tot;
    }
}

__global__ __launch_bounds__(256) void k_lcopy(const Task* tasks, const SegInfo* info, const uint32_t* idx2, uint32_t* idx) {
    const Task t = tasks[blockIdx.x];
    if (info[t.seg].axis < 0) return;
    for (int i = t.cs + threadIdx.x; i < t.ce; i += 256) idx[i] = idx2[i];
}

__global__ void k_lemit(const Seg* segs, int nseg, const SegInfo* info, BNode* nodes, int* counters, Seg* next_large,
                        Seg* small) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nseg) return;
    const Seg sg = segs[j];
    const int mid = info[j].mid;
    const int l = atomicAdd(&counters[0], 2);
    BNode& nd = nodes[sg.node];
    nd.left = l;
    nd.right = l + 1;
    nd.start = 0;
    nd.count = 0;
    const Seg ch[2] = {{l, sg.s, mid}, {l + 1, mid, sg.e}};
    for (int c = 0; c < 2; c++) {
        if (ch[c].e - ch[c].s > SMALL_MAX) next_large[atomicAdd(&counters[2], 1)] = ch[c];
        else small[atomicAdd(&counters[1], 1)] = ch[c];
    }
}

// ---------------------------------------------------------------- small nodes

// One wavefront builds the whole subtree of a node of at most SMALL_MAX
// primitives: its index range lives in LDS, each node's bounds reduce across
// the wave, bins fill with LDS atomics, lane 0 takes the SAH choice, and a
// ballot-ranked stable partition reorders the range.  The smaller child is
// built next and the larger one stacked, so the stack holds at most
// log2(SMALL_MAX) + 1 entries.
__global__ __launch_bounds__(64) void k_small(const Seg* segs, const WbBox* pb, const float4* cen, uint32_t* idx,
                                              BNode* nodes, int* counters, int max_leaf) {
    __shared__ uint32_t li[SMALL_MAX], lt[SMALL_MAX];
    __shared__ uint32_t bins[BIN_WORDS];
    __shared__ int stk[SMALL_STACK][3];
    __shared__ int bc[4];
    const int lane = threadIdx.x;
    const Seg sg = segs[blockIdx.x];
    const int base = sg.s, n = sg.e - sg.s;
    for (int i = lane; i < n; i += 64) li[i] = idx[base + i];
    __syncthreads();
    int sp = 0, cs = 0, ce = n, cnode = sg.node;
    for (;;) {
        const int cnt = ce - cs;
        float v[12];
        init12(v);
        for (int i = cs + lane; i < ce; i += 64) {
            const uint32_t p = li[i];
            grow12(v, pb[p], cen[p]);
        }
        reduce12(v);
        WbBox box;
        for (int a = 0; a < 3; a++) { box.mn[a] = v[a]; box.mx[a] = v[3 + a]; }
        if (cnt <= max_leaf) {
            if (lane == 0) {
                BNode& nd = nodes[cnode];
                nd.box = box;
                nd.left = nd.right = -1;
                nd.start = base + cs;
                nd.count = cnt;
            }
            if (sp == 0) break;
            sp--;
            cs = stk[sp][0];
            ce = stk[sp][1];
            cnode = stk[sp][2];
            continue;
        }
        float lo[3], k[3];
        bool ok[3];
        axis_scales(v + 6, v + 9, ok, lo, k);
        for (int w = lane; w < BIN_WORDS; w += 64) bins[w] = init_word(12 + w);
        __syncthreads();
        for (int i = cs + lane; i < ce; i += 64) {
            const uint32_t p = li[i];
            bin_add(bins, pb[p], cen[p], ok, lo, k);
        }
        __syncthreads();
        if (lane == 0) {
            int axis, bin, nleft;
            pick_split(bins, v + 6, v + 9, axis, bin, nleft);
            bc[0] = axis;
            bc[1] = bin;
            bc[2] = nleft;
            bc[3] = atomicAdd(&counters[0], 2);
        }
        __syncthreads();
        const int axis = bc[0], bin = bc[1], nleft = bc[2], l = bc[3];
        int mid;
        if (axis < 0) {
            mid = cs + cnt / 2;  // all centroids coincide: split by index
        } else {
            const float la = lo[axis], ka = k[axis];
            const uint64_t below = (1ull << lane) - 1ull;
            int run_l = 0, run_r = 0;
            for (int b0 = cs; b0 < ce; b0 += 64) {
                const int i = b0 + lane;
                const bool valid = i < ce;
                const uint32_t p = valid ? li[i] : 0u;
                const bool left = valid && wb_bin_of(comp(cen[p], axis), la, ka) < bin;
                const uint64_t ml = __ballot(left), mv = __ballot(valid);
                const uint64_t mr = mv & ~ml;
                if (valid) {
                    if (left) lt[cs + run_l + __popcll(ml & below)] = p;
                    else lt[cs + nleft + run_r + __popcll(mr & below)] = p;
                }
                run_l += __popcll(ml);
                run_r += __popcll(mr);
            }
            __syncthreads();
            for (int i = cs + lane; i < ce; i += 64) li[i] = lt[i];
            __syncthreads();
            mid = cs + nleft;
        }
        if (lane == 0) {
            BNode& nd = nodes[cnode];
            nd.box = box;
            nd.left = l;
            nd.right = l + 1;
            nd.start = 0;
            nd.count = 0;
        }
        const bool left_first = (mid - cs) <= (ce - mid);
        const int ps = left_first ? mid : cs, pe = left_first ? ce : mid, pn = left_first ? l + 1 : l;
        if (sp >= SMALL_STACK) {  // cannot happen (the stacked child is the larger half)
            if (lane == 0) atomicOr(&counters[3], 1);
            break;
        }
        if (lane == 0) {
            stk[sp][0] = ps;
            stk[sp][1] = pe;
            stk[sp][2] = pn;
        }
        sp++;
        if (left_first) { ce = mid; cnode = l; }
        else { cs = mid; cnode = l + 1; }
        __syncthreads();
    }
    __syncthreads();
    for (int i = lane; i < n; i += 64) idx[base + i] = li[i];
}

// ---------------------------------------------------------------- collapse

__device__ __forceinline__ void node_view(const BNode* nodes, int i, WbBox& box, int& l, int& r, int& cnt) {
    const BNode& b = nodes[i];
    box = b.box;
    l = b.left;
    r = b.right;
    cnt = b.count;
}

__global__ void k_wplan(const int* wl, int nw, const BNode* nodes, int* wch, uint32_t* ni, uint32_t* np) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    int ch[WBVH_WIDTH];
    const int nch = wb_collapse(wl[w], [&](int i, WbBox& box, int& l, int& r, int& cnt) { node_view(nodes, i, box, l, r, cnt); }, ch);
    // internal children take the first slots, leaves the next ones
    uint32_t a = 0, b = 0;
    for (int k = 0; k < nch; k++)
        if (nodes[ch[k]].count == 0) wch[w * WBVH_WIDTH + a++] = ch[k];
    int at = (int)a;
    for (int k = 0; k < nch; k++) {
        const int c = nodes[ch[k]].count;
        if (c > 0) {
            b += (uint32_t)c;
            wch[w * WBVH_WIDTH + at++] = ch[k];
        }
    }
    for (int k = nch; k < WBVH_WIDTH; k++) wch[w * WBVH_WIDTH + k] = -1;
    ni[w] = a;
    np[w] = b;
}

__global__ void k_wemit(const int* wl, int nw, const BNode* nodes, const int* wch, const uint32_t* ci, const uint32_t* pi,
                        uint32_t level_base, uint32_t prim_run, const uint32_t* idx, WNode* out, uint32_t* prims,
                        int* wl_next, uint8_t* n_int, int* counters) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    WNode wn;
    memset(&wn, 0, sizeof(wn));
    memset(wn.meta, 0xFF, sizeof(wn.meta));
    const uint32_t next_base = level_base + (uint32_t)nw;
    wn.child_base = next_base + ci[w];
    wn.prim_base = prim_run + pi[w];
    WbBox cb[WBVH_WIDTH];
    int nch = 0, k_int = 0, off = 0;
    for (int k = 0; k < WBVH_WIDTH; k++) {
        const int c = wch[w * WBVH_WIDTH + k];
        if (c < 0) break;
        nch++;
        const BNode& b = nodes[c];
        cb[k] = b.box;
        if (b.count > 0) {
            wn.meta[k] = (uint8_t)((b.count << 5) | off);
            for (int j = 0; j < b.count; j++) prims[wn.prim_base + off + j] = idx[b.start + j];
            off += b.count;
        } else {
            wn.meta[k] = (uint8_t)k_int;
            wl_next[ci[w] + k_int] = c;
            k_int++;
        }
    }
    wn.slots = (uint8_t)(nch | (k_int << 4));
    if (!wb_quantize(wn, cb, nch)) atomicOr(&counters[3], 2);
    out[level_base + w] = wn;
    n_int[level_base + w] = (uint8_t)k_int;
}

// worst-case traversal stack entries: need(v) = [internal(v) >= 2] + max need(child)
__global__ void k_need(const WNode* nodes, const uint8_t* n_int, int* need, int begin, int count) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= count) return;
    const int v = begin + w;
    const WNode& n = nodes[v];
    int best = 0;
    for (int k = 0; k < WBVH_WIDTH; k++) {
        const uint8_t mt = n.meta[k];
        if (mt == WBVH_EMPTY || (mt >> 5) != 0) continue;
        best = max(best, need[n.child_base + (mt & 31)]);
    }
    need[v] = (n_int[v] >= 2 ? 1 : 0) + best;
}

// ---------------------------------------------------------------- host side

struct DevPool {
    std::vector<void*> ptrs;
    ~DevPool() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    T* get(size_t n) {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T*)p;
    }
};

inline int blocks(long n, int t) { return (int)std::max<long>(1, (n + t - 1) / t); }

struct Scanner {
    uint32_t* sums = nullptr;
    uint32_t* total = nullptr;
    // exclusive prefix sum of n words (in and out may be the same array);
    // the total is left in *total (device)
    void run(const uint32_t* in, uint32_t* out, int n, hipStream_t s) {
        if (n <= 0) {
            (void)hipMemsetAsync(total, 0, sizeof(uint32_t), s);
            return;
        }
        const int nb = blocks(n, SCAN_BLK);
        k_scan_local<<<nb, SCAN_T, 0, s>>>(in, out, sums, n);
        k_scan_sums<<<1, SCAN_T, 0, s>>>(sums, nb, total);
        k_scan_add<<<blocks(n, 256), 256, 0, s>>>(out, sums, n);
    }
};

#define HIP_TRY(x)                                                        \
    do {                                                                  \
        hipError_t e_ = (x);                                              \
        if (e_ != hipSuccess) {                                           \
            err = std::string(#x) + ": " + hipGetErrorString(e_);         \
            return false;                                                 \
        }                                                                 \
    } while (0)

}  // namespace

void wbvh_triangle_bounds(const void* triangles, int n, float* box, hipStream_t stream) {
    if (n > 0) k_tri_bounds<<<blocks(n, 256), 256, 0, stream>>>((const Triangle*)triangles, n, box);
}

bool wbvh_build_device(const float* box, int n, const unsigned char* skip, int max_leaf, hipStream_t s, WBvhDevice& out,
                       std::string& err) {
    out = WBvhDevice();
    if (n <= 0) return true;
    if (max_leaf < 1) max_leaf = 1;
    if (max_leaf > 32 / WBVH_WIDTH) max_leaf = 32 / WBVH_WIDTH;  // leaf offsets fit 5 bits (wbvh.h)
    DevPool pool;
    const size_t cap_large = (size_t)n / SMALL_MAX + 2;
    const size_t cap_tasks = (size_t)n / CHUNK + cap_large + 2;
    WbBox* pb = pool.get<WbBox>(n);
    float4* cen = pool.get<float4>(n);
    uint32_t* flag = pool.get<uint32_t>(n);
    uint32_t* pos = pool.get<uint32_t>(n);
    uint32_t* idx = pool.get<uint32_t>(n);
    uint32_t* idx2 = pool.get<uint32_t>(n);
    BNode* nodes = pool.get<BNode>(2 * (size_t)n);
    int* counters = pool.get<int>(4);  // binary nodes, small segments, next large segments, error bits
    Seg* small = pool.get<Seg>(n);
    Seg* large = pool.get<Seg>(cap_large);
    Seg* next_large = pool.get<Seg>(cap_large);
    uint32_t* acc = pool.get<uint32_t>(cap_large * ACC_WORDS);
    SegInfo* info = pool.get<SegInfo>(cap_large);
    Task* tasks = pool.get<Task>(cap_tasks);
    uint32_t* task_left = pool.get<uint32_t>(cap_tasks);
    Scanner sc;
    sc.sums = pool.get<uint32_t>((size_t)blocks(2L * n + (long)cap_tasks, SCAN_BLK) + 1);
    sc.total = pool.get<uint32_t>(1);
    if (!pb || !cen || !flag || !pos || !idx || !idx2 || !nodes || !counters || !small || !large || !next_large || !acc ||
        !info || !tasks || !task_left || !sc.sums || !sc.total) {
        err = "BVH build: device allocation failed";
        return false;
    }

    // active primitives in input order
    k_prep<<<blocks(n, 256), 256, 0, s>>>(box, n, skip, pb, cen, flag);
    sc.run(flag, pos, n, s);
    k_compact<<<blocks(n, 256), 256, 0, s>>>(flag, pos, n, idx);
    uint32_t m32 = 0;
    HIP_TRY(hipMemcpyAsync(&m32, sc.total, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const int m = (int)m32;
    if (m == 0) return true;

    // top levels
    int ctr[4] = {1, 0, 0, 0};
    std::vector<Seg> lv;
    const Seg root = {0, 0, m};
    if (m > SMALL_MAX) lv.push_back(root);
    else ctr[1] = 1;
    HIP_TRY(hipMemcpyAsync(small, &root, sizeof(Seg), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(counters, ctr, sizeof(ctr), hipMemcpyHostToDevice, s));
    std::vector<Task> tv;
    while (!lv.empty()) {
        const int nseg = (int)lv.size();
        tv.clear();
        for (int j = 0; j < nseg; j++) {
            const int first = (int)tv.size();
            for (int c = lv[j].s; c < lv[j].e; c += CHUNK) tv.push_back({j, c, std::min(c + CHUNK, lv[j].e), first});
        }
        const int nt = (int)tv.size();
        HIP_TRY(hipMemcpyAsync(large, lv.data(), nseg * sizeof(Seg), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(tasks, tv.data(), nt * sizeof(Task), hipMemcpyHostToDevice, s));
        k_acc_init<<<blocks((long)nseg * ACC_WORDS, 256), 256, 0, s>>>(acc, nseg);
        k_lbox<<<nt, 256, 0, s>>>(tasks, large, idx, pb, cen, acc);
        k_lbin<<<nt, 256, 0, s>>>(tasks, idx, pb, cen, acc);
        k_lsah<<<blocks(nseg, 64), 64, 0, s>>>(large, nseg, acc, info, nodes);
        k_lcount<<<nt, 256, 0, s>>>(tasks, info, idx, cen, task_left);
        sc.run(task_left, task_left, nt, s);
        k_lscatter<<<nt, 256, 0, s>>>(tasks, large, info, task_left, idx, cen, idx2);
        k_lcopy<<<nt, 256, 0, s>>>(tasks, info, idx2, idx);
        k_lemit<<<blocks(nseg, 64), 64, 0, s>>>(large, nseg, info, nodes, counters, next_large, small);
        HIP_TRY(hipMemcpyAsync(ctr, counters, sizeof(ctr), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        lv.resize(ctr[2]);
        if (ctr[2] > 0) {
            HIP_TRY(hipMemcpyAsync(lv.data(), next_large, ctr[2] * sizeof(Seg), hipMemcpyDeviceToHost, s));
            ctr[2] = 0;
            HIP_TRY(hipMemcpyAsync(&counters[2], &ctr[2], sizeof(int), hipMemcpyHostToDevice, s));
            HIP_TRY(hipStreamSynchronize(s));
        }
    }
    // subtrees of small nodes
    if (ctr[1] > 0) k_small<<<ctr[1], 64, 0, s>>>(small, pb, cen, idx, nodes, counters, max_leaf);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(ctr, counters, sizeof(ctr), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (ctr[3]) { err = "BVH build: small-node stack overflow"; return false; }
    const int n_bin = ctr[0];

    // collapse, one wide level at a time
    WNode* wn = pool.get<WNode>(n_bin);
    uint8_t* n_int = pool.get<uint8_t>(n_bin);
    int* wl[2] = {pool.get<int>(n_bin), pool.get<int>(n_bin)};
    int* wch = pool.get<int>((size_t)n_bin * WBVH_WIDTH);
    uint32_t* ni = pool.get<uint32_t>(n_bin);
    uint32_t* np = pool.get<uint32_t>(n_bin);
    uint32_t* prims = nullptr;
    if (hipMalloc((void**)&prims, (size_t)m * sizeof(uint32_t)) != hipSuccess) prims = nullptr;
    if (!wn || !n_int || !wl[0] || !wl[1] || !wch || !ni || !np || !prims) {
        if (prims) (void)hipFree(prims);
        err = "BVH build: device allocation failed";
        return false;
    }
    const int zero = 0;
    HIP_TRY(hipMemcpyAsync(wl[0], &zero, sizeof(int), hipMemcpyHostToDevice, s));
    std::vector<std::pair<int, int>> levels;  // (first wide node, count)
    uint32_t level_base = 0, prim_run = 0;
    int nw = 1, cur = 0;
    bool ok = true;
    while (nw > 0 && ok) {
        k_wplan<<<blocks(nw, 64), 64, 0, s>>>(wl[cur], nw, nodes, wch, ni, np);
        uint32_t tot[2];
        sc.run(ni, ni, nw, s);
        ok = hipMemcpyAsync(&tot[0], sc.total, 4, hipMemcpyDeviceToHost, s) == hipSuccess;
        sc.run(np, np, nw, s);
        ok = ok && hipMemcpyAsync(&tot[1], sc.total, 4, hipMemcpyDeviceToHost, s) == hipSuccess;
        k_wemit<<<blocks(nw, 64), 64, 0, s>>>(wl[cur], nw, nodes, wch, ni, np, level_base, prim_run, idx, wn, prims,
                                              wl[cur ^ 1], n_int, counters);
        ok = ok && hipStreamSynchronize(s) == hipSuccess;
        levels.push_back({(int)level_base, nw});
        level_base += (uint32_t)nw;
        prim_run += tot[1];
        nw = (int)tot[0];
        cur ^= 1;
    }
    const int n_wide = (int)level_base;
    int* need = (int*)wch;  // n_bin * 8 ints >= n_wide
    for (int li = (int)levels.size() - 1; li >= 0 && ok; li--)
        k_need<<<blocks(levels[li].second, 64), 64, 0, s>>>(wn, n_int, need, levels[li].first, levels[li].second);
    int need0 = 0;
    ok = ok && hipMemcpyAsync(&need0, need, sizeof(int), hipMemcpyDeviceToHost, s) == hipSuccess;
    ok = ok && hipMemcpyAsync(ctr, counters, sizeof(ctr), hipMemcpyDeviceToHost, s) == hipSuccess;
    ok = ok && hipStreamSynchronize(s) == hipSuccess;
    WNode* final_nodes = nullptr;
    if (ok && hipMalloc((void**)&final_nodes, (size_t)n_wide * sizeof(WNode)) != hipSuccess) final_nodes = nullptr;
    ok = ok && final_nodes &&
         hipMemcpyAsync(final_nodes, wn, (size_t)n_wide * sizeof(WNode), hipMemcpyDeviceToDevice, s) == hipSuccess &&
         hipStreamSynchronize(s) == hipSuccess;
    if (!ok || ctr[3] || prim_run != (uint32_t)m) {
        if (final_nodes) (void)hipFree(final_nodes);
        (void)hipFree(prims);
        err = !ok ? "BVH build: device error" : (ctr[3] & 2) ? "wbvh: cannot quantise child box" : "BVH build: inconsistent tree";
        return false;
    }
    out.nodes = final_nodes;
    out.prims = prims;
    out.n_nodes = n_wide;
    out.n_prims = m;
    out.max_stack = need0;
    out.depth = (int)levels.size();
    return true;
}
