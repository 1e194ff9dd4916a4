// dev_trace.h — wide-BVH traversal and scene queries of the HIP path.
//
// Semantics follow the reference's scene queries (integrator/raycast.h:14-97
// over the per-category traversals bvh.h:605-922):
//   * a primitive is a candidate when its own bounds pass the reference's
//     slab test (the gate every leaf passes in the reference traversal);
//   * closest hit within a category; equal distances go to the lowest
//     primitive index (the reference's winner depends on its DFS order);
//   * across categories the earlier category wins ties (strict <);
//   * the opaque query returns the closest hit of the first category that
//     has any hit (planes last).
// The traversal visits a superset of the nodes the reference visits: node
// boxes decode conservatively, the slab arithmetic is the reference's
// (b - o) * (1/d), and closest-distance culling keeps a relative margin of
// 2^-7 so a primitive whose computed distance undershoots its box entry by
// rounding is still tested.
#pragma once
#include "dev_prims.h"

// Stack entries kept in LDS per lane, and the waves-per-SIMD target of the
// traversal kernels (0: compiler's choice); build-time tuning knobs.
#ifndef LDS_DEPTH
#define LDS_DEPTH 6  // the 1M-triangle mesh's worst-case stack; 8: 1.5 % slower frame
#endif
#ifndef TRACE_MIN_WAVES
#define TRACE_MIN_WAVES 6
#endif
#define SPILL_DEPTH (128 - LDS_DEPTH)
#ifndef TRACE_BLOCK
#define TRACE_BLOCK 256
#endif
#if TRACE_MIN_WAVES > 0
#define TRACE_BOUNDS __launch_bounds__(TRACE_BLOCK, TRACE_MIN_WAVES)
#else
#define TRACE_BOUNDS __launch_bounds__(TRACE_BLOCK)
#endif
// the instantiations for every primitive kind (FULL) may take another target
#ifndef TRACE_MIN_WAVES_FULL
#define TRACE_MIN_WAVES_FULL TRACE_MIN_WAVES
#endif
#if TRACE_MIN_WAVES > 0 && TRACE_MIN_WAVES_FULL > 0
#define TRACE_BOUNDS_OF(FULL) __launch_bounds__(TRACE_BLOCK, (FULL) ? TRACE_MIN_WAVES_FULL : TRACE_MIN_WAVES)
#else
#define TRACE_BOUNDS_OF(FULL) TRACE_BOUNDS
#endif
// The shading kernels run 1024-thread blocks: their queue reservations take
// one device atomic per block (block_reserve), and each light's shadow
// requests of a block land in one contiguous run of the queue.  On the
// 1M-triangle frame, against 256-thread blocks: shading 1.83 -> 1.37 ms and
// the shadow kernel, tracing the longer coherent runs, 5.35 -> 4.95 ms (512:
// 1.41 and 5.11 ms).  The instantiations for every primitive kind (FULL)
// hold more registers and run 512-thread blocks (analytic 10k workload
// 95.6 ms per frame against 98.8 ms with 1024).  Their traversal stack is laid
// out in groups of TRACE_BLOCK lanes (SHEEN_STACK_N), so the stack code keeps
// its stride.
#ifndef SHADE_BLOCK
#define SHADE_BLOCK 1024
#endif
#ifndef SHADE_BLOCK_FULL
#define SHADE_BLOCK_FULL 512
#endif
#define SHADE_BLOCK_OF(FULL) ((FULL) ? SHADE_BLOCK_FULL : SHADE_BLOCK)
static_assert(SHADE_BLOCK <= 1024 && SHADE_BLOCK_FULL <= 1024, "block_reserve holds up to 16 waves");
#ifndef SHADE_MIN_WAVES
#define SHADE_MIN_WAVES 0
#endif
#if SHADE_MIN_WAVES > 0
#define SHADE_BOUNDS(FULL) __launch_bounds__(SHADE_BLOCK_OF(FULL), SHADE_MIN_WAVES)
#else
#define SHADE_BOUNDS(FULL) __launch_bounds__(SHADE_BLOCK_OF(FULL))
#endif
#define ADJOINT_BOUNDS __launch_bounds__(TRACE_BLOCK)
#ifndef BOX_FAST
#define BOX_FAST 1
#endif
#define CULL_MARGIN 0.0078125f

// Per-lane traversal stack: the top LDS_DEPTH entries live in LDS (one
// column per lane, conflict-free), older entries spill to a private array
// owned by the kernel.  An entry is a group of sibling nodes still to visit:
// x = first internal child of their parent, y = bit mask of their ranks
// (low 8 bits) | a lower bound of their entry distances (the float bits of
// the nearest one with the low 8 bits cleared, which rounds toward zero).  The struct holds only pointers and counters so it
// stays in registers; the stack pointer itself is a local of traverse().
struct TStack {
    uint2* lds;    // &shared[tid], stride TRACE_BLOCK
    uint2* spill;  // private array of SPILL_DEPTH entries
    uint32_t spills;
    uint32_t* list;            // this wave's LEAF_LIST entries of pending triangle tests
    unsigned long long* keys;  // this wave's 64 best-hit keys, one per lane
};

// Triangle tests of one node visit shared across the wave (see trav_step):
// entries of the pending-test list per wave, and the slot bits of an entry
// (owner lane above them).
#ifndef LEAF_BATCH
#define LEAF_BATCH 1
#endif
#define LEAF_LIST 256
#define LEAF_SLOT_BITS 26
SHD void stack_push(TStack& st, int& sp, uint2 e) {
    const uint32_t slot = (uint32_t)sp % LDS_DEPTH;
    if (sp >= LDS_DEPTH) {
        st.spill[sp - LDS_DEPTH] = st.lds[slot * TRACE_BLOCK];
        st.spills++;
    }
    st.lds[slot * TRACE_BLOCK] = e;
    sp++;
}
SHD uint2 stack_pop(TStack& st, int& sp) {
    sp--;
    const uint32_t slot = (uint32_t)sp % LDS_DEPTH;
    uint2 e = st.lds[slot * TRACE_BLOCK];
    if (sp >= LDS_DEPTH) st.lds[slot * TRACE_BLOCK] = st.spill[sp - LDS_DEPTH];
    return e;
}
// The top entry (always in LDS) read and rewritten in place.
SHD uint2 stack_top(const TStack& st, int sp) { return st.lds[(uint32_t)(sp - 1) % LDS_DEPTH * TRACE_BLOCK]; }
SHD void stack_set_top(TStack& st, int sp, uint2 e) { st.lds[(uint32_t)(sp - 1) % LDS_DEPTH * TRACE_BLOCK] = e; }

struct TraceStats {
    uint32_t nodes, prims;
    uint32_t witers;  // node-visit iterations of the wave, counted by its lowest active lane
};

// Leaf test of one primitive reference: reference gate, then the distance.
// `smax`: the traversal's running best (hits beyond it cannot be taken).
template <int CAT>
SHD bool leaf_dist(const DScene& S, uint32_t slot, const DRay& r, v3 inv, float& t, int& idx, float smax) {
    if (CAT == C_TRI) {
        float4 a = S.tri_leaf[3 * slot], b = S.tri_leaf[3 * slot + 1], c = S.tri_leaf[3 * slot + 2];
        idx = __float_as_int(a.w);
        // both tests must pass; the gate runs only for a hit (most tests miss)
        if (!t_triangle(r, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(c.x, c.y, c.z), t)) return false;
        float mn[3] = {fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z))};
        float mx[3] = {fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z))};
        return ref_slab(r, inv, mn, mx);
    }
    idx = (int)S.prim_ids[CAT][slot];
    const float* bb = S.bounds[CAT] + 6 * (size_t)idx;
    if (!ref_slab(r, inv, bb, bb + 3)) return false;
    DHit h;
    switch (CAT) {
        case C_SPHERE: {
            const Object& o = S.objects[idx];
            return t_sphere(r, cv3(o.param1), o.param2, t);
        }
        case C_STRAND:
            if (!hit_strand(r, S.strands[idx], S.sweep_points, h, smax)) return false;
            break;
        case C_HF:
            if (!hit_hf(r, S.hfs[idx], h, smax, false)) return false;
            break;
        case C_SE:
            if (!hit_se(r, S.ses[idx], h, smax, false)) return false;
            break;
        case C_LATHE:
            if (!hit_lathe(r, S.lathes[idx], S.lathe_points, h, smax)) return false;
            break;
        case C_PRISM:
            if (!hit_prism(r, S.prisms[idx], S.prism_points, h)) return false;
            break;
        default:
            return false;
    }
    t = h.t;
    return true;
}

// State of one in-flight traversal of one category's wide BVH: the node to
// visit next, the stack depth, the running best (ct: acceptance bound,
// exclusive unless a same-category tie; ci/ch: best index / found), and the
// ray's box-test margins (E, fast: see trav_begin).
struct TravState {
    uint32_t node;
    int sp;
    float ct;
    int ci;
    bool ch;
    bool fast;
    v3 E;
};

// Margins of the distance-space box test (trav_step) for one ray.  A plane q
// of axis a decodes at t = q*B_a + A_a with B_a = s_a*inv_a (exact: s_a is a
// power of two) and A_a = (o_a - ray_a)*inv_a; rounding makes this differ
// from the exact-path value ((q*s+o) - ray)*inv by at most
// u |inv_a| (7|o_a| + 6|ray_a| + 4 Q s_a) + 2u E_a (u = 2^-24, Q = 2048).
// The scene's box_k[cat][a] bounds 2|o_a| + 2 Q s_a over the nodes of the
// category's BVH, so E_a = 2^-20 |inv_a| (box_k[cat][a] + |ray_a|) (16u
// times a bound of twice the terms) covers every node; the 2^-100 floor covers a result or
// operand flushed from the denormal range.  Near planes move by -E_a and far
// planes by +E_a, so the test passes every box the exact path passes.  The
// margin is per axis: an axis the ray nearly parallels has huge |inv| and
// planes far off in t, and must not widen the other axes.  A ray whose
// margins or offsets could overflow (a zero direction component, huge
// coordinates) takes the exact path (fast = false; the whole wave does).
SHD void trav_begin(const DScene& S, int cat, const DRay& r, v3 inv, TravState& T) {
    const float* K = S.box_k[cat];
    const float xa = fabsf(inv.x) * (K[0] + fabsf(r.o.x));
    const float ya = fabsf(inv.y) * (K[1] + fabsf(r.o.y));
    const float za = fabsf(inv.z) * (K[2] + fabsf(r.o.z));
    T.fast = xa < 0x1p120f && ya < 0x1p120f && za < 0x1p120f;  // false for NaN too
    T.E = mk3(fmaf(xa, 0x1p-20f, 0x1p-100f), fmaf(ya, 0x1p-20f, 0x1p-100f), fmaf(za, 0x1p-20f, 0x1p-100f));
}

// Second smallest of {a, b, c} when a <= b: the running second-nearest key.
SHD uint32_t umed3(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// CAT_MESH: the category (sphere or triangle) is chosen at run time.
#define CAT_MESH (-1)

template <int CAT>
SHD bool leaf_test(const DScene& S, int cat, uint32_t slot, const DRay& r, v3 inv, float& t, int& idx, float smax) {
    if constexpr (CAT == CAT_MESH) {
        if (cat == C_TRI) return leaf_dist<C_TRI>(S, slot, r, inv, t, idx, smax);
        return leaf_dist<C_SPHERE>(S, slot, r, inv, t, idx, smax);
    } else {
        return leaf_dist<CAT>(S, slot, r, inv, t, idx, smax);
    }
}

// One node visit: box tests of the eight children, the primitives of every
// leaf child hit, then the next node (nearest internal child, else the
// nearest stacked node still within the bound).  Returns true when the
// traversal is over (stack exhausted, or with `any` the first accepted hit).
// A traversal either keeps the closest hit (NoLeafVisit) or hands every leaf
// primitive it reaches to a visitor (LV::active), leaving the bound at the
// ray's t_max so that every leaf whose box the segment meets is reached.
struct NoLeafVisit {
    static constexpr bool active = false;
    SHD void operator()(uint32_t) {}
};
template <int CAT, bool COUNT, class LV = NoLeafVisit>
SHD bool trav_step(const DScene& S, int cat, bool any, const DRay& r, v3 inv, TravState& T, TStack& st,
                   TraceStats& ts, bool share = false, LV* lv = nullptr) {
    const WNode* nodes;
    if constexpr (CAT == CAT_MESH) nodes = (cat == C_TRI) ? S.bvh[C_TRI] : S.bvh[C_SPHERE];
    else nodes = S.bvh[CAT];
    const uint4* np = reinterpret_cast<const uint4*>(nodes + T.node);
    // Each axis's two plane rows in near/far order for this ray, chosen by the
    // sign bit of inv in the load address (no per-node selects); the exact
    // path below is symmetric in the two rows.
    // The signs are packed into one word that is opaque to the compiler, so
    // the six row offsets are derived per visit rather than hoisted out of
    // the traversal loop and kept (spilled) across it.
    uint32_t z = (__float_as_uint(inv.x) >> 31) | ((__float_as_uint(inv.y) >> 31) << 1) |
                 ((__float_as_uint(inv.z) >> 31) << 2);
    asm volatile("" : "+v"(z));
    const uint32_t zx = z & 1u, zy = (z >> 1) & 1u, zz = z >> 2;
    const uint4 h0 = np[0], h1 = np[1];
    const uint4 nx = np[2 + zx], fx = np[3 - zx], ny = np[4 + zy], fy = np[5 - zy], nz = np[6 + zz], fz = np[7 - zz];
    if (COUNT) ts.nodes++;
    const float ox = __uint_as_float(h0.x), oy = __uint_as_float(h0.y), oz = __uint_as_float(h0.z);
    const float sx = __uint_as_float((h0.w & 0xFFu) << 23);
    const float sy = __uint_as_float(((h0.w >> 8) & 0xFFu) << 23);
    const float sz = __uint_as_float(((h0.w >> 16) & 0xFFu) << 23);
    // slots 0..ni-1: internal children (node child_base + slot); ni..nc-1: leaves
    const uint32_t nc = (h0.w >> 24) & 15u, ni = h0.w >> 28;
    const uint32_t child_base = h1.x, prim_base = h1.y;
    const uint64_t meta = (uint64_t)h1.z | ((uint64_t)h1.w << 32);
    const float bound = T.ct + T.ct * CULL_MARGIN;
    // hit children as slot bits; the two nearest internal children hit as
    // keys (entry distance bits with the low 3 bits replaced by the slot;
    // distances are >= t_min > 0 for every query that culls by them)
    uint32_t hitbits = 0, k0 = 0xFFFFFFFFu, k1 = 0xFFFFFFFFu;
    // child plane i of a row: the fp16 pattern of an integer 0..WBVH_QMAX
    auto word = [](uint4 row, int i) -> uint32_t {
        const int w = i >> 1;
        return (w == 0) ? row.x : (w == 1) ? row.y : (w == 2) ? row.z : row.w;
    };
    auto g = [&](uint4 row, int i) -> float {
        const uint32_t d = word(row, i) >> (16 * (i & 1));
        return (float)__builtin_bit_cast(_Float16, (uint16_t)(d & 0xFFFFu));
    };
    // fmaf(plane, b, a) of an fp16 plane: one mixed-precision FMA
    // (v_fma_mix_f32, the half taken straight from the word), product and sum
    // fused and rounded once as fmaf of the exact float
    auto gfma = [&](uint4 row, int i, float b, float a) -> float { return fmaf(g(row, i), b, a); };
    auto take = [&](int i, float tn, float tf) {
        const bool hit = tn <= tf;
        hitbits |= hit ? (1u << i) : 0u;
        const uint32_t key = (hit && (uint32_t)i < ni) ? ((__float_as_uint(tn) & ~7u) | (uint32_t)i) : 0xFFFFFFFFu;
        k1 = umed3(k0, k1, key);
        k0 = min(k0, key);
    };
    if (BOX_FAST && __ballot(!T.fast) == 0ull) {
        // distance-space decode with the ray's margins (trav_begin)
        const float Ax = (ox - r.o.x) * inv.x, Ay = (oy - r.o.y) * inv.y, Az = (oz - r.o.z) * inv.z;
        const float Bx = sx * inv.x, By = sy * inv.y, Bz = sz * inv.z;
        const float Anx = Ax - T.E.x, Afx = Ax + T.E.x, Any = Ay - T.E.y, Afy = Ay + T.E.y;
        const float Anz = Az - T.E.z, Afz = Az + T.E.z;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float tn = fmaxf(fmaxf(fmaxf(gfma(nx, i, Bx, Anx), gfma(ny, i, By, Any)), gfma(nz, i, Bz, Anz)), r.tmin);
            const float tf = fminf(fminf(fminf(gfma(fx, i, Bx, Afx), gfma(fy, i, By, Afy)), gfma(fz, i, Bz, Afz)), bound);
            take(i, tn, tf);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float lx = fmaf(g(nx, i), sx, ox), hx = fmaf(g(fx, i), sx, ox);
            const float ly = fmaf(g(ny, i), sy, oy), hy = fmaf(g(fy, i), sy, oy);
            const float lz = fmaf(g(nz, i), sz, oz), hz = fmaf(g(fz, i), sz, oz);
            const float ax = (lx - r.o.x) * inv.x, bx = (hx - r.o.x) * inv.x;
            const float ay = (ly - r.o.y) * inv.y, by = (hy - r.o.y) * inv.y;
            const float az = (lz - r.o.z) * inv.z, bz = (hz - r.o.z) * inv.z;
            const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), r.tmin));
            const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), bound));
            take(i, tn, tf);
        }
    }
    const uint32_t inner = (1u << ni) - 1u;
    uint32_t leafmask = hitbits & ((1u << nc) - 1u) & ~inner;
    // Triangle leaves: when some lane has two or more triangles to test, the
    // wave's pending tests are listed in LDS and dealt to its active lanes,
    // each testing another lane's ray (fetched by cross-lane permute) and
    // merging into that lane's best-hit key (t bits, index) with an LDS
    // 64-bit atomic min.  The minimum over the tests of one lane equals the
    // sequential loop's result: a hit replaces the running best when
    // (t, index) is lexicographically smaller, and with no best yet only a
    // strictly smaller t qualifies (the initial key carries index 0).  With
    // `any`, a lane whose key improved stops, as the loop stops at its first
    // accepted hit (the caller then only needs that some hit was found).
    if constexpr (CAT == C_TRI) {
        // `share`: every lane traces the same scene (a batch's lanes may hold
        // rays of different scenes, whose leaf arrays differ)
        if (LEAF_BATCH && share) {
            const uint32_t lane = __lane_id();
            const uint64_t lt = (1ull << lane) - 1ull;
            // tests of the leaf slots in `take`
            auto tests = [&](uint32_t take) -> uint32_t {
                uint32_t n = 0;
                for (uint32_t lm = take; lm; lm &= lm - 1) n += ((uint32_t)(meta >> (8 * __builtin_ctz(lm))) & 0xFFu) >> 5;
                return n;
            };
            // offset of this lane's n tests among the wave's, and the wave's total
            auto prefix = [&](uint32_t n, uint32_t& off, uint32_t& N) {
                off = 0;
                N = 0;
#pragma unroll
                for (int b = 0; b < 6; b++) {
                    const uint64_t mb = __ballot((n >> b) & 1u);
                    // lanes below this one with bit b set (v_mbcnt)
                    off += __builtin_amdgcn_mbcnt_hi((uint32_t)(mb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mb, 0u)) << b;
                    N += (uint32_t)__popcll(mb) << b;
                }
            };
            // Rounds: every pending leaf at once when the wave's tests fit the
            // list, else each lane's lowest pending leaf (a bottom node holds
            // up to 8 leaves); single tests are left to the lanes themselves
            // (the common case of a visit with no leaf, or one test per lane,
            // costs one ballot).
            for (;;) {
                const uint64_t active = __ballot(1);
                uint32_t take = leafmask, off, N;
                uint32_t n = tests(take);
                if (__ballot(n >= 2) == 0ull) break;
                prefix(n, off, N);
                if (N > LEAF_LIST) {
                    take = leafmask & (0u - leafmask);
                    n = tests(take);
                    prefix(n, off, N);
                    if (N > LEAF_LIST) break;
                }
                for (uint32_t lm = take, j = off; lm; lm &= lm - 1) {
                    const uint32_t m = (uint32_t)(meta >> (8 * __builtin_ctz(lm))) & 0xFFu;
                    for (uint32_t k = 0; k < (m >> 5); k++) st.list[j++] = (lane << LEAF_SLOT_BITS) | (prim_base + (m & 31u) + k);
                }
                const unsigned long long key0 =
                    ((unsigned long long)__float_as_uint(T.ct) << 32) | (uint32_t)(T.ch ? T.ci : 0);
                st.keys[lane] = key0;
                if (COUNT) ts.prims += n;
                __builtin_amdgcn_wave_barrier();
                const uint32_t A = (uint32_t)__popcll(active), q = (uint32_t)__popcll(active & lt);
                for (uint32_t base = 0; base < N; base += A) {
                    const uint32_t g = base + q;
                    const uint32_t e = g < N ? st.list[g] : (lane << LEAF_SLOT_BITS);
                    const int owner = (int)(e >> LEAF_SLOT_BITS);
                    DRay ro;
                    ro.o = mk3(__shfl(r.o.x, owner), __shfl(r.o.y, owner), __shfl(r.o.z, owner));
                    ro.d = mk3(__shfl(r.d.x, owner), __shfl(r.d.y, owner), __shfl(r.d.z, owner));
                    ro.tmin = __shfl(r.tmin, owner);
                    ro.tmax = __shfl(r.tmax, owner);
                    const v3 io = mk3(__shfl(inv.x, owner), __shfl(inv.y, owner), __shfl(inv.z, owner));
                    if (g < N) {
                        float t;
                        int idx;
                        if (leaf_dist<C_TRI>(S, e & ((1u << LEAF_SLOT_BITS) - 1u), ro, io, t, idx, ro.tmax))
                            atomicMin(&st.keys[owner], ((unsigned long long)__float_as_uint(t) << 32) | (uint32_t)idx);
                    }
                }
                __builtin_amdgcn_wave_barrier();
                const unsigned long long key = st.keys[lane];
                leafmask &= ~take;
                if (key < key0) {
                    T.ct = __uint_as_float((uint32_t)(key >> 32));
                    T.ci = (int)(uint32_t)key;
                    T.ch = true;
                    if (any) return true;
                }
                if (__ballot(leafmask != 0u) == 0ull) break;
            }
        }
    }
    if constexpr (LV::active) {
        for (; leafmask; leafmask &= leafmask - 1) {
            const uint32_t m = (uint32_t)(meta >> (8 * __builtin_ctz(leafmask))) & 0xFFu;
            for (uint32_t k = 0; k < (m >> 5); k++) (*lv)(prim_base + (m & 31u) + k);
        }
    }
    while (leafmask) {
        const int i = __builtin_ctz(leafmask);
        leafmask &= leafmask - 1;
        const uint32_t m = (uint32_t)(meta >> (8 * i)) & 0xFFu;
        const uint32_t cnt = m >> 5, off = m & 31u;
        for (uint32_t k = 0; k < cnt; k++) {
            float t;
            int idx;
            if (COUNT) ts.prims++;
            if (leaf_test<CAT>(S, cat, prim_base + off + k, r, inv, t, idx, T.ct)) {
                if (t < T.ct || (t == T.ct && T.ch && idx < T.ci)) {
                    T.ct = t;
                    T.ci = idx;
                    T.ch = true;
                    if (any) return true;
                }
            }
        }
    }
    // nearest internal child next; the other hit internal children become
    // one stack entry bounded by the nearest of them (k1)
    if (k0 != 0xFFFFFFFFu) {
        const uint32_t rest = hitbits & inner & ~(1u << (k0 & 7u));
        if (rest && T.sp < LDS_DEPTH + SPILL_DEPTH) stack_push(st, T.sp, make_uint2(child_base, (k1 & ~0xFFu) | rest));
        T.node = child_base + (k0 & 7u);
        return false;
    }
    const float lim = T.ct + T.ct * CULL_MARGIN;
    while (T.sp > 0) {
        const uint2 e = stack_top(st, T.sp);
        if (__uint_as_float(e.y & ~0xFFu) > lim) {
            (void)stack_pop(st, T.sp);
            continue;
        }
        const uint32_t m = e.y & 0xFFu;
        const uint32_t rest = m & (m - 1u);
        if (rest) stack_set_top(st, T.sp, make_uint2(e.x, (e.y & ~0xFFu) | rest));
        else (void)stack_pop(st, T.sp);
        T.node = e.x + (uint32_t)__builtin_ctz(m);
        return false;
    }
    return true;
}

// Closest (or, with ANY, first) hit of one category's wide BVH.
// cat_t: in = acceptance bound (exclusive unless a same-category tie), out =
// best distance; cat_idx/cat_hit as the running best of this category.
template <int CAT, bool ANY, bool COUNT>
SHD void traverse(const DScene& S, const DRay& r, v3 inv, float& cat_t, int& cat_idx, bool& cat_hit, TStack& st,
                  TraceStats& ts) {
    TravState T;
    T.node = 0; T.sp = 0; T.ct = cat_t; T.ci = cat_idx; T.ch = cat_hit;
    trav_begin(S, CAT, r, inv, T);
    // decided once per traversal: the scene descriptor may be a per-lane
    // reference into global memory, read again on every node visit otherwise
    const bool share = CAT == C_TRI && S.wave_one_scene && S.count[C_TRI] < (1 << LEAF_SLOT_BITS);
    for (;;) {
        if (COUNT) ts.witers += (__lane_id() == (uint32_t)(__ffsll((unsigned long long)__ballot(1)) - 1)) ? 1u : 0u;
        if (trav_step<CAT, COUNT>(S, CAT, ANY, r, inv, T, st, ts, share)) break;
    }
    cat_t = T.ct; cat_idx = T.ci; cat_hit = T.ch;
}

// CSG tree evaluation (reference integrator/csg.h:113-186) with an explicit
// post-order stack.  Returns the hit distance and the interval record.
// Entry and exit distances of the ray's line through a CSG node's box
// widened by 2^-8 (L1 distance from the ray origin to the box centre + the
// half extents).  Every span the node's subtree can return lies inside
// [tn, tf]: a sphere span needs a discriminant >= 0, which the float
// arithmetic can report for lines up to about 2^-9.5 (|o - c| + r) outside
// the sphere and then places both chord ends within about 2^-10.5 |o - c| of
// the closest approach; a superellipsoid span starts with its own box test;
// intersections and differences return parts of their operands' spans.
// Returns false for a box the line misses; a NaN in the test (a line on a
// widened face) counts as a miss only where the line lies on the margin.
SHD bool csg_line_box(const DRay& r, v3 inv, const float* b, float& tn, float& tf) {
    const float hx = 0.5f * (b[3] - b[0]), hy = 0.5f * (b[4] - b[1]), hz = 0.5f * (b[5] - b[2]);
    const float cx = b[0] + hx, cy = b[1] + hy, cz = b[2] + hz;
    const float pad = 0.00390625f * (fabsf(r.o.x - cx) + fabsf(r.o.y - cy) + fabsf(r.o.z - cz) + hx + hy + hz) + 1e-6f;
    if (!(pad < 3.0e38f)) {  // unbounded or NaN box: no culling
        tn = -__builtin_huge_valf();
        tf = __builtin_huge_valf();
        return true;
    }
    const float ax = (b[0] - pad - r.o.x) * inv.x, bx = (b[3] + pad - r.o.x) * inv.x;
    const float ay = (b[1] - pad - r.o.y) * inv.y, by = (b[4] + pad - r.o.y) * inv.y;
    const float az = (b[2] - pad - r.o.z) * inv.z, bz = (b[5] + pad - r.o.z) * inv.z;
    tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
    tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
    return !(tn > tf);
}

// Spans of the CSG subtree rooted at `root`, evaluated with `frames`
// frames of the post-order stack (the root's frame included; a child that
// would need more is not evaluated and contributes nothing).  Subtrees whose
// widened box the line misses, or meets only outside [t0, t1], are skipped:
// each leaf span is clipped to [t0, t1] and would be empty, and so would
// every combination of them.
SHD void csg_eval(const DScene& S, const DRay& r, v3 inv, float t0, float t1, int root, int frames, CSpan* ret, int& nret) {
    auto visit = [&](int node) {
        if (!S.csg_box) return true;
        float tn, tf;
        if (!csg_line_box(r, inv, S.csg_box + 6 * (size_t)node, tn, tf)) return false;
        return !(tn > t1 || tf < t0);
    };
    nret = 0;
    if (frames < 1 || !visit(root)) return;
    struct Frame { int node; int state; int nA; CSpan A[8]; };
    Frame fr[CSG_MAX_DEPTH];
    int top = 0;
    fr[0].node = root; fr[0].state = 0; fr[0].nA = 0;
    while (top >= 0) {
        Frame& f = fr[top];
        const CSGNode& n = S.csg[f.node];
        if (n.leaf_kind >= 0) {
            nret = 0;
            CSpan iv;
            iv.te = -1e30f; iv.tx = -1e30f;
            bool ok = false;
            if (n.leaf_kind == 0) {
                const Object& o = S.objects[n.leaf_index];
                ok = sphere_span(r, cv3(o.param1), o.param2, t0, t1, iv.te, iv.tx);
            } else if (n.leaf_kind == 1) {
                ok = se_span(r, S.ses[n.leaf_index], t0, t1, iv.te, iv.tx);
            }
            if (ok) {
                iv.mat = (n.leaf_kind == 0) ? S.objects[n.leaf_index].material_id
                       : (n.leaf_kind == 1) ? S.ses[n.leaf_index].material_id : 0;
                iv.leaf = n.leaf_index;
                iv.kind = n.leaf_kind;
                ret[nret++] = iv;
            }
            top--;
            continue;
        }
        if (f.state == 0) {  // evaluate left
            if (n.op == 1 && !(n.right >= 0 && top + 1 < frames && visit(n.right))) {
                // an intersection whose right operand contributes nothing
                // (its box missed, or no frame left for it) is empty, so the
                // left subtree is not evaluated either
                nret = 0;
                top--;
                continue;
            }
            f.state = 1;
            if (n.left >= 0 && top + 1 < frames && visit(n.left)) {
                fr[top + 1].node = n.left; fr[top + 1].state = 0; fr[top + 1].nA = 0;
                top++;
            } else {
                nret = 0;
            }
            continue;
        }
        if (f.state == 1) {  // left result in ret; evaluate right
            if ((n.op == 1 || n.op == 2) && nret == 0) {
                // an intersection or difference with an empty left operand is
                // empty (csg_combine returns no span whatever the right one
                // holds), so the right subtree is not evaluated
                top--;
                continue;
            }
            f.nA = nret;
            for (int i = 0; i < nret; i++) f.A[i] = ret[i];
            f.state = 2;
            if (n.right >= 0 && top + 1 < frames && visit(n.right)) {
                fr[top + 1].node = n.right; fr[top + 1].state = 0; fr[top + 1].nA = 0;
                top++;
            } else {
                nret = 0;
            }
            continue;
        }
        // combine left (f.A) with right (ret)
        CSpan B[8];
        int nB = nret;
        for (int i = 0; i < nB; i++) B[i] = ret[i];
        csg_combine(f.A, f.nA, B, nB, n.op, ret, nret);
        top--;
    }
}

// CSG hit of a ray (reference integrator/csg.h:113-186): the first span, by
// entry, whose exit lies beyond t0; the caller uses its entry (t) and its
// material, leaf and kind.
//
// When the root is a union, the unions reachable from it through unions
// only merge the spans of the subtrees below them: sorted by entry (stable,
// so equal entries keep the in-order position of their subtree, then their
// position in its list), overlapping spans fused into groups headed by their
// first span, at most 8 groups kept (the farthest dropped).  Every span
// starts at or after t0, so the only spans whose exit does not pass t0 are
// points at t0, and the chosen group's head is the first span, by (entry,
// subtree rank, position), whose entry equals the smallest entry E of the
// spans that end beyond t0 — a point span at t0 counts when E = t0.  The
// subtrees are therefore visited near-first through the spine hierarchy and
// one whose box entry lies beyond the best E found so far is skipped.
SHD bool csg_query(const DScene& S, const DRay& r, float t0, float t1, CSpan& best_iv, float& t_out, TStack& st) {
    const v3 inv = mk3(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
    CSpan ret[8];
    int nret = 0;
    if (!S.csg_spine) {
        csg_eval(S, r, inv, t0, t1, S.csg_root, CSG_MAX_DEPTH, ret, nret);
        float best = 1e30f;
        int bi = -1;
        for (int i = 0; i < nret; i++)
            if (ret[i].te < best && ret[i].tx > t0) { best = ret[i].te; bi = i; }
        if (bi < 0) return false;
        float t = fmaxf(ret[bi].te, t0);
        if (t < t0 || t > t1) return false;
        best_iv = ret[bi];
        t_out = t;
        return true;
    }
    // best span ending beyond t0 (E, rank, pos) and best point span at t0
    float E = 1e30f;
    int er = 0x7FFFFFFF, ep = 0, dr = 0x7FFFFFFF, dp = 0;
    CSpan e_iv, d_iv;
    e_iv.te = e_iv.tx = 0.0f; e_iv.mat = e_iv.leaf = e_iv.kind = 0;
    d_iv = e_iv;
    auto take = [&](int4 sub) {
        csg_eval(S, r, inv, t0, t1, sub.x, CSG_MAX_DEPTH - sub.z, ret, nret);
        for (int p = 0; p < nret; p++) {
            const CSpan& sp = ret[p];
            if (!(sp.te < 1e30f)) continue;
            if (sp.tx > t0) {
                if (sp.te < E || (sp.te == E && (sub.y < er || (sub.y == er && p < ep)))) { E = sp.te; er = sub.y; ep = p; e_iv = sp; }
            } else if (sp.te == t0 && (sub.y < dr || (sub.y == dr && p < dp))) {
                dr = sub.y; dp = p; d_iv = sp;
            }
        }
    };
    for (int k = 0; k < S.spine_always; k++) take(S.spine_sub[k]);
    if (S.spine_n_nodes > 0) {
        // the spine's stack of {node, entry distance bits} in the traversal
        // stack (LDS first; the category traversals are over), its spill
        // count left as the traversals had it
        const uint32_t spills0 = st.spills;
        int sp = 0;
        const CsgSpineNode* N = S.spine_nodes;
        auto node_box = [&](int i, float& tn, float& tf) {
            const CsgSpineNode& cn = N[i];
            const float bx[6] = {cn.lo[0], cn.lo[1], cn.lo[2], cn.hi[0], cn.hi[1], cn.hi[2]};
            return csg_line_box(r, inv, bx, tn, tf) && !(tn > t1 || tf < t0);
        };
        float tn0, tf0;
        if (node_box(0, tn0, tf0)) stack_push(st, sp, make_uint2(0u, __float_as_uint(tn0)));
        // the hierarchy is balanced (median splits): at most log2(subtrees) + 1
        // entries are ever stacked, far below 32
        while (sp > 0) {
            const uint2 se = stack_pop(st, sp);
            const int i = (int)se.x;
            if (__uint_as_float(se.y) > E) continue;
            const CsgSpineNode nd = N[i];
            if (nd.a < 0) {
                for (int k = 0; k < nd.b; k++) take(S.spine_sub[-nd.a - 1 + k]);
                continue;
            }
            float tnc[2];
            bool hit[2];
            const int ch[2] = {nd.a, nd.b};
            for (int c = 0; c < 2; c++) {
                float a0, a1;
                hit[c] = node_box(ch[c], a0, a1) && !(a0 > E);
                tnc[c] = a0;
            }
            // nearer child popped first
            const int first = (hit[0] && hit[1] && tnc[1] < tnc[0]) ? 1 : 0;
            for (int q = 1; q >= 0; q--) {
                const int c = q == 0 ? first : 1 - first;
                if (!hit[c] || sp >= 32) continue;
                stack_push(st, sp, make_uint2((uint32_t)ch[c], __float_as_uint(tnc[c])));
            }
        }
        st.spills = spills0;
    }
    if (!(E < 1e30f)) return false;
    CSpan head = e_iv;
    if (E == t0 && (dr < er || (dr == er && dp < ep))) head = d_iv;
    const float t = fmaxf(head.te, t0);
    if (t < t0 || t > t1) return false;
    best_iv = head;
    t_out = t;
    return true;
}

struct QHit {
    float t;
    int kind;   // -1 miss
    int idx;
    int carry;  // object_id carried into a CSG hit
};

// Lathes of a scene with at most LATHE_SCAN_MAX of them are tested without
// their tree, in one fixed order (`lathe_order`, widest profile first), so
// the lathe and frustum in hand are the same in every lane of a wave and are
// read through the scalar cache; the frusta come from the scene's records
// (lathe_frustum).  The result is the traversal's: both take the
// lexicographic minimum of (distance, index) over the lathes whose reference
// bounds the ray passes (leaf_dist's gate), with the running best as the
// frusta's bound (hit_lathe's `smax`), and the tree culls little for
// lathes, whose reference bounds span +-1e30 in x and z.  With ANY (no
// lathe material passes shadow rays) a lane stops at its first hit nearer
// than its bound, which the traversal would take as well.
#ifndef LATHE_SCAN_MAX
#define LATHE_SCAN_MAX 64
#endif
template <bool ANY>
SHD bool lathe_frusta_hit(const DScene& S, const DRay& r, const ConeYRay& cr, const Lathe& l, float smax, float& t) {
    bool any = false;
    float closest = r.tmax;
    DHit tmp;
    for (int i = 0; i < l.point_count - 1; i++) {
        const int g = __builtin_amdgcn_readfirstlane(l.point_offset + i);
        const LatheFrustum f = sload(S.lathe_frusta + g);
        if (f.mode == 1) continue;
        bool hit;
        if (f.mode == 0) {
            hit = hit_cone_y_rec(r, cr, f, tmp, fminf(smax, closest));
        } else {
            const LathePoint p0 = sload(S.lathe_points + g), p1 = sload(S.lathe_points + g + 1);
            hit = hit_cone_y_div(r, p0.y, p0.radius, p1.y, p1.radius, tmp, fminf(smax, closest));
        }
        if (hit && tmp.t < closest) {
            closest = tmp.t;
            any = true;
            if (ANY && closest < smax) break;
        }
    }
    t = closest;
    return any;
}
template <bool ANY, bool COUNT>
SHD void lathe_scan(const DScene& S, const DRay& r, v3 inv, float& ct, int& ci, bool& ch, TraceStats& ts) {
    const ConeYRay cr = cone_y_ray(r);
    for (int j = 0; j < S.lathe_scan; j++) {
        const int idx = __builtin_amdgcn_readfirstlane(sload(S.lathe_order + __builtin_amdgcn_readfirstlane(j)));
        const float* bb = S.bounds[C_LATHE] + 6 * (size_t)idx;
        const float mn[3] = {sload(bb), sload(bb + 1), sload(bb + 2)};
        const float mx[3] = {sload(bb + 3), sload(bb + 4), sload(bb + 5)};
        if (!ref_slab(r, inv, mn, mx)) continue;
        // The lathe is skipped when the ray's segment [t_min, running best]
        // lies wholly above or below the y range of its frusta (lathe_yr): then
        // every frustum's axial test would skip it (tp = +-(y - y0) beyond [0, L]
        // at both ends).  The margin, 2^-16 of the magnitudes involved, exceeds
        // by far the rounding that separates the float y here from the float tp
        // of those tests (a few units of 2^-24 of the same magnitudes).
        {
            const float2 yr = sload(S.lathe_yr + idx);
            const float ya = r.o.y + r.tmin * r.d.y, yb = r.o.y + ct * r.d.y;
            const float ym = fmaxf(fabsf(yr.x), fabsf(yr.y)) + fabsf(r.o.y);
            const float ma = 0x1p-16f * (ym + fabsf(r.tmin * r.d.y)), mb = 0x1p-16f * (ym + fabsf(ct * r.d.y));
            if ((ya > yr.y + ma && yb > yr.y + mb) || (ya < yr.x - ma && yb < yr.x - mb)) continue;
        }
        const Lathe l = sload(S.lathes + idx);
        if (l.point_count < 2) continue;
        if (COUNT) ts.prims++;
        float t;
        if (!lathe_frusta_hit<ANY>(S, r, cr, l, ct, t)) continue;
        if (t < ct || (t == ct && ch && idx < ci)) {
            ct = t;
            ci = idx;
            ch = true;
            if (ANY) return;
        }
    }
}
template <int CAT, bool ANY, bool COUNT>
SHD void cat_traverse(const DScene& S, const DRay& r, v3 inv, float& ct, int& ci, bool& ch, TStack& st,
                      TraceStats& ts) {
    // the scan reads the lathes through the scalar cache: every active lane
    // must hold this descriptor (a batch wave whose lanes trace several
    // scenes traverses the trees)
    if (CAT == C_LATHE && S.lathe_scan > 0 && S.wave_one_scene) lathe_scan<ANY, COUNT>(S, r, inv, ct, ci, ch, ts);
    else traverse<CAT, ANY, COUNT>(S, r, inv, ct, ci, ch, st, ts);
}

template <int CAT, bool COUNT>
SHD void closest_cat(const DScene& S, const DRay& r, v3 inv, float init_t, QHit& q, float& best, int& last_obj,
                     TStack& st, TraceStats& ts) {
    if (S.count[CAT] == 0) return;
    float ct = init_t;
    int ci = -1;
    bool ch = false;
    cat_traverse<CAT, false, COUNT>(S, r, inv, ct, ci, ch, st, ts);
    if (ch) {
        last_obj = (CAT == C_SPHERE) ? ci : -(ci + 1);
        if (ct < best) { best = ct; q.t = ct; q.kind = CAT; q.idx = ci; }
    }
}

// A CSG hit's record carries the leaf of the winning span (index in the low
// 30 bits, leaf kind above: 0 sphere, 1 superellipsoid), which with the hit
// distance fixes everything the shading needs: spans are copied whole, so a
// span's material is always its leaf's (csg_eval).
SHD int csg_hit_leaf(const CSpan& iv) { return (int)(((uint32_t)iv.kind << 30) | ((uint32_t)iv.leaf & 0x3FFFFFFFu)); }
SHD int csg_leaf_material(const DScene& S, int packed) {
    const int leaf = packed & 0x3FFFFFFF, kind = (int)((uint32_t)packed >> 30);
    return kind == 0 ? S.objects[leaf].material_id : kind == 1 ? S.ses[leaf].material_id : 0;
}

// scene_intersect (reference integrator/raycast.h:14-70).  FULL = false
// compiles only the sphere, triangle and plane categories (scenes without
// other kinds use that instantiation).
template <bool COUNT, bool FULL = true>
SHD void query_closest(const DScene& S, const DRay& r, QHit& q, TStack& st, TraceStats& ts) {
    const v3 inv = mk3(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
    q.t = 0.0f; q.kind = -1; q.idx = -1; q.carry = 0;
    float best = r.tmax;
    int last_obj = 0;
    // With a CSG tree over a scene with triangles the CSG hit inherits the
    // object id of the last category that hit at all, so those categories
    // must not be culled by the running best.
    const bool exact_carry = (S.csg_root >= 0) && (S.num_triangles > 0);
#define SHEEN_INIT (exact_carry ? r.tmax : best)
    closest_cat<C_SPHERE, COUNT>(S, r, inv, SHEEN_INIT, q, best, last_obj, st, ts);
    if (FULL) closest_cat<C_STRAND, COUNT>(S, r, inv, SHEEN_INIT, q, best, last_obj, st, ts);
    closest_cat<C_TRI, COUNT>(S, r, inv, SHEEN_INIT, q, best, last_obj, st, ts);
    if (FULL) {
        closest_cat<C_HF, COUNT>(S, r, inv, SHEEN_INIT, q, best, last_obj, st, ts);
        closest_cat<C_SE, COUNT>(S, r, inv, SHEEN_INIT, q, best, last_obj, st, ts);
        closest_cat<C_LATHE, COUNT>(S, r, inv, SHEEN_INIT, q, best, last_obj, st, ts);
        closest_cat<C_PRISM, COUNT>(S, r, inv, SHEEN_INIT, q, best, last_obj, st, ts);
    }
#undef SHEEN_INIT
    if (FULL && S.csg_root >= 0) {
        CSpan iv;
        float t;
        if (csg_query(S, r, r.tmin, r.tmax, iv, t, st) && t < best) {
            best = t; q.t = t; q.kind = C_CSG; q.idx = csg_hit_leaf(iv); q.carry = last_obj;
        }
    }
    for (int i = 0; i < S.num_planes; i++) {
        DHit h;
        const Plane& p = S.planes[i];
        if (hit_plane(r, cv3(p.normal), p.d, h) && h.t < best) { best = h.t; q.t = h.t; q.kind = C_PLANE; q.idx = i; }
    }
}

template <int CAT, bool COUNT>
SHD bool opaque_cat(const DScene& S, const DRay& r, v3 inv, QHit& q, TStack& st, TraceStats& ts) {
    if (S.count[CAT] == 0) return false;
    float ct = r.tmax;
    int ci = -1;
    bool ch = false;
    if (S.cat_opaque[CAT]) cat_traverse<CAT, true, COUNT>(S, r, inv, ct, ci, ch, st, ts);
    else cat_traverse<CAT, false, COUNT>(S, r, inv, ct, ci, ch, st, ts);
    if (ch) { q.t = ct; q.kind = CAT; q.idx = ci; }
    return ch;
}

// scene_intersect_opaque (reference integrator/raycast.h:74-97).  For a
// category whose materials can never pass a shadow ray the first hit found
// decides (the closest hit of that category would block as well).
template <bool COUNT, bool FULL = true>
SHD bool query_opaque(const DScene& S, const DRay& r, QHit& q, TStack& st, TraceStats& ts) {
    const v3 inv = mk3(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
    q.kind = -1; q.idx = -1; q.carry = 0; q.t = 0.0f;
    if (opaque_cat<C_SPHERE, COUNT>(S, r, inv, q, st, ts)) return true;
    if (opaque_cat<C_TRI, COUNT>(S, r, inv, q, st, ts)) return true;
    auto csg = [&]() {
        CSpan iv;
        float t;
        if (!csg_query(S, r, r.tmin, r.tmax, iv, t, st)) return false;
        q.t = t; q.kind = C_CSG; q.idx = csg_hit_leaf(iv); q.carry = 0;
        return true;
    };
    if (FULL) {
        if (opaque_cat<C_HF, COUNT>(S, r, inv, q, st, ts)) return true;
        if (opaque_cat<C_SE, COUNT>(S, r, inv, q, st, ts)) return true;
        if (S.opaque_tail) {
            // lathes, prisms and the CSG tree all block every shadow ray they
            // hit, so whichever of them reports a hit first, the pass is
            // blocked as when the first of them in the reference's order
            // reports it: the cheaper ones are asked first
            if (opaque_cat<C_PRISM, COUNT>(S, r, inv, q, st, ts)) return true;
            if (S.csg_root >= 0 && csg()) return true;
            if (opaque_cat<C_LATHE, COUNT>(S, r, inv, q, st, ts)) return true;
        } else {
            if (opaque_cat<C_LATHE, COUNT>(S, r, inv, q, st, ts)) return true;
            if (opaque_cat<C_PRISM, COUNT>(S, r, inv, q, st, ts)) return true;
            if (S.csg_root >= 0 && csg()) return true;
        }
    }
    bool any = false;
    float closest = r.tmax;
    for (int i = 0; i < S.num_planes; i++) {
        DHit h;
        const Plane& p = S.planes[i];
        if (hit_plane(r, cv3(p.normal), p.d, h) && h.t < closest) { closest = h.t; q.t = h.t; q.kind = C_PLANE; q.idx = i; any = true; }
    }
    return any;
}

// Material of a hit without rebuilding the record.
template <bool FULL = true>
SHD int hit_material(const DScene& S, const QHit& q, const DRay& r) {
    switch (q.kind) {
        case C_SPHERE: return S.objects[q.idx].material_id;
        case C_STRAND: return S.strands[q.idx].material_id;
        case C_TRI: return S.triangles[q.idx].material_id;
        case C_HF: return S.hfs[q.idx].material_id;
        case C_SE: return S.ses[q.idx].material_id;
        case C_LATHE: return S.lathes[q.idx].material_id;
        case C_PRISM: return S.prisms[q.idx].material_id;
        case C_CSG:
            if (!FULL) return 0;
            return csg_leaf_material(S, q.idx);
        case C_PLANE: return S.planes[q.idx].material_id;
    }
    return 0;
}

// Full hit record of the winning primitive: its test is run again on the
// same ray, which reproduces the distance and every attribute exactly.
template <bool FULL = true>
SHD void rebuild_hit(const DScene& S, const DRay& r, const QHit& q, DHit& h) {
    h.kind = q.kind;
    h.index = q.idx;
    h.tangent = mk3(0, 0, 0);
    h.uvx = 0.0f; h.uvy = 0.0f;
    switch (q.kind) {
        case C_SPHERE: {
            const Object& o = S.objects[q.idx];
            hit_sphere(r, cv3(o.param1), o.param2, h);
            h.material_id = o.material_id; h.object_id = q.idx;
            break;
        }
        case C_STRAND:
            // every frustum solved: a cap hit's tangent is the last frustum hit's
            if (FULL) hit_strand(r, S.strands[q.idx], S.sweep_points, h, r.tmax, false);
            h.object_id = -(q.idx + 1);
            break;
        case C_TRI:
            hit_triangle(r, S.triangles[q.idx], h);
            h.material_id = S.triangles[q.idx].material_id; h.object_id = -(q.idx + 1);
            break;
        case C_HF:
            // the traced distance; the march is not repeated
            if (FULL) hf_attrs(r, S.hfs[q.idx], q.t, r.o + q.t * r.d, h);
            h.material_id = S.hfs[q.idx].material_id; h.object_id = -(q.idx + 1);
            break;
        case C_SE:
            if (FULL) se_attrs(r, S.ses[q.idx], q.t, h);  // the traced entry; no new search
            h.material_id = S.ses[q.idx].material_id; h.object_id = -(q.idx + 1);
            break;
        case C_LATHE:
            if (FULL) hit_lathe(r, S.lathes[q.idx], S.lathe_points, h, q.t);  // frusta past the hit are not needed
            h.object_id = -(q.idx + 1);
            break;
        case C_PRISM:
            if (FULL) hit_prism(r, S.prisms[q.idx], S.prism_points, h);
            h.object_id = -(q.idx + 1);
            break;
        case C_CSG: {
            if (!FULL) break;
            // the query's winning span, carried in the hit (csg_hit_leaf)
            const float t = q.t;
            const int leaf = q.idx & 0x3FFFFFFF, kind = (int)((uint32_t)q.idx >> 30);
            h.t = t;
            h.material_id = csg_leaf_material(S, q.idx);
            h.point = r.o + t * r.d;
            if (kind == 0) {
                h.normal = normalize(h.point - cv3(S.objects[leaf].param1));
                h.tangent = mk3(0, 0, 0);
            } else if (kind == 1) {
                h.normal = normalize(se_gradient(S.ses[leaf], h.point));
                h.tangent = mk3(1, 0, 0);
            } else {
                h.normal = mk3(0, 1, 0);
                h.tangent = mk3(1, 0, 0);
            }
            h.uvx = h.point.x; h.uvy = h.point.y;
            if (dot(h.normal, r.d) > 0.0f) h.normal = -h.normal;
            h.object_id = q.carry;
            break;
        }
        case C_PLANE: {
            const Plane& p = S.planes[q.idx];
            hit_plane(r, cv3(p.normal), p.d, h);
            h.material_id = p.material_id; h.object_id = -1 - q.idx;
            break;
        }
    }
}

// Optical depth of all strands along a shadow ray, every strand in index
// order (the restatement of the reference's sum, bvh.h:928-960, that the
// BVH-driven sum below reproduces).
SHD float strands_depth_linear(const DScene& S, const DRay& r) {
    const v3 inv = mk3(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
    float tau = 0.0f;
    for (int i = 0; i < S.count[C_STRAND]; i++) {
        const float* bb = S.bounds[C_STRAND] + 6 * (size_t)i;
        if (!ref_slab(r, inv, bb, bb + 3)) continue;
        tau += strand_depth(r, S.strands[i], S.sweep_points, 6.0f);
        if (tau > 6.0f) return tau;
    }
    return tau;
}

// The strands' nonzero optical depths met along a ray, at most STRAND_HITS
// of them, kept sorted by strand index (unrolled insertion, no indexed
// registers); `over` when more were met.
#define STRAND_HITS 4
struct StrandTau {
    static constexpr bool active = true;
    const DScene* S;
    const DRay* r;
    v3 inv;
    uint32_t n = 0;
    bool over = false;
    int id[STRAND_HITS];
    float tau[STRAND_HITS];
    SHD void operator()(uint32_t slot) {
        const int i = (int)S->prim_ids[C_STRAND][slot];
        const float* bb = S->bounds[C_STRAND] + 6 * (size_t)i;
        if (!ref_slab(*r, inv, bb, bb + 3)) return;
        const float d = strand_depth(*r, S->strands[i], S->sweep_points, 6.0f);
        if (d == 0.0f) return;  // a zero term leaves every partial sum unchanged
        if (n == STRAND_HITS) { over = true; return; }
        int ci = i;
        float cd = d;
#pragma unroll
        for (int k = 0; k < STRAND_HITS; k++) {
            if ((uint32_t)k < n && id[k] > ci) {  // shift the larger entry up
                const int ti = id[k];
                const float td = tau[k];
                id[k] = ci; tau[k] = cd;
                ci = ti; cd = td;
            }
            if ((uint32_t)k == n) { id[k] = ci; tau[k] = cd; }
        }
        n++;
    }
};

// Optical depth through the strands (reference strand_bvh_tau, bvh.h:928-960,
// with sigma 6 and the stop past 6): the sum over the strands whose bounds
// pass the slab test, in strand index order.  The strand BVH finds them; a
// strand the ray passes outside adds exactly 0, so summing only the nonzero
// terms in index order gives the index-order sum over all of them bit for
// bit.  A ray that meets more than STRAND_HITS strands takes the linear sum.
// The slab test (bvh.h:578-600) checks only the x/y interval against
// [t_min, t_max] and the z interval against the x/y one, so it accepts boxes
// that the ray's line meets outside [t_min, t_max] (and strand_depth then
// counts from t = 0); the walk therefore tests the node boxes against the
// whole line, which contains every box the slab test accepts.
SHD float strands_depth(const DScene& S, const DRay& r, TStack& st) {
    if (S.count[C_STRAND] == 0) return 0.0f;
    StrandTau v;
    v.S = &S;
    v.r = &r;
    v.inv = mk3(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
    DRay line = r;
    line.tmin = -__builtin_huge_valf();
    line.tmax = __builtin_huge_valf();
    TravState T;
    T.node = 0; T.sp = 0; T.ct = __builtin_huge_valf(); T.ci = 0; T.ch = false;
    trav_begin(S, C_STRAND, line, v.inv, T);
    TraceStats ts{0, 0};
    while (!trav_step<C_STRAND, false, StrandTau>(S, C_STRAND, false, line, v.inv, T, st, ts, false, &v)) {
    }
    if (v.over) return strands_depth_linear(S, r);
    float tau = 0.0f;
#pragma unroll
    for (int k = 0; k < STRAND_HITS; k++) {
        if ((uint32_t)k < v.n) {
            tau += v.tau[k];
            if (tau > 6.0f) return tau;
        }
    }
    return tau;
}
