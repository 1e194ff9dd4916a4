// kernels.h — wavefront kernels of one frame chunk.
//
// A chunk is a contiguous range of sample slots.  Slot g maps to 8x8 pixel
// tiles so that one wave64 holds the same sample of 64 neighbouring pixels:
//   lane = g & 63, s = (g >> 6) % ns, tile = (g >> 6) / ns.
// Per chunk: primary trace -> primary shading (shadow requests, reflection
// and first refraction rays) -> up to three rounds of secondary trace and
// shading -> shadow trace -> per-pixel accumulation in the reference's order
// (render_kernel.cu:491-813).
#pragma once
#include "dev_shade.h"

struct FrameDesc {
    int width, height, ns;
    int tiles_x;
    uint32_t g0;       // one scene: first slot of the chunk; a batch: the scene's first slot
    uint32_t n_slots;  // one scene: slots in the chunk (multiple of 64*ns); a batch: the scene's slots
    uint32_t out0;     // first pixel of the image in the output buffer
};

// Sample slot -> pixel / sample index.
SHD bool slot_pixel(const FrameDesc& F, uint32_t g, int& x, int& y, int& s) {
    uint32_t lane = g & 63u, blk = g >> 6;
    s = (int)(blk % (uint32_t)F.ns);
    uint32_t tile = blk / (uint32_t)F.ns;
    x = (int)((tile % (uint32_t)F.tiles_x) * 8u + (lane & 7u));
    y = (int)((tile / (uint32_t)F.tiles_x) * 8u + (lane >> 3));
    return x < F.width && y < F.height;
}

// Camera ray of one sample (reference render_kernel.cu:491-521 with
// sampling/aa.h:25-71).
SHD DRay camera_ray(const CameraData& cam, int x, int y, int s, int ns, int W, int H, float& pix_angle,
                    uint32_t& seed) {
    uint32_t js = (uint32_t(x) * 73856093u) ^ (uint32_t(y) * 19349663u) ^ (uint32_t(s) * 83492791u);
    float ox, oy;
    if (ns <= 4) {
        float cx = (s == 0 || s == 2) ? 0.25f : 0.75f;
        float cy = (s < 2) ? 0.25f : 0.75f;
        float jx = rand01(js) - 0.5f;
        float jy = rand01(js ^ 0x9e3779b9u) - 0.5f;
        ox = cx + jx * 0.5f;
        oy = cy + jy * 0.5f;
    } else {
        int side = 2;
        while (side * side < ns) side++;
        int cx = s % side, cy = s / side;
        float inv = 1.0f / (float)side;
        ox = ((float)cx + rand01(js)) * inv;
        oy = ((float)cy + rand01(js ^ 0x9e3779b9u)) * inv;
    }
    float u = ((float)x + ox) / (float)W * 2.0f - 1.0f;
    float v = 1.0f - ((float)y + oy) / (float)H * 2.0f;
    seed = mix32(js);
    float tan_h = 1.0f, tan_x = 1.0f, tan_y = 1.0f;
    if (cam.focal_length > 0.0f) {
        tan_h = 0.5f * cam.sensor_width / cam.focal_length;
        if (W >= H) { tan_x = tan_h; tan_y = tan_h * (float)H / (float)W; }
        else { tan_y = tan_h; tan_x = tan_h * (float)W / (float)H; }
    }
    DRay r;
    r.o = cv3(cam.position);
    r.d = normalize(cv3(cam.forward) + cv3(cam.right) * (u * tan_x + cam.shift_x * tan_h) +
                    cv3(cam.up) * (v * tan_y + cam.shift_y * tan_h));
    r.tmin = 0.001f;
    r.tmax = 1e10f;
    pix_angle = 2.0f * tan_x / (float)W;
    return r;
}

struct SampleBufs {
    int4* hit_p;        // [slots] primary hits
    float4* samp;       // [slots] x: reflect weight, y: transmission, z: chain weight, w: flags
    float4* beam_add;   // [slots * n_beam] or null
    float4* media_p;    // [slots] or null: inscatter xyz, exp(-tau) w
    float4* media_r;    // [slots] or null
    int n_beam;
};
#define SF_VALID 1
#define SF_HIT 2
#define SF_REFL 4
#define SF_REFR 8
#define SF_EV2 16

#define SORT_BINS_MAX 64
// Secondary ray queue entry (reflection or refraction face), 32 bytes.
struct SecRay {
    float ox, oy, oz, w;
    float dx, dy, dz;
    uint32_t meta;  // bits 0..27 slot, 28 refraction, 29 entering, 30..31 face
};

// An event count that many waves add to: CTR_LINES partial sums, one
// 128-byte line apart, each wave adding to the line of its block.
// Device-scope atomics on one address are serialised past the per-XCD L2s,
// and waves that finish together (the empty tail of a grid-stride launch)
// queued on a single counter: the three secondary trace launches of the
// 1M-triangle frame spent 0.23 ms on it.  The host adds the partial sums.
#define CTR_LINES 16
struct CtrSum {
    unsigned long long line[CTR_LINES][16];
};
inline unsigned long long ctr_total(const CtrSum& c) {
    unsigned long long t = 0;
    for (int i = 0; i < CTR_LINES; i++) t += c.line[i][0];
    return t;
}

struct Counters {
    uint32_t sec_n[4];       // secondary queues (ping-pong) sizes
    uint32_t cont_n[2];      // shadow continuations queued for the second and third pass
    uint32_t pad[26];        // (the shadow queue counts are per light: EventBufs::shadow_nl)
    uint32_t sort_tot[SORT_BINS_MAX];  // per-bin totals of the secondary-ray sort (k_sec_scan)
    CtrSum shadow_passes;
    CtrSum nodes, prims, snodes, sprims, pnodes, pprims;
    CtrSum reflection, refraction, extra_shadow;
    CtrSum spills, sspills;
    CtrSum switers;  // node-visit iterations of the shadow kernel's waves
    CtrSum cnodes, cprims;  // the shadow continuation kernels
};

SHD uint32_t wave_sum(uint32_t v) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
SHD void wave_add64(unsigned long long* dst, uint32_t v) {
    uint32_t t = wave_sum(v);
    if (__lane_id() == __ffsll((unsigned long long)__ballot(1)) - 1) atomicAdd(dst, (unsigned long long)t);
}
// A wave's total into its block's line of `dst`; a zero total adds nothing.
SHD void wave_add64(CtrSum& dst, uint32_t v) {
    const uint32_t t = wave_sum(v);
    if (t != 0u && __lane_id() == __ffsll((unsigned long long)__ballot(1)) - 1)
        atomicAdd(&dst.line[blockIdx.x % CTR_LINES][0], (unsigned long long)t);
}

// One pass of the shadow test (reference render_kernel.cu:243-268): the
// opaque query from r; 1 when nothing is hit, 0 when the hit blocks, 2 when
// it is glass the ray passes (r.o then moved past it, t_max kept).
template <bool COUNT, bool FULL>
SHD int shadow_pass(const DScene& S, DRay& r, TStack& st, uint32_t& passes, TraceStats& ts) {
    passes++;
    QHit h;
    if (!query_opaque<COUNT, FULL>(S, r, h, st, ts)) return 1;
    const Material& m = S.materials[hit_material<FULL>(S, h, r)];
    if (m.transmission > 0.5f) {
        DHit full;
        rebuild_hit<FULL>(S, r, h, full);
        if (transmission_at(S, full.uvx, full.uvy, m) > 0.5f) {
            r.o = full.point + r.d * 0.002f;
            return 2;
        }
    }
    return 0;
}

// Shadow test of one request (reference render_kernel.cu:237-280) from its
// pass `first` on: at most three passes in all; after three glass passes the
// ray is unblocked.
template <bool COUNT, bool FULL>
SHD float shadow_rest(const DScene& S, DRay& r, const DRay& orig, uint32_t slot, int first, TStack& st,
                      uint32_t& passes, TraceStats& ts) {
    for (int pass = first; pass < 3; pass++) {
        const int res = shadow_pass<COUNT, FULL>(S, r, st, passes, ts);
        if (res == 0) return 0.0f;
        if (res == 1) break;
    }
    if (FULL && (slot & 0x80000000u)) return exp_r(-strands_depth(S, orig, st));
    return 1.0f;
}
template <bool COUNT, bool FULL>
SHD float shadow_eval(const DScene& S, const ShadowReq& q, TStack& st, uint32_t& passes, TraceStats& ts) {
    DRay r;
    r.o = mk3(q.ox, q.oy, q.oz);
    r.d = mk3(q.dx, q.dy, q.dz);
    r.tmin = 0.001f;
    r.tmax = q.tmax;
    const DRay orig = r;
    return shadow_rest<COUNT, FULL>(S, r, orig, q.slot, 0, st, passes, ts);
}

// Occlusion query used by beams and media (scene_intersect_opaque only).
SHD bool occluded(const DScene& S, const DRay& r, TStack& st, QHit& h) {
    TraceStats ts{0, 0};
    return query_opaque<false>(S, r, h, st, ts);
}

// Spot-light beams of one primary sample (reference render_kernel.cu:560-652).
// out: n_out entries (the largest beam count of the scenes sharing the work
// buffers); entries past this scene's beams are written as zero.
SHD void beams_of(const DScene& S, const DRay& ray, bool hit, float hit_t, uint32_t seed, float4* out, int n_out,
                  TStack& st, uint32_t& passes) {
    int k_out = 0;
    for (int i = 0; i < S.num_lights; i++) {
        const Light& light = S.lights[i];
        if (light.type != 3 || light.beam <= 0.0f) continue;
        v3 acc = mk3(0, 0, 0);
        float t_end = hit ? fminf(hit_t, 30.0f) : 30.0f;
        if (t_end > 0.0f) {
            const int STEPS = 6;
            float dt = t_end / (float)STEPS;
            float cos_half = cos_r(light.angle);
            float blend = fminf(fmaxf(light.blend, 0.0f), 1.0f);
            float cos_soft = cos_r(light.angle * (1.0f - blend));
            for (int k = 0; k < STEPS; k++) {
                float t = ((float)k + 0.5f) * dt;
                v3 p = ray.o + ray.d * t;
                v3 vv = p - cv3(light.position);
                float d2 = fmaxf(dot(vv, vv), 1.0e-4f);
                float c = dot(vv, cv3(light.direction)) * (1.0f / sqrtf(d2));
                if (c <= cos_half) continue;
                float w = 1.0f;
                if (cos_soft > cos_half + 1.0e-4f) w = fminf((c - cos_half) / (cos_soft - cos_half), 1.0f);
                float dens = light.beam * light.intensity * 1.0e-3f / (1.0f + 0.25f * d2);
                v3 target = cv3(light.position);
                if (light.radius > 0.0f) {
                    v3 axis = normalize(-vv);
                    v3 up = (fabsf(axis.y) < 0.9f) ? mk3(0, 1, 0) : mk3(1, 0, 0);
                    v3 t1 = normalize(cross(axis, up));
                    v3 t2 = cross(axis, t1);
                    float rr = sqrtf(rand01(seed ^ (uint32_t)(uint32_t(k) * 2654435761u) ^ 0x85ebca6bu)) * light.radius;
                    float ph = 6.2831853f * rand01(seed ^ (uint32_t)(uint32_t(k) * 2246822519u) ^ 0xc2b2ae35u);
                    target = cv3(light.position) + t1 * (cos_r(ph) * rr) + t2 * (sin_r(ph) * rr);
                }
                v3 to_l = target - p;
                float ldist = sqrtf(fmaxf(dot(to_l, to_l), 1.0e-8f));
                v3 ldir = to_l / ldist;
                float remain = ldist * 0.999f;
                DRay sr;
                sr.o = p; sr.d = ldir; sr.tmin = 0.002f; sr.tmax = remain;
                bool blocked = false;
                for (int pass = 0; pass < 3; pass++) {
                    QHit h;
                    passes++;
                    if (!occluded(S, sr, st, h)) { blocked = false; break; }
                    DHit full;
                    const Material& m = S.materials[hit_material(S, h, sr)];
                    if (m.transmission > 0.5f) {
                        rebuild_hit(S, sr, h, full);
                        remain -= full.t + 0.002f;
                        if (remain <= 0.0f) { blocked = false; break; }
                        sr.o = full.point + ldir * 0.002f;
                        sr.tmax = remain;
                        continue;
                    }
                    blocked = true;
                    break;
                }
                if (blocked) continue;
                acc = acc + cv3(light.color) * (dens * w * dt);
            }
        }
        out[k_out++] = make_float4(acc.x, acc.y, acc.z, 0.0f);
    }
    for (; k_out < n_out; k_out++) out[k_out] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

SHD bool ray_box(const DRay& r, v3 c, v3 h, float& t0, float& t1) {  // render_kernel.cu:302-325
    t0 = -1.0e30f; t1 = 1.0e30f;
    const float o[3] = {r.o.x, r.o.y, r.o.z}, d[3] = {r.d.x, r.d.y, r.d.z};
    const float cc[3] = {c.x, c.y, c.z}, hh[3] = {h.x, h.y, h.z};
    for (int a = 0; a < 3; a++) {
        if (fabsf(d[a]) < 1.0e-8f) {
            if (o[a] < cc[a] - hh[a] || o[a] > cc[a] + hh[a]) return false;
            continue;
        }
        float inv = 1.0f / d[a];
        float tn = (cc[a] - hh[a] - o[a]) * inv;
        float tf = (cc[a] + hh[a] - o[a]) * inv;
        if (tn > tf) { float t = tn; tn = tf; tf = t; }
        t0 = fmaxf(t0, tn);
        t1 = fminf(t1, tf);
        if (t0 > t1) return false;
    }
    return true;
}

// Single-scatter media along a ray (reference render_kernel.cu:339-436).
SHD float4 media_of(const DScene& S, const DRay& ray, float t_surface, uint32_t seed, TStack& st, uint32_t& passes) {
    v3 in_scatter = mk3(0, 0, 0);
    float tau_total = 0.0f;
    for (int m = 0; m < S.num_media; m++) {
        const Medium& med = S.media[m];
        if (med.samples <= 0) continue;
        float m0, m1;
        if (!ray_box(ray, cv3(med.center), cv3(med.half), m0, m1)) continue;
        m0 = fmaxf(m0, ray.tmin);
        m1 = fminf(m1, t_surface);
        if (m1 <= m0) continue;
        int n = med.samples > 64 ? 64 : med.samples;
        float dt = (m1 - m0) / (float)n;
        float tau = 0.0f;
        for (int k = 0; k < n; k++) {
            float jitter = rand01(seed ^ (uint32_t)(uint32_t(m) * 9781u) ^ (uint32_t)(uint32_t(k) * 6271u));
            float t = m0 + ((float)k + jitter) * dt;
            v3 p = ray.o + ray.d * t;
            v3 radiance = mk3(0, 0, 0);
            for (int i = 0; i < S.num_lights; i++) {
                const Light& light = S.lights[i];
                if (!(light.flags & 1)) continue;
                v3 ldir, lrad;
                float ldist;
                if (light.type == 0) {
                    ldir = sun_dir(light, seed ^ (uint32_t)(uint32_t(i) * 2654435761u) ^ (uint32_t)(uint32_t(k) * 2246822519u));
                    ldist = 1.0e9f;
                    lrad = cv3(light.color) * light.intensity;
                } else {
                    v3 tl = cv3(light.position) - p;
                    float d2 = fmaxf(dot(tl, tl), 1.0e-4f);
                    ldist = sqrtf(d2);
                    ldir = tl / ldist;
                    lrad = cv3(light.color) * (light.intensity / (12.5663706f * d2));
                    if (light.type == 3) {
                        float c = dot(-ldir, cv3(light.direction));
                        float cos_half = cos_r(light.angle);
                        if (c <= cos_half) continue;
                        float blend = fminf(fmaxf(light.blend, 0.0f), 1.0f);
                        float cos_soft = cos_r(light.angle * (1.0f - blend));
                        if (cos_soft > cos_half + 1.0e-4f) lrad = lrad * fminf((c - cos_half) / (cos_soft - cos_half), 1.0f);
                    }
                }
                DRay sr;
                sr.o = p; sr.d = ldir; sr.tmin = 0.002f; sr.tmax = ldist;
                QHit h;
                passes++;
                if (occluded(S, sr, st, h)) continue;
                radiance = radiance + lrad;
            }
            v3 scatter = cv3(med.sigma_s) * radiance * 0.07957747f;
            in_scatter = in_scatter + scatter * (exp_r(-tau) * dt);
            tau += med.sigma_t * dt;
        }
        tau_total += tau;
    }
    return make_float4(in_scatter.x, in_scatter.y, in_scatter.z, exp_r(-tau_total));
}

// First/next refraction face (reference render_kernel.cu:750-765).
SHD void refract_from(const DScene& S, const DHit& gh, v3 d, bool entering, float& w, v3& o_out, v3& td) {
    const Material& sm = S.materials[gh.material_id];
    v3 n = gh.normal;
    float ior = fmaxf(ior_at(S, gh.uvx, gh.uvy, sm), 1.0f);
    float eta = entering ? (1.0f / ior) : ior;
    float cosi = fmaxf(-dot(d, n), 1.0e-4f);
    float k = 1.0f - eta * eta * (1.0f - cosi * cosi);
    td = (k < 0.0f) ? normalize(d + 2.0f * cosi * n) : normalize(eta * d + (eta * cosi - sqrtf(k)) * n);
    w *= transmission_at(S, gh.uvx, gh.uvy, sm);
    o_out = gh.point + td * 0.002f;
}

// Where a chunk's samples come from.  OneSrc: one scene, passed as the
// kernel argument, chunk slots [F.g0, F.g0 + F.n_slots) of its slot space.
// ManySrc: a batch; the scenes' slot ranges are concatenated (each a
// multiple of 64, so a wave's 64 slots belong to one scene), the chunk is
// [g0, g0 + n) of that space, and a slot's scene is found by binary search
// over the scenes' first slots: first over every step-th of them, copied
// into LDS when the kernel starts (prepare), then over at most step of them
// in global memory (step = 1 for batches of up to SCENE_LDS scenes).
// Kernels use:
//   prepare(lds)    first statement of every kernel (a block barrier for ManySrc);
//   n_slots()       slots of the chunk;
//   find(i)         scene of chunk slot i (i < n_slots());
//   scene(k), frame(k), local(k, i)  the scene, its frame, i in its slot space;
//   pix_count(), pix_first(p, i0)    k_accumulate's enumeration of pixels.
#define SCENE_LDS 512
#define SCENE_TABLE(src)                          \
    __shared__ uint32_t s_scene_first[SCENE_LDS]; \
    src.prepare(s_scene_first)
struct OneSrc {
    DScene S;
    FrameDesc F;
    SHD void prepare(uint32_t*) {}
    SHD uint32_t n_slots() const { return F.n_slots; }
    SHD int find(uint32_t) const { return 0; }
    SHD int find_wave(uint32_t) const { return 0; }
    SHD const DScene& scene(int) const { return S; }
    SHD const DScene& trace_scene(int) const { return S; }
    SHD const FrameDesc& frame(int) const { return F; }
    SHD uint32_t local(int, uint32_t i) const { return F.g0 + i; }
    SHD uint32_t pix_count() const { return F.n_slots / (uint32_t)F.ns; }
    // pixel p of the chunk: lane p&63 of tile p>>6; its samples at i0 + s*64
    SHD bool pix_first(uint32_t p, uint32_t& i0) const {
        i0 = (p >> 6) * 64u * (uint32_t)F.ns + (p & 63u);
        return true;
    }
};
struct ManySrc {
    const DScene* S;
    const DScene* S1;  // the same descriptors with wave_one_scene set
    const FrameDesc* F;
    const uint32_t* base;  // [n + 1] first slot of each scene, then the total
    int n;
    uint32_t g0, count;    // the chunk
    int cap;               // entries of the LDS table, 1..SCENE_LDS
    const uint32_t* tab;   // LDS: base[j * step] for j < m
    int step, m;
    SHD void prepare(uint32_t* lds) {
        step = (n + cap - 1) / cap;
        m = (n + step - 1) / step;
        for (int j = (int)threadIdx.x; j < m; j += (int)blockDim.x) lds[j] = base[j * step];
        __syncthreads();
        tab = lds;
    }
    SHD uint32_t n_slots() const { return count; }
    SHD int find(uint32_t i) const {
        const uint32_t g = g0 + i;
        int lo = 0, hi = m - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tab[mid] <= g) lo = mid; else hi = mid - 1;
        }
        int a = lo * step, b = min(a + step - 1, n - 1);
        while (a < b) {
            const int mid = (a + b + 1) >> 1;
            if (base[mid] <= g) a = mid; else b = mid - 1;
        }
        return a;
    }
    // every lane of the wave holds a slot of the same scene
    SHD int find_wave(uint32_t i) const { return __builtin_amdgcn_readfirstlane(find(i)); }
    SHD const DScene& scene(int k) const { return S[k]; }
    // descriptor for the traversal: the one that lets the wave share its
    // triangle tests when every active lane traces scene k (checked per call)
    SHD const DScene& trace_scene(int k) const {
        const int k0 = __builtin_amdgcn_readfirstlane(k);
        return __ballot(k == k0) == __ballot(1) ? S1[k0] : S[k];
    }
    SHD const FrameDesc& frame(int k) const { return F[k]; }
    SHD uint32_t local(int k, uint32_t i) const { return g0 + i - base[k]; }
    SHD uint32_t pix_count() const { return count; }
    SHD bool pix_first(uint32_t p, uint32_t& i0) const {
        i0 = p;
        const int k = find(p);
        return ((local(k, p) >> 6) % (uint32_t)F[k].ns) == 0u;
    }
};

// Every kernel that traverses declares its LDS columns and its private
// spill array, then builds the stack view over them.
// SHEEN_STACK_N: the same for blocks of NT threads (a multiple of
// TRACE_BLOCK), the LDS columns in groups of TRACE_BLOCK lanes so that every
// lane's entries keep the TRACE_BLOCK stride the stack code assumes.
#define SHEEN_STACK(st) SHEEN_STACK_N(st, TRACE_BLOCK)
#define SHEEN_STACK_N(st, NT)                                                                    \
    static_assert((NT) % TRACE_BLOCK == 0, "stack layout: blocks of TRACE_BLOCK lanes");        \
    __shared__ uint2 s_stack[LDS_DEPTH * (NT)];                                                  \
    __shared__ uint32_t s_leaf_list[(NT) / 64][LEAF_LIST];                                       \
    __shared__ unsigned long long s_leaf_keys[NT];                                               \
    uint2 s_spill[SPILL_DEPTH];                                                                  \
    TStack st;                                                                                   \
    st.lds = s_stack + (threadIdx.x / TRACE_BLOCK) * (LDS_DEPTH * TRACE_BLOCK) + threadIdx.x % TRACE_BLOCK; \
    st.spill = s_spill;                                                                          \
    st.spills = 0;                                                                               \
    st.list = s_leaf_list[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)];                     \
    st.keys = s_leaf_keys + __builtin_amdgcn_readfirstlane(threadIdx.x & ~63u)

// ---------------------------------------------------------------------------
// Hit record of the wavefront buffers: {t bits, index, carried object id, kind}.
SHD int4 pack_hit(const QHit& q) {
    return make_int4(__float_as_int(q.t), q.idx, q.carry, q.kind);
}

template <bool COUNT, bool FULL, class SRC>
__global__ TRACE_BOUNDS_OF(FULL) void k_primary(SRC src, int4* hit_p, Counters* C) {
    SCENE_TABLE(src);
    SHEEN_STACK(st);
    TraceStats ts{0, 0};
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t n = src.n_slots();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int sid = src.find_wave(i);
        const DScene& S = src.trace_scene(sid);
        const FrameDesc& F = src.frame(sid);
        int x, y, s;
        int4 out = make_int4(0, -1, 0, -1);
        if (slot_pixel(F, src.local(sid, i), x, y, s)) {
            float pa;
            uint32_t seed;
            DRay r = camera_ray(S.cam, x, y, s, F.ns, F.width, F.height, pa, seed);
            QHit q;
            query_closest<COUNT, FULL>(S, r, q, st, ts);
            out = pack_hit(q);
        }
        hit_p[i] = out;
    }
    if (COUNT) {
        wave_add64(C->nodes, ts.nodes);
        wave_add64(C->prims, ts.prims);
        wave_add64(C->pnodes, ts.nodes);
        wave_add64(C->pprims, ts.prims);
        wave_add64(C->spills, st.spills);
    }
}

SHD QHit unpack_hit(int4 h) {
    QHit q;
    q.t = __int_as_float(h.x);
    q.idx = h.y;
    q.carry = h.z;
    q.kind = h.w;
    return q;
}

template <bool FULL, class SRC>
__global__ SHADE_BOUNDS(FULL) void k_shade_primary(SRC src, SampleBufs SB, EventBufs EB, SecRay* sec_q, Counters* C) {
    SCENE_TABLE(src);
    SHEEN_STACK_N(st, SHADE_BLOCK_OF(FULL));
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t passes = 0;
    const uint32_t n = src.n_slots();
    // loop bound rounded up to whole waves so every lane reaches the queue ballots
    const uint32_t n_iter = (n + stride - 1) / stride;
    for (uint32_t it = 0; it < n_iter; it++) {
        const uint32_t i = it * stride + blockIdx.x * blockDim.x + threadIdx.x;
        const int sid = src.find_wave(i < n ? i : n - 1);
        const DScene& S = src.scene(sid);
        const FrameDesc& F = src.frame(sid);
        int x = 0, y = 0, s = 0;
        const bool valid = i < n && slot_pixel(F, src.local(sid, i), x, y, s);
        float pa = 0.0f;
        uint32_t seed = 0;
        DRay r;
        QHit q;
        q.kind = -1;
        if (valid) {
            r = camera_ray(S.cam, x, y, s, F.ns, F.width, F.height, pa, seed);
            q = unpack_hit(SB.hit_p[i]);
        }
        const bool hit = valid && q.kind >= 0;
        DHit rec;
        if (hit) rebuild_hit<FULL>(S, r, q, rec);
        if (valid && !hit) sky_event(S, r.d, ev_of(EB, i, 0), EB);
        shade_event(S, hit, rec, -r.d, hit ? rec.t * pa : 0.0f, seed, ev_of(EB, i, 0), EB);
        float4 sp = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(0));
        int flags = valid ? SF_VALID : 0;
        bool want_r = false, want_t = false;
        SecRay rr, tr;
        if (hit) {
            flags |= SF_HIT;
            const Material& pm = S.materials[rec.material_id];
            if (pm.reflect > 0.0f) {  // render_kernel.cu:672-680
                want_r = true;
                flags |= SF_REFL;
                sp.x = fminf(pm.reflect, 1.0f);
                v3 rd = r.d - rec.normal * (2.0f * dot(r.d, rec.normal));
                v3 o = rec.point + rec.normal * 0.002f;
                v3 d = normalize(rd);
                rr.ox = o.x; rr.oy = o.y; rr.oz = o.z; rr.w = 0.0f;
                rr.dx = d.x; rr.dy = d.y; rr.dz = d.z;
                rr.meta = i;
            }
            float gtr = transmission_at(S, rec.uvx, rec.uvy, pm);  // render_kernel.cu:739
            if (gtr > 0.0f) {
                want_t = true;
                flags |= SF_REFR;
                sp.y = gtr;
                float w = 1.0f;
                v3 o, td;
                refract_from(S, rec, r.d, true, w, o, td);
                tr.ox = o.x; tr.oy = o.y; tr.oz = o.z; tr.w = w;
                tr.dx = td.x; tr.dy = td.y; tr.dz = td.z;
                tr.meta = i | (1u << 28) | (1u << 29);
            }
        }
        if (FULL && valid && SB.beam_add)
            beams_of(S, r, hit, hit ? rec.t : 0.0f, seed, SB.beam_add + (size_t)i * SB.n_beam, SB.n_beam, st, passes);
        if (FULL && valid && SB.media_p) SB.media_p[i] = media_of(S, r, hit ? rec.t : 1.0e9f, seed, st, passes);
        if (EB.order & ORDER_SEC) {
            uint32_t qa, qb;
            block_reserve_ab(&C->sec_n[0], want_r, want_t, qa, qb);
            if (want_r) sec_q[qa] = rr;
            if (want_t) sec_q[qb] = tr;
        } else {
            const uint32_t a = block_reserve(&C->sec_n[0], want_r, want_t);
            if (want_r) sec_q[a] = rr;
            if (want_t) sec_q[a + (want_r ? 1u : 0u)] = tr;
        }
        if (i < n) {
            sp.w = __int_as_float(flags);
            SB.samp[i] = sp;
        }
    }
    wave_add64(C->extra_shadow, passes);
}

template <bool COUNT, bool FULL, class SRC>
__global__ TRACE_BOUNDS_OF(FULL) void k_trace_sec(SRC src, const SecRay* q_in, int which, int4* hits, Counters* C) {
    SCENE_TABLE(src);
    SHEEN_STACK(st);
    TraceStats ts{0, 0};
    const uint32_t n = C->sec_n[which];
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t nrefl = 0, nrefr = 0;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        SecRay e = q_in[j];
        const DScene& S = src.trace_scene(src.find(e.meta & 0x0FFFFFFFu));
        DRay r;
        r.o = mk3(e.ox, e.oy, e.oz);
        r.d = mk3(e.dx, e.dy, e.dz);
        r.tmin = 0.001f;
        r.tmax = 1.0e10f;
        QHit q;
        query_closest<COUNT, FULL>(S, r, q, st, ts);
        hits[j] = pack_hit(q);
        if (e.meta & (1u << 28)) nrefr++; else nrefl++;
    }
    wave_add64(C->reflection, nrefl);
    wave_add64(C->refraction, nrefr);
    if (COUNT) {
        wave_add64(C->nodes, ts.nodes);
        wave_add64(C->prims, ts.prims);
    }
}

template <bool FULL, class SRC>
__global__ SHADE_BOUNDS(FULL) void k_shade_sec(SRC src, SampleBufs SB, EventBufs EB, const SecRay* q_in, int which,
                                         const int4* hits, SecRay* q_out, int which_out, Counters* C) {
    SCENE_TABLE(src);
    SHEEN_STACK_N(st, SHADE_BLOCK_OF(FULL));
    const uint32_t n = C->sec_n[which];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t n_iter = (n + stride - 1) / stride;
    uint32_t passes = 0;
    for (uint32_t it = 0; it < n_iter; it++) {
        const uint32_t j = it * stride + blockIdx.x * blockDim.x + threadIdx.x;
        const bool valid = j < n;
        SecRay e;
        QHit q;
        q.kind = -1;
        uint32_t i = 0;
        bool refr = false;
        DRay r;
        if (valid) {
            e = q_in[j];
            q = unpack_hit(hits[j]);
            i = e.meta & 0x0FFFFFFFu;
            refr = (e.meta >> 28) & 1u;
            r.o = mk3(e.ox, e.oy, e.oz);
            r.d = mk3(e.dx, e.dy, e.dz);
            r.tmin = 0.001f;
            r.tmax = 1.0e10f;
        }
        const int sid = src.find(i);
        const DScene& S = src.scene(sid);
        const FrameDesc& F = src.frame(sid);
        float tan_x = 1.0f;  // pixel footprint angle of the scene's camera
        if (S.cam.focal_length > 0.0f) {
            float tan_h = 0.5f * S.cam.sensor_width / S.cam.focal_length;
            tan_x = (F.width >= F.height) ? tan_h : tan_h * (float)F.width / (float)F.height;
        }
        const float pa = 2.0f * tan_x / (float)F.width;
        int x = 0, y = 0, s = 0;
        uint32_t seed = 0;
        if (valid) {
            slot_pixel(F, src.local(sid, i), x, y, s);
            seed = mix32((uint32_t(x) * 73856093u) ^ (uint32_t(y) * 19349663u) ^ (uint32_t(s) * 83492791u));
        }
        const bool hit = valid && q.kind >= 0;
        DHit h;
        if (hit) rebuild_hit<FULL>(S, r, q, h);
        bool lit = false;
        uint32_t ev = 0;
        bool want_next = false;
        SecRay nx;
        if (valid && !refr) {  // reflection: render_kernel.cu:682-711
            ev = ev_of(EB, i, 1);
            if (hit) lit = true;
            else sky_event(S, r.d, ev, EB);
            if (FULL && SB.media_r) SB.media_r[i] = media_of(S, r, hit ? h.t : 1.0e9f, seed, st, passes);
        } else if (valid) {  // refraction face: render_kernel.cu:767-802
            ev = ev_of(EB, i, 2);
            const uint32_t face = e.meta >> 30;
            const bool entering = (e.meta >> 29) & 1u;
            float w = e.w;
            if (!hit) {
                sky_event(S, r.d, ev, EB);
                float4 sp = SB.samp[i];
                sp.z = w;
                sp.w = __int_as_float(__float_as_int(sp.w) | SF_EV2);
                SB.samp[i] = sp;
            } else if (transmission_at(S, h.uvx, h.uvy, S.materials[h.material_id]) > 0.0f) {
                if (face < 2) {
                    v3 o, td;
                    refract_from(S, h, r.d, !entering, w, o, td);
                    nx.ox = o.x; nx.oy = o.y; nx.oz = o.z; nx.w = w;
                    nx.dx = td.x; nx.dy = td.y; nx.dz = td.z;
                    nx.meta = i | (1u << 28) | ((!entering) ? (1u << 29) : 0u) | ((face + 1) << 30);
                    want_next = true;
                }
            } else {
                lit = true;
                float4 sp = SB.samp[i];
                sp.z = w;
                sp.w = __int_as_float(__float_as_int(sp.w) | SF_EV2);
                SB.samp[i] = sp;
            }
        }
        shade_event(S, lit, h, -r.d, lit ? h.t * pa : 0.0f, seed, ev, EB);
        const uint32_t a = block_reserve(&C->sec_n[which_out], want_next, false);
        if (want_next) q_out[a] = nx;
    }
    wave_add64(C->extra_shadow, passes);
}

template <bool COUNT, bool FULL, class SRC>
__global__ TRACE_BOUNDS_OF(FULL) void k_shadow(SRC src, EventBufs EB, Counters* C, int part) {
    SCENE_TABLE(src);
    SHEEN_STACK(st);
    TraceStats ts{0, 0};
    // The queue is EB.runs runs (one per light, or one shared by the lights),
    // traced run after run.  part 0: every queued request; 1: those of the
    // primary shading; 2: the rest.  The FULL instantiation reads one run
    // with no search in its loop (its traversal ran 7% slower on the
    // analytic 10k frame with one).
    const int R = EB.runs;
    uint32_t lo0 = 0, n = 0;
    if (FULL) {
        lo0 = (part == 2) ? EB.shadow_nl[1] : 0u;
        n = (part == 1) ? EB.shadow_nl[1] : EB.shadow_nl[0];
    } else {
        for (int l = 0; l < R; l++) {
            const uint32_t lo = (part == 2) ? EB.shadow_nl[R + l] : 0u, hi = (part == 1) ? EB.shadow_nl[R + l] : EB.shadow_nl[l];
            n += hi - lo;
        }
    }
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t passes = 0;
    for (uint32_t j = lo0 + blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        uint32_t at = j;
        if (!FULL) {
            uint32_t e = j;
            for (int l = 0; l < R; l++) {
                const uint32_t lo = (part == 2) ? EB.shadow_nl[R + l] : 0u, hi = (part == 1) ? EB.shadow_nl[R + l] : EB.shadow_nl[l];
                if (e < hi - lo) { at = (uint32_t)l * EB.seg + lo + e; break; }
                e -= hi - lo;
            }
        }
        ShadowReq q = EB.shadow_q[at];
        // the request's index of its (event, light) entry gives the sample slot
        const DScene& S = src.trace_scene(src.find(ev_slot(EB, q.slot & 0x7FFFFFFFu)));
        if (!FULL && EB.cont_cap) {
            // first pass here; a ray that passes glass continues in
            // k_shadow_cont, in waves of such rays (queue full: here)
            DRay r;
            r.o = mk3(q.ox, q.oy, q.oz);
            r.d = mk3(q.dx, q.dy, q.dz);
            r.tmin = 0.001f;
            r.tmax = q.tmax;
            const int res = shadow_pass<COUNT, false>(S, r, st, passes, ts);
            const uint32_t k = wave_reserve(&C->cont_n[0], res == 2);
            if (res != 2) {
                EB.trans[q.slot & 0x7FFFFFFFu] = (res == 1) ? 1.0f : 0.0f;
            } else if (k < EB.cont_cap) {
                // rebuilt from the ray (the same values: only the origin moved)
                // so that only the slot stays live across the traversal
                ShadowReq c;
                c.ox = r.o.x; c.oy = r.o.y; c.oz = r.o.z; c.tmax = r.tmax;
                c.dx = r.d.x; c.dy = r.d.y; c.dz = r.d.z; c.slot = q.slot;
                EB.cont[0][k] = c;
            } else {
                EB.trans[q.slot & 0x7FFFFFFFu] = shadow_rest<COUNT, false>(S, r, r, q.slot, 1, st, passes, ts);
            }
            continue;
        }
        float tr = shadow_eval<COUNT, FULL>(S, q, st, passes, ts);
        EB.trans[q.slot & 0x7FFFFFFFu] = tr;
    }
    wave_add64(C->shadow_passes, passes);
    if (COUNT) {
        wave_add64(C->snodes, ts.nodes);
        wave_add64(C->sprims, ts.prims);
        wave_add64(C->sspills, st.spills);
        wave_add64(C->switers, ts.witers);
    }
}

// Second (stage 0) and third (stage 1) shadow pass of the requests whose
// earlier pass went through glass, queued by k_shadow / the stage before
// (spheres, triangles and planes only).  Tracing them in waves of their own
// keeps the first-pass waves from idling while a few of their lanes cross the
// glass; each request still sees the same passes from the same origins.
template <bool COUNT, class SRC>
__global__ TRACE_BOUNDS void k_shadow_cont(SRC src, EventBufs EB, Counters* C, int stage) {
    SCENE_TABLE(src);
    SHEEN_STACK(st);
    TraceStats ts{0, 0, 0};
    const uint32_t n = min(C->cont_n[stage], EB.cont_cap);
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t passes = 0;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        const ShadowReq q = EB.cont[stage][j];
        const uint32_t slot = q.slot & 0x7FFFFFFFu;
        const DScene& S = src.trace_scene(src.find(ev_slot(EB, slot)));
        DRay r;
        r.o = mk3(q.ox, q.oy, q.oz);
        r.d = mk3(q.dx, q.dy, q.dz);
        r.tmin = 0.001f;
        r.tmax = q.tmax;
        const int res = shadow_pass<COUNT, false>(S, r, st, passes, ts);
        const uint32_t k = wave_reserve(&C->cont_n[1], stage == 0 && res == 2);
        if (res != 2) {
            EB.trans[slot] = (res == 1) ? 1.0f : 0.0f;
        } else if (stage == 1) {
            EB.trans[slot] = 1.0f;  // three glass passes: unblocked
        } else if (k < EB.cont_cap) {
            ShadowReq c;  // as in k_shadow
            c.ox = r.o.x; c.oy = r.o.y; c.oz = r.o.z; c.tmax = r.tmax;
            c.dx = r.d.x; c.dy = r.d.y; c.dz = r.d.z; c.slot = q.slot;
            EB.cont[1][k] = c;
        } else {
            EB.trans[slot] = shadow_rest<COUNT, false>(S, r, r, q.slot, 2, st, passes, ts);
        }
    }
    wave_add64(C->shadow_passes, passes);
    if (COUNT) {
        wave_add64(C->cnodes, ts.nodes);
        wave_add64(C->cprims, ts.prims);
        wave_add64(C->sspills, st.spills);
    }
}

// ---------------------------------------------------------------------------
// Secondary-ray sorting.  Before a round's trace, the queue is reordered by a
// bin key so that a wave holds rays of one kind (reflection or refraction)
// travelling into one cell of the direction sphere: mode 1, the 8 octants
// (16 bins); mode 2, also the dominant axis (24 cells, 64 bins, 48 used).
// The order inside a bin is the queue's own (stable), which is tile-major,
// so a wave's rays also start close together.  Every entry carries its slot,
// so the order changes no result.  Three passes over tiles of SORT_TILE
// entries: per-tile bin counts (k_sec_hist), an exclusive scan of each bin's
// counts over the tiles with the bin totals (k_sec_scan, one block per bin),
// and a stable scatter that ranks each tile's entries with wave ballots
// (k_sec_scatter).  The queue length is read on the device; the grids are
// sized on the queue's capacity and loop over the tiles in use.
#define SORT_BLOCK 256
#define SORT_ITEMS 4
#define SORT_TILE (SORT_BLOCK * SORT_ITEMS)
#define SORT_BINS SORT_BINS_MAX
#define SORT_BIN_BITS 6

SHD uint32_t sec_sort_key(const SecRay& e, int mode) {
    const uint32_t kind = (e.meta >> 28) & 1u;
    const uint32_t oct = (e.dx < 0.0f ? 4u : 0u) | (e.dy < 0.0f ? 2u : 0u) | (e.dz < 0.0f ? 1u : 0u);
    if (mode == 1) return (kind << 3) | oct;
    const float ax = fabsf(e.dx), ay = fabsf(e.dy), az = fabsf(e.dz);
    const uint32_t major = (ax >= ay && ax >= az) ? 0u : (ay >= az ? 1u : 2u);
    return (kind << 5) | (major << 3) | oct;
}
// Tiles in use for a queue of n entries.
SHD uint32_t sort_tiles(uint32_t n) { return (n + SORT_TILE - 1) / SORT_TILE; }

// hist[bin * max_tiles + tile] = entries of `tile` in `bin`.
__global__ __launch_bounds__(SORT_BLOCK) void k_sec_hist(const SecRay* q, const Counters* C, int which, int mode,
                                                         uint32_t* hist, uint32_t max_tiles) {
    __shared__ uint32_t h[SORT_BINS];
    const uint32_t n = C->sec_n[which];
    const uint32_t T = sort_tiles(n);
    for (uint32_t tile = blockIdx.x; tile < T; tile += gridDim.x) {
        for (uint32_t b = threadIdx.x; b < SORT_BINS; b += blockDim.x) h[b] = 0u;
        __syncthreads();
        for (int k = 0; k < SORT_ITEMS; k++) {
            const uint32_t j = tile * SORT_TILE + (uint32_t)k * SORT_BLOCK + threadIdx.x;
            if (j < n) atomicAdd(&h[sec_sort_key(q[j], mode)], 1u);
        }
        __syncthreads();
        for (uint32_t b = threadIdx.x; b < SORT_BINS; b += blockDim.x) hist[b * max_tiles + tile] = h[b];
        __syncthreads();
    }
}

// Block b: exclusive scan of bin b's tile counts in place; the bin's total
// to tot[b].
__global__ __launch_bounds__(1024) void k_sec_scan(const Counters* C, int which, uint32_t* hist, uint32_t max_tiles,
                                                   uint32_t* tot) {
    __shared__ uint32_t part[1024 / 64];
    __shared__ uint32_t carry;
    const uint32_t T = sort_tiles(C->sec_n[which]);
    uint32_t* row = hist + (size_t)blockIdx.x * max_tiles;
    if (threadIdx.x == 0) carry = 0u;
    __syncthreads();
    for (uint32_t t0 = 0; t0 < T; t0 += blockDim.x) {
        const uint32_t t = t0 + threadIdx.x;
        const uint32_t v = t < T ? row[t] : 0u;
        // inclusive scan within the wave, then across the block's waves
        uint32_t x = v;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if ((int)__lane_id() >= off) x += y;
        }
        const uint32_t w = threadIdx.x >> 6;
        if (__lane_id() == 63u) part[w] = x;
        __syncthreads();
        uint32_t before = carry;
        for (uint32_t k = 0; k < w; k++) before += part[k];
        if (t < T) row[t] = before + x - v;
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1) carry = before + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

// Stable scatter of the queue into `out` in bin order.
__global__ __launch_bounds__(SORT_BLOCK) void k_sec_scatter(const SecRay* q, const Counters* C, int which, int mode,
                                                            const uint32_t* hist, uint32_t max_tiles,
                                                            const uint32_t* tot, SecRay* out) {
    __shared__ uint32_t start[SORT_BINS];      // bin start + this tile's offset in the bin
    __shared__ uint32_t wc[SORT_BLOCK / 64][SORT_BINS];
    __shared__ uint32_t pre[SORT_BLOCK / 64][SORT_BINS];
    const uint32_t n = C->sec_n[which];
    const uint32_t T = sort_tiles(n);
    const uint32_t wave = threadIdx.x >> 6;
    const uint64_t lt = (1ull << __lane_id()) - 1ull;
    for (uint32_t tile = blockIdx.x; tile < T; tile += gridDim.x) {
        static_assert(SORT_BINS == 64, "one wave scans the bin totals");
        if (threadIdx.x < SORT_BINS) {  // wave 0: exclusive scan of the totals, plus this tile's offsets
            const uint32_t t = tot[threadIdx.x];
            uint32_t x = t;
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t y = __shfl_up(x, off);
                if ((int)__lane_id() >= off) x += y;
            }
            start[threadIdx.x] = x - t + hist[(size_t)threadIdx.x * max_tiles + tile];
        }
        for (int k = 0; k < SORT_ITEMS; k++) {
            const uint32_t j = tile * SORT_TILE + (uint32_t)k * SORT_BLOCK + threadIdx.x;
            const bool valid = j < n;
            SecRay e;
            uint32_t key = SORT_BINS;  // invalid lanes match no valid lane's key
            if (valid) {
                e = q[j];
                key = sec_sort_key(e, mode);
            }
            // lanes of the wave with the same key, by ballots over its bits
            uint64_t same = __ballot(valid);
            if (!valid) same = ~same;
            for (int bit = 0; bit < SORT_BIN_BITS; bit++) {
                const uint64_t ones = __ballot((key >> bit) & 1u);
                same &= ((key >> bit) & 1u) ? ones : ~ones;
            }
            for (uint32_t b = threadIdx.x; b < (SORT_BLOCK / 64) * SORT_BINS; b += blockDim.x) (&wc[0][0])[b] = 0u;
            __syncthreads();
            if (valid && (same & lt) == 0ull) wc[wave][key] = (uint32_t)__popcll(same);
            __syncthreads();
            for (uint32_t b = threadIdx.x; b < SORT_BINS; b += blockDim.x) {
                uint32_t acc = start[b];
                for (int w = 0; w < SORT_BLOCK / 64; w++) {
                    pre[w][b] = acc;
                    acc += wc[w][b];
                }
                start[b] = acc;
            }
            __syncthreads();
            if (valid) out[pre[wave][key] + (uint32_t)__popcll(same & lt)] = e;
        }
        __syncthreads();
    }
}

// Per-pixel accumulation (reference render_kernel.cu:536-813 blend order).
template <class SRC>
__global__ __launch_bounds__(256) void k_accumulate(SRC src, SampleBufs SB, EventBufs EB, float4* frame) {
    SCENE_TABLE(src);
    const uint32_t n_pix = src.pix_count();
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n_pix; p += stride) {
        uint32_t i0;
        if (!src.pix_first(p, i0)) continue;
        const int sid = src.find(i0);
        const DScene& S = src.scene(sid);
        const FrameDesc& F = src.frame(sid);
        int x = 0, y = 0, s0 = 0;
        if (!slot_pixel(F, src.local(sid, i0), x, y, s0)) continue;
        v3 acc = mk3(0, 0, 0);
        for (int s = 0; s < F.ns; s++) {
            const uint32_t i = i0 + (uint32_t)s * 64u;
            const float4 sp = SB.samp[i];
            const int flags = __float_as_int(sp.w);
            v3 c = event_value(S, ev_of(EB, i, 0), EB);
            for (int k = 0; k < SB.n_beam && SB.beam_add; k++) {
                float4 b = SB.beam_add[(size_t)i * SB.n_beam + k];
                c = c + mk3(b.x, b.y, b.z);
            }
            if (SB.media_p) {
                float4 m = SB.media_p[i];
                c = c * m.w + mk3(m.x, m.y, m.z);
            }
            if (flags & SF_REFL) {
                v3 rc = event_value(S, ev_of(EB, i, 1), EB);
                if (SB.media_r) {
                    float4 m = SB.media_r[i];
                    rc = rc * m.w + mk3(m.x, m.y, m.z);
                }
                const float w = sp.x;
                c = c * (1.0f - w) + rc * w;
            }
            if (flags & SF_REFR) {
                v3 through = mk3(0, 0, 0);
                if (flags & SF_EV2) through = through + sp.z * event_value(S, ev_of(EB, i, 2), EB);
                c = c * (1.0f - fminf(sp.y, 1.0f)) + through;
            }
            acc = acc + c;
        }
        v3 f = acc * (1.0f / (float)F.ns);
        frame[(size_t)F.out0 + (size_t)y * F.width + x] = make_float4(f.x, f.y, f.z, 1.0f);
    }
}

// Marks the end of the primary shading's shadow requests in each run of the queue.
__global__ void k_snap_shadow(uint32_t* shadow_nl, int runs) {
    for (int l = threadIdx.x; l < runs; l += blockDim.x) shadow_nl[runs + l] = shadow_nl[l];
}
__global__ void k_reset(Counters* C, uint32_t* shadow_nl, int runs) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        for (int k = 0; k < 4; k++) C->sec_n[k] = 0;
        C->cont_n[0] = C->cont_n[1] = 0;
    }
    for (int l = threadIdx.x; l < runs; l += blockDim.x) shadow_nl[l] = 0;
}
// Every counter of a frame to zero (a kernel, so that captured graphs hold no
// memset nodes).
__global__ void k_zero_counters(Counters* C) {
    uint4* w = reinterpret_cast<uint4*>(C);
    static_assert(sizeof(Counters) % 16 == 0, "counters clear in 16-byte words");
    for (unsigned int i = threadIdx.x; i < sizeof(Counters) / 16u; i += blockDim.x) w[i] = make_uint4(0u, 0u, 0u, 0u);
}
__global__ void k_clear_sec(Counters* C, int which) {
    if (threadIdx.x == 0 && blockIdx.x == 0) C->sec_n[which] = 0;
}

// Primary hits of a frame chunk for parity checks: kind/index/t per sample.
__global__ void k_export_hits(FrameDesc F, const int4* hit_p, int* kind, int* index, float* t) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < F.n_slots; i += stride) {
        int x, y, s;
        if (!slot_pixel(F, F.g0 + i, x, y, s)) continue;
        const size_t o = ((size_t)y * F.width + x) * F.ns + s;
        const int4 h = hit_p[i];
        kind[o] = h.w;
        index[o] = h.w < 0 ? -1 : h.w == C_CSG ? 0 : h.y;  // a CSG hit reports index 0 (its record holds the leaf)
        t[o] = h.w < 0 ? 0.0f : __int_as_float(h.x);
    }
}
