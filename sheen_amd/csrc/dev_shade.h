// dev_shade.h — shading of the HIP path (reference integrator/shading.h and
// render_kernel.cu:13-298), split for the wavefront: a shading event writes
// its unshadowed terms and one shadow-ray request per light; the shadow
// kernel later supplies each request's transmittance (0 when blocked), and
// the accumulate kernel sums them in the reference's order.
#pragma once
#include "dev_trace.h"

SHD bool tex_ok(const DScene& S, int id) { return id >= 0 && id < S.num_textures && S.textures[id].pixels; }
SHD float tex_lod(const TextureRef& t, float uv_fp) { return log2_r(fmaxf(uv_fp * (float)t.width, 1.0f)); }

// Bilinear tap of a texture's mip chain (reference integrator/shading.h:26-58):
// nearest level to lod, v flipped, coordinates wrapped.
struct TexTap {
    uint32_t e00, e10, e01, e11;  // element offsets (texel * 3) in the chain
    float tx, ty;
};
SHD TexTap tex_tap(const TextureRef& t, float ux, float uy, float lod) {
    int level = (int)floorf(lod + 0.5f);
    if (level < 0) level = 0;
    if (level >= t.mips) level = t.mips - 1;
    int w = t.width, h = t.height;
    uint32_t off = 0;
    for (int i = 0; i < level; i++) {
        off += (uint32_t)w * h * 3u;
        w = (w > 1) ? w / 2 : 1;
        h = (h > 1) ? h / 2 : 1;
    }
    float u = ux - floorf(ux);
    float v = 1.0f - (uy - floorf(uy));
    float fx = u * (float)(w - 1);
    float fy = v * (float)(h - 1);
    int x0 = (int)fx, y0 = (int)fy;
    int x1 = (x0 + 1) % w;
    int y1 = (y0 + 1) % h;
    TexTap p;
    p.tx = fx - (float)x0;
    p.ty = fy - (float)y0;
    p.e00 = off + (uint32_t)(y0 * w + x0) * 3u;
    p.e10 = off + (uint32_t)(y0 * w + x1) * 3u;
    p.e01 = off + (uint32_t)(y1 * w + x0) * 3u;
    p.e11 = off + (uint32_t)(y1 * w + x1) * 3u;
    return p;
}
SHD v3 tex_blend(v3 c00, v3 c10, v3 c01, v3 c11, float tx, float ty) {
    v3 a = c00 * (1.0f - tx) + c10 * tx;
    v3 b = c01 * (1.0f - tx) + c11 * tx;
    return a * (1.0f - ty) + b * ty;
}
// Texture `id` (tex_ok) at uv and lod: 8-bit texels / 255, or the float chain.
SHD v3 tex_sample(const DScene& S, int id, float ux, float uy, float lod) {
    const TextureRef& t = S.textures[id];
    const TexTap p = tex_tap(t, ux, uy, lod);
    if (S.tex_f) {
        const float* f = S.tex_f + S.tex_f_off[id];
        auto at = [&](uint32_t e) { return mk3(f[e], f[e + 1], f[e + 2]); };
        return tex_blend(at(p.e00), at(p.e10), at(p.e01), at(p.e11), p.tx, p.ty);
    }
    auto at = [&](uint32_t e) {
        const unsigned char* q = t.pixels + e;
        return mk3(q[0] / 255.0f, q[1] / 255.0f, q[2] / 255.0f);
    };
    return tex_blend(at(p.e00), at(p.e10), at(p.e01), at(p.e11), p.tx, p.ty);
}

struct ShadeIn {
    v3 point, normal, tangent, view;
    float uvx, uvy, uv_fp;
};

SHD v3 albedo_of(const DScene& S, const Material& m, const ShadeIn& in) {
    if (tex_ok(S, m.texture_id)) {
        return tex_sample(S, m.texture_id, in.uvx, in.uvy, tex_lod(S.textures[m.texture_id], in.uv_fp));
    }
    return cv3(m.diffuse_color);
}
SHD float rough_to_shin(float r) {  // shading.h:85-89
    float rr = fmaxf(r, 0.02f);
    float r2 = rr * rr;
    return fminf(fmaxf(2.0f / (r2 * r2) - 2.0f, 1.0f), 4096.0f);
}
SHD float transmission_at(const DScene& S, float uvx, float uvy, const Material& m) {  // shading.h:92-99
    if (m.transmission <= 0.0f) return 0.0f;
    if (tex_ok(S, m.mask_id)) return m.transmission * tex_sample(S, m.mask_id, uvx, uvy, 0.0f).x;
    return m.transmission;
}
SHD float ior_at(const DScene& S, float uvx, float uvy, const Material& m) {  // shading.h:102-108
    if (tex_ok(S, m.mask_id)) return 1.0f + tex_sample(S, m.mask_id, uvx, uvy, 0.0f).y;
    return m.ior;
}
SHD v3 shade_hair(const Material& m, const ShadeIn& in, v3 L, v3 rad) {  // shading.h:112-152
    v3 T = in.tangent, N = in.normal, V = in.view;
    float tl = dot(T, L);
    float sin_tl = sqrtf(fmaxf(1.0f - tl * tl, 0.0f));
    float back = fmaxf(-tl, 0.0f);
    float diffuse = sin_tl + m.hair_backlight * back * back;
    v3 result = cv3(m.diffuse_color) * rad * diffuse * 0.75f;
    float w_r = m.hair_spec_r, w_t = m.hair_spec_t, e_r = m.hair_exp_r, e_t = m.hair_exp_t;
    if (w_r <= 0.0f && w_t <= 0.0f) { w_r = m.specular; e_r = m.shininess; }
    v3 H = normalize(L + V);
    if (w_r > 0.0f) {
        v3 Tr = normalize(T + m.hair_shift * N);
        float sr = dot(Tr, H);
        float rr = pow_r(sqrtf(fmaxf(1.0f - sr * sr, 0.0f)), fmaxf(e_r, 1.0f));
        result = result + rad * w_r * rr;
    }
    if (w_t > 0.0f) {
        v3 Tt = normalize(T - m.hair_shift * N);
        float st = dot(Tt, H);
        float rt = pow_r(sqrtf(fmaxf(1.0f - st * st, 0.0f)), fmaxf(e_t, 1.0f));
        result = result + rad * w_t * rt;
    }
    return result;
}
SHD v3 shade_direct(const DScene& S, const Material& m, const ShadeIn& in, v3 L, v3 rad, float shin) {  // shading.h:158-175
    if (m.anisotropy > 0.0f) return shade_hair(m, in, L, rad);
    v3 base = albedo_of(S, m, in);
    float diff = fmaxf(dot(in.normal, L), 0.0f);
    v3 result = base * rad * diff;
    if (m.specular > 0.0f && diff > 0.0f) {
        v3 hv = normalize(L + in.view);
        float spec = pow_r(fmaxf(dot(in.normal, hv), 0.0f), shin);
        result = result + rad * m.specular * spec;
    }
    return result;
}
SHD v3 sun_dir(const Light& light, uint32_t seed) {  // render_kernel.cu:13-24
    v3 base = -cv3(light.direction);
    if (light.angle <= 0.0f) return base;
    v3 up = (fabsf(base.y) < 0.9f) ? mk3(0, 1, 0) : mk3(1, 0, 0);
    v3 t1 = normalize(cross(base, up));
    v3 t2 = cross(base, t1);
    float r = sqrtf(rand01(seed ^ 0x68e31da4u)) * tan_r(light.angle);
    float phi = 6.2831853f * rand01(seed ^ 0xb5297a4du);
    return normalize(base + t1 * (cos_r(phi) * r) + t2 * (sin_r(phi) * r));
}
SHD v3 sky(v3 d) {  // render_kernel.cu:329-334
    float t = 0.5f * (d.y + 1.0f);
    return mk3((1.0f - t) * 1.0f + t * 0.5f, (1.0f - t) * 1.0f + t * 0.7f, (1.0f - t) * 1.0f + t * 1.0f);
}

// One light's sample for a shading point (reference render_kernel.cu:150-253):
// direction, radiance, shadow target and distance bound; skip = no
// contribution (back of an area light, outside a spot cone).  Radiance is
// linear in `intensity`, which the caller passes (the light's own, or 1 for
// the derivative).
struct LightSample {
    v3 Ld, rad, target;
    float t_max;
    bool skip;
};
SHD LightSample light_sample(const DScene& S, const Light& light, float intensity, uint32_t ls, v3 point) {
    v3 Ld, radiance, target;
    float t_max = 1e10f;
    bool skip = false;
    if (light.type == 0) {
        Ld = sun_dir(light, ls);
        radiance = cv3(light.color) * intensity;
        target = point + Ld * 1.0e6f;
    } else {
        v3 sp = cv3(light.position);
        if (light.type == 2) {
            float ru = rand01(ls) * 2.0f - 1.0f;
            float rv = rand01(ls ^ 0x9e3779b9u) * 2.0f - 1.0f;
            sp = cv3(light.position) + cv3(light.u_axis) * ru + cv3(light.v_axis) * rv;
        }
        v3 tl = sp - point;
        float dist2 = fmaxf(dot(tl, tl), 1.0e-4f);
        float dist = sqrtf(dist2);
        Ld = tl / dist;
        if ((light.type == 1 || light.type == 3) && light.radius > 0.0f) {
            v3 up = (fabsf(Ld.y) < 0.9f) ? mk3(0, 1, 0) : mk3(1, 0, 0);
            v3 t1 = normalize(cross(Ld, up));
            v3 t2 = cross(Ld, t1);
            float r = sqrtf(rand01(ls ^ 0x85ebca6bu)) * light.radius;
            float phi = 6.2831853f * rand01(ls ^ 0xc2b2ae35u);
            sp = sp + t1 * (cos_r(phi) * r) + t2 * (sin_r(phi) * r);
            tl = sp - point;
            dist2 = fmaxf(dot(tl, tl), 1.0e-4f);
            dist = sqrtf(dist2);
            Ld = tl / dist;
        }
        float atten = intensity / (12.5663706f * dist2);
        v3 gobo = mk3(1, 1, 1);
        if (light.type == 2) {
            v3 n = normalize(cross(cv3(light.u_axis), cv3(light.v_axis)));
            float cos_l = dot(n, Ld);
            if (cos_l <= 0.0f) skip = true;
            else atten *= cos_l;
        } else if (light.type == 3) {
            v3 ts = normalize(point - sp);
            float c = dot(ts, cv3(light.direction));
            float cos_half = cos_r(light.angle);
            if (c <= cos_half) skip = true;
            else {
                float blend = fminf(fmaxf(light.blend, 0.0f), 1.0f);
                float cos_soft = cos_r(light.angle * (1.0f - blend));
                if (cos_soft > cos_half + 1.0e-4f) atten *= fminf((c - cos_half) / (cos_soft - cos_half), 1.0f);
                if (tex_ok(S, light.texture_id)) {
                    v3 axis = cv3(light.direction);
                    v3 up = (fabsf(axis.y) < 0.9f) ? mk3(0, 1, 0) : mk3(1, 0, 0);
                    v3 t1 = normalize(cross(axis, up));
                    v3 t2 = cross(axis, t1);
                    float tan_half = tan_r(fmaxf(light.angle, 1.0e-3f));
                    float u = 0.5f + 0.5f * dot(ts, t1) / tan_half;
                    float v = 0.5f + 0.5f * dot(ts, t2) / tan_half;
                    gobo = tex_sample(S, light.texture_id, u, v, 0.0f);
                }
            }
        }
        radiance = cv3(light.color) * atten * gobo;
        target = sp;
        t_max = dist * 0.999f;
    }
    LightSample out;
    out.Ld = Ld; out.rad = radiance; out.target = target; out.t_max = t_max; out.skip = skip;
    return out;
}

// Photon hash grid (reference render_kernel.h:55-111).
SHD int photon_cell(int ix, int iy, int iz, int num_cells) {
    uint32_t h = (uint32_t(ix) * 73856093u) ^ (uint32_t(iy) * 19349663u) ^ (uint32_t(iz) * 83492791u);
    h = mix32(h);
    return (int)(h & (uint32_t)(num_cells - 1));
}
SHD v3 photon_gather(const PhotonMapView& m, v3 p, v3 n, float radius, int max_samples, int* count_out) {
    if (!m.photons || m.num_cells <= 0 || m.cell <= 0.0f) { if (count_out) *count_out = 0; return mk3(0, 0, 0); }
    int x0 = (int)floorf((p.x - radius) / m.cell), x1 = (int)floorf((p.x + radius) / m.cell);
    int y0 = (int)floorf((p.y - radius) / m.cell), y1 = (int)floorf((p.y + radius) / m.cell);
    int z0 = (int)floorf((p.z - radius) / m.cell), z1 = (int)floorf((p.z + radius) / m.cell);
    float r2 = radius * radius;
    v3 acc = mk3(0, 0, 0);
    int taken = 0;
    int cells[8];
    int ncell = 0;
    for (int iz = z0; iz <= z1; iz++)
        for (int iy = y0; iy <= y1; iy++)
            for (int ix = x0; ix <= x1; ix++) {
                int c = photon_cell(ix, iy, iz, m.num_cells);
                bool seen = false;
                for (int q = 0; q < ncell; q++) if (cells[q] == c) { seen = true; break; }
                if (!seen && ncell < 8) cells[ncell++] = c;
            }
    for (int q = 0; q < ncell && taken < max_samples; q++) {
        int c = cells[q];
        int start = m.cell_start[c], cnt = m.cell_count[c];
        for (int k = 0; k < cnt && taken < max_samples; k++) {
            const Photon& ph = m.photons[start + k];
            v3 d = cv3(ph.position) - p;
            if (dot(d, d) > r2) continue;
            if (dot(cv3(ph.normal), n) < 0.1f) continue;
            acc = acc + cv3(ph.flux);
            taken++;
        }
    }
    if (count_out) *count_out = taken;
    return acc;
}

// One wave reserves `want` slots of a queue with a single atomic.
SHD uint32_t wave_reserve(uint32_t* counter, bool want) {
    const uint64_t mask = __ballot(want);
    if (mask == 0) return 0xFFFFFFFFu;
    const int lane = (int)__lane_id();
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return want ? base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull)) : 0xFFFFFFFFu;
}

// Block-wide reservation of up to two queue entries per thread (flags a, b)
// with one atomic per block: device-scope atomics are performed past the
// per-XCD L2, so a per-wave atomic for every light and ray kind serialises the
// shading kernels.  Every thread of the block must make the call at the same
// point (block-uniform control flow).  A thread's entries are base (if a) and
// base + a (if b); returns 0xFFFFFFFF when it asks for none.
SHD uint32_t block_reserve(uint32_t* counter, bool a, bool b) {
    __shared__ uint32_t s_res[1024 / 64 + 1];  // up to 1024-thread blocks
    const uint64_t ma = __ballot(a), mb = __ballot(b);
    const int lane = (int)__lane_id();
    const int wave = (int)(threadIdx.x >> 6);
    const int nw = (int)(blockDim.x >> 6);
    const uint64_t lt = (1ull << lane) - 1ull;
    if (lane == 0) s_res[wave] = (uint32_t)(__popcll(ma) + __popcll(mb));
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < nw; w++) {
            const uint32_t c = s_res[w];
            s_res[w] = tot;
            tot += c;
        }
        s_res[nw] = tot ? atomicAdd(counter, tot) : 0u;
    }
    __syncthreads();
    const uint32_t base = s_res[nw] + s_res[wave] + (uint32_t)(__popcll(ma & lt) + __popcll(mb & lt));
    __syncthreads();  // s_res is reused by the next call
    return (a || b) ? base : 0xFFFFFFFFu;
}

// Block-wide reservation of up to two queue entries per thread, one of
// class A (flag a) and one of class B (flag b), with one device atomic per
// block: the block's entries form one contiguous range holding every class-A
// entry first, then every class-B entry, each class in thread order.  The
// per-(class, wave) counts are scanned by the first wave.  Every thread of
// the block must make the call at the same point.  ra / rb: the entries
// (0xFFFFFFFF where not asked for).
SHD void block_reserve_ab(uint32_t* counter, bool a, bool b, uint32_t& ra, uint32_t& rb) {
    __shared__ uint32_t s_cnt[2 * (1024 / 64)];  // up to 1024-thread blocks
    const uint64_t ma = __ballot(a), mb = __ballot(b);
    const int lane = (int)__lane_id();
    const int wave = (int)(threadIdx.x >> 6);
    const int nw = (int)(blockDim.x >> 6);
    const uint64_t lt = (1ull << lane) - 1ull;
    if (lane == 0) {
        s_cnt[wave] = (uint32_t)__popcll(ma);
        s_cnt[nw + wave] = (uint32_t)__popcll(mb);
    }
    __syncthreads();
    if (wave == 0) {
        const uint32_t own = lane < 2 * nw ? s_cnt[lane] : 0u;
        uint32_t v = own;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t t = __shfl_up(v, off);
            if (lane >= off) v += t;
        }
        const uint32_t tot = __shfl(v, 63);
        uint32_t base = 0;
        if (lane == 0 && tot) base = atomicAdd(counter, tot);
        base = __shfl(base, 0);
        if (lane < 2 * nw) s_cnt[lane] = base + v - own;
    }
    __syncthreads();
    ra = a ? s_cnt[wave] + (uint32_t)__popcll(ma & lt) : 0xFFFFFFFFu;
    rb = b ? s_cnt[nw + wave] + (uint32_t)__popcll(mb & lt) : 0xFFFFFFFFu;
    __syncthreads();  // s_cnt is reused by the next call
}

// Shadow-ray request: 32 bytes.
struct ShadowReq {
    float ox, oy, oz, tmax;
    float dx, dy, dz;
    uint32_t slot;  // event light slot; bit 31: strand optical depth needed
};

// Per-event outputs of the wavefront.  An event is one of a sample slot's
// three shading points (kind 0 primary, 1 reflection, 2 refraction), stored
// in kind-major planes of `cap` slots: event ev = kind * cap + slot, so the
// lanes of a wave (consecutive slots) touch consecutive entries.  Per-light
// entries are light-major planes of 3 * cap events: ev + l * 3 * cap; a
// shadow request carries that index.
struct EventBufs {
    float4* base;     // [3 cap]: unshadowed term (lit) or sky colour; w = event kind bits
    float4* caustic;  // [3 cap] or null
    float4* contrib;  // [L][3 cap]: shade_direct of each light
    float* trans;     // [L][3 cap]: transmittance, written by the shadow kernel
    ShadowReq* shadow_q;
    ShadowReq* cont[2];   // requests continuing past glass into their second / third pass
    uint32_t cont_cap;    // entries of each (0: every pass in the first shadow kernel)
    uint32_t* shadow_nl;  // [2 runs]: requests queued per run, then their count after the primary shading
    uint32_t seg;         // queue entries per run (3 cap): run r's requests at r * seg
    int runs;             // queue runs: L (light l's requests are run l) or 1 (one queue shared by the lights)
    int order;            // queue order (ORDER_* bits): thread order when clear
    int L;         // lights per event in contrib/trans (a batch: the most of any scene)
    uint32_t cap;  // sample slots of the work buffers (3 * cap * max(L, 1) < 2^31)
};
// EventBufs::order bits.  ORDER_SHADOW: a shading block's shadow requests of
// one light queue those leaving the surface first, then those entering it
// (dot(normal, direction) < 0: the ray crosses the object and traces
// longer), so that a wave of the shadow kernel holds rays of one kind.
// ORDER_SEC: a primary-shading block queues its reflection rays first, then
// its refraction rays.  Results are written per sample slot, so the order
// changes no image.
#define ORDER_SHADOW 1
#define ORDER_SEC 2
// SKIP_DARK: a shadow request whose unshadowed contribution is exactly zero
// (a light behind the shading normal of a non-hair material: max(n.l, 0) = 0
// and no specular without it) is not traced; the pixel adds 0 x
// transmittance either way, so the image is the same, but fewer rays are
// traced than the reference traces (off by default: the reference's ray
// counts are the unit of work).
#define ORDER_SKIP_DARK 4
SHD uint32_t ev_of(const EventBufs& B, uint32_t slot, uint32_t kind) { return kind * B.cap + slot; }
SHD uint32_t ev_light(const EventBufs& B, uint32_t ev, int l) { return ev + (uint32_t)l * 3u * B.cap; }
// sample slot of a per-light entry (a shadow request's index)
SHD uint32_t ev_slot(const EventBufs& B, uint32_t e) { return e % B.cap; }
// event kind bits (stored in base.w as int bits)
#define EV_NONE 0
#define EV_SKY 1
#define EV_LIT 2
#define EV_DEBUG 4      // material debug colours; light index in bits 8..
#define EV_CAUSTIC 8

// Shading inputs of a hit (reference render_kernel.cu:30-146): material,
// footprint and tangent frame of a triangle hit, normal map, shininess from
// the mask texture.
SHD void shade_setup(const DScene& S, const DHit& rec, v3 view_dir, float footprint, Material& mat, ShadeIn& in,
                     float& shininess) {
    mat = S.materials[rec.material_id];
    in.point = rec.point; in.normal = rec.normal; in.tangent = rec.tangent;
    in.uvx = rec.uvx; in.uvy = rec.uvy; in.view = view_dir;
    float uv_fp = footprint;
    v3 tan_u = mk3(0, 0, 0), tan_v = mk3(0, 0, 0);
    if (rec.object_id < 0 && S.num_triangles > 0) {  // render_kernel.cu:82-114
        int ti = -rec.object_id - 1;
        if (ti >= 0 && ti < S.num_triangles) {
            const Triangle& tr = S.triangles[ti];
            v3 e1 = cv3(tr.v1) - cv3(tr.v0), e2 = cv3(tr.v2) - cv3(tr.v0);
            float d1x = tr.uv1.x - tr.uv0.x, d1y = tr.uv1.y - tr.uv0.y;
            float d2x = tr.uv2.x - tr.uv0.x, d2y = tr.uv2.y - tr.uv0.y;
            float w1 = sqrtf(dot(e1, e1)), w2 = sqrtf(dot(e2, e2));
            float u1 = sqrtf(d1x * d1x + d1y * d1y);
            float u2 = sqrtf(d2x * d2x + d2y * d2y);
            float a1 = (w1 > 1.0e-6f) ? u1 / w1 : 0.0f;
            float a2 = (w2 > 1.0e-6f) ? u2 / w2 : 0.0f;
            float dens = (a1 > a2) ? a1 : a2;
            if (dens > 1.0e-6f) uv_fp = footprint * dens;
            float det = d1x * d2y - d2x * d1y;
            if (fabsf(det) > 1.0e-8f) {
                v3 du = (e1 * d2y - e2 * d1y) / det;
                v3 dv = (e2 * d1x - e1 * d2x) / det;
                v3 n0 = rec.normal;
                du = du - n0 * dot(du, n0);
                dv = dv - n0 * dot(dv, n0);
                float lu = dot(du, du);
                if (lu > 1.0e-12f) {
                    tan_u = du * (1.0f / sqrtf(lu));
                    dv = dv - tan_u * dot(dv, tan_u);
                    float lv = dot(dv, dv);
                    if (lv > 1.0e-12f) tan_v = dv * (1.0f / sqrtf(lv));
                }
            }
        }
    }
    in.uv_fp = uv_fp;
    if (tex_ok(S, mat.normal_id)) {  // render_kernel.cu:119-136
        const TextureRef& nt = S.textures[mat.normal_id];
        v3 ntex = tex_sample(S, mat.normal_id, in.uvx, in.uvy, tex_lod(nt, in.uv_fp));
        float nx = (ntex.x * 2.0f - 1.0f) * mat.normal_strength;
        float ny = (ntex.y * 2.0f - 1.0f) * mat.normal_strength;
        v3 n0 = in.normal, t1, t2;
        if (dot(tan_u, tan_u) > 0.5f && dot(tan_v, tan_v) > 0.5f) { t1 = tan_u; t2 = tan_v; }
        else {
            v3 up = (fabsf(n0.y) < 0.9f) ? mk3(0, 1, 0) : mk3(1, 0, 0);
            t1 = normalize(cross(up, n0));
            t2 = cross(n0, t1);
        }
        in.normal = normalize(n0 + t1 * nx + t2 * ny);
    }
    shininess = mat.shininess;
    if (tex_ok(S, mat.mask_id)) {
        const TextureRef& mt = S.textures[mat.mask_id];
        shininess = rough_to_shin(tex_sample(S, mat.mask_id, in.uvx, in.uvy, tex_lod(mt, in.uv_fp)).z);
    }
}

// compute_lighting (reference render_kernel.cu:26-298) for one event.
// Writes the unshadowed term, one contribution per light and queues one
// shadow request per light that reaches the shadow test.  `active` lanes
// only; every lane of the wave must call this (it reserves queue slots).
SHD void shade_event(const DScene& S, bool active, const DHit& rec, v3 view_dir, float footprint, uint32_t seed,
                     uint32_t ev, const EventBufs& B) {
    const int L = B.L;  // block-uniform trip count: the loop reserves queue slots
    Material mat;
    ShadeIn in;
    v3 base = mk3(0, 0, 0), result = mk3(0, 0, 0);
    float shininess = 0.0f;
    if (active) {
        shade_setup(S, rec, view_dir, footprint, mat, in, shininess);
        base = albedo_of(S, mat, in);
        result = base * mat.ambient + cv3(mat.emission_color) * mat.emission;
    }
    const bool debug = active && mat.shininess < 0.0f;
    const bool want_tau = active && mat.anisotropy <= 0.0f && S.count[C_STRAND] > 0;
    bool debug_taken = false;
    int debug_light = -1;
    for (int i = 0; i < L; i++) {
        bool emit = false;
        v3 contrib = mk3(0, 0, 0);
        v3 so = mk3(0, 0, 0), sd = mk3(0, 0, 1);
        float stmax = 0.0f;
        const bool has_light = active && i < S.num_lights;
        if (has_light && !(debug && debug_taken)) {
            const LightSample ls = light_sample(S, S.lights[i], S.lights[i].intensity, seed ^ (uint32_t)(uint32_t(i) * 2654435761u),
                                               rec.point);
            const v3 Ld = ls.Ld, radiance = ls.rad, target = ls.target;
            const float t_max = ls.t_max;
            const bool skip = ls.skip;
            if (!skip) {
                emit = true;
                so = rec.point + rec.normal * 0.002f;
                sd = normalize(target - so);
                stmax = t_max;
                if (debug) { debug_taken = true; debug_light = i; }
                else {
                    contrib = shade_direct(S, mat, in, Ld, radiance, shininess);
                    if ((B.order & ORDER_SKIP_DARK) && contrib.x == 0.0f && contrib.y == 0.0f && contrib.z == 0.0f)
                        emit = false;
                }
            }
        }
        if (has_light) {
            B.contrib[ev_light(B, ev, i)] = make_float4(contrib.x, contrib.y, contrib.z, 0.0f);
            B.trans[ev_light(B, ev, i)] = 0.0f;
        }
        const int run = B.runs > 1 ? i : 0;
        uint32_t q;
        if (B.order & ORDER_SHADOW) {
            const bool into = dot(rec.normal, sd) < 0.0f;
            uint32_t qa, qb;
            block_reserve_ab(B.shadow_nl + run, emit && !into, emit && into, qa, qb);
            q = (uint32_t)run * B.seg + (into ? qb : qa);
        } else {
            q = (uint32_t)run * B.seg + block_reserve(B.shadow_nl + run, emit, false);
        }
        if (emit) {
            ShadowReq r;
            r.ox = so.x; r.oy = so.y; r.oz = so.z; r.tmax = stmax;
            r.dx = sd.x; r.dy = sd.y; r.dz = sd.z;
            r.slot = ev_light(B, ev, i) | ((want_tau && !debug) ? 0x80000000u : 0u);
            B.shadow_q[q] = r;
        }
    }
    if (!active) return;
    int kind = EV_LIT;
    if (debug_taken) kind |= EV_DEBUG | (debug_light << 8);
    if (B.caustic) {
        v3 cterm = mk3(0, 0, 0);
        if (S.pm.photons && S.caustics > 0.0f && !debug_taken) {  // render_kernel.cu:288-296
            float r = S.pm.cell * 0.5f;
            v3 flux = photon_gather(S.pm, rec.point, rec.normal, r, 128, nullptr);
            cterm = base * flux * (S.caustics / (9.8696044f * r * r));
            kind |= EV_CAUSTIC;
        }
        B.caustic[ev] = make_float4(cterm.x, cterm.y, cterm.z, 0.0f);
    }
    B.base[ev] = make_float4(result.x, result.y, result.z, __int_as_float(kind));
}

SHD void sky_event(const DScene& S, v3 d, uint32_t ev, const EventBufs& B) {
    v3 c = sky(d);
    B.base[ev] = make_float4(c.x, c.y, c.z, __int_as_float(EV_SKY));
    (void)S;
}

// Value of an event in the reference's summation order.
SHD v3 event_value(const DScene& S, uint32_t ev, const EventBufs& B) {
    float4 b = B.base[ev];
    int kind = __float_as_int(b.w);
    v3 r = mk3(b.x, b.y, b.z);
    if (!(kind & EV_LIT)) return r;
    if (kind & EV_DEBUG) {
        int li = kind >> 8;
        return (B.trans[ev_light(B, ev, li)] > 0.0f) ? mk3(0, 1, 0) : mk3(1, 0, 0);
    }
    for (int i = 0; i < S.num_lights; i++) {
        float tr = B.trans[ev_light(B, ev, i)];
        if (tr != 0.0f) {
            float4 c = B.contrib[ev_light(B, ev, i)];
            r = r + mk3(c.x, c.y, c.z) * tr;
        }
    }
    if (kind & EV_CAUSTIC) {
        float4 c = B.caustic[ev];
        r = r + mk3(c.x, c.y, c.z);
    }
    return r;
}
