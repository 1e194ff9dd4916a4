// dev_adjoint.h — differentiable shading of primary hits.
//
// Parameters: per-material albedo (diffuse_color) and emitted radiance
// (emission_color * emission), per-light intensity, and the level-0 texels of
// every texture as floats.  The forward is the ordinary wavefront render with
// those values substituted (float texel chains whose mips are 2x2 means).
// The adjoint replays the frame (primary trace, shading, shadow rays with the
// same RNG streams) and differentiates the radiance of each primary hit with
// visibility held fixed: for fixed geometry and shadow transmittance the
// pixel is linear in every parameter group.  Per sample (render_kernel.cu:
// 26-298 restricted to surfaces without reflection, refraction, hair and
// caustics):
//   c = base * ambient + E + sum_l tr_l * (base * rad_l * diff_l + rad_l * spec * s_l)
// with base = albedo or the bilinear texture tap, rad_l = intensity_l * q_l.
#pragma once
#include "kernels.h"

struct AdjGrads {
    float* albedo;     // [materials * 3]
    float* emission;   // [materials * 3]
    float* intensity;  // [lights]
    float* tex;        // gradient of every float texel chain (tex_f layout)
};

// Writes a parameter set into the device materials and lights.
__global__ void k_adj_set(Material* mats, int M, Light* lights, int L, const float* albedo, const float* emission,
                          const float* intensity) {
    const int stride = gridDim.x * blockDim.x;
    const int n = M > L ? M : L;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (i < M) {
            mats[i].diffuse_color = {albedo[3 * i], albedo[3 * i + 1], albedo[3 * i + 2]};
            mats[i].emission_color = {emission[3 * i], emission[3 * i + 1], emission[3 * i + 2]};
            mats[i].emission = 1.0f;
        }
        if (i < L) lights[i].intensity = intensity[i];
    }
}

// The scene's own parameter values.
__global__ void k_adj_get(const Material* mats, int M, const Light* lights, int L, float* albedo, float* emission,
                          float* intensity) {
    const int stride = gridDim.x * blockDim.x;
    const int n = M > L ? M : L;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (i < M) {
            const Material& m = mats[i];
            albedo[3 * i] = m.diffuse_color.x; albedo[3 * i + 1] = m.diffuse_color.y; albedo[3 * i + 2] = m.diffuse_color.z;
            emission[3 * i] = m.emission_color.x * m.emission;
            emission[3 * i + 1] = m.emission_color.y * m.emission;
            emission[3 * i + 2] = m.emission_color.z * m.emission;
        }
        if (i < L) intensity[i] = lights[i].intensity;
    }
}

__global__ void k_bytes_to_float(const unsigned char* px, uint32_t n, float* out) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = px[i] / 255.0f;
}

__global__ void k_fill_f(float* p, size_t n, float v) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// One float mip level from the one below (the 8-bit chain's 2x2 rule with
// clamped odd edges, as a mean instead of rounded integers):
// parent = ((a + b) + (c + d)) * 0.25.
__global__ void k_mip_f(float* chain, uint32_t src, int sw, int sh, uint32_t dst, int nw, int nh) {
    const int n = nw * nh;
    const int stride = gridDim.x * blockDim.x;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        const int y = p / nw, x = p - y * nw;
        const int ya = min(2 * y, sh - 1), yb = min(2 * y + 1, sh - 1);
        const int xa = min(2 * x, sw - 1), xb = min(2 * x + 1, sw - 1);
        for (int c = 0; c < 3; c++) {
            const float a = chain[src + (uint32_t)(ya * sw + xa) * 3u + c], b = chain[src + (uint32_t)(ya * sw + xb) * 3u + c];
            const float d = chain[src + (uint32_t)(yb * sw + xa) * 3u + c], e = chain[src + (uint32_t)(yb * sw + xb) * 3u + c];
            chain[dst + (uint32_t)p * 3u + c] = ((a + b) + (d + e)) * 0.25f;
        }
    }
}

// Adjoint of k_mip_f: each parent hands a quarter of its gradient to each of
// its four (clamped) children.  For a level at least 2 wide and tall every
// child has one parent; a 1-wide level repeats its column (the child gets
// both quarters), so plain adds are race free.
__global__ void k_mip_adj(float* g, uint32_t src, int sw, int sh, uint32_t dst, int nw, int nh) {
    const int n = nw * nh;
    const int stride = gridDim.x * blockDim.x;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        const int y = p / nw, x = p - y * nw;
        const int ya = min(2 * y, sh - 1), yb = min(2 * y + 1, sh - 1);
        const int xa = min(2 * x, sw - 1), xb = min(2 * x + 1, sw - 1);
        for (int c = 0; c < 3; c++) {
            const float q = g[dst + (uint32_t)p * 3u + c] * 0.25f;
            g[src + (uint32_t)(ya * sw + xa) * 3u + c] += q;
            g[src + (uint32_t)(ya * sw + xb) * 3u + c] += q;
            g[src + (uint32_t)(yb * sw + xa) * 3u + c] += q;
            g[src + (uint32_t)(yb * sw + xb) * 3u + c] += q;
        }
    }
}

SHD float wave_sum_f(float v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Per-wave accumulator table: lane j holds entry j (a key and a 3-float sum)
// in registers.  Wave-reduced sums for a key land in its entry, so a key that
// every wave meets in every grid-stride iteration (the ground material, a
// light) takes one atomic per wave at the end instead of one per iteration.
struct WaveTable {
    uint32_t key;  // 0xFFFFFFFF = free
    float x, y, z;
};
SHD void table_init(WaveTable& t) { t.key = 0xFFFFFFFFu; t.x = t.y = t.z = 0.0f; }

// Adds a wave-uniform (k, x, y, z) to the table; a full table hands it to
// dst[k .. k+2] directly.
SHD void table_add(WaveTable& t, float* dst, uint32_t k, float x, float y, float z) {
    const int lane = (int)__lane_id();
    const uint64_t hit = __ballot(t.key == k);
    int slot = -1;
    if (hit) slot = __ffsll((unsigned long long)hit) - 1;
    else {
        const uint64_t free_ = __ballot(t.key == 0xFFFFFFFFu);
        if (free_) slot = __ffsll((unsigned long long)free_) - 1;
    }
    if (slot >= 0) {
        if (lane == slot) { t.key = k; t.x += x; t.y += y; t.z += z; }
    } else if (lane == 0) {
        if (x != 0.0f) unsafeAtomicAdd(dst + k, x);
        if (y != 0.0f) unsafeAtomicAdd(dst + k + 1, y);
        if (z != 0.0f) unsafeAtomicAdd(dst + k + 2, z);
    }
}
SHD void table_flush(const WaveTable& t, float* dst) {
    if (t.key == 0xFFFFFFFFu) return;
    if (t.x != 0.0f) unsafeAtomicAdd(dst + t.key, t.x);
    if (t.y != 0.0f) unsafeAtomicAdd(dst + t.key + 1, t.y);
    if (t.z != 0.0f) unsafeAtomicAdd(dst + t.key + 2, t.z);
}

// Adds val (3 floats) at dst[e .. e+2] for active lanes.  Lanes sharing an
// address are summed first (all distinct addresses when `exhaustive`, else
// the first four, the rest adding directly: addresses met by one lane see
// little contention) and the sums go through the wave's table.  Every lane
// of the wave must call it.
SHD void wave_add(WaveTable& t, float* dst, bool active, uint32_t e, v3 val, bool exhaustive) {
    uint64_t pending = __ballot(active);
    for (int round = 0; pending && (exhaustive || round < 4); round++) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t k = __shfl(e, leader);
        const bool mine = ((pending >> __lane_id()) & 1ull) && e == k;
        const float x = wave_sum_f(mine ? val.x : 0.0f);
        const float y = wave_sum_f(mine ? val.y : 0.0f);
        const float z = wave_sum_f(mine ? val.z : 0.0f);
        table_add(t, dst, k, x, y, z);
        pending &= ~__ballot(mine);
    }
    if ((pending >> __lane_id()) & 1ull) {
        if (val.x != 0.0f) unsafeAtomicAdd(dst + e, val.x);
        if (val.y != 0.0f) unsafeAtomicAdd(dst + e + 1, val.y);
        if (val.z != 0.0f) unsafeAtomicAdd(dst + e + 2, val.z);
    }
}

// Texel gradients of one 8x8 tile, summed in LDS before they go out: an
// open-addressed table of ADJ_TEX_SLOTS element keys (the chain offset of a
// texel's first channel) with three float sums each.  Every sample of a tile
// taps a small, overlapping patch of texels (the bilinear 2x2 of each of the
// tile's samples), so a texel that tens of samples touch takes one
// device-scope atomic per channel instead of one per sample.
#define ADJ_TEX_BITS 10
#define ADJ_TEX_SLOTS (1 << ADJ_TEX_BITS)
#define ADJ_TEX_PROBES 32
struct TexSums {
    uint32_t* key;  // 0xFFFFFFFF = free
    float* x;
    float* y;
    float* z;
};
SHD void tex_sums_add(const TexSums& T, float* dst, bool active, uint32_t e, v3 v) {
    if (!active || (v.x == 0.0f && v.y == 0.0f && v.z == 0.0f)) return;
    uint32_t h = (e * 2654435761u) >> (32 - ADJ_TEX_BITS);
    for (int p = 0; p < ADJ_TEX_PROBES; p++) {
        const uint32_t old = atomicCAS(&T.key[h], 0xFFFFFFFFu, e);
        if (old == 0xFFFFFFFFu || old == e) {
            atomicAdd(&T.x[h], v.x);
            atomicAdd(&T.y[h], v.y);
            atomicAdd(&T.z[h], v.z);
            return;
        }
        h = (h + 1u) & (ADJ_TEX_SLOTS - 1u);
    }
    // table crowded: straight to the gradient buffer
    if (v.x != 0.0f) unsafeAtomicAdd(dst + e, v.x);
    if (v.y != 0.0f) unsafeAtomicAdd(dst + e + 1, v.y);
    if (v.z != 0.0f) unsafeAtomicAdd(dst + e + 2, v.z);
}
// Every entry out to dst and the table cleared; block-wide (barriers inside).
SHD void tex_sums_flush(const TexSums& T, float* dst) {
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < ADJ_TEX_SLOTS; j += blockDim.x) {
        const uint32_t k = T.key[j];
        if (k == 0xFFFFFFFFu) continue;
        if (T.x[j] != 0.0f) unsafeAtomicAdd(dst + k, T.x[j]);
        if (T.y[j] != 0.0f) unsafeAtomicAdd(dst + k + 1, T.y[j]);
        if (T.z[j] != 0.0f) unsafeAtomicAdd(dst + k + 2, T.z[j]);
        T.key[j] = 0xFFFFFFFFu;
        T.x[j] = T.y[j] = T.z[j] = 0.0f;
    }
    __syncthreads();
}

// Gradient of sum_pixels <grad_img, pixel> for one chunk of primary samples.
// Requires the chunk's primary hits (k_primary) and transmittances
// (k_shade_primary + k_shadow) of the replayed frame.  A block takes whole
// 8x8 tiles (the tile's 64 * ns slots are contiguous), so that its texel
// table sums every sample of the tile before it is flushed.
__global__ ADJOINT_BOUNDS void k_adjoint(DScene S, FrameDesc F, const int4* hit_p, EventBufs EB, const float* grad_img,
                                       AdjGrads G) {
    __shared__ uint32_t s_key[ADJ_TEX_SLOTS];
    __shared__ float s_x[ADJ_TEX_SLOTS], s_y[ADJ_TEX_SLOTS], s_z[ADJ_TEX_SLOTS];
    const TexSums T{s_key, s_x, s_y, s_z};
    for (uint32_t j = threadIdx.x; j < ADJ_TEX_SLOTS; j += blockDim.x) {
        s_key[j] = 0xFFFFFFFFu;
        s_x[j] = s_y[j] = s_z[j] = 0.0f;
    }
    __syncthreads();
    const uint32_t tile_slots = 64u * (uint32_t)F.ns;
    const uint32_t n_tiles = F.n_slots / tile_slots;
    // whole waves (the block) reach every ballot: a uniform count per tile
    const uint32_t per_tile = (tile_slots + blockDim.x - 1) / blockDim.x;
    const int L = S.num_lights;
    const float inv_ns = 1.0f / (float)F.ns;
    WaveTable t_em, t_alb, t_int;
    table_init(t_em); table_init(t_alb); table_init(t_int);
    for (uint32_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    for (uint32_t it = 0; it < per_tile; it++) {
        const uint32_t o = it * blockDim.x + threadIdx.x;
        const uint32_t i = tile * tile_slots + o;
        int x = 0, y = 0, s = 0;
        const bool valid = o < tile_slots && slot_pixel(F, F.g0 + i, x, y, s);
        float pa = 0.0f;
        uint32_t seed = 0;
        DRay r;
        QHit q;
        q.kind = -1;
        v3 g = mk3(0, 0, 0);
        if (valid) {
            r = camera_ray(S.cam, x, y, s, F.ns, F.width, F.height, pa, seed);
            q = unpack_hit(hit_p[i]);
            const float* gp = grad_img + ((size_t)y * F.width + x) * 4;
            g = mk3(gp[0], gp[1], gp[2]) * inv_ns;
        }
        const bool hit = valid && q.kind >= 0;
        DHit rec;
        Material mat;
        ShadeIn in;
        float shin = 0.0f;
        v3 base = mk3(0, 0, 0), dbase = mk3(0, 0, 0);
        if (hit) {
            rebuild_hit<false>(S, r, q, rec);
            shade_setup(S, rec, -r.d, rec.t * pa, mat, in, shin);
            base = albedo_of(S, mat, in);
            dbase = mk3(mat.ambient, mat.ambient, mat.ambient);
        }
        const uint32_t ev = ev_of(EB, i, 0);
        for (int l = 0; l < L; l++) {
            float dI = 0.0f;
            if (hit) {
                const Light& light = S.lights[l];
                const LightSample u = light_sample(S, light, 1.0f, seed ^ (uint32_t)(uint32_t(l) * 2654435761u), rec.point);
                const float tr = u.skip ? 0.0f : EB.trans[ev_light(EB, ev, l)];
                if (tr != 0.0f) {
                    const float diff = fmaxf(dot(in.normal, u.Ld), 0.0f);
                    float sp = 0.0f;
                    if (mat.specular > 0.0f && diff > 0.0f)
                        sp = pow_r(fmaxf(dot(in.normal, normalize(u.Ld + in.view)), 0.0f), shin);
                    dbase = dbase + u.rad * (light.intensity * diff * tr);
                    const v3 dc = (base * u.rad * diff + u.rad * (mat.specular * sp)) * tr;
                    dI = dot(g, dc);
                }
            }
            dI = wave_sum_f(dI);
            table_add(t_int, G.intensity, (uint32_t)l, dI, 0.0f, 0.0f);
        }
        const bool textured = hit && tex_ok(S, mat.texture_id);
        if (textured) {
            const TextureRef& t = S.textures[mat.texture_id];
            const TexTap p = tex_tap(t, in.uvx, in.uvy, tex_lod(t, in.uv_fp));
            const v3 gb = g * dbase;
            const uint32_t base_e = S.tex_f_off[mat.texture_id];  // the texture's chain offset: keys are absolute
            tex_sums_add(T, G.tex, true, base_e + p.e00, gb * ((1.0f - p.tx) * (1.0f - p.ty)));
            tex_sums_add(T, G.tex, true, base_e + p.e10, gb * (p.tx * (1.0f - p.ty)));
            tex_sums_add(T, G.tex, true, base_e + p.e01, gb * ((1.0f - p.tx) * p.ty));
            tex_sums_add(T, G.tex, true, base_e + p.e11, gb * (p.tx * p.ty));
        }
        const uint32_t mk = hit ? 3u * (uint32_t)rec.material_id : 0u;
        wave_add(t_em, G.emission, hit, mk, g, false);
        wave_add(t_alb, G.albedo, hit && !textured, mk, g * dbase, false);
    }
        tex_sums_flush(T, G.tex);
    }
    table_flush(t_em, G.emission);
    table_flush(t_alb, G.albedo);
    table_flush(t_int, G.intensity);
}
