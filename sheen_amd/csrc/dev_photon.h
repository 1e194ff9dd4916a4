// dev_photon.h — photon emission and hash-grid map of the caustic path
// (reference render_kernel.cu:820-1045).  Emission is one thread per photon
// index; the map is deterministic: stored photons keep emission-index order
// (prefix sum instead of an atomic slot) and a stable radix sort by cell
// keeps that order inside each cell, where the reference's atomic scatter
// order varies from run to run.
#pragma once
#include "kernels.h"

struct PhotonCounters {
    unsigned long long rays;
};

__global__ __launch_bounds__(TRACE_BLOCK) void k_photon_emit(DScene S, int per_light, uint32_t total, Photon* out,
                                                             uint32_t* flag, PhotonCounters* pc) {
    SHEEN_STACK(st);
    TraceStats ts{0, 0};
    uint32_t rays = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        flag[idx] = 0;
        const Light& light = S.lights[idx / (uint32_t)per_light];
        if (light.type == 0) continue;  // suns emit nothing (render_kernel.cu:828)
        const uint32_t sd = mix32(0x5eed1234u + idx * 2654435761u);
        const float u1 = rand01(sd), u2 = rand01(sd ^ 0x9e3779b9u), u3 = rand01(sd ^ 0x85ebca6bu);
        v3 pos = cv3(light.position), dir;
        if (light.type == 3) {
            float cos_half = cos_r(light.angle);
            float ct = cos_half + (1.0f - cos_half) * u1;
            float stn = sqrtf(fmaxf(1.0f - ct * ct, 0.0f));
            float phi = 6.2831853f * u2;
            v3 ax = cv3(light.direction);
            v3 up = (fabsf(ax.y) < 0.9f) ? mk3(0, 1, 0) : mk3(1, 0, 0);
            v3 t1 = normalize(cross(ax, up));
            v3 t2 = cross(ax, t1);
            dir = normalize(ax * ct + t1 * (cos_r(phi) * stn) + t2 * (sin_r(phi) * stn));
        } else if (light.type == 2) {
            pos = cv3(light.position) + cv3(light.u_axis) * (u1 * 2.0f - 1.0f) + cv3(light.v_axis) * (u2 * 2.0f - 1.0f);
            v3 n = normalize(cross(cv3(light.u_axis), cv3(light.v_axis)));
            v3 up = (fabsf(n.y) < 0.9f) ? mk3(0, 1, 0) : mk3(1, 0, 0);
            v3 t1 = normalize(cross(n, up));
            v3 t2 = cross(n, t1);
            float cz = sqrtf(fmaxf(1.0f - u3, 0.0f));
            float r = sqrtf(u3);
            float phi = 6.2831853f * rand01(sd ^ 0xc2b2ae35u);
            dir = normalize(n * cz + t1 * (cos_r(phi) * r) + t2 * (sin_r(phi) * r));
        } else {
            float cz = 2.0f * u1 - 1.0f;
            float stn = sqrtf(fmaxf(1.0f - cz * cz, 0.0f));
            float phi = 6.2831853f * u2;
            dir = mk3(stn * cos_r(phi), cz, stn * sin_r(phi));
        }
        v3 flux = cv3(light.color) * (light.intensity / (float)per_light);
        DRay ray;
        ray.o = pos; ray.d = dir; ray.tmin = 0.001f; ray.tmax = 1.0e10f;
        for (int bounce = 0; bounce < 4; bounce++) {
            QHit q;
            rays++;
            query_closest<false, true>(S, ray, q, st, ts);
            if (q.kind < 0) break;
            DHit h;
            rebuild_hit(S, ray, q, h);
            const Material& m = S.materials[h.material_id];
            if (m.transmission > 0.5f) {
                v3 n = h.normal;
                float ior = fmaxf(m.ior, 1.0f);
                float eta = 1.0f / ior;
                float cosi = fmaxf(-dot(ray.d, n), 1.0e-4f);
                float k = 1.0f - eta * eta * (1.0f - cosi * cosi);
                v3 td = (k < 0.0f) ? normalize(ray.d + 2.0f * cosi * n) : normalize(eta * ray.d + (eta * cosi - sqrtf(k)) * n);
                flux = flux * m.transmission;
                ray.o = h.point + td * 0.002f;
                ray.d = td;
                continue;
            }
            if (m.reflect > 0.5f) {
                v3 rd = ray.d - h.normal * (2.0f * dot(ray.d, h.normal));
                ray.o = h.point + h.normal * 0.002f;
                ray.d = normalize(rd);
                continue;
            }
            Photon p;
            p.position = {h.point.x, h.point.y, h.point.z};
            p.direction = {-ray.d.x, -ray.d.y, -ray.d.z};
            p.normal = {h.normal.x, h.normal.y, h.normal.z};
            p.flux = {flux.x, flux.y, flux.z};
            out[idx] = p;
            flag[idx] = 1;
            break;
        }
    }
    wave_add64(&pc->rays, rays);
}

__global__ void k_photon_compact(const Photon* all, const uint32_t* flag, const uint32_t* pos, uint32_t total,
                                 uint32_t cap, Photon* stored) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride)
        if (flag[i] && pos[i] < cap) stored[pos[i]] = all[i];
}

// Cell of a stored photon (reference render_kernel.cu:963-966).
__global__ void k_photon_keys(const Photon* p, uint32_t n, float cell, int num_cells, uint32_t* keys, uint32_t* vals,
                              int* counts) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        int ix = (int)floorf(p[i].position.x / cell);
        int iy = (int)floorf(p[i].position.y / cell);
        int iz = (int)floorf(p[i].position.z / cell);
        int c = photon_cell(ix, iy, iz, num_cells);
        keys[i] = (uint32_t)c;
        vals[i] = i;
        atomicAdd(&counts[c], 1);
    }
}

__global__ void k_photon_gather_sorted(const Photon* src, const uint32_t* order, uint32_t n, Photon* dst) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[order[i]];
}

__global__ void k_zero_i32(int* p, uint32_t n) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

// Gather probes for map-vs-brute-force checks (reference render_kernel.cu:981-990, 1092-1107).
__global__ void k_photon_probe(PhotonMapView m, const Photon* src, uint32_t n, int stride_p, float radius, int max_s,
                               int* counts, int brute) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t probes = (n + stride_p - 1) / stride_p;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < probes; i += stride) {
        const Photon& p0 = src[(size_t)i * stride_p];
        int taken = 0;
        if (brute) {
            const float r2 = radius * radius;
            for (uint32_t j = 0; j < n && taken < max_s; j++) {
                v3 d = cv3(src[j].position) - cv3(p0.position);
                if (dot(d, d) > r2) continue;
                if (dot(cv3(src[j].normal), cv3(p0.normal)) < 0.1f) continue;
                taken++;
            }
        } else {
            photon_gather(m, cv3(p0.position), cv3(p0.normal), radius, max_s, &taken);
        }
        counts[i] = taken;
    }
}
