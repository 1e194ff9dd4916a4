// sheen_oracle.cpp — CPU oracle for the Sheen render hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain scalar C++, the
// algorithm of the reference's render path so that the HIP implementation can
// be checked against it.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load liboracle.so, and only as the checker or as the
// timed CPU baseline — never as a product code path.
//
// Every function cites the reference file:line it follows.  The restatement
// is templated on the arithmetic type R: R=float is the parity oracle (same
// operation order as the reference, contraction off, correctly rounded
// sqrt/div), R=double is the FP64 oracle used for RMSE estimates.
//
// Defined behaviour where the reference reads uninitialised memory (none of
// these are pinned by any reference result):
//   * HitRecord.uv of sphere, strand, lathe and prism hits: (0,0);
//     plane hits: (x, z) of the hit point.
//   * HitRecord.object_id of a CSG hit: the object_id of the last earlier
//     category that reported a hit in the same scene_intersect call, else 0.
//   * normalize(v) = v * (1/sqrt(v.v)) (the reference uses rsqrtf).
// Closest-hit ties (equal t within one category) go to the lowest primitive
// index; the reference's winner depends on its BVH's DFS order.
//
// Pinning (tests/test_oracle_*.py, tests/test_reference_stats.py):
// aa_hash against its published inverse; sample layout and camera against
// closed forms; sphere/plane/triangle/cone frusta (equal and unequal radii)/
// superellipsoid/heightfield/prism/CSG hits against analytic solutions;
// BVH closest hit against brute force; photon hash-map gather against brute
// force; mips and sRGB against closed forms; normal maps, gobos, hair, media,
// beams and the caustic term against their closed forms; point, disk-
// jittered, area and spot lights, Blinn-Phong, mask-map roughness,
// transmission and IOR, emission, sky, the reflection blend, slab refraction
// and total internal reflection, the glass passes of shadow rays, strand
// optical depth and the procedural height-field noise against closed forms,
// each also shown to fail with the planted mistakes it targets (the enum
// Plant below); the reference's recorded emission_demo pixel values.
//
// Parity unpinned: no function.  The reference's own transcendentals
// (CUDA's single-precision library) are closed; see sin_r below.

#include "sheen_c_api.h"
#include "sheen_math.h"  // the product's transcendentals: the parity boundary is shared, not restated

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------------------
// Planted mistakes.  The pin tests check every pin twice: the oracle must
// pass it, and the oracle with a plausible mistake planted must fail it, so
// that a pin is shown to be able to catch such a mistake.  The mistakes are
// compiled only into liboracle_plant.so (-DORACLE_PLANTS), where
// oracle_plant(k) switches mistake k on; liboracle.so, the checker and CPU
// baseline, contains none of them (PLANT(k) is the constant false there).
// ---------------------------------------------------------------------------
enum Plant {
    P_NONE = 0,
    P_POINT_DIST = 1,      // point/area/spot attenuation I/(4 pi d) instead of d^2
    P_NO_COS = 2,          // Lambert term without n.l (only its sign kept)
    P_TWO_PI = 3,          // attenuation over 2 pi d^2
    P_DISK_LINEAR = 4,     // disk jitter radius u R instead of sqrt(u) R
    P_AREA_NO_COS = 5,     // area light without its cos_l factor
    P_AREA_SIDE = 6,       // area light lit from the other side
    P_AREA_HALF = 7,       // area sample u in [0, 1] instead of [-1, 1]
    P_SPOT_RAMP = 8,       // spot edge ramp over (1 - cos_half)
    P_SPOT_HALF_ANGLE = 9, // spot cone of half the angle
    P_SPEC_NL = 10,        // Blinn-Phong lobe on n.l instead of n.h
    P_SPEC_NO_GUARD = 11,  // specular added on surfaces facing away from the light
    P_ROUGH_R2 = 12,       // shininess 2/r^2 - 2
    P_IOR_RED = 13,        // mask IOR from the R channel
    P_TRANS_NO_MASK = 14,  // mask transmission ignored
    P_SKY_X = 15,          // sky gradient on dir.x
    P_SKY_SWAP = 16,       // sky zenith colour channels reversed
    P_REFL_NO_CLAMP = 17,  // reflection weight not clamped to 1
    P_REFL_SWAP = 18,      // reflection blend weights exchanged
    P_ETA_FIXED = 19,      // refraction eta 1/ior on every face
    P_TIR_SIGN = 20,       // total internal reflection about the wrong side
    P_W_ONCE = 21,         // refraction weight taken from the first face only
    P_SHADOW_TMAX = 22,    // glass shadow passes shorten t_max
    P_SHADOW_4 = 23,       // four glass passes per shadow ray
    P_CHORD_HALF = 24,     // strand chord without its factor 2
    P_RADIUS_REV = 25,     // strand radius interpolated from the wrong end
    P_TAU_NO_STOP = 26,    // strand optical depth summed past 6
    P_HF_HASH = 27,        // height-field hash with another z multiplier
    P_HF_DETAIL = 28,      // height-field detail octave seed + 777
    P_HF_LINEAR = 29,      // value noise without the smoothstep
    P_HF_OCT_SEED = 30,    // fBm octave seed step 113
    P_SIGMA = 31,          // strand extinction 5 instead of 6
    P_SHADOW_THRESH = 32,  // glass passes shadow rays from transmission 0.3
    P_EMISSION_LIT = 33,   // emission scaled by the ambient term
    P_BEAM_NO_SHADOW = 34, // beam steps lit without their shadow rays (the reference's beam before
                           // it traced them, ROADMAP.md:101-111)
};
#ifdef ORACLE_PLANTS
static int g_plant = 0;
#define PLANT(k) (g_plant == (k))
#else
#define PLANT(k) false
#endif

// ---------------------------------------------------------------------------
// Vector arithmetic (reference types.h:306-329)
// ---------------------------------------------------------------------------
template <class R> struct V3 { R x, y, z; };
template <class R> inline V3<R> mk(R x, R y, R z) { return {x, y, z}; }
template <class R> inline V3<R> cv(const sheen_f3& f) { return {R(f.x), R(f.y), R(f.z)}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R b) { return {a.x * b, a.y * b, a.z * b}; }
template <class R> inline V3<R> operator*(R b, V3<R> a) { return {a.x * b, a.y * b, a.z * b}; }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline V3<R> operator/(V3<R> a, R b) { return {a.x / b, a.y / b, a.z / b}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline V3<R> normalize(V3<R> v) {
    R inv = R(1) / std::sqrt(dot(v, v));
    return {v.x * inv, v.y * inv, v.z * inv};
}
// Transcendentals of shading, light sampling, media and photons (reference
// render_kernel.cu, shading.h; the reference calls its library's
// single-precision functions, whose last bits no other library reproduces).
// For R = float: sheen_math.h's double series rounded once to float, the
// correctly rounded value except in rare hard cases; the HIP path evaluates
// the same operations, so the two agree bit for bit.  sheen_math.h is pinned
// against the host libm in tests/test_sheen_math.py.  For R = double: the
// double functions of the host libm.  Superellipsoid tests and normals keep
// powf (hit distances).
inline float sin_r(float x) { return shm::sinf_r(x); }
inline float cos_r(float x) { return shm::cosf_r(x); }
inline float tan_r(float x) { return shm::tanf_r(x); }
inline float exp_r(float x) { return shm::expf_r(x); }
inline float log2_r(float x) { return shm::log2f_r(x); }
inline float pow_r(float x, float y) { return shm::powf_r(x, y); }
inline double sin_r(double x) { return std::sin(x); }
inline double cos_r(double x) { return std::cos(x); }
inline double tan_r(double x) { return std::tan(x); }
inline double exp_r(double x) { return std::exp(x); }
inline double log2_r(double x) { return std::log2(x); }
inline double pow_r(double x, double y) { return std::pow(x, y); }
template <class R> inline R fmin_(R a, R b) { return std::fmin(a, b); }
template <class R> inline R fmax_(R a, R b) { return std::fmax(a, b); }

// ---------------------------------------------------------------------------
// Sampling (reference sampling/aa.h:15-71)
// ---------------------------------------------------------------------------
inline uint32_t aa_hash(uint32_t x) {  // aa.h:32-37
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
inline float aa_rand(uint32_t seed) {  // aa.h:39-41
    return float(aa_hash(seed) & 0xffffffu) / 16777216.0f;
}
inline void aa_sample_offset(int s, float& ox, float& oy) {  // aa.h:15-22
    switch (s) {
        case 0: ox = 0.25f; oy = 0.25f; break;
        case 1: ox = 0.75f; oy = 0.25f; break;
        case 2: ox = 0.25f; oy = 0.75f; break;
        default: ox = 0.75f; oy = 0.75f; break;
    }
}
inline uint32_t jitter_seed(int s, int x, int y) {  // aa.h:47-48, 66-67 (signed wrap = uint32 wrap)
    return (uint32_t(x) * 73856093u) ^ (uint32_t(y) * 19349663u) ^ (uint32_t(s) * 83492791u);
}
inline void aa_sample_offset_n(int s, int n, int x, int y, float& ox, float& oy) {  // aa.h:57-71
    uint32_t seed = jitter_seed(s, x, y);
    if (n <= 4) {  // aa.h:45-52
        float cx, cy;
        aa_sample_offset(s, cx, cy);
        float jx = aa_rand(seed) - 0.5f;
        float jy = aa_rand(seed ^ 0x9e3779b9u) - 0.5f;
        ox = cx + jx * 0.5f;
        oy = cy + jy * 0.5f;
        return;
    }
    int side = 2;
    while (side * side < n) side++;
    int cx = s % side, cy = s / side;
    float inv = 1.0f / (float)side;
    float jx = aa_rand(seed);
    float jy = aa_rand(seed ^ 0x9e3779b9u);
    ox = ((float)cx + jx) * inv;
    oy = ((float)cy + jy) * inv;
}
template <class R> inline void aa_uv(int x, int y, int w, int h, float ox, float oy, R& u, R& v) {  // aa.h:25-29
    u = (R(x) + R(ox)) / R(w) * R(2.0f) - R(1.0f);
    v = R(1.0f) - (R(y) + R(oy)) / R(h) * R(2.0f);
}
inline uint32_t sample_seed(int x, int y, int s) {  // render_kernel.cu:494-496
    return aa_hash(uint32_t(x) * 73856093u ^ uint32_t(y) * 19349663u ^ uint32_t(s) * 83492791u);
}

// ---------------------------------------------------------------------------
// Ray / hit
// ---------------------------------------------------------------------------
template <class R> struct Ray { V3<R> o, d; R tmin, tmax; };
enum Kind { K_SPHERE = 0, K_STRAND = 1, K_TRI = 2, K_HF = 3, K_SE = 4, K_LATHE = 5, K_PRISM = 6, K_CSG = 7, K_PLANE = 8 };

template <class R> struct Hit {
    R t = 0;
    int material_id = 0;
    int object_id = 0;
    int kind = -1;
    int index = -1;
    V3<R> point{0, 0, 0}, normal{0, 0, 0}, tangent{0, 0, 0};
    R uvx = 0, uvy = 0;
};

// ---------------------------------------------------------------------------
// Scene copy (host) with category BVHs
// ---------------------------------------------------------------------------
// fpx/dpx: float texel chains set by oracle_set_params (level 0 given, mips
// as 2x2 means ((a + b) + (c + d)) * 0.25 in float and in double); when
// use_float is set they replace the 8-bit chain.
struct Tex {
    std::vector<unsigned char> px;
    std::vector<float> fpx;
    std::vector<double> dpx;
    int w = 0, h = 0, mips = 1;
    bool valid = false, use_float = false;
};
template <class R> inline const R* float_chain(const Tex& t);
template <> inline const float* float_chain<float>(const Tex& t) { return t.fpx.data(); }
template <> inline const double* float_chain<double>(const Tex& t) { return t.dpx.data(); }
template <class R> inline void build_float_mips(std::vector<R>& c, int w, int h) {
    size_t src = 0;
    int sw = w, sh = h;
    while (sw > 1 || sh > 1) {
        const int nw = sw > 1 ? sw / 2 : 1, nh = sh > 1 ? sh / 2 : 1;
        const size_t dst = src + (size_t)sw * sh * 3;
        for (int y = 0; y < nh; y++) {
            const int ya = std::min(2 * y, sh - 1), yb = std::min(2 * y + 1, sh - 1);
            for (int x = 0; x < nw; x++) {
                const int xa = std::min(2 * x, sw - 1), xb = std::min(2 * x + 1, sw - 1);
                for (int k = 0; k < 3; k++) {
                    const R a = c[src + ((size_t)ya * sw + xa) * 3 + k], b = c[src + ((size_t)ya * sw + xb) * 3 + k];
                    const R d = c[src + ((size_t)yb * sw + xa) * 3 + k], e = c[src + ((size_t)yb * sw + xb) * 3 + k];
                    c[dst + ((size_t)y * nw + x) * 3 + k] = ((a + b) + (d + e)) * R(0.25);
                }
            }
        }
        src = dst; sw = nw; sh = nh;
    }
}

// reference scene.h:230-261 (2x2 box, (s+2)/4, levels contiguous)
inline void build_mips(Tex& t) {
    t.mips = 1;
    if (t.w < 2 && t.h < 2) return;
    std::vector<unsigned char> chain(t.px);
    int sw = t.w, sh = t.h;
    size_t src = 0;
    while (sw > 1 || sh > 1) {
        int nw = sw > 1 ? sw / 2 : 1, nh = sh > 1 ? sh / 2 : 1;
        size_t dst = chain.size();
        chain.resize(dst + (size_t)nw * nh * 3);
        for (int y = 0; y < nh; y++) {
            int y0 = (2 * y) < sh ? 2 * y : sh - 1;
            int y1 = (2 * y + 1) < sh ? 2 * y + 1 : sh - 1;
            for (int x = 0; x < nw; x++) {
                int x0 = (2 * x) < sw ? 2 * x : sw - 1;
                int x1 = (2 * x + 1) < sw ? 2 * x + 1 : sw - 1;
                for (int c = 0; c < 3; c++) {
                    int s = chain[src + ((size_t)y0 * sw + x0) * 3 + c] + chain[src + ((size_t)y0 * sw + x1) * 3 + c] +
                            chain[src + ((size_t)y1 * sw + x0) * 3 + c] + chain[src + ((size_t)y1 * sw + x1) * 3 + c];
                    chain[dst + ((size_t)y * nw + x) * 3 + c] = (unsigned char)((s + 2) / 4);
                }
            }
        }
        src = dst; sw = nw; sh = nh;
        t.mips++;
    }
    t.px.swap(chain);
}

struct Box { float mn[3], mx[3]; };
struct BNode { float mn[3], mx[3]; int left, right; };  // leaf: left = ~prim

struct Photons {
    std::vector<Photon> sorted;
    std::vector<int> cell_start, cell_count;
    int num_cells = 0;
    float cell = 0.f;
    bool valid() const { return !sorted.empty() && num_cells > 0 && cell > 0.f; }
};

struct Scene {
    SheenScene s{};
    std::vector<Tex> tex;
    std::vector<Material> own_mats;  // set by oracle_set_params
    std::vector<Light> own_lights;
    const Material* orig_mats = nullptr;  // the caller's arrays
    const Light* orig_lights = nullptr;
    std::vector<BNode> bvh[7];  // sphere, strand, tri, hf, se, lathe, prism
    int count[7] = {0, 0, 0, 0, 0, 0, 0};
    Photons pm;
    int photon_emitted = 0, photon_stored = 0;
    uint64_t photon_rays = 0;
};

// Primitive bounds (reference bvh.h:24-66, 235-242, 304-331, 394-407).
inline Box bounds_of(const Scene& S, int cat, int i) {
    Box b;
    auto set = [&](float a0, float a1, float a2, float b0, float b1, float b2) {
        b.mn[0] = a0; b.mn[1] = a1; b.mn[2] = a2; b.mx[0] = b0; b.mx[1] = b1; b.mx[2] = b2;
    };
    const SheenScene& s = S.s;
    switch (cat) {
        case K_SPHERE: {
            const Object& o = s.objects[i];
            // the builder folds the single leaf's bound into {1e10,-1e10} (bvh.h:80-89)
            set(std::fmin(1e10f, o.param1.x - o.param2), std::fmin(1e10f, o.param1.y - o.param2),
                std::fmin(1e10f, o.param1.z - o.param2), std::fmax(-1e10f, o.param1.x + o.param2),
                std::fmax(-1e10f, o.param1.y + o.param2), std::fmax(-1e10f, o.param1.z + o.param2));
            break;
        }
        case K_STRAND: {
            const SweepStrand& st = s.sweep_strands[i];
            float a0 = 1e10f, a1 = 1e10f, a2 = 1e10f, b0 = -1e10f, b1 = -1e10f, b2 = -1e10f;
            for (int k = 0; k < st.point_count; k++) {
                const SweepPoint& p = s.sweep_points[st.point_offset + k];
                float r = p.radius;
                a0 = std::fmin(a0, p.position.x - r); a1 = std::fmin(a1, p.position.y - r); a2 = std::fmin(a2, p.position.z - r);
                b0 = std::fmax(b0, p.position.x + r); b1 = std::fmax(b1, p.position.y + r); b2 = std::fmax(b2, p.position.z + r);
            }
            set(std::fmin(1e10f, a0), std::fmin(1e10f, a1), std::fmin(1e10f, a2),
                std::fmax(-1e10f, b0), std::fmax(-1e10f, b1), std::fmax(-1e10f, b2));
            break;
        }
        case K_TRI: {
            const Triangle& t = s.triangles[i];
            set(std::fmin(t.v0.x, std::fmin(t.v1.x, t.v2.x)), std::fmin(t.v0.y, std::fmin(t.v1.y, t.v2.y)),
                std::fmin(t.v0.z, std::fmin(t.v1.z, t.v2.z)), std::fmax(t.v0.x, std::fmax(t.v1.x, t.v2.x)),
                std::fmax(t.v0.y, std::fmax(t.v1.y, t.v2.y)), std::fmax(t.v0.z, std::fmax(t.v1.z, t.v2.z)));
            break;
        }
        case K_HF: {
            const HeightField& h = s.heightfields[i];
            set(h.center.x - 0.5f * h.size_x, h.center.y, h.center.z - 0.5f * h.size_z,
                h.center.x + 0.5f * h.size_x, h.center.y + h.height, h.center.z + 0.5f * h.size_z);
            break;
        }
        case K_SE: {
            const Superellipsoid& e = s.superellipsoids[i];
            set(e.center.x - e.a, e.center.y - e.b, e.center.z - e.c, e.center.x + e.a, e.center.y + e.b, e.center.z + e.c);
            break;
        }
        case K_LATHE: {
            const Lathe& l = s.lathes[i];
            float a0 = l.center.x - 1e30f, a1 = 1e30f, a2 = l.center.z - 1e30f;
            float b0 = l.center.x + 1e30f, b1 = -1e30f, b2 = l.center.z + 1e30f;
            for (int k = 0; k < l.point_count; k++) {
                const LathePoint& p = s.lathe_points[l.point_offset + k];
                float r = p.radius;
                if (r < 0.0f) continue;
                a0 = std::fmin(a0, l.center.x - r); a1 = std::fmin(a1, l.center.y + p.y - r); a2 = std::fmin(a2, l.center.z - r);
                b0 = std::fmax(b0, l.center.x + r); b1 = std::fmax(b1, l.center.y + p.y + r); b2 = std::fmax(b2, l.center.z + r);
            }
            if (a0 > b0) { a0 = a1 = a2 = b0 = b1 = b2 = 0.0f; }
            set(a0, a1, a2, b0, b1, b2);
            break;
        }
        case K_PRISM: {
            const Prism& p = s.prisms[i];
            float a0 = 1e30f, a1 = p.center.y + p.y_min, a2 = 1e30f;
            float b0 = -1e30f, b1 = p.center.y + p.y_max, b2 = -1e30f;
            for (int k = 0; k < p.point_count; k++) {
                const PrismPoint& q = s.prism_points[p.point_offset + k];
                a0 = std::fmin(a0, p.center.x + q.x); b0 = std::fmax(b0, p.center.x + q.x);
                a2 = std::fmin(a2, p.center.z + q.z); b2 = std::fmax(b2, p.center.z + q.z);
            }
            if (a0 > b0) { a0 = 0; a1 = p.center.y; a2 = 0; b0 = 0; b1 = p.center.y; b2 = 0; }
            set(a0, a1, a2, b0, b1, b2);
            break;
        }
    }
    return b;
}

// Object-median binary BVH with one primitive per leaf (reference bvh.h:71-576).
inline int build_rec(std::vector<BNode>& nodes, const std::vector<Box>& bx, std::vector<int>& idx, int start, int end) {
    int ni = (int)nodes.size();
    nodes.push_back(BNode());
    BNode n;
    for (int a = 0; a < 3; a++) { n.mn[a] = 1e30f; n.mx[a] = -1e30f; }
    float cmn[3] = {1e30f, 1e30f, 1e30f}, cmx[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = start; i < end; i++) {
        const Box& b = bx[idx[i]];
        for (int a = 0; a < 3; a++) {
            n.mn[a] = std::fmin(n.mn[a], b.mn[a]);
            n.mx[a] = std::fmax(n.mx[a], b.mx[a]);
            float c = (b.mn[a] + b.mx[a]) * 0.5f;
            cmn[a] = std::fmin(cmn[a], c);
            cmx[a] = std::fmax(cmx[a], c);
        }
    }
    int count = end - start;
    if (count == 1) {
        n.left = n.right = ~idx[start];
        nodes[ni] = n;
        return ni;
    }
    float ex = cmx[0] - cmn[0], ey = cmx[1] - cmn[1], ez = cmx[2] - cmn[2];
    int axis = 0;
    if (ey > ex && ey > ez) axis = 1;
    if (ez > ex && ez > ey) axis = 2;
    int mid = start + count / 2;
    std::nth_element(idx.begin() + start, idx.begin() + mid, idx.begin() + end, [&](int a, int b) {
        return (bx[a].mn[axis] + bx[a].mx[axis]) * 0.5f < (bx[b].mn[axis] + bx[b].mx[axis]) * 0.5f;
    });
    int l = build_rec(nodes, bx, idx, start, mid);
    int r = build_rec(nodes, bx, idx, mid, end);
    n.left = l; n.right = r;
    nodes[ni] = n;
    return ni;
}

// reference bvh.h:578-600 (slab test, culls against t_min/t_max only)
template <class R> inline bool intersect_aabb(const Ray<R>& ray, const float* mn, const float* mx) {
    R invDirX = R(1) / ray.d.x;
    R tminX = (R(mn[0]) - ray.o.x) * invDirX;
    R tmaxX = (R(mx[0]) - ray.o.x) * invDirX;
    if (tminX > tmaxX) std::swap(tminX, tmaxX);
    R invDirY = R(1) / ray.d.y;
    R tminY = (R(mn[1]) - ray.o.y) * invDirY;
    R tmaxY = (R(mx[1]) - ray.o.y) * invDirY;
    if (tminY > tmaxY) std::swap(tminY, tmaxY);
    if (tminX > tmaxY || tminY > tmaxX) return false;
    if (tminY > tminX) tminX = tminY;
    if (tmaxY < tmaxX) tmaxX = tmaxY;
    R invDirZ = R(1) / ray.d.z;
    R tminZ = (R(mn[2]) - ray.o.z) * invDirZ;
    R tmaxZ = (R(mx[2]) - ray.o.z) * invDirZ;
    if (tminZ > tmaxZ) std::swap(tminZ, tmaxZ);
    if (tminX > tmaxZ || tminZ > tmaxX) return false;
    return (tmaxX >= ray.tmin) && (tminX <= ray.tmax);
}

// ---------------------------------------------------------------------------
// Primitives
// ---------------------------------------------------------------------------
// reference primitives/sphere.h:30-53
template <class R> inline bool intersect_sphere(const Ray<R>& ray, V3<R> c, R radius, Hit<R>& rec) {
    V3<R> oc = ray.o - c;
    R a = dot(ray.d, ray.d);
    R b = R(2.0f) * dot(oc, ray.d);
    R cc = dot(oc, oc) - radius * radius;
    R disc = b * b - R(4.0f) * a * cc;
    if (disc < R(0)) return false;
    R sq = std::sqrt(disc);
    R t = (-b - sq) / (R(2.0f) * a);
    if (t < ray.tmin || t > ray.tmax) {
        t = (-b + sq) / (R(2.0f) * a);
        if (t < ray.tmin || t > ray.tmax) return false;
    }
    rec.t = t;
    rec.point = ray.o + t * ray.d;
    rec.normal = normalize(rec.point - c);
    rec.tangent = mk<R>(0, 0, 0);
    rec.uvx = 0; rec.uvy = 0;
    return true;
}
// reference primitives/sphere.h:8-25
template <class R> inline bool sphere_intervals(const Ray<R>& ray, V3<R> c, R radius, R t0, R t1, R& te, R& tx) {
    V3<R> oc = ray.o - c;
    R a = dot(ray.d, ray.d);
    R b = R(2.0f) * dot(oc, ray.d);
    R cc = dot(oc, oc) - radius * radius;
    R disc = b * b - R(4.0f) * a * cc;
    if (disc < R(0)) return false;
    R sq = std::sqrt(disc);
    R ta = (-b - sq) / (R(2.0f) * a);
    R tb = (-b + sq) / (R(2.0f) * a);
    if (tb < t0 || ta > t1) return false;
    te = fmax_(ta, t0);
    tx = fmin_(tb, t1);
    return te <= tx;
}
// reference primitives/plane.h:6-24
template <class R> inline bool intersect_plane(const Ray<R>& ray, V3<R> n, R d, Hit<R>& rec) {
    R denom = dot(n, ray.d);
    if (std::fabs(denom) > R(1e-6f)) {
        R t = -(dot(ray.o, n) + d) / denom;
        if (t >= ray.tmin && t <= ray.tmax) {
            rec.t = t;
            rec.point = ray.o + t * ray.d;
            rec.normal = n;
            rec.tangent = mk<R>(0, 0, 0);
            if (dot(rec.normal, ray.d) > R(0)) rec.normal = -rec.normal;
            rec.uvx = rec.point.x; rec.uvy = rec.point.z;
            return true;
        }
    }
    return false;
}
// reference primitives/triangle.h:7-46
template <class R> inline bool intersect_triangle(const Ray<R>& ray, const Triangle& tri, Hit<R>& rec) {
    V3<R> v0 = cv<R>(tri.v0), v1 = cv<R>(tri.v1), v2 = cv<R>(tri.v2);
    V3<R> e1 = v1 - v0, e2 = v2 - v0;
    V3<R> pvec = cross(ray.d, e2);
    R det = dot(e1, pvec);
    if (std::fabs(det) < R(1e-8f)) return false;
    R inv_det = R(1) / det;
    V3<R> tvec = ray.o - v0;
    R u = dot(tvec, pvec) * inv_det;
    if (u < R(0) || u > R(1)) return false;
    V3<R> qvec = cross(tvec, e1);
    R v = dot(ray.d, qvec) * inv_det;
    if (v < R(0) || u + v > R(1)) return false;
    R t = dot(e2, qvec) * inv_det;
    if (t < ray.tmin || t > ray.tmax) return false;
    rec.t = t;
    rec.point = ray.o + t * ray.d;
    R w = R(1) - u - v;
    V3<R> nv = cv<R>(tri.n0) * w + cv<R>(tri.n1) * u + cv<R>(tri.n2) * v;
    R len2 = dot(nv, nv);
    if (len2 > R(1e-12f)) rec.normal = normalize(nv);
    else rec.normal = normalize(cross(e1, e2));
    rec.tangent = normalize(e1);
    rec.uvx = R(tri.uv0.x) * w + R(tri.uv1.x) * u + R(tri.uv2.x) * v;
    rec.uvy = R(tri.uv0.y) * w + R(tri.uv1.y) * u + R(tri.uv2.y) * v;
    if (dot(rec.normal, ray.d) > R(0)) rec.normal = -rec.normal;
    return true;
}
// reference primitives/cone.h:11-73
template <class R> inline bool intersect_cone_segment(const Ray<R>& ray, V3<R> p0, R r0, V3<R> p1, R r1, Hit<R>& rec) {
    V3<R> axis = p1 - p0;
    R L2 = dot(axis, axis);
    if (L2 < R(1e-12f)) return false;
    R L = std::sqrt(L2);
    V3<R> da = axis / L;
    R dr = r1 - r0;
    R k = dr / L;
    V3<R> Q = ray.o - p0;
    R u = dot(ray.d, da);
    R v = dot(Q, da);
    V3<R> Dp = ray.d - u * da;
    V3<R> Qp = Q - v * da;
    R A = dot(Dp, Dp) - k * k * u * u;
    R B = R(2.0f) * (dot(Dp, Qp) - k * r0 * u - k * k * v * u);
    R C = dot(Qp, Qp) - k * k * v * v - R(2.0f) * k * r0 * v - r0 * r0;
    R disc = B * B - R(4.0f) * A * C;
    if (disc < R(0)) return false;
    R sq = std::sqrt(disc);
    R inv2A = R(1) / (R(2.0f) * A);
    R s1 = (-B - sq) * inv2A;
    R s2 = (-B + sq) * inv2A;
    for (int i = 0; i < 2; i++) {
        R s = (i == 0) ? s1 : s2;
        if (s < ray.tmin || s > ray.tmax) continue;
        R tp = v + s * u;
        if (tp < R(0) || tp > L) continue;
        rec.t = s;
        rec.point = ray.o + s * ray.d;
        V3<R> radial = Qp + s * Dp;
        rec.normal = normalize(radial);
        rec.tangent = da;
        if (dot(rec.normal, ray.d) > R(0)) rec.normal = -rec.normal;
        return true;
    }
    return false;
}
// reference primitives/cone.h:78-121
template <class R> inline bool intersect_hemisphere(const Ray<R>& ray, V3<R> c, R radius, V3<R> axis, Hit<R>& rec) {
    V3<R> oc = ray.o - c;
    R a = dot(ray.d, ray.d);
    R b = R(2.0f) * dot(oc, ray.d);
    R cc = dot(oc, oc) - radius * radius;
    R disc = b * b - R(4.0f) * a * cc;
    if (disc < R(0)) return false;
    R sq = std::sqrt(disc);
    R inv2a = R(1) / (R(2.0f) * a);
    R t1 = (-b - sq) * inv2a;
    R t2 = (-b + sq) * inv2a;
    for (int i = 0; i < 2; i++) {
        R t = (i == 0) ? t1 : t2;
        if (t < ray.tmin || t > ray.tmax) continue;
        V3<R> p = ray.o + t * ray.d;
        V3<R> n = normalize(p - c);
        if (dot(n, axis) < R(0)) continue;
        rec.t = t;
        rec.point = p;
        rec.normal = n;
        if (dot(rec.normal, ray.d) > R(0)) rec.normal = -rec.normal;
        return true;
    }
    return false;
}
// reference primitives/sweep.h:13-92.  intersect_hemisphere leaves the tangent
// unset, so a cap hit carries the tangent of the last successful test of the
// same call; the first test of a call starts from (0,0,0).
template <class R> inline bool intersect_sphere_sweep(const Ray<R>& ray, const SweepStrand& st, const SweepPoint* pts, Hit<R>& rec) {
    if (st.point_count < 2) return false;
    bool any = false;
    R closest = ray.tmax;
    Hit<R> tmp;
    tmp.tangent = mk<R>(0, 0, 0);
    int base = st.point_offset, count = st.point_count;
    auto P = [&](int i) { return cv<R>(pts[base + i].position); };
    auto Rr = [&](int i) { return R(pts[base + i].radius); };
    auto take = [&]() {
        if (tmp.t < closest) { rec = tmp; rec.material_id = st.material_id; rec.uvx = 0; rec.uvy = 0; closest = tmp.t; any = true; }
    };
    {
        V3<R> ax = normalize(P(0) - P(1));
        if (intersect_hemisphere(ray, P(0), Rr(0), ax, tmp)) take();
    }
    for (int i = 0; i < count - 1; i++) {
        if (intersect_cone_segment(ray, P(i), Rr(i), P(i + 1), Rr(i + 1), tmp)) take();
        if (i < count - 2) {
            V3<R> ax = normalize(P(i + 2) - P(i + 1));
            if (intersect_hemisphere(ray, P(i + 1), Rr(i + 1), ax, tmp)) take();
        }
    }
    {
        V3<R> ax = normalize(P(count - 1) - P(count - 2));
        if (intersect_hemisphere(ray, P(count - 1), Rr(count - 1), ax, tmp)) take();
    }
    return any;
}
// reference primitives/sweep.h:97-148
template <class R> inline R strand_segment_tau(const Ray<R>& ray, const SweepPoint& p0, const SweepPoint& p1, R sigma) {
    V3<R> u = cv<R>(p1.position) - cv<R>(p0.position);
    R A = dot(u, u);
    if (A < R(1e-12f)) return R(0);
    V3<R> v = ray.d;
    V3<R> w = ray.o - cv<R>(p0.position);
    R B = dot(u, v), C = dot(v, v), D = dot(u, w), E = dot(v, w);
    R denom = A * C - B * B;
    R s = (denom > R(1e-9f)) ? (B * E - C * D) / denom : R(0);
    s = fmin_(fmax_(s, R(0)), R(1));
    R t = (B * s + E) / C;
    if (t < R(0)) {
        t = R(0);
        s = fmin_(fmax_(-D / A, R(0)), R(1));
    }
    if (t > ray.tmax) return R(0);
    V3<R> closest = w + u * s - v * t;
    R d2 = dot(closest, closest);
    R r = PLANT(P_RADIUS_REV) ? R(p1.radius) + (R(p0.radius) - R(p1.radius)) * s
                              : R(p0.radius) + (R(p1.radius) - R(p0.radius)) * s;
    R r2 = r * r;
    if (d2 >= r2) return R(0);
    R chord = (PLANT(P_CHORD_HALF) ? R(1.0f) : R(2.0f)) * std::sqrt(fmax_(r2 - d2, R(0)));
    return sigma * chord;
}
template <class R> inline R strand_tau(const Ray<R>& ray, const SweepStrand& st, const SweepPoint* pts, R sigma) {
    R tau = 0;
    for (int i = 0; i < st.point_count - 1; i++) {
        tau += strand_segment_tau(ray, pts[st.point_offset + i], pts[st.point_offset + i + 1], sigma);
        if (tau > R(6.0f) && !PLANT(P_TAU_NO_STOP)) break;
    }
    return tau;
}
// reference primitives/superellipsoid.h:12-28
template <class R> inline R se_f(const Superellipsoid& s, V3<R> p) {
    R qx = (p.x - R(s.center.x)) / R(s.a);
    R qy = (p.y - R(s.center.y)) / R(s.b);
    R qz = (p.z - R(s.center.z)) / R(s.c);
    return std::pow(std::fabs(qx), R(s.n)) + std::pow(std::fabs(qy), R(s.n)) + std::pow(std::fabs(qz), R(s.n));
}
template <class R> inline V3<R> se_grad(const Superellipsoid& s, V3<R> p) {
    R qx = (p.x - R(s.center.x)) / R(s.a);
    R qy = (p.y - R(s.center.y)) / R(s.b);
    R qz = (p.z - R(s.center.z)) / R(s.c);
    R n = R(s.n);
    R gx = n * std::pow(std::fabs(qx), n - R(1.0f)) * (qx >= R(0) ? R(1) : R(-1)) / R(s.a);
    R gy = n * std::pow(std::fabs(qy), n - R(1.0f)) * (qy >= R(0) ? R(1) : R(-1)) / R(s.b);
    R gz = n * std::pow(std::fabs(qz), n - R(1.0f)) * (qz >= R(0) ? R(1) : R(-1)) / R(s.c);
    return {gx, gy, gz};
}
// box entry/exit on the SE bounds (superellipsoid.h:48-63 and 131-144 share this form)
template <class R> inline void se_box(const Ray<R>& ray, const Superellipsoid& se, R& tA, R& tB) {
    R bmnx = R(se.center.x - se.a), bmny = R(se.center.y - se.b), bmnz = R(se.center.z - se.c);
    R bmxx = R(se.center.x + se.a), bmxy = R(se.center.y + se.b), bmxz = R(se.center.z + se.c);
    R invDx = R(1) / ray.d.x;
    R tx0 = (bmnx - ray.o.x) * invDx, tx1 = (bmxx - ray.o.x) * invDx;
    if (tx0 > tx1) std::swap(tx0, tx1);
    R invDy = R(1) / ray.d.y;
    R ty0 = (bmny - ray.o.y) * invDy, ty1 = (bmxy - ray.o.y) * invDy;
    if (ty0 > ty1) std::swap(ty0, ty1);
    R invDz = R(1) / ray.d.z;
    R tz0 = (bmnz - ray.o.z) * invDz, tz1 = (bmxz - ray.o.z) * invDz;
    if (tz0 > tz1) std::swap(tz0, tz1);
    tA = fmax_(fmax_(tx0, ty0), tz0);
    tB = fmin_(fmin_(tx1, ty1), tz1);
}
// reference primitives/superellipsoid.h:41-117
template <class R> inline bool se_intervals(const Ray<R>& ray, const Superellipsoid& se, R t0, R t1, R& te, R& tx) {
    {
        R tA, tB;
        se_box(ray, se, tA, tB);
        if (tB < t0 || tA > t1 || tB < tA) return false;
        t0 = fmax_(t0, tA);
        t1 = fmin_(t1, tB);
    }
    const R EPS = R(1.0f);
    const int NS = 32;
    const R STEP = (t1 - t0) / R(NS);
    R t_prev = t0;
    R F_prev = se_f(se, ray.o + t_prev * ray.d);
    bool inside = (F_prev <= EPS);
    te = R(-1e30f); tx = R(-1e30f);
    bool fe = false, fx = false;
    if (inside) { te = t0; fe = true; }
    for (int i = 1; i <= NS && !fx; i++) {
        R t = fmin_(t0 + STEP * R(i), t1);
        R F = se_f(se, ray.o + t * ray.d);
        if (inside && F > EPS) {
            R lo = t_prev, hi = t;
            for (int b = 0; b < 20; b++) {
                R mid = R(0.5f) * (lo + hi);
                if (se_f(se, ray.o + mid * ray.d) > EPS) hi = mid; else lo = mid;
            }
            tx = R(0.5f) * (lo + hi);
            if (tx > t1) tx = t1;
            fx = true;
            break;
        }
        t_prev = t;
        F_prev = F;
        if (!inside && F <= EPS) {
            R lo = t_prev, hi = t;
            for (int b = 0; b < 20; b++) {
                R mid = R(0.5f) * (lo + hi);
                if (se_f(se, ray.o + mid * ray.d) <= EPS) hi = mid; else lo = mid;
            }
            te = R(0.5f) * (lo + hi);
            if (te < t0) te = t0;
            fe = true;
            inside = true;
        }
    }
    (void)F_prev;
    if (!fe) return false;
    if (!fx) tx = t1;
    return te <= tx;
}
// reference primitives/superellipsoid.h:123-161
template <class R> inline bool intersect_superellipsoid(const Ray<R>& ray, const Superellipsoid& se, Hit<R>& rec) {
    R t0, t1;
    se_box(ray, se, t0, t1);
    if (t1 < ray.tmin || t0 > ray.tmax || t1 < t0) return false;
    t0 = fmax_(t0, ray.tmin);
    t1 = fmin_(t1, ray.tmax);
    R te, tx;
    if (!se_intervals(ray, se, t0, t1, te, tx)) return false;
    R t = te;
    V3<R> P = ray.o + t * ray.d;
    rec.t = t;
    rec.point = P;
    rec.normal = normalize(se_grad(se, P));
    rec.tangent = mk<R>(1, 0, 0);
    rec.uvx = P.x; rec.uvy = P.y;
    if (dot(rec.normal, ray.d) > R(0)) rec.normal = -rec.normal;
    return true;
}
// reference primitives/heightfield.h:13-18 (signed wrap restated in uint32)
inline float hf_cell_hash(int xi, int zi, int seed) {
    uint32_t h = uint32_t(xi) * 374761393u + uint32_t(zi) * (PLANT(P_HF_HASH) ? 668265261u : 668265263u) +
                 uint32_t(seed) * 974634097u;
    h = (h ^ (h >> 13)) * 1274126177u;
    h = h ^ (h >> 16);
    return float(h & 0xffffffu) / 16777216.0f;
}
template <class R> inline R hf_value_noise(R fx, R fz, int seed) {  // heightfield.h:21-36
    int x0 = (int)std::floor(fx), z0 = (int)std::floor(fz);
    R tx = fx - R(x0), tz = fz - R(z0);
    if (!PLANT(P_HF_LINEAR)) {
        tx = tx * tx * (R(3.0f) - R(2.0f) * tx);
        tz = tz * tz * (R(3.0f) - R(2.0f) * tz);
    }
    R v00 = R(hf_cell_hash(x0, z0, seed));
    R v10 = R(hf_cell_hash(x0 + 1, z0, seed));
    R v01 = R(hf_cell_hash(x0, z0 + 1, seed));
    R v11 = R(hf_cell_hash(x0 + 1, z0 + 1, seed));
    R a = v00 + (v10 - v00) * tx;
    R b = v01 + (v11 - v01) * tx;
    return a + (b - a) * tz;
}
template <class R> inline R hf_fbm(R x, R z, R freq, int octaves, int seed) {  // heightfield.h:40-51
    R fx = x * freq, fz = z * freq;
    R amp = R(0.5f), sum = R(0), norm = R(0);
    for (int o = 0; o < octaves; o++) {
        sum += amp * hf_value_noise(fx, fz, seed + o * (PLANT(P_HF_OCT_SEED) ? 113 : 131));
        norm += amp;
        amp *= R(0.5f);
        fx *= R(2.0f); fz *= R(2.0f);
    }
    return sum / norm;
}
template <class R> inline R hf_image_height(const HeightField& hf, R x, R z) {  // heightfield.h:57-80
    R lx = (x - R(hf.center.x)) / R(hf.size_x) + R(0.5f);
    R lz = (z - R(hf.center.z)) / R(hf.size_z) + R(0.5f);
    if (lx < R(0) || lx > R(1) || lz < R(0) || lz > R(1)) return R(0);
    if (hf.img_w < 1 || hf.img_h < 1) return R(0);
    R fu = lx * R(hf.img_w - 1);
    R fv = lz * R(hf.img_h - 1);
    int u0 = (int)fu, v0 = (int)fv;
    int u1 = (u0 < hf.img_w - 1) ? u0 + 1 : u0;
    int v1 = (v0 < hf.img_h - 1) ? v0 + 1 : v0;
    R tx = fu - R(u0), ty = fv - R(v0);
    auto pix = [&](int u, int v) { return R(hf.image[v * hf.img_w + u]) / R(255.0f); };
    R h00 = pix(u0, v0), h10 = pix(u1, v0), h01 = pix(u0, v1), h11 = pix(u1, v1);
    R h = h00 * (R(1) - tx) * (R(1) - ty) + h10 * tx * (R(1) - ty) + h01 * (R(1) - tx) * ty + h11 * tx * ty;
    return h * R(hf.height);
}
template <class R> inline R hf_eval(const HeightField& hf, R x, R z) {  // heightfield.h:84-98
    R lx = (x - R(hf.center.x)) / R(hf.size_x) + R(0.5f);
    R lz = (z - R(hf.center.z)) / R(hf.size_z) + R(0.5f);
    if (lx < R(0) || lx > R(1) || lz < R(0) || lz > R(1)) return R(0);
    if (hf.image) return hf_image_height(hf, x, z);
    R t = hf_fbm(lx, lz, R(hf.freq), hf.octaves, hf.seed);
    R detail = hf_value_noise(lx * R(7.0f), lz * R(7.0f), hf.seed + (PLANT(P_HF_DETAIL) ? 777 : 7777)) * R(0.06f);
    t = fmin_(t + detail - R(0.03f), R(1.0f));
    return t * R(hf.height);
}
// reference primitives/heightfield.h:103-183
template <class R> inline bool intersect_heightfield(const Ray<R>& ray, const HeightField& hf, Hit<R>& rec) {
    R bmnx = R(hf.center.x - 0.5f * hf.size_x), bmny = R(hf.center.y), bmnz = R(hf.center.z - 0.5f * hf.size_z);
    R bmxx = R(hf.center.x + 0.5f * hf.size_x), bmxy = R(hf.center.y + hf.height), bmxz = R(hf.center.z + 0.5f * hf.size_z);
    R t0, t1;
    {
        R invDx = R(1) / ray.d.x, invDy = R(1) / ray.d.y, invDz = R(1) / ray.d.z;
        R tx0 = (bmnx - ray.o.x) * invDx, tx1 = (bmxx - ray.o.x) * invDx;
        R ty0 = (bmny - ray.o.y) * invDy, ty1 = (bmxy - ray.o.y) * invDy;
        R tz0 = (bmnz - ray.o.z) * invDz, tz1 = (bmxz - ray.o.z) * invDz;
        if (tx0 > tx1) std::swap(tx0, tx1);
        if (ty0 > ty1) std::swap(ty0, ty1);
        if (tz0 > tz1) std::swap(tz0, tz1);
        t0 = fmax_(fmax_(tx0, ty0), tz0);
        t1 = fmin_(fmin_(tx1, ty1), tz1);
        if (t1 < ray.tmin || t0 > ray.tmax || t1 < t0) return false;
        t0 = fmax_(t0, ray.tmin);
        t1 = fmin_(t1, ray.tmax);
    }
    R diag = std::sqrt(R(hf.size_x) * R(hf.size_x) + R(hf.size_z) * R(hf.size_z));
    R step = fmax_(diag / R(64.0f), R(1e-4f));
    R t = t0;
    bool inside = false;
    for (int i = 0; i < 256; i++) {
        V3<R> p = ray.o + t * ray.d;
        R hgt = hf_eval(hf, p.x, p.z) + R(hf.center.y);
        bool iy = (p.y <= hgt);
        if (iy != inside) {
            R lo = t - step, hi = t;
            if (lo < t0) lo = t0;
            for (int b = 0; b < 10; b++) {
                R mid = R(0.5f) * (lo + hi);
                V3<R> pm = ray.o + mid * ray.d;
                R hm = hf_eval(hf, pm.x, pm.z) + R(hf.center.y);
                if ((pm.y <= hm) == iy) hi = mid; else lo = mid;
            }
            R tt = R(0.5f) * (lo + hi);
            V3<R> P = ray.o + tt * ray.d;
            if (tt >= ray.tmin && tt <= ray.tmax) {
                rec.t = tt;
                rec.point = P;
                R e = fmax_(R(hf.size_x), R(hf.size_z)) / R(256.0f);
                R hxp = hf_eval(hf, P.x + e, P.z);
                R hxm = hf_eval(hf, P.x - e, P.z);
                R hzp = hf_eval(hf, P.x, P.z + e);
                R hzm = hf_eval(hf, P.x, P.z - e);
                rec.normal = normalize(mk<R>(-(hxp - hxm) / (R(2.0f) * e), R(1.0f), -(hzp - hzm) / (R(2.0f) * e)));
                rec.tangent = mk<R>(1, 0, 0);
                rec.uvx = P.x; rec.uvy = P.z;
                if (dot(rec.normal, ray.d) > R(0)) rec.normal = -rec.normal;
                return true;
            }
            break;
        }
        inside = iy;
        t += step;
        if (t > t1) break;
    }
    return false;
}
// reference primitives/lathe.h:14-55 (segments on the Y axis; centre ignored)
template <class R> inline bool intersect_lathe(const Ray<R>& ray, const Lathe& l, const LathePoint* pts, Hit<R>& rec) {
    if (l.point_count < 2) return false;
    bool any = false;
    R closest = ray.tmax;
    Hit<R> tmp;
    for (int i = 0; i < l.point_count - 1; i++) {
        const LathePoint& p0 = pts[l.point_offset + i];
        const LathePoint& p1 = pts[l.point_offset + i + 1];
        if (p0.radius <= 0.0f && p1.radius <= 0.0f) continue;
        if (intersect_cone_segment(ray, mk<R>(0, R(p0.y), 0), R(p0.radius), mk<R>(0, R(p1.y), 0), R(p1.radius), tmp)) {
            if (tmp.t < closest) {
                tmp.point = ray.o + tmp.t * ray.d;
                rec = tmp;
                rec.material_id = l.material_id;
                rec.uvx = 0; rec.uvy = 0;
                closest = tmp.t;
                any = true;
            }
        }
    }
    return any;
}
// reference primitives/prism.h:18-73
inline bool prism_in_contour(float x, float z, const PrismPoint* pts, int base, int count) {
    bool inside = false;
    int j = count - 1;
    for (int i = 0; i < count; i++) {
        const PrismPoint& pi = pts[base + i];
        const PrismPoint& pj = pts[base + j];
        if ((pi.z > z) != (pj.z > z)) {
            float xint = (pj.x - pi.x) * (z - pi.z) / (pj.z - pi.z) + pi.x;
            if (x < xint) inside = !inside;
        }
        j = i;
    }
    return inside;
}
inline bool prism_point_inside(float x, float z, const Prism& p, const PrismPoint* all) {
    int base = p.point_offset, cnt = p.point_count, i = 0;
    bool inside = false;
    while (i < cnt) {
        int start = i, j = i + 1;
        while (j < cnt) {
            if (all[base + j].x == all[base + start].x && all[base + j].z == all[base + start].z) break;
            j++;
        }
        int clen = j - start;
        if (clen >= 3 && prism_in_contour(x, z, all, base + start, clen)) inside = !inside;
        i = (j < cnt) ? j + 1 : j;
    }
    return inside;
}
// reference primitives/prism.h:76-134 (evaluated in float: the contour test is float-only)
template <class R> inline bool intersect_prism(const Ray<R>& ray, const Prism& p, const PrismPoint* pts, Hit<R>& rec) {
    R y0 = R(p.center.y + p.y_min);
    R y1 = R(p.center.y + p.y_max);
    R ty0 = (y0 - ray.o.y) / ray.d.y;
    R ty1 = (y1 - ray.o.y) / ray.d.y;
    if (ty0 > ty1) std::swap(ty0, ty1);
    R tmin = fmax_(ty0, ray.tmin);
    R tmax = fmin_(ty1, ray.tmax);
    if (tmin > tmax) return false;
    bool any = false;
    R closest = ray.tmax;
    Hit<R> temp;
    for (int face = 0; face < 2; face++) {
        R yf = (face == 0) ? y0 : y1;
        R t = (yf - ray.o.y) / ray.d.y;
        if (t < ray.tmin || t > ray.tmax) continue;
        V3<R> P = ray.o + t * ray.d;
        if (!prism_point_inside(float(P.x - R(p.center.x)), float(P.z - R(p.center.z)), p, pts)) continue;
        temp.t = t; temp.point = P;
        temp.normal = mk<R>(0, (face == 0) ? R(-1) : R(1), 0);
        if (dot(temp.normal, ray.d) > R(0)) temp.normal = -temp.normal;
        temp.tangent = mk<R>(1, 0, 0);
        if (t < closest) { rec = temp; rec.material_id = p.material_id; rec.uvx = 0; rec.uvy = 0; closest = t; any = true; }
    }
    R dx = ray.d.x, dy = ray.d.y, dz = ray.d.z, ox = ray.o.x, oz = ray.o.z;
    int base = p.point_offset, cnt = p.point_count;
    for (int i = 0; i < cnt; i++) {
        int j = (i + 1) % cnt;
        const PrismPoint& pi = pts[base + i];
        const PrismPoint& pj = pts[base + j];
        if (pi.x == pj.x && pi.z == pj.z) continue;
        R ax = R(pi.x + p.center.x), az = R(pi.z + p.center.z);
        R bx = R(pj.x + p.center.x), bz = R(pj.z + p.center.z);
        R ex = bx - ax, ez = bz - az;
        R denom = ex * dz - ez * dx;
        if (std::fabs(denom) < R(1e-9f)) continue;
        R s = (dx * (az - oz) - dz * (ax - ox)) / denom;
        if (s < R(-0.001f) || s > R(1.001f)) continue;
        R t;
        if (std::fabs(dx) > R(1e-9f)) t = (ax + s * ex - ox) / dx;
        else if (std::fabs(dz) > R(1e-9f)) t = (az + s * ez - oz) / dz;
        else continue;
        if (t < ray.tmin || t > ray.tmax) continue;
        R yy = ray.o.y + dy * t;
        if (yy < y0 - R(0.001f) || yy > y1 + R(0.001f)) continue;
        temp.t = t; temp.point = ray.o + t * ray.d;
        temp.normal = normalize(mk<R>(ez, R(0), -ex));
        temp.tangent = mk<R>(0, 1, 0);
        if (dot(temp.normal, ray.d) > R(0)) temp.normal = -temp.normal;
        if (t < closest) { rec = temp; rec.material_id = p.material_id; rec.uvx = 0; rec.uvy = 0; closest = t; any = true; }
    }
    return any;
}

// ---------------------------------------------------------------------------
// CSG (reference integrator/csg.h:22-186), recursive as in the reference
// ---------------------------------------------------------------------------
template <class R> struct CIv { R te, tx; int mat, leaf, kind; };
template <class R> inline bool csg_leaf(const Ray<R>& ray, int kind, int index, const SheenScene& s, R t0, R t1, CIv<R>& out) {
    out.te = R(-1e30f); out.tx = R(-1e30f);
    if (kind == 0) {
        const Object& o = s.objects[index];
        return sphere_intervals(ray, cv<R>(o.param1), R(o.param2), t0, t1, out.te, out.tx);
    } else if (kind == 1) {
        return se_intervals(ray, s.superellipsoids[index], t0, t1, out.te, out.tx);
    }
    return false;
}
template <class R> inline void csg_merge(const CIv<R>* A, int nA, const CIv<R>* B, int nB, int op, CIv<R>* out, int& nOut) {
    nOut = 0;
    if (op == 0) {
        CIv<R> tmp[16];
        int n = 0;
        for (int i = 0; i < nA && n < 16; i++) tmp[n++] = A[i];
        for (int i = 0; i < nB && n < 16; i++) tmp[n++] = B[i];
        for (int i = 1; i < n; i++) {
            CIv<R> key = tmp[i];
            int j = i - 1;
            while (j >= 0 && tmp[j].te > key.te) { tmp[j + 1] = tmp[j]; j--; }
            tmp[j + 1] = key;
        }
        for (int i = 0; i < n && nOut < 8; i++) {
            if (nOut == 0 || tmp[i].te > out[nOut - 1].tx) out[nOut++] = tmp[i];
            else if (tmp[i].tx > out[nOut - 1].tx) out[nOut - 1].tx = tmp[i].tx;
        }
    } else if (op == 1) {
        for (int i = 0; i < nA; i++)
            for (int j = 0; j < nB; j++) {
                R t0 = fmax_(A[i].te, B[j].te), t1 = fmin_(A[i].tx, B[j].tx);
                if (t0 < t1 && nOut < 8) { out[nOut] = A[i]; out[nOut].te = t0; out[nOut].tx = t1; nOut++; }
            }
    } else if (op == 2) {
        for (int i = 0; i < nA; i++) {
            R cur = A[i].te;
            for (int j = 0; j < nB; j++) {
                R b0 = B[j].te, b1 = B[j].tx;
                if (b1 <= cur) continue;
                if (b0 >= A[i].tx) break;
                if (b0 > cur && nOut < 8) { out[nOut] = A[i]; out[nOut].te = cur; out[nOut].tx = b0; nOut++; }
                cur = fmax_(cur, b1);
                if (cur >= A[i].tx) break;
            }
            if (cur < A[i].tx && nOut < 8) { out[nOut] = A[i]; out[nOut].te = cur; out[nOut].tx = A[i].tx; nOut++; }
        }
    }
}
template <class R> inline void csg_collect(const SheenScene& s, int root, const Ray<R>& ray, R t0, R t1, CIv<R>* out, int& nOut) {
    const CSGNode& n = s.csg_nodes[root];
    if (n.leaf_kind >= 0) {
        nOut = 0;
        CIv<R> iv;
        if (csg_leaf(ray, n.leaf_kind, n.leaf_index, s, t0, t1, iv)) {
            iv.mat = (n.leaf_kind == 0) ? s.objects[n.leaf_index].material_id
                   : (n.leaf_kind == 1) ? s.superellipsoids[n.leaf_index].material_id : 0;
            iv.leaf = n.leaf_index;
            iv.kind = n.leaf_kind;
            out[nOut++] = iv;
        }
        return;
    }
    CIv<R> A[8], B[8];
    int nA = 0, nB = 0;
    if (n.left >= 0) csg_collect(s, n.left, ray, t0, t1, A, nA);
    if (n.right >= 0) csg_collect(s, n.right, ray, t0, t1, B, nB);
    csg_merge(A, nA, B, nB, n.op, out, nOut);
}
template <class R> inline bool intersect_csg(const Ray<R>& ray, const SheenScene& s, int root, R t0, R t1, Hit<R>& rec) {
    CIv<R> iv[8];
    int count = 0;
    csg_collect(s, root, ray, t0, t1, iv, count);
    R best = R(1e30f);
    int bi = -1;
    for (int i = 0; i < count; i++)
        if (iv[i].te < best && iv[i].tx > t0) { best = iv[i].te; bi = i; }
    if (bi < 0) return false;
    const CIv<R>& h = iv[bi];
    R t = fmax_(h.te, t0);
    if (t < t0 || t > t1) return false;
    rec.t = t;
    rec.material_id = h.mat;
    rec.point = ray.o + t * ray.d;
    if (h.kind == 0) {
        rec.normal = normalize(rec.point - cv<R>(s.objects[h.leaf].param1));
        rec.tangent = mk<R>(0, 0, 0);
    } else if (h.kind == 1) {
        rec.normal = normalize(se_grad(s.superellipsoids[h.leaf], rec.point));
        rec.tangent = mk<R>(1, 0, 0);
    } else {
        rec.normal = mk<R>(0, 1, 0);
        rec.tangent = mk<R>(1, 0, 0);
    }
    rec.uvx = rec.point.x; rec.uvy = rec.point.y;
    if (dot(rec.normal, ray.d) > R(0)) rec.normal = -rec.normal;
    return true;
}

// ---------------------------------------------------------------------------
// Scene queries (reference bvh.h:605-922, integrator/raycast.h:14-97)
// ---------------------------------------------------------------------------
struct Counters { uint64_t primary = 0, shadow = 0, reflection = 0, refraction = 0, photon = 0; };

template <class R> inline bool prim_hit(const Scene& S, int cat, int i, const Ray<R>& ray, Hit<R>& h) {
    const SheenScene& s = S.s;
    switch (cat) {
        case K_SPHERE: {
            const Object& o = s.objects[i];
            if (o.type != 0) return false;
            if (!intersect_sphere(ray, cv<R>(o.param1), R(o.param2), h)) return false;
            h.material_id = o.material_id; h.object_id = i;
            return true;
        }
        case K_STRAND:
            if (!intersect_sphere_sweep(ray, s.sweep_strands[i], s.sweep_points, h)) return false;
            h.object_id = -(i + 1);
            return true;
        case K_TRI:
            if (!intersect_triangle(ray, s.triangles[i], h)) return false;
            h.material_id = s.triangles[i].material_id; h.object_id = -(i + 1);
            return true;
        case K_HF:
            if (!intersect_heightfield(ray, s.heightfields[i], h)) return false;
            h.material_id = s.heightfields[i].material_id; h.object_id = -(i + 1);
            return true;
        case K_SE:
            if (!intersect_superellipsoid(ray, s.superellipsoids[i], h)) return false;
            h.material_id = s.superellipsoids[i].material_id; h.object_id = -(i + 1);
            return true;
        case K_LATHE:
            if (!intersect_lathe(ray, s.lathes[i], s.lathe_points, h)) return false;
            h.object_id = -(i + 1);
            return true;
        case K_PRISM:
            if (!intersect_prism(ray, s.prisms[i], s.prism_points, h)) return false;
            h.object_id = -(i + 1);
            return true;
    }
    return false;
}

// Closest hit within one category: every primitive whose own bounds pass the
// reference slab test is a candidate (the set the reference's DFS reaches);
// ties at equal t go to the lowest index.
template <class R> inline bool category_closest(const Scene& S, int cat, const Ray<R>& ray, Hit<R>& rec) {
    const std::vector<BNode>& nodes = S.bvh[cat];
    if (nodes.empty()) return false;
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    bool any = false;
    R best = ray.tmax;
    int best_idx = -1;
    Hit<R> tmp;
    while (sp > 0) {
        const BNode& n = nodes[stack[--sp]];
        if (!intersect_aabb(ray, n.mn, n.mx)) continue;
        if (n.left < 0) {
            int pi = ~n.left;
            if (prim_hit(S, cat, pi, ray, tmp)) {
                if (tmp.t < best || (tmp.t == best && any && pi < best_idx)) {
                    rec = tmp; rec.kind = cat; rec.index = pi;
                    best = tmp.t; best_idx = pi; any = true;
                }
            }
        } else {
            stack[sp++] = n.left;
            stack[sp++] = n.right;
        }
    }
    return any;
}
template <class R> inline bool planes_closest(const Scene& S, const Ray<R>& ray, Hit<R>& rec) {  // intersect.h:11-34
    bool any = false;
    R closest = ray.tmax;
    Hit<R> tmp;
    for (int i = 0; i < S.s.num_planes; i++) {
        const Plane& p = S.s.planes[i];
        if (intersect_plane(ray, cv<R>(p.normal), R(p.d), tmp) && tmp.t < closest) {
            rec = tmp; rec.material_id = p.material_id; rec.object_id = -1 - i;
            rec.kind = K_PLANE; rec.index = i;
            closest = tmp.t; any = true;
        }
    }
    return any;
}

// reference integrator/raycast.h:14-70
template <class R> inline bool scene_intersect(const Scene& S, const Ray<R>& ray, Hit<R>& rec) {
    R best = ray.tmax;
    bool hit = false;
    int carry_obj = 0;
    Hit<R> tmp;
    static const int order[7] = {K_SPHERE, K_STRAND, K_TRI, K_HF, K_SE, K_LATHE, K_PRISM};
    for (int c : order) {
        if (category_closest(S, c, ray, tmp)) {
            carry_obj = tmp.object_id;
            if (tmp.t < best) { best = tmp.t; rec = tmp; hit = true; }
        }
    }
    if (S.s.csg_root >= 0) {
        if (intersect_csg(ray, S.s, S.s.csg_root, ray.tmin, ray.tmax, tmp)) {
            tmp.object_id = carry_obj; tmp.kind = K_CSG; tmp.index = 0;
            if (tmp.t < best) { best = tmp.t; rec = tmp; hit = true; }
        }
    }
    if (planes_closest(S, ray, tmp) && tmp.t < best) { best = tmp.t; rec = tmp; hit = true; }
    return hit;
}
// reference integrator/raycast.h:74-97
template <class R> inline bool scene_intersect_opaque(const Scene& S, const Ray<R>& ray, Hit<R>& rec) {
    static const int order[6] = {K_SPHERE, K_TRI, K_HF, K_SE, K_LATHE, K_PRISM};
    for (int c : order)
        if (category_closest(S, c, ray, rec)) return true;
    if (S.s.csg_root >= 0) {
        if (intersect_csg(ray, S.s, S.s.csg_root, ray.tmin, ray.tmax, rec)) {
            rec.object_id = 0; rec.kind = K_CSG; rec.index = 0;
            return true;
        }
    }
    return planes_closest(S, ray, rec);
}
// reference bvh.h:928-960: optical depth summed over the strands whose
// bounds pass the slab test, stopping once it exceeds 6.  The reference sums
// in its BVH's DFS order; here (and in the HIP path) strands are summed in
// index order, so the value after the early stop is defined independently of
// any tree.
template <class R> inline R strand_bvh_tau(const Scene& S, const Ray<R>& ray, R sigma) {
    R tau = 0;
    for (int i = 0; i < S.s.num_sweep_strands && S.s.sweep_strands; i++) {
        Box b = bounds_of(S, K_STRAND, i);
        if (!intersect_aabb(ray, b.mn, b.mx)) continue;
        tau += strand_tau(ray, S.s.sweep_strands[i], S.s.sweep_points, sigma);
        if (tau > R(6.0f) && !PLANT(P_TAU_NO_STOP)) return tau;
    }
    return tau;
}

// ---------------------------------------------------------------------------
// Shading (reference integrator/shading.h)
// ---------------------------------------------------------------------------
template <class R> inline R tex_lod(const Tex& t, R uv_fp) {  // shading.h:20-22
    return log2_r(fmax_(uv_fp * R(t.w), R(1.0f)));
}
template <class R> inline V3<R> sample_texture_lod(const Tex& t, R uvx, R uvy, R lod) {  // shading.h:26-58
    int level = (int)std::floor(lod + R(0.5f));
    if (level < 0) level = 0;
    if (level >= t.mips) level = t.mips - 1;
    int w = t.w, h = t.h;
    size_t off = 0;
    for (int i = 0; i < level; i++) {
        off += (size_t)w * h * 3;
        w = (w > 1) ? w / 2 : 1;
        h = (h > 1) ? h / 2 : 1;
    }
    const unsigned char* base = t.px.data() + off;
    const R* fbase = t.use_float ? float_chain<R>(t) + off : nullptr;
    R u = uvx - std::floor(uvx);
    R v = R(1.0f) - (uvy - std::floor(uvy));
    R fx = u * R(w - 1);
    R fy = v * R(h - 1);
    int x0 = (int)fx, y0 = (int)fy;
    int x1 = (x0 + 1) % w;
    int y1 = (y0 + 1) % h;
    R tx = fx - R(x0), ty = fy - R(y0);
    auto at = [&](int x, int y) {
        if (fbase) {
            const R* q = fbase + (size_t)(y * w + x) * 3;
            return mk<R>(q[0], q[1], q[2]);
        }
        const unsigned char* p = base + (size_t)(y * w + x) * 3;
        return mk<R>(R(p[0]) / R(255.0f), R(p[1]) / R(255.0f), R(p[2]) / R(255.0f));
    };
    V3<R> c00 = at(x0, y0), c10 = at(x1, y0), c01 = at(x0, y1), c11 = at(x1, y1);
    V3<R> a = c00 * (R(1.0f) - tx) + c10 * tx;
    V3<R> b = c01 * (R(1.0f) - tx) + c11 * tx;
    return a * (R(1.0f) - ty) + b * ty;
}
inline bool tex_ok(const Scene& S, int id) { return id >= 0 && id < (int)S.tex.size(); }

template <class R> struct ShadeIn { V3<R> point, normal, tangent, view; R uvx, uvy, uv_fp; };

template <class R> inline V3<R> albedo(const Scene& S, const Material& m, const ShadeIn<R>& in) {  // shading.h:64-71
    if (tex_ok(S, m.texture_id)) {
        const Tex& t = S.tex[m.texture_id];
        return sample_texture_lod(t, in.uvx, in.uvy, tex_lod(t, in.uv_fp));
    }
    return cv<R>(m.diffuse_color);
}
inline bool has_mask(const Scene& S, const Material& m) { return tex_ok(S, m.mask_id); }
template <class R> inline R roughness_to_shininess(R r) {  // shading.h:85-89
    R rr = fmax_(r, R(0.02f));
    R r2 = rr * rr;
    return fmin_(fmax_(R(2.0f) / (PLANT(P_ROUGH_R2) ? r2 : r2 * r2) - R(2.0f), R(1.0f)), R(4096.0f));
}
template <class R> inline R hit_transmission(const Scene& S, const Hit<R>& h, const Material& m) {  // shading.h:92-99
    if (m.transmission <= 0.0f) return R(0);
    if (has_mask(S, m) && !PLANT(P_TRANS_NO_MASK))
        return R(m.transmission) * sample_texture_lod(S.tex[m.mask_id], h.uvx, h.uvy, R(0)).x;
    return R(m.transmission);
}
template <class R> inline R hit_ior(const Scene& S, const Hit<R>& h, const Material& m) {  // shading.h:102-108
    if (has_mask(S, m)) {
        V3<R> mk_ = sample_texture_lod(S.tex[m.mask_id], h.uvx, h.uvy, R(0));
        return R(1.0f) + (PLANT(P_IOR_RED) ? mk_.x : mk_.y);
    }
    return R(m.ior);
}
template <class R> inline V3<R> shade_hair(const Material& m, const ShadeIn<R>& in, V3<R> L, V3<R> rad) {  // shading.h:112-152
    V3<R> T = in.tangent, N = in.normal, V = in.view;
    R tl = dot(T, L);
    R sin_tl = std::sqrt(fmax_(R(1.0f) - tl * tl, R(0)));
    R back = fmax_(-tl, R(0));
    R diffuse = sin_tl + R(m.hair_backlight) * back * back;
    V3<R> result = cv<R>(m.diffuse_color) * rad * diffuse * R(0.75f);
    R w_r = R(m.hair_spec_r), w_t = R(m.hair_spec_t), e_r = R(m.hair_exp_r), e_t = R(m.hair_exp_t);
    if (w_r <= R(0) && w_t <= R(0)) { w_r = R(m.specular); e_r = R(m.shininess); }
    V3<R> H = normalize(L + V);
    if (w_r > R(0)) {
        V3<R> Tr = normalize(T + R(m.hair_shift) * N);
        R sr = dot(Tr, H);
        R rr = pow_r(std::sqrt(fmax_(R(1.0f) - sr * sr, R(0))), fmax_(e_r, R(1.0f)));
        result = result + rad * w_r * rr;
    }
    if (w_t > R(0)) {
        V3<R> Tt = normalize(T - R(m.hair_shift) * N);
        R st = dot(Tt, H);
        R rt = pow_r(std::sqrt(fmax_(R(1.0f) - st * st, R(0))), fmax_(e_t, R(1.0f)));
        result = result + rad * w_t * rt;
    }
    return result;
}
template <class R> inline V3<R> shade_direct(const Scene& S, const Material& m, const ShadeIn<R>& in, V3<R> L, V3<R> rad, R shin) {  // shading.h:158-175
    if (m.anisotropy > 0.0f) return shade_hair(m, in, L, rad);
    V3<R> base = albedo(S, m, in);
    R diff = fmax_(dot(in.normal, L), R(0));
    if (PLANT(P_NO_COS)) diff = diff > R(0) ? R(1) : R(0);
    V3<R> result = base * rad * diff;
    if (m.specular > 0.0f && (diff > R(0) || PLANT(P_SPEC_NO_GUARD))) {
        V3<R> hv = normalize(L + in.view);
        R spec = pow_r(fmax_(dot(in.normal, PLANT(P_SPEC_NL) ? L : hv), R(0)), shin);
        result = result + rad * R(m.specular) * spec;
    }
    return result;
}

// ---------------------------------------------------------------------------
// Photon map (reference render_kernel.h:55-111, render_kernel.cu:820-1045)
// ---------------------------------------------------------------------------
inline int photon_cell_hash(int ix, int iy, int iz, int num_cells) {
    uint32_t h = (uint32_t(ix) * 73856093u) ^ (uint32_t(iy) * 19349663u) ^ (uint32_t(iz) * 83492791u);
    h ^= h >> 16; h *= 0x7feb352du;
    h ^= h >> 15; h *= 0x846ca68bu;
    h ^= h >> 16;
    return (int)(h & (uint32_t)(num_cells - 1));
}
inline int photon_cell_of(const Photon& p, float cell, int num_cells) {  // render_kernel.cu:963-966
    int ix = (int)std::floor(p.position.x / cell);
    int iy = (int)std::floor(p.position.y / cell);
    int iz = (int)std::floor(p.position.z / cell);
    return photon_cell_hash(ix, iy, iz, num_cells);
}
template <class R> inline V3<R> photon_gather(const Photons& m, V3<R> p, V3<R> n, R radius, int max_samples, int* count_out) {
    if (!m.valid()) { if (count_out) *count_out = 0; return mk<R>(0, 0, 0); }
    R cs = R(m.cell);
    int x0 = (int)std::floor((p.x - radius) / cs), x1 = (int)std::floor((p.x + radius) / cs);
    int y0 = (int)std::floor((p.y - radius) / cs), y1 = (int)std::floor((p.y + radius) / cs);
    int z0 = (int)std::floor((p.z - radius) / cs), z1 = (int)std::floor((p.z + radius) / cs);
    R r2 = radius * radius;
    V3<R> acc = mk<R>(0, 0, 0);
    int taken = 0;
    int cells[8];
    int ncell = 0;
    for (int iz = z0; iz <= z1; iz++)
        for (int iy = y0; iy <= y1; iy++)
            for (int ix = x0; ix <= x1; ix++) {
                int c = photon_cell_hash(ix, iy, iz, m.num_cells);
                bool seen = false;
                for (int q = 0; q < ncell; q++) if (cells[q] == c) { seen = true; break; }
                if (!seen && ncell < 8) cells[ncell++] = c;
            }
    for (int q = 0; q < ncell && taken < max_samples; q++) {
        int c = cells[q];
        int start = m.cell_start[c], cnt = m.cell_count[c];
        for (int k = 0; k < cnt && taken < max_samples; k++) {
            const Photon& ph = m.sorted[start + k];
            V3<R> d = cv<R>(ph.position) - p;
            if (dot(d, d) > r2) continue;
            if (dot(cv<R>(ph.normal), n) < R(0.1f)) continue;
            acc = acc + cv<R>(ph.flux);
            taken++;
        }
    }
    if (count_out) *count_out = taken;
    return acc;
}

// Photon emission for one photon index (reference render_kernel.cu:820-926).
// Returns true and fills out when the photon is stored.
inline bool photon_trace(const Scene& S, int idx, int per_light, Photon& out, uint64_t& rays) {
    const SheenScene& s = S.s;
    const Light& light = s.lights[idx / per_light];
    if (light.type == 0) return false;
    uint32_t sd = aa_hash(0x5eed1234u + uint32_t(idx) * 2654435761u);
    float u1 = aa_rand(sd), u2 = aa_rand(sd ^ 0x9e3779b9u), u3 = aa_rand(sd ^ 0x85ebca6bu);
    typedef float R;
    V3<R> pos = cv<R>(light.position), dir;
    if (light.type == 3) {
        R cos_half = cos_r(light.angle);
        R ct = cos_half + (1.0f - cos_half) * u1;
        R st = std::sqrt(fmax_(1.0f - ct * ct, 0.0f));
        R phi = 6.2831853f * u2;
        V3<R> ax = cv<R>(light.direction);
        V3<R> up = (std::fabs(ax.y) < 0.9f) ? mk<R>(0, 1, 0) : mk<R>(1, 0, 0);
        V3<R> t1 = normalize(cross(ax, up));
        V3<R> t2 = cross(ax, t1);
        dir = normalize(ax * ct + t1 * (cos_r(phi) * st) + t2 * (sin_r(phi) * st));
    } else if (light.type == 2) {
        pos = cv<R>(light.position) + cv<R>(light.u_axis) * (u1 * 2.0f - 1.0f) + cv<R>(light.v_axis) * (u2 * 2.0f - 1.0f);
        V3<R> n = normalize(cross(cv<R>(light.u_axis), cv<R>(light.v_axis)));
        V3<R> up = (std::fabs(n.y) < 0.9f) ? mk<R>(0, 1, 0) : mk<R>(1, 0, 0);
        V3<R> t1 = normalize(cross(n, up));
        V3<R> t2 = cross(n, t1);
        R cz = std::sqrt(fmax_(1.0f - u3, 0.0f));
        R r = std::sqrt(u3);
        R phi = 6.2831853f * aa_rand(sd ^ 0xc2b2ae35u);
        dir = normalize(n * cz + t1 * (cos_r(phi) * r) + t2 * (sin_r(phi) * r));
    } else {
        R cz = 2.0f * u1 - 1.0f;
        R st = std::sqrt(fmax_(1.0f - cz * cz, 0.0f));
        R phi = 6.2831853f * u2;
        dir = mk<R>(st * cos_r(phi), cz, st * sin_r(phi));
    }
    V3<R> flux = cv<R>(light.color) * (light.intensity / (float)per_light);
    Ray<R> ray{pos, dir, 0.001f, 1.0e10f};
    for (int bounce = 0; bounce < 4; bounce++) {
        Hit<R> hit;
        rays++;
        if (!scene_intersect(S, ray, hit)) break;
        const Material& m = s.materials[hit.material_id];
        if (m.transmission > 0.5f) {
            V3<R> n = hit.normal;
            R ior = fmax_(m.ior, 1.0f);
            R eta = 1.0f / ior;
            R cosi = fmax_(-dot(ray.d, n), 1.0e-4f);
            R k = 1.0f - eta * eta * (1.0f - cosi * cosi);
            V3<R> td = (k < 0.0f) ? normalize(ray.d + 2.0f * cosi * n) : normalize(eta * ray.d + (eta * cosi - std::sqrt(k)) * n);
            flux = flux * m.transmission;
            ray.o = hit.point + td * 0.002f;
            ray.d = td;
            continue;
        }
        if (m.reflect > 0.5f) {
            V3<R> rd = ray.d - hit.normal * (2.0f * dot(ray.d, hit.normal));
            ray.o = hit.point + hit.normal * 0.002f;
            ray.d = normalize(rd);
            continue;
        }
        out.position = {hit.point.x, hit.point.y, hit.point.z};
        out.direction = {-ray.d.x, -ray.d.y, -ray.d.z};
        out.normal = {hit.normal.x, hit.normal.y, hit.normal.z};
        out.flux = {flux.x, flux.y, flux.z};
        return true;
    }
    return false;
}

// Photon pass + deterministic hash-grid map: stored photons in emission-index
// order (the reference's atomicAdd slot order is run-dependent), counting sort
// by cell that keeps that order inside each cell.
inline void build_photons(Scene& S, int threads) {
    const SheenScene& s = S.s;
    if (s.photons <= 0 || s.num_lights <= 0) return;
    long long cap = (long long)s.photons * s.num_lights;
    if (cap > 4000000) cap = 4000000;  // sheen_c_api.cu:276-277
    int total = s.num_lights * s.photons;
    std::vector<Photon> ph(total);
    std::vector<char> ok(total, 0);
    uint64_t rays = 0;
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads) reduction(+ : rays)
    for (int i = 0; i < total; i++) {
        uint64_t r = 0;
        ok[i] = photon_trace(S, i, s.photons, ph[i], r) ? 1 : 0;
        rays += r;
    }
    std::vector<Photon> stored;
    for (int i = 0; i < total; i++)
        if (ok[i] && (long long)stored.size() < cap) stored.push_back(ph[i]);
    S.photon_emitted = total;
    S.photon_stored = (int)stored.size();
    S.photon_rays = rays;
    if (stored.empty()) return;
    float cell = s.photon_cell > 0.0f ? s.photon_cell : 0.25f;
    int nc = 1 << 18;
    S.pm.cell = cell;
    S.pm.num_cells = nc;
    S.pm.cell_count.assign(nc, 0);
    S.pm.cell_start.assign(nc + 1, 0);
    std::vector<int> cid(stored.size());
    for (size_t i = 0; i < stored.size(); i++) { cid[i] = photon_cell_of(stored[i], cell, nc); S.pm.cell_count[cid[i]]++; }
    int acc = 0;
    for (int c = 0; c < nc; c++) { S.pm.cell_start[c] = acc; acc += S.pm.cell_count[c]; }
    S.pm.cell_start[nc] = acc;
    std::vector<int> cur(S.pm.cell_start.begin(), S.pm.cell_start.end() - 1);
    S.pm.sorted.resize(stored.size());
    for (size_t i = 0; i < stored.size(); i++) S.pm.sorted[cur[cid[i]]++] = stored[i];
}

// ---------------------------------------------------------------------------
// Lighting (reference render_kernel.cu:13-298)
// ---------------------------------------------------------------------------
template <class R> inline V3<R> sun_dir(const Light& light, uint32_t seed) {  // render_kernel.cu:13-24
    V3<R> base = -cv<R>(light.direction);
    if (light.angle <= 0.0f) return base;
    V3<R> up = (std::fabs(base.y) < R(0.9f)) ? mk<R>(0, 1, 0) : mk<R>(1, 0, 0);
    V3<R> t1 = normalize(cross(base, up));
    V3<R> t2 = cross(base, t1);
    R r = std::sqrt(R(aa_rand(seed ^ 0x68e31da4u))) * tan_r(R(light.angle));
    R phi = R(6.2831853f) * R(aa_rand(seed ^ 0xb5297a4du));
    return normalize(base + t1 * (cos_r(phi) * r) + t2 * (sin_r(phi) * r));
}

template <class R> inline V3<R> compute_lighting(const Scene& S, const Hit<R>& rec, V3<R> view_dir, R footprint,
                                                 uint32_t seed, Counters& cnt) {
    const SheenScene& s = S.s;
    const Material& mat = s.materials[rec.material_id];
    ShadeIn<R> in;
    in.point = rec.point; in.normal = rec.normal; in.tangent = rec.tangent;
    in.uvx = rec.uvx; in.uvy = rec.uvy; in.view = view_dir;
    R uv_fp = footprint;
    V3<R> tan_u = mk<R>(0, 0, 0), tan_v = mk<R>(0, 0, 0);
    if (rec.object_id < 0 && s.triangles && s.num_triangles > 0) {  // render_kernel.cu:82-114
        int ti = -rec.object_id - 1;
        if (ti >= 0 && ti < s.num_triangles) {
            const Triangle& tr = s.triangles[ti];
            V3<R> e1 = cv<R>(tr.v1) - cv<R>(tr.v0), e2 = cv<R>(tr.v2) - cv<R>(tr.v0);
            R d1x = R(tr.uv1.x) - R(tr.uv0.x), d1y = R(tr.uv1.y) - R(tr.uv0.y);
            R d2x = R(tr.uv2.x) - R(tr.uv0.x), d2y = R(tr.uv2.y) - R(tr.uv0.y);
            R w1 = std::sqrt(dot(e1, e1)), w2 = std::sqrt(dot(e2, e2));
            R u1 = std::sqrt(d1x * d1x + d1y * d1y);
            R u2 = std::sqrt(d2x * d2x + d2y * d2y);
            R a1 = (w1 > R(1.0e-6f)) ? u1 / w1 : R(0);
            R a2 = (w2 > R(1.0e-6f)) ? u2 / w2 : R(0);
            R dens = (a1 > a2) ? a1 : a2;
            if (dens > R(1.0e-6f)) uv_fp = footprint * dens;
            R det = d1x * d2y - d2x * d1y;
            if (std::fabs(det) > R(1.0e-8f)) {
                V3<R> du = (e1 * d2y - e2 * d1y) / det;
                V3<R> dv = (e2 * d1x - e1 * d2x) / det;
                V3<R> n0 = rec.normal;
                du = du - n0 * dot(du, n0);
                dv = dv - n0 * dot(dv, n0);
                R lu = dot(du, du);
                if (lu > R(1.0e-12f)) {
                    tan_u = du * (R(1) / std::sqrt(lu));
                    dv = dv - tan_u * dot(dv, tan_u);
                    R lv = dot(dv, dv);
                    if (lv > R(1.0e-12f)) tan_v = dv * (R(1) / std::sqrt(lv));
                }
            }
        }
    }
    in.uv_fp = uv_fp;
    if (tex_ok(S, mat.normal_id)) {  // render_kernel.cu:119-136
        const Tex& nt = S.tex[mat.normal_id];
        V3<R> ntex = sample_texture_lod(nt, in.uvx, in.uvy, tex_lod(nt, in.uv_fp));
        R nx = (ntex.x * R(2.0f) - R(1.0f)) * R(mat.normal_strength);
        R ny = (ntex.y * R(2.0f) - R(1.0f)) * R(mat.normal_strength);
        V3<R> n0 = in.normal, t1, t2;
        if (dot(tan_u, tan_u) > R(0.5f) && dot(tan_v, tan_v) > R(0.5f)) { t1 = tan_u; t2 = tan_v; }
        else {
            V3<R> up = (std::fabs(n0.y) < R(0.9f)) ? mk<R>(0, 1, 0) : mk<R>(1, 0, 0);
            t1 = normalize(cross(up, n0));
            t2 = cross(n0, t1);
        }
        in.normal = normalize(n0 + t1 * nx + t2 * ny);
    }
    V3<R> base = albedo(S, mat, in);
    V3<R> result = base * R(mat.ambient) +
                   cv<R>(mat.emission_color) * (R(mat.emission) * (PLANT(P_EMISSION_LIT) ? R(mat.ambient) : R(1)));
    R shininess = R(mat.shininess);
    if (has_mask(S, mat)) {
        const Tex& mt = S.tex[mat.mask_id];
        V3<R> mask = sample_texture_lod(mt, in.uvx, in.uvy, tex_lod(mt, in.uv_fp));
        shininess = roughness_to_shininess(mask.z);
    }
    for (int i = 0; i < s.num_lights; i++) {
        const Light& light = s.lights[i];
        uint32_t ls = seed ^ (uint32_t)(uint32_t(i) * 2654435761u);
        V3<R> L, radiance, target;
        R t_max = R(1e10f);
        if (light.type == 0) {
            L = sun_dir<R>(light, ls);
            radiance = cv<R>(light.color) * R(light.intensity);
            target = rec.point + L * R(1.0e6f);
        } else {
            V3<R> sp = cv<R>(light.position);
            if (light.type == 2) {
                R ru = R(aa_rand(ls)) * R(2.0f) - R(1.0f);
                R rv = R(aa_rand(ls ^ 0x9e3779b9u)) * R(2.0f) - R(1.0f);
                if (PLANT(P_AREA_HALF)) { ru = R(aa_rand(ls)); rv = R(aa_rand(ls ^ 0x9e3779b9u)); }
                sp = cv<R>(light.position) + cv<R>(light.u_axis) * ru + cv<R>(light.v_axis) * rv;
            }
            V3<R> tl = sp - rec.point;
            R dist2 = fmax_(dot(tl, tl), R(1.0e-4f));
            R dist = std::sqrt(dist2);
            L = tl / dist;
            if ((light.type == 1 || light.type == 3) && light.radius > 0.0f) {
                V3<R> up = (std::fabs(L.y) < R(0.9f)) ? mk<R>(0, 1, 0) : mk<R>(1, 0, 0);
                V3<R> t1 = normalize(cross(L, up));
                V3<R> t2 = cross(L, t1);
                R r = std::sqrt(R(aa_rand(ls ^ 0x85ebca6bu))) * R(light.radius);
                if (PLANT(P_DISK_LINEAR)) r = R(aa_rand(ls ^ 0x85ebca6bu)) * R(light.radius);
                R phi = R(6.2831853f) * R(aa_rand(ls ^ 0xc2b2ae35u));
                sp = sp + t1 * (cos_r(phi) * r) + t2 * (sin_r(phi) * r);
                tl = sp - rec.point;
                dist2 = fmax_(dot(tl, tl), R(1.0e-4f));
                dist = std::sqrt(dist2);
                L = tl / dist;
            }
            R atten = R(light.intensity) / ((PLANT(P_TWO_PI) ? R(6.2831853f) : R(12.5663706f)) *
                                            (PLANT(P_POINT_DIST) ? dist : dist2));
            V3<R> gobo = mk<R>(1, 1, 1);
            if (light.type == 2) {
                V3<R> n = normalize(cross(cv<R>(light.u_axis), cv<R>(light.v_axis)));
                R cos_l = dot(n, L);
                if (PLANT(P_AREA_SIDE)) cos_l = -cos_l;
                if (cos_l <= R(0)) continue;
                if (!PLANT(P_AREA_NO_COS)) atten *= cos_l;
            } else if (light.type == 3) {
                V3<R> ts = normalize(rec.point - sp);
                R c = dot(ts, cv<R>(light.direction));
                R cos_half = cos_r(R(light.angle) * (PLANT(P_SPOT_HALF_ANGLE) ? R(0.5f) : R(1.0f)));
                if (c <= cos_half) continue;
                R blend = fmin_(fmax_(R(light.blend), R(0)), R(1));
                R cos_soft = cos_r(R(light.angle) * (R(1.0f) - blend));
                if (cos_soft > cos_half + R(1.0e-4f))
                    atten *= fmin_((c - cos_half) / ((PLANT(P_SPOT_RAMP) ? R(1.0f) : cos_soft) - cos_half), R(1.0f));
                if (tex_ok(S, light.texture_id)) {
                    V3<R> axis = cv<R>(light.direction);
                    V3<R> up = (std::fabs(axis.y) < R(0.9f)) ? mk<R>(0, 1, 0) : mk<R>(1, 0, 0);
                    V3<R> t1 = normalize(cross(axis, up));
                    V3<R> t2 = cross(axis, t1);
                    R tan_half = tan_r(fmax_(R(light.angle), R(1.0e-3f)));
                    R u = R(0.5f) + R(0.5f) * dot(ts, t1) / tan_half;
                    R v = R(0.5f) + R(0.5f) * dot(ts, t2) / tan_half;
                    gobo = sample_texture_lod(S.tex[light.texture_id], u, v, R(0));
                }
            }
            radiance = cv<R>(light.color) * atten * gobo;
            target = sp;
            t_max = dist * R(0.999f);
        }
        Ray<R> shadow;
        shadow.o = rec.point + rec.normal * R(0.002f);
        shadow.d = normalize(target - shadow.o);
        shadow.tmin = R(0.001f);
        shadow.tmax = t_max;
        Ray<R> sray = shadow;
        bool blocked = false;
        for (int pass = 0; pass < (PLANT(P_SHADOW_4) ? 4 : 3); pass++) {  // render_kernel.cu:247-268
            Hit<R> sh;
            cnt.shadow++;
            if (!scene_intersect_opaque(S, sray, sh)) { blocked = false; break; }
            if (hit_transmission(S, sh, s.materials[sh.material_id]) > (PLANT(P_SHADOW_THRESH) ? R(0.3f) : R(0.5f))) {
                if (PLANT(P_SHADOW_TMAX)) sray.tmax -= sh.t + R(0.002f);
                sray.o = sh.point + sray.d * R(0.002f);
                continue;
            }
            blocked = true;
            break;
        }
        if (mat.shininess < 0.0f) return blocked ? mk<R>(1, 0, 0) : mk<R>(0, 1, 0);
        if (blocked) continue;
        R transmittance = R(1.0f);
        if (mat.anisotropy <= 0.0f) {
            R tau = strand_bvh_tau(S, shadow, PLANT(P_SIGMA) ? R(5.0f) : R(6.0f));
            transmittance = exp_r(-tau);
        }
        result = result + shade_direct(S, mat, in, L, radiance, shininess) * transmittance;
    }
    if (S.pm.valid() && s.caustics > 0.0f) {  // render_kernel.cu:288-296
        R r = R(S.pm.cell) * R(0.5f);
        V3<R> flux = photon_gather(S.pm, rec.point, rec.normal, r, 128, (int*)nullptr);
        result = result + base * flux * (R(s.caustics) / (R(9.8696044f) * r * r));
    }
    return result;
}

template <class R> inline bool ray_box(const Ray<R>& ray, V3<R> c, V3<R> h, R& t0, R& t1) {  // render_kernel.cu:302-325
    t0 = R(-1.0e30f); t1 = R(1.0e30f);
    const R eps = R(1.0e-8f);
    R o[3] = {ray.o.x, ray.o.y, ray.o.z}, d[3] = {ray.d.x, ray.d.y, ray.d.z};
    R cc[3] = {c.x, c.y, c.z}, hh[3] = {h.x, h.y, h.z};
    for (int a = 0; a < 3; a++) {
        if (std::fabs(d[a]) < eps) {
            if (o[a] < cc[a] - hh[a] || o[a] > cc[a] + hh[a]) return false;
            continue;
        }
        R inv = R(1.0f) / d[a];
        R tn = (cc[a] - hh[a] - o[a]) * inv;
        R tf = (cc[a] + hh[a] - o[a]) * inv;
        if (tn > tf) std::swap(tn, tf);
        t0 = fmax_(t0, tn);
        t1 = fmin_(t1, tf);
        if (t0 > t1) return false;
    }
    return true;
}
template <class R> inline V3<R> sky_color(V3<R> d) {  // render_kernel.cu:329-334
    R t = R(0.5f) * ((PLANT(P_SKY_X) ? d.x : d.y) + R(1.0f));
    if (PLANT(P_SKY_SWAP))
        return mk<R>((R(1.0f) - t) * R(1.0f) + t * R(1.0f), (R(1.0f) - t) * R(1.0f) + t * R(0.7f), (R(1.0f) - t) * R(1.0f) + t * R(0.5f));
    return mk<R>((R(1.0f) - t) * R(1.0f) + t * R(0.5f), (R(1.0f) - t) * R(1.0f) + t * R(0.7f), (R(1.0f) - t) * R(1.0f) + t * R(1.0f));
}
// reference render_kernel.cu:339-436
template <class R> inline V3<R> media_march(const Scene& S, const Ray<R>& ray, R t_surface, uint32_t seed, R& tau_out, Counters& cnt) {
    const SheenScene& s = S.s;
    V3<R> in_scatter = mk<R>(0, 0, 0);
    R tau_total = 0;
    for (int m = 0; m < s.num_media; m++) {
        const Medium& med = s.media[m];
        if (med.samples <= 0) continue;
        R m0, m1;
        if (!ray_box(ray, cv<R>(med.center), cv<R>(med.half), m0, m1)) continue;
        m0 = fmax_(m0, ray.tmin);
        m1 = fmin_(m1, t_surface);
        if (m1 <= m0) continue;
        int n = med.samples > 64 ? 64 : med.samples;
        R dt = (m1 - m0) / R(n);
        R tau = 0;
        for (int k = 0; k < n; k++) {
            R jitter = R(aa_rand(seed ^ (uint32_t)(uint32_t(m) * 9781u) ^ (uint32_t)(uint32_t(k) * 6271u)));
            R t = m0 + (R(k) + jitter) * dt;
            V3<R> p = ray.o + ray.d * t;
            V3<R> radiance = mk<R>(0, 0, 0);
            for (int i = 0; i < s.num_lights; i++) {
                const Light& light = s.lights[i];
                if (!(light.flags & 1)) continue;
                V3<R> ldir, lrad;
                R ldist;
                if (light.type == 0) {
                    ldir = sun_dir<R>(light, seed ^ (uint32_t)(uint32_t(i) * 2654435761u) ^ (uint32_t)(uint32_t(k) * 2246822519u));
                    ldist = R(1.0e9f);
                    lrad = cv<R>(light.color) * R(light.intensity);
                } else {
                    V3<R> tl = cv<R>(light.position) - p;
                    R d2 = fmax_(dot(tl, tl), R(1.0e-4f));
                    ldist = std::sqrt(d2);
                    ldir = tl / ldist;
                    lrad = cv<R>(light.color) * (R(light.intensity) / (R(12.5663706f) * d2));
                    if (light.type == 3) {
                        R c = dot(-ldir, cv<R>(light.direction));
                        R cos_half = cos_r(R(light.angle));
                        if (c <= cos_half) continue;
                        R blend = fmin_(fmax_(R(light.blend), R(0)), R(1));
                        R cos_soft = cos_r(R(light.angle) * (R(1.0f) - blend));
                        if (cos_soft > cos_half + R(1.0e-4f)) lrad = lrad * fmin_((c - cos_half) / (cos_soft - cos_half), R(1.0f));
                    }
                }
                Ray<R> sr{p, ldir, R(0.002f), ldist};
                Hit<R> srec;
                cnt.shadow++;
                if (scene_intersect_opaque(S, sr, srec)) continue;
                radiance = radiance + lrad;
            }
            V3<R> scatter = cv<R>(med.sigma_s) * radiance * R(0.07957747f);
            in_scatter = in_scatter + scatter * (exp_r(-tau) * dt);
            tau += R(med.sigma_t) * dt;
        }
        tau_total += tau;
    }
    tau_out = tau_total;
    return in_scatter;
}

// One pixel sample: reference render_kernel.cu:491-809 (loop body).
template <class R> inline V3<R> render_sample(const Scene& S, int x, int y, int sidx, int ns, int width, int height,
                                              Counters& cnt, int* hk, int* hi, float* ht) {
    const SheenScene& s = S.s;
    const CameraData& cam = s.camera;
    float ox, oy;
    aa_sample_offset_n(sidx, ns, x, y, ox, oy);
    R u, v;
    aa_uv<R>(x, y, width, height, ox, oy, u, v);
    uint32_t seed = sample_seed(x, y, sidx);
    R tan_h = 1, tan_x = 1, tan_y = 1;
    if (cam.focal_length > 0.0f) {
        tan_h = R(0.5f) * R(cam.sensor_width) / R(cam.focal_length);
        if (width >= height) { tan_x = tan_h; tan_y = tan_h * R(height) / R(width); }
        else { tan_y = tan_h; tan_x = tan_h * R(width) / R(height); }
    }
    Ray<R> ray;
    ray.o = cv<R>(cam.position);
    ray.d = normalize(cv<R>(cam.forward) + cv<R>(cam.right) * (u * tan_x + R(cam.shift_x) * tan_h) +
                      cv<R>(cam.up) * (v * tan_y + R(cam.shift_y) * tan_h));
    ray.tmin = R(0.001f);
    ray.tmax = R(1e10f);
    const R pix_angle = R(2.0f) * tan_x / R(width);
    Hit<R> rec;
    cnt.primary++;
    bool hit = scene_intersect(S, ray, rec);
    if (hk) { *hk = hit ? rec.kind : -1; *hi = hit ? rec.index : -1; *ht = hit ? float(rec.t) : 0.0f; }
    V3<R> color = hit ? compute_lighting(S, rec, -ray.d, rec.t * pix_angle, seed, cnt) : sky_color(ray.d);
    for (int i = 0; i < s.num_lights; i++) {  // beams, render_kernel.cu:560-652
        const Light& light = s.lights[i];
        if (light.type != 3 || light.beam <= 0.0f) continue;
        R t_end = hit ? fmin_(rec.t, R(30.0f)) : R(30.0f);
        if (t_end <= R(0)) continue;
        const int STEPS = 6;
        R dt = t_end / R(STEPS);
        R cos_half = cos_r(R(light.angle));
        R blend = fmin_(fmax_(R(light.blend), R(0)), R(1));
        R cos_soft = cos_r(R(light.angle) * (R(1.0f) - blend));
        V3<R> acc = mk<R>(0, 0, 0);
        for (int k = 0; k < STEPS; k++) {
            R t = (R(k) + R(0.5f)) * dt;
            V3<R> p = ray.o + ray.d * t;
            V3<R> vv = p - cv<R>(light.position);
            R d2 = fmax_(dot(vv, vv), R(1.0e-4f));
            R c = dot(vv, cv<R>(light.direction)) * (R(1) / std::sqrt(d2));
            if (c <= cos_half) continue;
            R w = R(1.0f);
            if (cos_soft > cos_half + R(1.0e-4f)) w = fmin_((c - cos_half) / (cos_soft - cos_half), R(1.0f));
            R dens = R(light.beam) * R(light.intensity) * R(1.0e-3f) / (R(1.0f) + R(0.25f) * d2);
            V3<R> target = cv<R>(light.position);
            if (light.radius > 0.0f) {
                V3<R> axis = normalize(-vv);
                V3<R> up = (std::fabs(axis.y) < R(0.9f)) ? mk<R>(0, 1, 0) : mk<R>(1, 0, 0);
                V3<R> t1 = normalize(cross(axis, up));
                V3<R> t2 = cross(axis, t1);
                R rr = std::sqrt(R(aa_rand(seed ^ (uint32_t)(uint32_t(k) * 2654435761u) ^ 0x85ebca6bu))) * R(light.radius);
                R ph = R(6.2831853f) * R(aa_rand(seed ^ (uint32_t)(uint32_t(k) * 2246822519u) ^ 0xc2b2ae35u));
                target = cv<R>(light.position) + t1 * (cos_r(ph) * rr) + t2 * (sin_r(ph) * rr);
            }
            V3<R> to_l = target - p;
            R ldist = std::sqrt(fmax_(dot(to_l, to_l), R(1.0e-8f)));
            V3<R> ldir = to_l / ldist;
            R remain = ldist * R(0.999f);
            Ray<R> sray{p, ldir, R(0.002f), remain};
            bool blocked = false;
            for (int pass = 0; pass < (PLANT(P_BEAM_NO_SHADOW) ? 0 : 3); pass++) {
                Hit<R> srec;
                cnt.shadow++;
                if (!scene_intersect_opaque(S, sray, srec)) { blocked = false; break; }
                if (s.materials[srec.material_id].transmission > 0.5f) {
                    remain -= srec.t + R(0.002f);
                    if (remain <= R(0)) { blocked = false; break; }
                    sray.o = srec.point + ldir * R(0.002f);
                    sray.tmax = remain;
                    continue;
                }
                blocked = true;
                break;
            }
            if (blocked) continue;
            acc = acc + cv<R>(light.color) * (dens * w * dt);
        }
        color = color + acc;
    }
    if (s.num_media > 0) {
        R tau = 0;
        V3<R> ins = media_march(S, ray, hit ? rec.t : R(1.0e9f), seed, tau, cnt);
        color = color * exp_r(-tau) + ins;
    }
    if (hit) {  // reflection, render_kernel.cu:672-733
        const Material& pm = s.materials[rec.material_id];
        if (pm.reflect > 0.0f) {
            V3<R> rd = ray.d - rec.normal * (R(2.0f) * dot(ray.d, rec.normal));
            Ray<R> rr{rec.point + rec.normal * R(0.002f), normalize(rd), R(0.001f), R(1.0e10f)};
            Hit<R> rrec;
            cnt.reflection++;
            bool rhit = scene_intersect(S, rr, rrec);
            V3<R> rcol = rhit ? compute_lighting(S, rrec, -rr.d, rrec.t * pix_angle, seed, cnt) : sky_color(rr.d);
            if (s.num_media > 0) {
                R rtau = 0;
                V3<R> rins = media_march(S, rr, rhit ? rrec.t : R(1.0e9f), seed, rtau, cnt);
                rcol = rcol * exp_r(-rtau) + rins;
            }
            R w = PLANT(P_REFL_NO_CLAMP) ? R(pm.reflect) : fmin_(R(pm.reflect), R(1.0f));
            color = PLANT(P_REFL_SWAP) ? color * w + rcol * (R(1.0f) - w) : color * (R(1.0f) - w) + rcol * w;
        }
    }
    if (hit) {  // refraction, render_kernel.cu:736-807
        const Material& gm = s.materials[rec.material_id];
        R gtr = hit_transmission(S, rec, gm);
        if (gtr > R(0)) {
            V3<R> through = mk<R>(0, 0, 0);
            R w = R(1.0f);
            Ray<R> gr = ray;
            Hit<R> gh = rec;
            bool entering = true;
            for (int b = 0; b < 3; b++) {
                const Material& sm = s.materials[gh.material_id];
                V3<R> n = gh.normal, d = gr.d;
                R ior = fmax_(hit_ior(S, gh, sm), R(1.0f));
                R eta = (entering || PLANT(P_ETA_FIXED)) ? (R(1.0f) / ior) : ior;
                R cosi = fmax_(-dot(d, n), R(1.0e-4f));
                R k = R(1.0f) - eta * eta * (R(1.0f) - cosi * cosi);
                V3<R> td = (k < R(0)) ? normalize(d + (PLANT(P_TIR_SIGN) ? R(-2.0f) : R(2.0f)) * cosi * n)
                                      : normalize(eta * d + (eta * cosi - std::sqrt(k)) * n);
                if (b == 0 || !PLANT(P_W_ONCE)) w *= hit_transmission(S, gh, sm);
                gr.o = gh.point + td * R(0.002f);
                gr.d = td;
                gr.tmin = R(0.001f);
                gr.tmax = R(1.0e10f);
                Hit<R> nh;
                cnt.refraction++;
                if (!scene_intersect(S, gr, nh)) { through = through + w * sky_color(td); break; }
                if (hit_transmission(S, nh, s.materials[nh.material_id]) > R(0)) { entering = !entering; gh = nh; continue; }
                through = through + w * compute_lighting(S, nh, -td, nh.t * pix_angle, seed, cnt);
                break;
            }
            color = color * (R(1.0f) - fmin_(gtr, R(1.0f))) + through;
        }
    }
    return color;
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C entry points (used by tests and by bench.py's CPU baseline only)
// ---------------------------------------------------------------------------
using namespace orc;

extern "C" {

typedef struct OracleStats {
    uint64_t primary_rays, shadow_rays, reflection_rays, refraction_rays, photon_rays, total_rays, samples;
    int photons_emitted, photons_stored;
} OracleStats;

void* oracle_scene_create(const SheenScene* in, int threads) {
    if (!in) return nullptr;
    Scene* S = new Scene();
    S->s = *in;
    if (threads < 1) threads = 1;
    for (int i = 0; i < in->num_textures; i++) {
        Tex t;
        t.w = in->texture_widths ? in->texture_widths[i] : 0;
        t.h = in->texture_heights ? in->texture_heights[i] : 0;
        if (in->texture_pixels && in->texture_pixels[i] && t.w > 0 && t.h > 0) {
            t.px.assign(in->texture_pixels[i], in->texture_pixels[i] + (size_t)t.w * t.h * 3);
            build_mips(t);
            t.valid = true;
        }
        S->tex.push_back(std::move(t));
    }
    int counts[7] = {in->num_objects, in->num_sweep_strands, in->num_triangles, in->num_heightfields,
                     in->num_superellipsoids, in->num_lathes, in->num_prisms};
    for (int c = 0; c < 7; c++) {
        std::vector<int> idx;
        for (int i = 0; i < counts[c]; i++) {
            if (c == K_SPHERE && in->objects[i].type != 0) continue;
            if (c == K_SE && in->superellipsoids[i]._pad0 == 1) continue;  // CSG-only (scene.h:99-104 se_hidden)
            idx.push_back(i);
        }
        S->count[c] = counts[c];
        if (idx.empty()) continue;
        std::vector<Box> bx(counts[c]);
        for (int i : idx) bx[i] = bounds_of(*S, c, i);
        S->bvh[c].reserve(2 * idx.size());
        build_rec(S->bvh[c], bx, idx, 0, (int)idx.size());
    }
    build_photons(*S, threads);
    return S;
}

void oracle_scene_destroy(void* h) { delete (Scene*)h; }

// Parameter substitution for gradient checks (the differentiable entry
// points' forward): albedo -> diffuse_color, emission -> emission_color with
// emission 1, intensity -> Light.intensity, texels (nullable) -> float chains.
int oracle_set_params(void* h, const float* albedo, const float* emission, const float* intensity, const float* texels) {
    Scene* S = (Scene*)h;
    SheenScene& s = S->s;
    if (!S->orig_mats) { S->orig_mats = s.materials; S->orig_lights = s.lights; }
    S->own_mats.assign(S->orig_mats, S->orig_mats + s.num_materials);
    for (int i = 0; i < s.num_materials; i++) {
        Material& m = S->own_mats[i];
        if (albedo) m.diffuse_color = {albedo[3 * i], albedo[3 * i + 1], albedo[3 * i + 2]};
        if (emission) { m.emission_color = {emission[3 * i], emission[3 * i + 1], emission[3 * i + 2]}; m.emission = 1.0f; }
    }
    s.materials = S->own_mats.data();
    if (s.lights && s.num_lights > 0) {
        S->own_lights.assign(S->orig_lights, S->orig_lights + s.num_lights);
        for (int i = 0; i < s.num_lights; i++)
            if (intensity) S->own_lights[i].intensity = intensity[i];
        s.lights = S->own_lights.data();
    }
    if (texels) {
        size_t p = 0;
        for (Tex& t : S->tex) {
            if (!t.valid) continue;
            const size_t n0 = (size_t)t.w * t.h * 3;
            t.fpx.assign(t.px.size(), 0.0f);
            t.dpx.assign(t.px.size(), 0.0);
            for (size_t k = 0; k < n0; k++) { t.fpx[k] = texels[p + k]; t.dpx[k] = texels[p + k]; }
            build_float_mips(t.fpx, t.w, t.h);
            build_float_mips(t.dpx, t.w, t.h);
            t.use_float = true;
            p += n0;
        }
    }
    return 0;
}

int oracle_photon_stats(void* h, int* emitted, int* stored) {
    Scene* S = (Scene*)h;
    if (emitted) *emitted = S->photon_emitted;
    if (stored) *stored = S->photon_stored;
    return 0;
}

// Copies the sorted photon array (cell order) and the cell tables.
int oracle_photon_map(void* h, Photon* photons, int* cell_start, int* cell_count) {
    Scene* S = (Scene*)h;
    if (!S->pm.valid()) return 0;
    if (photons) std::memcpy(photons, S->pm.sorted.data(), S->pm.sorted.size() * sizeof(Photon));
    if (cell_start) std::memcpy(cell_start, S->pm.cell_start.data(), S->pm.num_cells * sizeof(int));
    if (cell_count) std::memcpy(cell_count, S->pm.cell_count.data(), S->pm.num_cells * sizeof(int));
    return (int)S->pm.sorted.size();
}

// Gather probes (reference render_kernel.cu:981-990): photons taken within
// radius around stored photon stride*i, through the hash map.
int oracle_photon_probe(void* h, int stride, float radius, int max_samples, int* counts) {
    Scene* S = (Scene*)h;
    if (!S->pm.valid() || stride <= 0) return 0;
    int n = (int)S->pm.sorted.size();
    int probes = (n + stride - 1) / stride;
    for (int i = 0; i < probes; i++) {
        const Photon& p = S->pm.sorted[(size_t)i * stride];
        int taken = 0;
        photon_gather<float>(S->pm, cv<float>(p.position), cv<float>(p.normal), radius, max_samples, &taken);
        counts[i] = taken;
    }
    return probes;
}
// Brute-force counterpart (reference render_kernel.cu:1092-1107).
int oracle_photon_probe_brute(void* h, int stride, float radius, int max_samples, int* counts) {
    Scene* S = (Scene*)h;
    if (!S->pm.valid() || stride <= 0) return 0;
    int n = (int)S->pm.sorted.size();
    int probes = (n + stride - 1) / stride;
    float r2 = radius * radius;
    for (int i = 0; i < probes; i++) {
        const Photon& p0 = S->pm.sorted[(size_t)i * stride];
        int taken = 0;
        for (int j = 0; j < n && taken < max_samples; j++) {
            V3<float> d = cv<float>(S->pm.sorted[j].position) - cv<float>(p0.position);
            if (dot(d, d) > r2) continue;
            if (dot(cv<float>(S->pm.sorted[j].normal), cv<float>(p0.normal)) < 0.1f) continue;
            taken++;
        }
        counts[i] = taken;
    }
    return probes;
}

// Renders rows [y0, y1) of a frame.  out_rgba (w*h*4) and the per-sample hit
// arrays (w*h*aa, index (y*w+x)*aa+s) may be NULL.  fp64 selects the double
// restatement.  Returns 0, or -1 on bad arguments.
int oracle_render(void* h, int width, int height, int aa, int y0, int y1, int threads, int fp64,
                  float* out_rgba, int* hit_kind, int* hit_index, float* hit_t, OracleStats* stats) {
    Scene* S = (Scene*)h;
    if (!S || width <= 0 || height <= 0 || aa < 1 || S->s.num_materials <= 0) return -1;
    int ns = aa > 64 ? 64 : aa;
    if (y0 < 0) y0 = 0;
    if (y1 > height) y1 = height;
    if (threads < 1) threads = 1;
    std::atomic<uint64_t> cp{0}, cs{0}, cr{0}, cf{0};
#pragma omp parallel num_threads(threads)
    {
        Counters c;
#pragma omp for schedule(dynamic, 1)
        for (int y = y0; y < y1; y++) {
            for (int x = 0; x < width; x++) {
                double acc[3] = {0, 0, 0};
                float accf[3] = {0, 0, 0};
                for (int s = 0; s < ns; s++) {
                    size_t si = ((size_t)y * width + x) * ns + s;
                    int* hk = hit_kind ? hit_kind + si : nullptr;
                    int hkd, hid;
                    float htd;
                    if (fp64) {
                        V3<double> col = render_sample<double>(*S, x, y, s, ns, width, height, c, hk ? hk : &hkd,
                                                                hit_index ? hit_index + si : &hid, hit_t ? hit_t + si : &htd);
                        acc[0] += col.x; acc[1] += col.y; acc[2] += col.z;
                    } else {
                        V3<float> col = render_sample<float>(*S, x, y, s, ns, width, height, c, hk ? hk : &hkd,
                                                              hit_index ? hit_index + si : &hid, hit_t ? hit_t + si : &htd);
                        accf[0] += col.x; accf[1] += col.y; accf[2] += col.z;
                    }
                }
                if (out_rgba) {
                    float* o = out_rgba + ((size_t)y * width + x) * 4;
                    if (fp64) {
                        double inv = 1.0 / double(ns);
                        o[0] = float(acc[0] * inv); o[1] = float(acc[1] * inv); o[2] = float(acc[2] * inv);
                    } else {
                        float inv = 1.0f / float(ns);
                        o[0] = accf[0] * inv; o[1] = accf[1] * inv; o[2] = accf[2] * inv;
                    }
                    o[3] = 1.0f;
                }
            }
        }
        cp += c.primary; cs += c.shadow; cr += c.reflection; cf += c.refraction;
    }
    if (stats) {
        stats->primary_rays = cp; stats->shadow_rays = cs; stats->reflection_rays = cr; stats->refraction_rays = cf;
        stats->photon_rays = S->photon_rays;
        stats->total_rays = cp + cs + cr + cf;
        stats->samples = (uint64_t)(y1 - y0) * width * ns;
        stats->photons_emitted = S->photon_emitted;
        stats->photons_stored = S->photon_stored;
    }
    return 0;
}

// Reference radiance of a window of the width x height image: every pixel
// averages its sub x sub sub-pixels of the (sub*width) x (sub*height) image,
// each rendered with aa samples (FP64 arithmetic, or the FP32 parity path),
// i.e. sub*sub*aa samples spread over the pixel's footprint (4096 with
// sub = 8, aa = 64).  out: double RGB, cw*ch*3, row-major over the window.
int oracle_render_reference(void* h, int width, int height, int x0, int y0, int cw, int ch, int sub, int aa,
                            int threads, int fp64, double* out_rgb) {
    Scene* S = (Scene*)h;
    if (!S || width <= 0 || height <= 0 || sub < 1 || aa < 1 || aa > 64 || cw <= 0 || ch <= 0 || !out_rgb) return -1;
    if (x0 < 0 || y0 < 0 || x0 + cw > width || y0 + ch > height) return -1;
    if (threads < 1) threads = 1;
    const int W = width * sub, H = height * sub;
#pragma omp parallel num_threads(threads)
    {
        Counters c;
#pragma omp for schedule(dynamic, 1) collapse(2)
        for (int j = 0; j < ch; j++) {
            for (int i = 0; i < cw; i++) {
                double acc[3] = {0, 0, 0};
                for (int sy = 0; sy < sub; sy++) {
                    for (int sx = 0; sx < sub; sx++) {
                        const int X = (x0 + i) * sub + sx, Y = (y0 + j) * sub + sy;
                        for (int s = 0; s < aa; s++) {
                            int hk, hi;
                            float ht;
                            if (fp64) {
                                V3<double> col = render_sample<double>(*S, X, Y, s, aa, W, H, c, &hk, &hi, &ht);
                                acc[0] += col.x; acc[1] += col.y; acc[2] += col.z;
                            } else {
                                V3<float> col = render_sample<float>(*S, X, Y, s, aa, W, H, c, &hk, &hi, &ht);
                                acc[0] += col.x; acc[1] += col.y; acc[2] += col.z;
                            }
                        }
                    }
                }
                const double inv = 1.0 / (double(sub) * sub * aa);
                double* o = out_rgb + ((size_t)j * cw + i) * 3;
                o[0] = acc[0] * inv; o[1] = acc[1] * inv; o[2] = acc[2] * inv;
            }
        }
    }
    return 0;
}

// Closest hit of arbitrary rays (scene_intersect).  For per-primitive tests.
int oracle_trace(void* h, int n, const float* orig, const float* dir, float tmin, float tmax, int opaque,
                 int* kind, int* index, float* t, float* point, float* normal, int* material) {
    Scene* S = (Scene*)h;
    for (int i = 0; i < n; i++) {
        Ray<float> r{mk(orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]), mk(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]), tmin, tmax};
        Hit<float> hh;
        bool ok = opaque ? scene_intersect_opaque(*S, r, hh) : scene_intersect(*S, r, hh);
        kind[i] = ok ? hh.kind : -1;
        index[i] = ok ? hh.index : -1;
        t[i] = ok ? hh.t : 0.0f;
        if (point) { point[3 * i] = hh.point.x; point[3 * i + 1] = hh.point.y; point[3 * i + 2] = hh.point.z; }
        if (normal) { normal[3 * i] = hh.normal.x; normal[3 * i + 1] = hh.normal.y; normal[3 * i + 2] = hh.normal.z; }
        if (material) material[i] = ok ? hh.material_id : -1;
    }
    return 0;
}

// Brute-force closest hit over one category, without any BVH: every
// primitive is tested (the reference's leaf slab gate included).
int oracle_trace_brute(void* h, int n, const float* orig, const float* dir, float tmin, float tmax, int cat,
                       int* index, float* t) {
    Scene* S = (Scene*)h;
    for (int i = 0; i < n; i++) {
        Ray<float> r{mk(orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]), mk(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]), tmin, tmax};
        float best = tmax;
        int bi = -1;
        Hit<float> tmp;
        for (int p = 0; p < S->count[cat]; p++) {
            Box b = bounds_of(*S, cat, p);
            if (!intersect_aabb(r, b.mn, b.mx)) continue;
            if (!prim_hit(*S, cat, p, r, tmp)) continue;
            if (tmp.t < best || (tmp.t == best && bi >= 0 && p < bi)) { best = tmp.t; bi = p; }
        }
        index[i] = bi;
        t[i] = bi >= 0 ? best : 0.0f;
    }
    return 0;
}

uint32_t oracle_aa_hash(uint32_t x) { return aa_hash(x); }
float oracle_aa_rand(uint32_t x) { return aa_rand(x); }
void oracle_sample_offset(int s, int n, int x, int y, float* ox, float* oy) { aa_sample_offset_n(s, n, x, y, *ox, *oy); }
float oracle_hf_eval(const HeightField* hf, float x, float z) { return hf_eval<float>(*hf, x, z); }
float oracle_hf_cell_hash(int xi, int zi, int seed) { return hf_cell_hash(xi, zi, seed); }
// sheen_math.h's float functions (0 sin, 1 cos, 2 tan, 3 exp, 4 log2, 5 pow).
float oracle_math(int fn, float x, float y) {
    switch (fn) {
        case 0: return shm::sinf_r(x);
        case 1: return shm::cosf_r(x);
        case 2: return shm::tanf_r(x);
        case 3: return shm::expf_r(x);
        case 4: return shm::log2f_r(x);
        default: return shm::powf_r(x, y);
    }
}
float oracle_hf_value_noise(float fx, float fz, int seed) { return hf_value_noise<float>(fx, fz, seed); }
float oracle_hf_fbm(float x, float z, float freq, int octaves, int seed) { return hf_fbm<float>(x, z, freq, octaves, seed); }

#ifdef ORACLE_PLANTS
// Switches planted mistake k on (0: none).  Only in liboracle_plant.so.
void oracle_plant(int k) { g_plant = k; }
#endif
int oracle_photon_cell_hash(int ix, int iy, int iz, int num_cells) { return photon_cell_hash(ix, iy, iz, num_cells); }

// Mip chain of an RGB texture (scene.h:230-261); returns bytes written.
int oracle_build_mips(const unsigned char* px, int w, int h, unsigned char* out, int* mips) {
    Tex t;
    t.w = w; t.h = h;
    t.px.assign(px, px + (size_t)w * h * 3);
    build_mips(t);
    std::memcpy(out, t.px.data(), t.px.size());
    *mips = t.mips;
    return (int)t.px.size();
}
// Bilinear mip sample (shading.h:26-58) of texture id of a scene.
int oracle_sample_texture(void* h, int tex, float u, float v, float lod, float* rgb) {
    Scene* S = (Scene*)h;
    if (!tex_ok(*S, tex)) return -1;
    V3<float> c = sample_texture_lod(S->tex[tex], u, v, lod);
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
    return 0;
}

}  // extern "C"
