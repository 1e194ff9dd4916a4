// scenes.cpp — seeded synthetic workloads (declared in include/sheen_scenes.h).
//
// Generators are deterministic functions of (config, seed, detail): every
// random draw comes from SplitMix64 streams, every shared vertex of the
// geodesic mesh is computed from a canonical key so adjacent faces agree
// bit-for-bit (no cracks).
#include "sheen_scenes.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace {

struct Rng {  // SplitMix64
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    float uni() { return float(next() >> 40) / 16777216.0f; }  // [0,1)
    float uni(float a, float b) { return a + (b - a) * uni(); }
};

inline sheen_f3 f3(float x, float y, float z) { return {x, y, z}; }
inline sheen_f3 add(sheen_f3 a, sheen_f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline sheen_f3 sub(sheen_f3 a, sheen_f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline sheen_f3 mul(sheen_f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot3(sheen_f3 a, sheen_f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline sheen_f3 cross3(sheen_f3 a, sheen_f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline sheen_f3 norm3(sheen_f3 a) { float l = std::sqrt(dot3(a, a)); return l > 0 ? mul(a, 1.0f / l) : a; }

struct Store {
    SheenScene scene{};
    std::vector<Object> objects;
    std::vector<SweepPoint> sweep_points;
    std::vector<SweepStrand> strands;
    std::vector<Triangle> tris;
    std::vector<HeightField> hfs;
    std::vector<Superellipsoid> ses;
    std::vector<LathePoint> lathe_points;
    std::vector<Lathe> lathes;
    std::vector<PrismPoint> prism_points;
    std::vector<Prism> prisms;
    std::vector<CSGNode> csg;
    int csg_root = -1;
    std::vector<Plane> planes;
    std::vector<Material> mats;
    std::vector<std::vector<unsigned char>> tex;
    std::vector<int> tex_w, tex_h;
    std::vector<const unsigned char*> tex_ptr;
    std::vector<std::vector<unsigned char>> hf_images;
    std::vector<Light> lights;
    std::vector<Medium> media;
    int photons = 0;
    float photon_cell = 0.0f, caustics = 0.0f;
    CameraData cam{};
    int view_w = 0, view_h = 0, view_aa = 0;  // a generator's own image size and spp (0 = the params')

    void finalize() {
        SheenScene& s = scene;
        std::memset(&s, 0, sizeof(s));
        s.camera = cam;
        s.objects = objects.empty() ? nullptr : objects.data(); s.num_objects = (int)objects.size();
        s.sweep_points = sweep_points.empty() ? nullptr : sweep_points.data(); s.num_sweep_points = (int)sweep_points.size();
        s.sweep_strands = strands.empty() ? nullptr : strands.data(); s.num_sweep_strands = (int)strands.size();
        s.triangles = tris.empty() ? nullptr : tris.data(); s.num_triangles = (int)tris.size();
        s.heightfields = hfs.empty() ? nullptr : hfs.data(); s.num_heightfields = (int)hfs.size();
        s.superellipsoids = ses.empty() ? nullptr : ses.data(); s.num_superellipsoids = (int)ses.size();
        s.lathe_points = lathe_points.empty() ? nullptr : lathe_points.data(); s.num_lathe_points = (int)lathe_points.size();
        s.lathes = lathes.empty() ? nullptr : lathes.data(); s.num_lathes = (int)lathes.size();
        s.prism_points = prism_points.empty() ? nullptr : prism_points.data(); s.num_prism_points = (int)prism_points.size();
        s.prisms = prisms.empty() ? nullptr : prisms.data(); s.num_prisms = (int)prisms.size();
        s.csg_nodes = csg.empty() ? nullptr : csg.data(); s.num_csg_nodes = (int)csg.size(); s.csg_root = csg_root;
        s.planes = planes.empty() ? nullptr : planes.data(); s.num_planes = (int)planes.size();
        s.materials = mats.data(); s.num_materials = (int)mats.size();
        tex_ptr.clear();
        for (auto& t : tex) tex_ptr.push_back(t.data());
        s.texture_pixels = tex_ptr.empty() ? nullptr : tex_ptr.data();
        s.texture_widths = tex_w.empty() ? nullptr : tex_w.data();
        s.texture_heights = tex_h.empty() ? nullptr : tex_h.data();
        s.num_textures = (int)tex.size();
        s.lights = lights.empty() ? nullptr : lights.data(); s.num_lights = (int)lights.size();
        s.photons = photons; s.photon_cell = photon_cell; s.caustics = caustics;
        s.media = media.empty() ? nullptr : media.data(); s.num_media = (int)media.size();
    }

    // Material with the reference's add_material defaults (scene.h:303-330).
    int material(sheen_f3 c, float spec, float shin, float amb) {
        Material m;
        std::memset(&m, 0, sizeof(m));
        m.diffuse_color = c; m.specular = spec; m.shininess = shin; m.ambient = amb;
        m.hair_shift = 0.1f; m.texture_id = -1; m.mask_id = -1; m.normal_id = -1;
        m.ior = 1.45f; m.normal_strength = 1.0f;
        mats.push_back(m);
        return (int)mats.size() - 1;
    }
    void sphere(sheen_f3 c, float r, int mat) {
        Object o; std::memset(&o, 0, sizeof(o));
        o.type = 0; o.material_id = mat; o.param1 = c; o.param2 = r;
        objects.push_back(o);
    }
    void plane(sheen_f3 n, float d, int mat) {
        Plane p; std::memset(&p, 0, sizeof(p));
        p.normal = n; p.d = d; p.material_id = mat;
        planes.push_back(p);
    }
    static Light base_light() {  // reference scene.h:391-408
        Light l; std::memset(&l, 0, sizeof(l));
        l.direction = f3(0, 0, -1); l.color = f3(1, 1, 1); l.intensity = 1.0f;
        l.flags = 1; l.texture_id = -1;
        return l;
    }
    void point_light(sheen_f3 p, sheen_f3 c, float power, float radius) {
        Light l = base_light(); l.type = 1; l.position = p; l.color = c; l.intensity = power; l.radius = radius;
        lights.push_back(l);
    }
    void area_light(sheen_f3 p, sheen_f3 u, sheen_f3 v, sheen_f3 c, float power) {
        Light l = base_light(); l.type = 2; l.position = p; l.u_axis = u; l.v_axis = v; l.color = c; l.intensity = power;
        lights.push_back(l);
    }
    void sun_light(sheen_f3 dir, sheen_f3 c, float intensity, float angle) {
        Light l = base_light(); l.type = 0; l.direction = norm3(dir); l.color = c; l.intensity = intensity; l.angle = angle;
        lights.push_back(l);
    }
    void spot_light(sheen_f3 p, sheen_f3 dir, sheen_f3 c, float power, float half, float blend, float radius, float beam) {
        Light l = base_light(); l.type = 3; l.position = p; l.direction = norm3(dir); l.color = c; l.intensity = power;
        l.angle = half; l.blend = blend; l.radius = radius; l.beam = beam;
        lights.push_back(l);
    }
    void tri(sheen_f3 a, sheen_f3 b, sheen_f3 c, sheen_f3 na, sheen_f3 nb, sheen_f3 nc, int mat,
             sheen_f2 ta = {0, 0}, sheen_f2 tb = {0, 0}, sheen_f2 tc = {0, 0}) {
        Triangle t; std::memset(&t, 0, sizeof(t));
        t.v0 = a; t.v1 = b; t.v2 = c; t.n0 = na; t.n1 = nb; t.n2 = nc; t.uv0 = ta; t.uv1 = tb; t.uv2 = tc;
        t.material_id = mat;
        tris.push_back(t);
    }
    // 12-triangle cube with per-face UVs (reference scene.h:271-301 layout).
    void cube(sheen_f3 c, float s, int mat) {
        sheen_f3 v[8];
        int k = 0;
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                for (int l = 0; l < 2; l++) v[k++] = f3(c.x + (i ? s : -s), c.y + (j ? s : -s), c.z + (l ? s : -s));
        static const int f[12][3] = {{0, 1, 3}, {0, 3, 2}, {4, 6, 7}, {4, 7, 5}, {0, 4, 5}, {0, 5, 1},
                                     {2, 3, 7}, {2, 7, 6}, {0, 2, 6}, {0, 6, 4}, {1, 5, 7}, {1, 7, 3}};
        static const int axis[6] = {0, 0, 1, 1, 2, 2};
        auto uv = [&](sheen_f3 p, int a) {
            float u = (a == 0) ? (p.z - (c.z - s)) : (p.x - (c.x - s));
            float w = (a == 1) ? (p.z - (c.z - s)) : (p.y - (c.y - s));
            return sheen_f2{u / (2.0f * s), w / (2.0f * s)};
        };
        sheen_f3 z = f3(0, 0, 0);
        for (int i = 0; i < 12; i++) {
            int ax = axis[i / 2];
            sheen_f3 a = v[f[i][0]], b = v[f[i][1]], d = v[f[i][2]];
            tri(a, b, d, z, z, z, mat, uv(a, ax), uv(b, ax), uv(d, ax));
        }
    }
    int texture(int w, int h, std::vector<unsigned char> px) {
        tex.push_back(std::move(px)); tex_w.push_back(w); tex_h.push_back(h);
        return (int)tex.size() - 1;
    }
};

// Pinhole camera looking from pos at target with a horizontal FOV (degrees),
// expressed in the reference's Blender-style lens terms (36 mm sensor).
CameraData look_at(sheen_f3 pos, sheen_f3 target, float hfov_deg) {
    CameraData c; std::memset(&c, 0, sizeof(c));
    sheen_f3 fwd = norm3(sub(target, pos));
    sheen_f3 wup = f3(0, 1, 0);
    if (std::fabs(dot3(fwd, wup)) > 0.999f) wup = f3(0, 0, 1);
    sheen_f3 right = norm3(cross3(fwd, wup));
    sheen_f3 up = cross3(right, fwd);
    c.position = pos; c.forward = fwd; c.right = right; c.up = up;
    c.sensor_width = 36.0f;
    c.focal_length = 18.0f / std::tan(hfov_deg * 0.5f * 3.14159265358979f / 180.0f);
    return c;
}

// 3D value noise in [0,1] and its fBm (own definition for the workloads).
inline float lattice(int x, int y, int z, uint32_t seed) {
    uint32_t h = uint32_t(x) * 0x8da6b343u ^ uint32_t(y) * 0xd8163841u ^ uint32_t(z) * 0xcb1ab31fu ^ seed * 0x9e3779b9u;
    h ^= h >> 16; h *= 0x7feb352du; h ^= h >> 15; h *= 0x846ca68bu; h ^= h >> 16;
    return float(h & 0xffffffu) / 16777216.0f;
}
inline float smooth(float t) { return t * t * (3.0f - 2.0f * t); }
float value_noise3(float x, float y, float z, uint32_t seed) {
    int x0 = (int)std::floor(x), y0 = (int)std::floor(y), z0 = (int)std::floor(z);
    float tx = smooth(x - x0), ty = smooth(y - y0), tz = smooth(z - z0);
    float c[2][2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int k = 0; k < 2; k++) c[i][j][k] = lattice(x0 + i, y0 + j, z0 + k, seed);
    float a00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx, a10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    float a01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx, a11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    float b0 = a00 + (a10 - a00) * ty, b1 = a01 + (a11 - a01) * ty;
    return b0 + (b1 - b0) * tz;
}
float fbm3(sheen_f3 p, int octaves, uint32_t seed) {
    float amp = 0.5f, sum = 0.0f, norm = 0.0f, f = 2.0f;
    for (int o = 0; o < octaves; o++) {
        sum += amp * value_noise3(p.x * f, p.y * f, p.z * f, seed + 131u * o);
        norm += amp; amp *= 0.5f; f *= 2.0f;
    }
    return sum / norm;
}

// Icosahedral geodesic sphere of frequency n, displaced radially.
// Per triangle: 70% one of the diffuse materials, 15% mirror, 15% glass.
void geodesic_mesh(Store& st, int n, float radius, float amp, uint32_t seed, const std::vector<int>& diffuse,
                   int mirror, int glass, Rng& rng) {
    const float t = (1.0f + std::sqrt(5.0f)) * 0.5f;
    sheen_f3 V[12] = {f3(-1, t, 0), f3(1, t, 0), f3(-1, -t, 0), f3(1, -t, 0), f3(0, -1, t), f3(0, 1, t),
                      f3(0, -1, -t), f3(0, 1, -t), f3(t, 0, -1), f3(t, 0, 1), f3(-t, 0, -1), f3(-t, 0, 1)};
    for (auto& v : V) v = norm3(v);
    static const int F[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                                 {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                                 {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    // Canonical point on the unit icosahedron: edge points from the sorted
    // endpoint pair, face-interior points from the face.
    auto lerp_edge = [&](int a, int b, int k) {  // k steps from a towards b
        if (a > b) { std::swap(a, b); k = n - k; }
        float s = float(k) / float(n);
        return add(V[a], mul(sub(V[b], V[a]), s));
    };
    auto flat_point = [&](int f, int i, int j) {  // i along v0->v1, j along v0->v2
        int a = F[f][0], b = F[f][1], c = F[f][2];
        if (i == 0 && j == 0) return V[a];
        if (i == n) return V[b];
        if (j == n) return V[c];
        if (j == 0) return lerp_edge(a, b, i);
        if (i == 0) return lerp_edge(a, c, j);
        if (i + j == n) return lerp_edge(b, c, j);
        float si = float(i) / float(n), sj = float(j) / float(n);
        return add(V[a], add(mul(sub(V[b], V[a]), si), mul(sub(V[c], V[a]), sj)));
    };
    auto surf = [&](sheen_f3 dir) {  // displaced surface point along a unit direction
        float h = 2.0f * fbm3(dir, 6, seed) - 1.0f;
        return mul(dir, radius * (1.0f + amp * h));
    };
    auto vert = [&](sheen_f3 q, sheen_f3& pos, sheen_f3& nrm) {
        sheen_f3 d = norm3(q);
        pos = surf(d);
        sheen_f3 ta = norm3(cross3(std::fabs(d.y) < 0.9f ? f3(0, 1, 0) : f3(1, 0, 0), d));
        sheen_f3 tb = cross3(d, ta);
        const float e = 1.0f / (8.0f * float(n));
        sheen_f3 pa = surf(norm3(add(d, mul(ta, e)))), ma = surf(norm3(sub(d, mul(ta, e))));
        sheen_f3 pb = surf(norm3(add(d, mul(tb, e)))), mb = surf(norm3(sub(d, mul(tb, e))));
        nrm = norm3(cross3(sub(pa, ma), sub(pb, mb)));
        if (dot3(nrm, d) < 0) nrm = mul(nrm, -1.0f);
    };
    st.tris.reserve((size_t)20 * n * n);
    std::vector<sheen_f3> P((size_t)(n + 1) * (n + 2) / 2), N(P.size());
    auto id = [&](int i, int j) { return (size_t)j * (n + 1) - (size_t)j * (j - 1) / 2 + i; };
    for (int f = 0; f < 20; f++) {
        for (int j = 0; j <= n; j++)
            for (int i = 0; i + j <= n; i++) vert(flat_point(f, i, j), P[id(i, j)], N[id(i, j)]);
        for (int j = 0; j < n; j++)
            for (int i = 0; i + j < n; i++) {
                auto pick = [&]() {
                    float r = rng.uni();
                    if (r < 0.70f) return diffuse[(rng.next() >> 33) % diffuse.size()];
                    return r < 0.85f ? mirror : glass;
                };
                size_t a = id(i, j), b = id(i + 1, j), c = id(i, j + 1);
                st.tri(P[a], P[b], P[c], N[a], N[b], N[c], pick());
                if (i + j + 1 < n) {
                    size_t d = id(i + 1, j + 1);
                    st.tri(P[b], P[d], P[c], N[b], N[d], N[c], pick());
                }
            }
    }
}

std::vector<unsigned char> noise_texture(int w, int h, uint32_t seed) {
    std::vector<unsigned char> px((size_t)w * h * 3);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            for (int c = 0; c < 3; c++) {
                float v = 0.0f, amp = 0.5f, norm = 0.0f, f = 8.0f / float(w);
                for (int o = 0; o < 4; o++) {
                    v += amp * value_noise3(x * f, y * f, float(c) * 17.0f, seed + 97u * o);
                    norm += amp; amp *= 0.5f; f *= 2.0f;
                }
                px[((size_t)y * w + x) * 3 + c] = (unsigned char)std::lround(255.0f * (0.15f + 0.8f * v / norm));
            }
    return px;
}

void gen_cfg1(Store& st, const SheenSceneParams& p, Rng& rng) {
    st.cam = look_at(f3(0, 2, -15), f3(0, 0, 0), 45.0f);
    int n = p.detail > 0 ? p.detail : 64;
    for (int i = 0; i < n; i++) {
        sheen_f3 c = f3(rng.uni(-10, 10), rng.uni(-10, 10), rng.uni(-10, 10));
        float r = rng.uni(0.2f, 1.0f);
        sheen_f3 alb = f3(rng.uni(0.1f, 0.9f), rng.uni(0.1f, 0.9f), rng.uni(0.1f, 0.9f));
        int m = st.material(alb, 0.5f, 32.0f, 0.1f);
        if (rng.uni() < 0.25f) st.mats[m].reflect = 0.5f;
        st.sphere(c, r, m);
    }
    int g = st.material(f3(0.6f, 0.6f, 0.6f), 0.1f, 16.0f, 0.1f);
    st.plane(f3(0, 1, 0), 10.0f, g);
    st.point_light(f3(6.0f, 18.0f, -12.0f), f3(1, 1, 1), 6000.0f, 0.0f);
}

void gen_cfg2(Store& st, const SheenSceneParams& p, Rng& rng) {
    st.cam = look_at(f3(0, 0.4f, -3.2f), f3(0, 0, 0), 45.0f);
    std::vector<int> mats;
    for (int k = 0; k < 8; k++) {
        sheen_f3 alb = f3(rng.uni(0.15f, 0.9f), rng.uni(0.15f, 0.9f), rng.uni(0.15f, 0.9f));
        mats.push_back(st.material(alb, 0.4f, 48.0f, 0.08f));
    }
    int mirror = st.material(f3(0.9f, 0.9f, 0.9f), 0.6f, 128.0f, 0.02f);
    st.mats[mirror].reflect = 0.9f;
    int glass = st.material(f3(0.95f, 0.95f, 0.95f), 0.8f, 256.0f, 0.0f);
    st.mats[glass].transmission = 0.9f; st.mats[glass].reflect = 0.1f; st.mats[glass].ior = 1.5f;
    int f = p.detail > 0 ? p.detail : 229;
    geodesic_mesh(st, f, 1.15f, 0.15f, (uint32_t)(p.seed ^ (p.seed >> 32)), mats, mirror, glass, rng);
    st.point_light(f3(-3.0f, 3.5f, -3.0f), f3(1.0f, 0.95f, 0.9f), 300.0f, 0.0f);
    st.point_light(f3(3.5f, 1.0f, -2.5f), f3(0.7f, 0.8f, 1.0f), 150.0f, 0.0f);
    st.area_light(f3(0.0f, 4.0f, 0.0f), f3(1.0f, 0, 0), f3(0, 0, 1.0f), f3(1, 1, 1), 400.0f);
}

void gen_cfg5(Store& st, const SheenSceneParams& p, Rng& rng) {
    st.cam = look_at(f3(0, 6.0f, -14.0f), f3(0, 0, 2.0f), 50.0f);
    int n = p.detail > 0 ? p.detail : 10000;
    float sc = p.scale > 0 ? p.scale : 1.0f;
    int tex = st.texture(1024, 1024, noise_texture(1024, 1024, (uint32_t)p.seed * 7u + 1u));
    int g = st.material(f3(1, 1, 1), 0.0f, 16.0f, 0.05f);
    st.mats[g].texture_id = tex;
    st.plane(f3(0, 1, 0), 0.0f, g);
    // spheres scattered over a 30x30 patch, resting on the plane
    for (int i = 0; i < n; i++) {
        float r = rng.uni(0.08f, 0.3f);
        sheen_f3 c = f3(rng.uni(-15, 15), r, rng.uni(-8, 22));
        sheen_f3 alb = f3(rng.uni(0.05f, 0.95f), rng.uni(0.05f, 0.95f), rng.uni(0.05f, 0.95f));
        int m = st.material(alb, 0.3f, 32.0f, 0.05f);
        st.sphere(c, r, m);
    }
    for (int k = 0; k < 3; k++) {
        sheen_f3 dir = f3(rng.uni(-1, 1), -1.0f, rng.uni(-1, 1));
        float inten = rng.uni(0.5f, 2.0f);
        st.sun_light(dir, f3(1, 1, 1), inten * sc, 0.0f);
    }
}

// One scene of the uneven batch workload: triangle count log-uniform in
// [1e3, 2e5] (a displaced geodesic sphere; detail > 0 fixes it), 4-32 random
// spheres, a ground plane, a point light and a sun; square resolution drawn
// from {128, 256, 512, 1024} and 1-8 spp, all from the seed.
void gen_cfg4(Store& st, const SheenSceneParams& p, Rng& rng) {
    const float lo = std::log(1.0e3f), hi = std::log(2.0e5f);
    const double tris = p.detail > 0 ? (double)p.detail : std::exp((double)rng.uni(lo, hi));
    const int f = std::max(1, (int)std::lround(std::sqrt(tris / 20.0)));
    static const int res[4] = {128, 256, 512, 1024};
    st.view_w = st.view_h = res[rng.next() % 4];
    st.view_aa = 1 + (int)(rng.next() % 8);
    const float az = rng.uni(0.0f, 6.2831853f);
    st.cam = look_at(f3(3.4f * std::sin(az), rng.uni(0.3f, 1.5f), -3.4f * std::cos(az)), f3(0, 0, 0), 45.0f);
    std::vector<int> mats;
    for (int k = 0; k < 4; k++) {
        sheen_f3 alb = f3(rng.uni(0.15f, 0.9f), rng.uni(0.15f, 0.9f), rng.uni(0.15f, 0.9f));
        mats.push_back(st.material(alb, 0.4f, 48.0f, 0.08f));
    }
    int mirror = st.material(f3(0.9f, 0.9f, 0.9f), 0.6f, 128.0f, 0.02f);
    st.mats[mirror].reflect = 0.9f;
    int glass = st.material(f3(0.95f, 0.95f, 0.95f), 0.8f, 256.0f, 0.0f);
    st.mats[glass].transmission = 0.9f; st.mats[glass].reflect = 0.1f; st.mats[glass].ior = 1.5f;
    geodesic_mesh(st, f, 1.0f, 0.12f, (uint32_t)(p.seed ^ (p.seed >> 32)), mats, mirror, glass, rng);
    const int ns = 4 + (int)(rng.next() % 29);
    for (int i = 0; i < ns; i++) {
        const float r = rng.uni(0.08f, 0.3f);
        const float a = rng.uni(0.0f, 6.2831853f), d = rng.uni(1.4f, 3.0f);
        sheen_f3 alb = f3(rng.uni(0.1f, 0.9f), rng.uni(0.1f, 0.9f), rng.uni(0.1f, 0.9f));
        int m = st.material(alb, 0.3f, 32.0f, 0.05f);
        if (rng.uni() < 0.2f) st.mats[m].reflect = 0.6f;
        st.sphere(f3(d * std::cos(a), -1.2f + r, d * std::sin(a)), r, m);
    }
    int g = st.material(f3(0.6f, 0.6f, 0.6f), 0.1f, 16.0f, 0.1f);
    st.plane(f3(0, 1, 0), 1.2f, g);
    st.point_light(f3(rng.uni(-3.0f, 3.0f), 3.5f, rng.uni(-3.0f, 3.0f)), f3(1.0f, 0.95f, 0.9f), 250.0f, 0.0f);
    st.sun_light(f3(rng.uni(-0.6f, 0.6f), -1.0f, rng.uni(-0.6f, 0.6f)), f3(1, 1, 1), 0.8f, 0.0f);
}

// Parity scene 100: every light type, glass and mirror spheres, soft shadows.
void gen_p100(Store& st, const SheenSceneParams&, Rng& rng) {
    st.cam = look_at(f3(0.5f, 2.5f, -9.0f), f3(0, 0.5f, 0), 50.0f);
    int floor_m = st.material(f3(0.7f, 0.7f, 0.65f), 0.2f, 24.0f, 0.1f);
    st.plane(f3(0, 1, 0), 1.0f, floor_m);
    int back = st.material(f3(0.5f, 0.6f, 0.7f), 0.0f, 8.0f, 0.1f);
    st.plane(f3(0, 0, -1), 6.0f, back);
    for (int i = 0; i < 12; i++) {
        sheen_f3 alb = f3(rng.uni(0.1f, 0.9f), rng.uni(0.1f, 0.9f), rng.uni(0.1f, 0.9f));
        int m = st.material(alb, rng.uni(0.0f, 0.8f), rng.uni(4.0f, 128.0f), 0.1f);
        float r = rng.uni(0.3f, 0.9f);
        if (i % 4 == 1) st.mats[m].reflect = 0.7f;
        if (i % 4 == 2) { st.mats[m].transmission = 1.0f; st.mats[m].ior = 1.5f; st.mats[m].reflect = 0.1f; }
        if (i == 3) { st.mats[m].emission_color = f3(1.0f, 0.6f, 0.3f); st.mats[m].emission = 2.0f; }
        st.sphere(f3(rng.uni(-4, 4), -1.0f + r, rng.uni(-2, 4)), r, m);
    }
    st.point_light(f3(-3, 5, -4), f3(1, 0.9f, 0.8f), 900.0f, 0.3f);
    st.area_light(f3(2, 4.5f, 0), f3(0.8f, 0, 0), f3(0, 0, 0.8f), f3(0.8f, 0.9f, 1.0f), 500.0f);
    st.sun_light(f3(-0.4f, -1.0f, 0.3f), f3(1, 1, 1), 0.6f, 0.05f);
    st.spot_light(f3(0, 5, -3), f3(0, -1, 0.6f), f3(1, 1, 0.9f), 800.0f, 0.5f, 0.3f, 0.1f, 0.0f);
}

// Parity scene 101: triangle meshes (geodesic sphere, cubes) incl. glass.
void gen_p101(Store& st, const SheenSceneParams& p, Rng& rng) {
    st.cam = look_at(f3(0, 1.5f, -5.0f), f3(0, 0, 0), 50.0f);
    std::vector<int> diffuse;
    for (int k = 0; k < 4; k++) diffuse.push_back(st.material(f3(rng.uni(0.2f, 0.9f), rng.uni(0.2f, 0.9f), rng.uni(0.2f, 0.9f)), 0.4f, 32.0f, 0.1f));
    int mirror = st.material(f3(0.9f, 0.9f, 0.9f), 0.6f, 128.0f, 0.02f);
    st.mats[mirror].reflect = 0.9f;
    int glass = st.material(f3(0.95f, 0.95f, 0.95f), 0.8f, 256.0f, 0.0f);
    st.mats[glass].transmission = 0.9f; st.mats[glass].reflect = 0.1f; st.mats[glass].ior = 1.5f;
    geodesic_mesh(st, p.detail > 0 ? p.detail : 12, 1.0f, 0.15f, 77u, diffuse, mirror, glass, rng);
    int cm = st.material(f3(0.8f, 0.3f, 0.2f), 0.3f, 16.0f, 0.1f);
    st.cube(f3(-2.2f, -0.5f, 0.5f), 0.5f, cm);
    st.cube(f3(2.2f, -0.5f, 0.5f), 0.5f, glass);
    int floor_m = st.material(f3(0.6f, 0.6f, 0.6f), 0.1f, 16.0f, 0.1f);
    st.plane(f3(0, 1, 0), 1.2f, floor_m);
    st.point_light(f3(-3, 4, -3), f3(1, 1, 1), 400.0f, 0.0f);
    st.area_light(f3(2, 4, -1), f3(0.5f, 0, 0), f3(0, 0, 0.5f), f3(1, 1, 1), 300.0f);
}

// Parity scene 102: textures (albedo + mask + normal map) on cubes and plane.
void gen_p102(Store& st, const SheenSceneParams& p, Rng& rng) {
    (void)rng;
    st.cam = look_at(f3(0, 2.0f, -5.0f), f3(0, 0, 0), 50.0f);
    int w = 64;
    int t_alb = st.texture(w, w, noise_texture(w, w, 11u + (uint32_t)p.seed));
    int t_mask = st.texture(w, w, noise_texture(w, w, 23u + (uint32_t)p.seed));
    int t_nrm = st.texture(w, w, noise_texture(w, w, 37u + (uint32_t)p.seed));
    int m0 = st.material(f3(1, 1, 1), 0.5f, 32.0f, 0.1f);
    st.mats[m0].texture_id = t_alb;
    int m1 = st.material(f3(0.8f, 0.8f, 0.9f), 0.5f, 32.0f, 0.1f);
    st.mats[m1].mask_id = t_mask; st.mats[m1].transmission = 0.8f;
    int m2 = st.material(f3(0.7f, 0.5f, 0.3f), 0.6f, 64.0f, 0.1f);
    st.mats[m2].normal_id = t_nrm; st.mats[m2].normal_strength = 0.8f; st.mats[m2].texture_id = t_alb;
    st.cube(f3(-1.6f, 0, 0), 0.6f, m0);
    st.cube(f3(0, 0, 0.5f), 0.6f, m1);
    st.cube(f3(1.6f, 0, 0), 0.6f, m2);
    int g = st.material(f3(1, 1, 1), 0.0f, 8.0f, 0.1f);
    st.mats[g].texture_id = t_alb;
    st.plane(f3(0, 1, 0), 0.6f, g);
    st.point_light(f3(-2, 4, -4), f3(1, 1, 1), 500.0f, 0.0f);
}


// Cubic Bezier runs in groups of four control points, `subdiv` segments per
// cubic, first point once (host tessellation of the reference's
// lathe_bezier / prism_bezier, api/scene_io.cpp:66-102).
std::vector<sheen_f2> bezier_run(const std::vector<sheen_f2>& ctrl, int subdiv) {
    std::vector<sheen_f2> out;
    if (ctrl.size() < 4) return out;
    if (subdiv < 1) subdiv = 1;
    const int nseg = (int)ctrl.size() / 4;
    out.push_back(ctrl[0]);
    for (int s = 0; s < nseg; s++) {
        const sheen_f2 *P = &ctrl[4 * s];
        for (int k = 1; k <= subdiv; k++) {
            float t = (float)k / (float)subdiv, u = 1.0f - t;
            float w0 = u * u * u, w1 = 3.0f * u * u * t, w2 = 3.0f * u * t * t, w3 = t * t * t;
            out.push_back({w0 * P[0].x + w1 * P[1].x + w2 * P[2].x + w3 * P[3].x,
                           w0 * P[0].y + w1 * P[1].y + w2 * P[2].y + w3 * P[3].y});
        }
    }
    return out;
}

// Image-mode height map of seeded fBm (8-bit, row-major).
std::vector<unsigned char> fbm_image(int w, int h, uint32_t seed) {
    std::vector<unsigned char> px((size_t)w * h);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float v = fbm3(f3(x * 4.0f / w, 0.37f, y * 4.0f / h), 6, seed);
            px[(size_t)y * w + x] = (unsigned char)std::lround(255.0f * std::fmin(std::fmax(1.6f * (v - 0.2f), 0.0f), 1.0f));
        }
    return px;
}

// Analytic primitives + CSG + caustics (workload 3; `detail` primitives).
// Half the primitives are CSG-only sphere leaves (Object.type 1, which the
// sphere BVH skips) in depth-3 trees of 8 leaves under one balanced union
// root; the rest are spheres (10% glass), superellipsoids (exponent n = 2/e
// for e in U[0.2,2]), 2-point cone sweeps, multi-point sweeps, 12-point
// prisms, a few Bezier lathes (the reference ignores a lathe's centre, so
// they stand on the Y axis) and one fBm height map.
void gen_cfg3(Store& st, const SheenSceneParams& p, Rng& rng, bool image_hf) {
    const int N = p.detail > 0 ? p.detail : 10000;
    const float ext = N >= 2000 ? 24.0f : 8.0f;
    st.cam = look_at(f3(0.0f, 0.45f * ext + 2.0f, -1.35f * ext), f3(0.0f, 1.0f, 0.0f), 60.0f);
    std::vector<int> mats;
    for (int k = 0; k < 48; k++) {
        int m = st.material(f3(rng.uni(0.1f, 0.9f), rng.uni(0.1f, 0.9f), rng.uni(0.1f, 0.9f)), rng.uni(0.0f, 0.7f),
                            rng.uni(8.0f, 96.0f), 0.08f);
        if (k % 8 == 3) st.mats[m].reflect = 0.6f;
        mats.push_back(m);
    }
    int glass = st.material(f3(0.95f, 0.95f, 0.95f), 0.6f, 128.0f, 0.02f);
    st.mats[glass].transmission = 1.0f;
    st.mats[glass].ior = 1.5f;
    auto mat = [&]() { return mats[(rng.next() >> 33) % mats.size()]; };
    auto pos = [&](float lo, float hi) { return f3(rng.uni(-ext, ext), rng.uni(lo, hi), rng.uni(-ext, ext)); };
    // ground height map
    HeightField hf;
    std::memset(&hf, 0, sizeof(hf));
    hf.center = f3(0.0f, -0.5f, 0.0f);
    hf.size_x = hf.size_z = 2.6f * ext;
    hf.height = 1.5f;
    hf.freq = 6.0f;
    hf.octaves = 5;
    hf.seed = (int)(p.seed & 0x7fffffff);
    hf.material_id = st.material(f3(0.55f, 0.5f, 0.42f), 0.05f, 8.0f, 0.1f);
    if (image_hf) {
        const int R = 1024;
        st.hf_images.push_back(fbm_image(R, R, (uint32_t)p.seed * 3u + 5u));
        hf.image = st.hf_images.back().data();
        hf.img_w = hf.img_h = R;
        hf.seed = -1;
    }
    st.hfs.push_back(hf);
    int left = N - 1;
    // CSG: trees of 8 CSG-only spheres
    const int n_csg = (left / 2) / 8;
    std::vector<int> roots;
    for (int t = 0; t < n_csg; t++) {
        sheen_f3 c = pos(1.0f, 4.0f);
        int base = (int)st.csg.size();
        // 7 interior nodes (heap order) then 8 leaves
        for (int k = 0; k < 7; k++) {
            CSGNode n; std::memset(&n, 0, sizeof(n));
            float r = rng.uni();
            n.op = r < 0.5f ? 0 : (r < 0.75f ? 1 : 2);
            n.leaf_kind = -1;
            n.left = base + 2 * k + 1;
            n.right = base + 2 * k + 2;
            st.csg.push_back(n);
        }
        for (int k = 0; k < 8; k++) {
            Object o; std::memset(&o, 0, sizeof(o));
            o.type = 1;
            o.material_id = mat();
            o.param1 = add(c, f3(rng.uni(-0.6f, 0.6f), rng.uni(-0.6f, 0.6f), rng.uni(-0.6f, 0.6f)));
            o.param2 = rng.uni(0.3f, 0.8f);
            st.objects.push_back(o);
            CSGNode n; std::memset(&n, 0, sizeof(n));
            n.op = 0; n.leaf_kind = 0; n.leaf_index = (int)st.objects.size() - 1; n.left = n.right = -1;
            st.csg.push_back(n);
        }
        roots.push_back(base);
        left -= 8;
    }
    while (roots.size() > 1) {  // balanced union of the trees
        std::vector<int> next;
        for (size_t i = 0; i + 1 < roots.size(); i += 2) {
            CSGNode n; std::memset(&n, 0, sizeof(n));
            n.op = 0; n.leaf_kind = -1; n.left = roots[i]; n.right = roots[i + 1];
            st.csg.push_back(n);
            next.push_back((int)st.csg.size() - 1);
        }
        if (roots.size() % 2) next.push_back(roots.back());
        roots.swap(next);
    }
    st.csg_root = roots.empty() ? -1 : roots[0];
    // standalone primitives
    const int n_se = left / 5, n_cone = left / 10, n_sweep = left / 10, n_prism = left * 4 / 25;
    const int n_lathe = std::max(2, std::min(20, left / 250));
    for (int i = 0; i < n_se; i++) {
        Superellipsoid e; std::memset(&e, 0, sizeof(e));
        e.center = pos(0.8f, 4.0f);
        e.a = rng.uni(0.2f, 0.6f); e.b = rng.uni(0.2f, 0.6f); e.c = rng.uni(0.2f, 0.6f);
        e.n = 2.0f / rng.uni(0.2f, 2.0f);
        e.material_id = mat();
        st.ses.push_back(e);
    }
    auto strand = [&](int npts, bool cone) {
        SweepStrand s; std::memset(&s, 0, sizeof(s));
        s.point_offset = (int)st.sweep_points.size();
        s.point_count = npts;
        s.material_id = mat();
        sheen_f3 q = pos(0.5f, 4.0f);
        for (int k = 0; k < npts; k++) {
            SweepPoint sp; std::memset(&sp, 0, sizeof(sp));
            sp.position = q;
            sp.radius = cone ? (k == 0 ? rng.uni(0.2f, 0.5f) : rng.uni(0.02f, 0.1f)) : rng.uni(0.05f, 0.15f);
            st.sweep_points.push_back(sp);
            q = add(q, f3(rng.uni(-0.6f, 0.6f), rng.uni(0.2f, 0.8f), rng.uni(-0.6f, 0.6f)));
        }
        st.strands.push_back(s);
    };
    for (int i = 0; i < n_cone; i++) strand(2, true);
    for (int i = 0; i < n_sweep; i++) strand(4 + (int)(rng.next() % 3), false);
    for (int i = 0; i < n_prism; i++) {
        Prism pr; std::memset(&pr, 0, sizeof(pr));
        pr.center = pos(0.0f, 0.0f);
        pr.y_min = rng.uni(0.2f, 1.0f);
        pr.y_max = pr.y_min + rng.uni(0.2f, 1.5f);
        pr.point_offset = (int)st.prism_points.size();
        pr.point_count = 13;
        pr.material_id = mat();
        float r0 = rng.uni(0.2f, 0.6f), r1 = r0 * rng.uni(0.3f, 0.9f), ph = rng.uni(0.0f, 6.2831853f);
        PrismPoint first{};
        for (int k = 0; k < 12; k++) {
            float a = ph + 6.2831853f * k / 12.0f, rr = (k % 2) ? r1 : r0;
            PrismPoint q; std::memset(&q, 0, sizeof(q));
            q.x = rr * std::cos(a); q.z = rr * std::sin(a);
            if (k == 0) first = q;
            st.prism_points.push_back(q);
        }
        st.prism_points.push_back(first);
        st.prisms.push_back(pr);
    }
    for (int i = 0; i < n_lathe; i++) {
        std::vector<sheen_f2> ctrl;
        float y = 0.0f, r = rng.uni(0.3f, 0.8f) * (1.0f + i * 0.15f);
        for (int s = 0; s < 8; s++) {
            float y3 = y + rng.uni(0.15f, 0.35f), r3 = std::fmax(0.05f, r + rng.uni(-0.25f, 0.25f));
            ctrl.push_back({y, r});
            ctrl.push_back({y + (y3 - y) * 0.33f, r + rng.uni(-0.2f, 0.2f)});
            ctrl.push_back({y + (y3 - y) * 0.66f, r3 + rng.uni(-0.2f, 0.2f)});
            ctrl.push_back({y3, r3});
            y = y3; r = r3;
        }
        std::vector<sheen_f2> line = bezier_run(ctrl, 4);
        Lathe l; std::memset(&l, 0, sizeof(l));
        l.center = f3(0, 0, 0);
        l.point_offset = (int)st.lathe_points.size();
        l.point_count = (int)line.size();
        l.material_id = mat();
        for (auto& q : line) {
            LathePoint lp; std::memset(&lp, 0, sizeof(lp));
            lp.y = q.x; lp.radius = q.y;
            st.lathe_points.push_back(lp);
        }
        st.lathes.push_back(l);
    }
    left -= n_se + n_cone + n_sweep + n_prism + n_lathe;
    for (int i = 0; i < left; i++) {
        int m = (i % 10 == 0) ? glass : mat();
        st.sphere(pos(0.6f, 4.5f), rng.uni(0.15f, 0.7f), m);
    }
    st.spot_light(f3(0.0f, 0.9f * ext, -0.2f * ext), f3(0.0f, -1.0f, 0.2f), f3(1.0f, 0.95f, 0.85f),
                  120.0f * ext * ext, 1.2f, 0.3f, 0.1f, 0.0f);
    st.point_light(f3(-0.6f * ext, 0.5f * ext, -0.6f * ext), f3(0.7f, 0.8f, 1.0f), 40.0f * ext * ext, 0.0f);
    st.sun_light(f3(0.3f, -1.0f, 0.4f), f3(1.0f, 1.0f, 1.0f), 0.4f, 0.0f);
    // per light: the spot and the point light store about 46 % of what they
    // emit (the sun emits none), so 1,081,000 per light stores about 1 M
    // photons, the workload's map size
    st.photons = p.scale > 0 ? (int)p.scale : (N >= 2000 ? 1081000 : 20000);
    st.photon_cell = 0.25f;
    st.caustics = 0.3f;
}

// Parity scene 105: hair strands with optical-depth shadows, normal maps,
// spot gobo, sun with an angular radius, emission, a scattering medium and a
// visible spot beam.
void gen_p105(Store& st, const SheenSceneParams& p, Rng& rng) {
    st.cam = look_at(f3(0.0f, 1.6f, -4.5f), f3(0.0f, 0.6f, 0.0f), 55.0f);
    int floor_m = st.material(f3(0.6f, 0.6f, 0.6f), 0.1f, 16.0f, 0.1f);
    int tex = st.texture(32, 32, noise_texture(32, 32, 91u + (uint32_t)p.seed));
    st.mats[floor_m].normal_id = tex;
    st.mats[floor_m].normal_strength = 0.5f;
    st.plane(f3(0, 1, 0), 0.0f, floor_m);
    int skin = st.material(f3(0.75f, 0.6f, 0.5f), 0.25f, 24.0f, 0.15f);
    st.sphere(f3(0.0f, 0.8f, 0.5f), 0.5f, skin);
    int hair = st.material(f3(0.35f, 0.22f, 0.12f), 0.6f, 128.0f, 0.1f);
    Material& hm = st.mats[hair];
    hm.anisotropy = 1.0f; hm.hair_shift = 0.12f; hm.hair_spec_r = 1.2f; hm.hair_exp_r = 12.0f;
    hm.hair_spec_t = 0.5f; hm.hair_exp_t = 6.0f; hm.hair_backlight = 0.5f;
    for (int k = 0; k < 60; k++) {
        SweepStrand s; std::memset(&s, 0, sizeof(s));
        s.point_offset = (int)st.sweep_points.size();
        s.point_count = 6;
        s.material_id = hair;
        float a = rng.uni(0.0f, 6.2831853f), b = rng.uni(0.2f, 1.3f);
        sheen_f3 q = add(f3(0.0f, 0.8f, 0.5f), mul(f3(std::cos(a) * std::sin(b), std::cos(b), std::sin(a) * std::sin(b)), 0.52f));
        for (int j = 0; j < 6; j++) {
            SweepPoint sp; std::memset(&sp, 0, sizeof(sp));
            sp.position = q;
            sp.radius = 0.012f * (1.0f - 0.1f * j);
            st.sweep_points.push_back(sp);
            q = add(q, f3(std::cos(a) * 0.08f, -0.09f, std::sin(a) * 0.08f));
        }
        st.strands.push_back(s);
    }
    int lamp = st.material(f3(1, 1, 1), 0.0f, 8.0f, 0.0f);
    st.mats[lamp].emission_color = f3(1.0f, 0.8f, 0.5f);
    st.mats[lamp].emission = 3.0f;
    st.sphere(f3(1.4f, 0.25f, 0.2f), 0.25f, lamp);
    st.sun_light(f3(-0.3f, -1.0f, 0.5f), f3(1.0f, 0.95f, 0.9f), 0.5f, 0.04f);
    st.spot_light(f3(-1.2f, 2.6f, -0.6f), f3(0.5f, -1.0f, 0.4f), f3(1.0f, 1.0f, 0.9f), 300.0f, 0.6f, 0.3f, 0.05f, 0.8f);
    st.lights.back().texture_id = tex;
    Medium md; std::memset(&md, 0, sizeof(md));
    md.center = f3(-0.8f, 0.8f, 0.0f); md.half = f3(0.8f, 0.8f, 0.8f);
    md.sigma_s = f3(0.3f, 0.3f, 0.35f); md.sigma_t = 0.4f; md.samples = 8;
    st.media.push_back(md);
}

// Parity scene 103: the reference's scenes/caustics.sheen restated
// (glass sphere on a floor under a spot light, 20,000 photons, cell 0.25,
// caustic scale 0.3, legacy camera with focal_length 0).
void gen_p103(Store& st, const SheenSceneParams& p) {
    std::memset(&st.cam, 0, sizeof(st.cam));
    st.cam.position = f3(0.0f, 1.2f, 4.0f);
    st.cam.forward = f3(0.0f, -0.1f, -1.0f);
    st.cam.right = f3(1, 0, 0);
    st.cam.up = f3(0, 1, 0);
    int floor_m = st.material(f3(0.2f, 0.2f, 0.22f), 0.1f, 16.0f, 0.1f);
    int glass = st.material(f3(0.9f, 0.9f, 0.9f), 0.5f, 64.0f, 0.05f);
    st.plane(f3(0, 1, 0), 0.0f, floor_m);
    st.sphere(f3(0.0f, 1.0f, 0.0f), 0.5f, glass);
    st.mats[glass].transmission = 1.0f;
    st.mats[glass].ior = 1.45f;
    // spot: angle is the full cone (scene_io.cpp:615-624), direction normalised by the scene builder
    Light l = Store::base_light();
    l.type = 3;
    l.position = f3(0.0f, 3.0f, 2.0f);
    sheen_f3 d = f3(0.0f, -1.0f, -0.8f);
    float inv = 1.0f / std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
    l.direction = f3(d.x * inv, d.y * inv, d.z * inv);
    l.color = f3(1, 1, 1);
    l.intensity = 200.0f;
    l.angle = 0.5f * 0.5f;
    l.blend = 0.2f;
    l.radius = 0.05f;
    l.beam = 0.0f;
    st.lights.push_back(l);
    st.photons = p.detail > 0 ? p.detail : 20000;
    st.photon_cell = 0.25f;
    st.caustics = 0.3f;
}

}  // namespace

extern "C" {

int sheen_scene_defaults(int config, SheenSceneParams* out) {
    if (!out) return -1;
    std::memset(out, 0, sizeof(*out));
    out->config = config;
    out->seed = 0x5EED0000ull + (uint64_t)config;
    switch (config) {
        case 1: out->width = 256; out->height = 256; out->aa = 1; out->detail = 64; break;
        case 2: out->width = 1920; out->height = 1080; out->aa = 4; out->detail = 229; break;
        case 4: out->width = 256; out->height = 256; out->aa = 4; out->detail = 0; break;
        case 5: out->width = 512; out->height = 512; out->aa = 8; out->detail = 10000; out->scale = 1.0f; break;
        case 100: case 101: case 102: out->width = 160; out->height = 120; out->aa = 4; out->detail = 0; break;
        case 103: out->width = 320; out->height = 240; out->aa = 1; out->detail = 20000; break;
        case 3: out->width = 1280; out->height = 720; out->aa = 16; out->detail = 10000; break;
        case 104: out->width = 160; out->height = 90; out->aa = 4; out->detail = 300; break;
        case 105: out->width = 160; out->height = 120; out->aa = 4; out->detail = 0; break;
        default: return -1;
    }
    return 0;
}

void* sheen_scene_generate(const SheenSceneParams* p) {
    if (!p) return nullptr;
    SheenSceneParams q = *p;
    if (q.seed == 0) q.seed = 0x5EED0000ull + (uint64_t)q.config;
    std::unique_ptr<Store> st(new Store());
    Rng rng(q.seed);
    switch (q.config) {
        case 1: gen_cfg1(*st, q, rng); break;
        case 2: gen_cfg2(*st, q, rng); break;
        case 4: gen_cfg4(*st, q, rng); break;
        case 5: gen_cfg5(*st, q, rng); break;
        case 100: gen_p100(*st, q, rng); break;
        case 101: gen_p101(*st, q, rng); break;
        case 102: gen_p102(*st, q, rng); break;
        case 103: gen_p103(*st, q); break;
        case 3: gen_cfg3(*st, q, rng, true); break;
        case 104: gen_cfg3(*st, q, rng, false); break;
        case 105: gen_p105(*st, q, rng); break;
        default: return nullptr;
    }
    st->finalize();
    return st.release();
}

const SheenScene* sheen_scene_get(void* h) { return h ? &((Store*)h)->scene : nullptr; }

int sheen_scene_view(void* h, int* width, int* height, int* aa) {
    Store* s = (Store*)h;
    if (!s || !s->view_w) return 0;
    if (width) *width = s->view_w;
    if (height) *height = s->view_h;
    if (aa) *aa = s->view_aa;
    return 1;
}

Material* sheen_scene_materials(void* h, int* count) {
    Store* s = (Store*)h;
    if (count) *count = (int)s->mats.size();
    return s->mats.data();
}
Light* sheen_scene_lights(void* h, int* count) {
    Store* s = (Store*)h;
    if (count) *count = (int)s->lights.size();
    return s->lights.data();
}
unsigned char* sheen_scene_texture(void* h, int tex, int* w, int* hh) {
    Store* s = (Store*)h;
    if (tex < 0 || tex >= (int)s->tex.size()) return nullptr;
    if (w) *w = s->tex_w[tex];
    if (hh) *hh = s->tex_h[tex];
    return s->tex[tex].data();
}

void sheen_scene_free(void* h) { delete (Store*)h; }

}  // extern "C"
