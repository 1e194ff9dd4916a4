// sheen_file.cpp — the `.sheen` scene text format (reference
// api/scene_io.cpp:216-968 with the scene builders of scene.h:8-24, 160-330)
// into a SheenScene the C ABI renders.
//
// One directive per line, '#' starts a comment.  Positional directives take
// their numbers in order; keyword directives take key=value tokens (vectors
// as x,y,z).  Curves are tessellated on the host, before upload: sphere
// sweeps through Catmull-Rom segments, lathe and prism profiles through runs
// of cubic Bezier segments, so the kernels only see polylines.
//
// Differences from the reference loader, each deliberate:
//   * image heightfields point into buffers the loaded scene owns (the
//     reference keeps a pointer to a local vector: App. A-11);
//   * CSG-only superellipsoids (se_csg) carry _pad0 = 1, which the BVH
//     builders skip, because a SheenScene has no other way to say it;
//   * missing numbers read as 0 and malformed keyword values are an error
//     (the reference reads uninitialised values / throws from std::stof);
//   * the CSG stack belongs to one load (the reference's is a global).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "sheen_c_api.h"

namespace {

using F3 = sheen_f3;
inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
inline F3 add(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 mul(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
inline F3 cross3(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline F3 normalized(F3 v) {
    const float s = 1.0f / std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return mul(v, s);
}

// Catmull-Rom point between p1 and p2 (p0, p3 shape the tangents): the
// uniform spline 0.5 * [2p1 + (p2-p0)u + (2p0-5p1+4p2-p3)u^2 + (3p1-p0-3p2+p3)u^3]
// (reference scene.h:8-24, the POV-Ray sphere_sweep matrix).
inline float cr(float p0, float p1, float p2, float p3, float u) {
    const float u2 = u * u, u3 = u2 * u;
    return 0.5f * ((2.0f * p1) + (-p0 + p2) * u + (2.0f * p0 - 5.0f * p1 + 4.0f * p2 - p3) * u2 +
                   (-p0 + 3.0f * p1 - 3.0f * p2 + p3) * u3);
}

// Cubic Bezier runs: control points in groups of 4, each group one segment
// from its first to its fourth point; `subdiv` chords per segment; the first
// point once, then the subdiv points of each segment (scene_io.cpp:66-102).
struct P2 { float a, b; };
std::vector<P2> bezier_run(std::vector<P2> ctrl, int subdiv) {
    ctrl.resize(ctrl.size() / 4 * 4);
    std::vector<P2> out;
    if (ctrl.size() < 4) return out;
    if (subdiv < 1) subdiv = 1;
    out.push_back(ctrl[0]);
    for (size_t s = 0; s + 3 < ctrl.size(); s += 4) {
        for (int k = 1; k <= subdiv; k++) {
            const float t = (float)k / (float)subdiv, u = 1.0f - t;
            const float w0 = u * u * u, w1 = 3.0f * u * u * t, w2 = 3.0f * u * t * t, w3 = t * t * t;
            out.push_back({w0 * ctrl[s].a + w1 * ctrl[s + 1].a + w2 * ctrl[s + 2].a + w3 * ctrl[s + 3].a,
                           w0 * ctrl[s].b + w1 * ctrl[s + 1].b + w2 * ctrl[s + 2].b + w3 * ctrl[s + 3].b});
        }
    }
    return out;
}

// The scene being built, and everything its SheenScene points at.
struct Loaded {
    SheenScene scene{};
    CameraData cam{};
    std::vector<Object> objects;
    std::vector<Plane> planes;
    std::vector<Triangle> tris;
    std::vector<SweepPoint> sweep_points;
    std::vector<SweepStrand> strands;
    std::vector<Material> mats;
    std::vector<Light> lights;
    std::vector<Medium> media;
    std::vector<HeightField> hfs;
    std::vector<std::unique_ptr<std::vector<unsigned char>>> hf_images;
    std::vector<Superellipsoid> ses;
    std::vector<LathePoint> lathe_points;
    std::vector<Lathe> lathes;
    std::vector<PrismPoint> prism_points;
    std::vector<Prism> prisms;
    std::vector<CSGNode> csg;
    int csg_root = -1;
    std::vector<int> csg_stack;
    std::vector<std::vector<unsigned char>> tex;
    std::vector<int> tex_w, tex_h;
    std::vector<const unsigned char*> tex_ptr;
    int width = 800, height = 600, aa = 4, photons = 0;
    float photon_cell = 0.25f, caustics = 0.0f;
    std::string output = "scene_output.png";
    std::string dir;  // the file's directory: relative paths in it are tried there too

    void material(F3 c, float spec, float shin, float amb, float aniso = 0.0f, float shift = 0.1f, float sr = 0.0f,
                  float er = 0.0f, float st = 0.0f, float et = 0.0f, float back = 0.0f) {
        Material m;
        std::memset(&m, 0, sizeof(m));
        m.diffuse_color = c;
        m.specular = spec; m.shininess = shin; m.ambient = amb;
        m.anisotropy = aniso; m.hair_shift = shift;
        m.hair_spec_r = sr; m.hair_exp_r = er; m.hair_spec_t = st; m.hair_exp_t = et; m.hair_backlight = back;
        m.texture_id = -1; m.mask_id = -1; m.normal_id = -1;
        m.ior = 1.45f;
        m.normal_strength = 1.0f;
        mats.push_back(m);
    }
    void sphere(F3 c, float r, int m) {
        Object o;
        std::memset(&o, 0, sizeof(o));
        o.type = 0; o.material_id = m; o.param1 = c; o.param2 = r;
        objects.push_back(o);
    }
    void triangle(F3 a, F3 b, F3 c, F3 na, F3 nb, F3 nc, int m, sheen_f2 ta, sheen_f2 tb, sheen_f2 tc) {
        Triangle t;
        std::memset(&t, 0, sizeof(t));
        t.v0 = a; t.v1 = b; t.v2 = c; t.n0 = na; t.n1 = nb; t.n2 = nc;
        t.uv0 = ta; t.uv1 = tb; t.uv2 = tc;
        t.material_id = m;
        tris.push_back(t);
    }
    // 12 triangles, per-face UVs over [0,1]^2 (scene.h:271-301)
    void cube(F3 c, float s, int m) {
        F3 v[8];
        for (int k = 0; k < 8; k++)
            v[k] = f3(c.x + ((k >> 2) & 1 ? s : -s), c.y + ((k >> 1) & 1 ? s : -s), c.z + (k & 1 ? s : -s));
        static const int face[12][3] = {{0, 1, 3}, {0, 3, 2}, {4, 6, 7}, {4, 7, 5}, {0, 4, 5}, {0, 5, 1},
                                        {2, 3, 7}, {2, 7, 6}, {0, 2, 6}, {0, 6, 4}, {1, 5, 7}, {1, 7, 3}};
        const F3 z = f3(0, 0, 0);
        for (int i = 0; i < 12; i++) {
            const int axis = i / 4;  // faces 0-3: x, 4-7: y, 8-11: z
            auto uv = [&](F3 p) {
                const float u = (axis == 0) ? p.z - (c.z - s) : p.x - (c.x - s);
                const float w = (axis == 1) ? p.z - (c.z - s) : p.y - (c.y - s);
                return sheen_f2{u / (2.0f * s), w / (2.0f * s)};
            };
            const F3 a = v[face[i][0]], b = v[face[i][1]], d = v[face[i][2]];
            triangle(a, b, d, z, z, z, m, uv(a), uv(b), uv(d));
        }
    }
    void sweep_linear(const std::vector<SweepPoint>& pts, int m) {
        SweepStrand s;
        std::memset(&s, 0, sizeof(s));
        s.point_offset = (int)sweep_points.size();
        s.point_count = (int)pts.size();
        s.material_id = m;
        sweep_points.insert(sweep_points.end(), pts.begin(), pts.end());
        strands.push_back(s);
    }
    // Catmull-Rom through every control point, with mirrored phantom ends;
    // radii follow the same spline, floored at 1e-5 (scene.h:160-207).
    void sweep_smooth(const std::vector<SweepPoint>& c, int m, int subdiv) {
        const int n = (int)c.size();
        if (n < 2) return;
        std::vector<SweepPoint> p;
        SweepPoint head = c[0], tail = c[n - 1];
        head.position = add(c[0].position, sub(c[0].position, c[1].position));
        head.radius = c[0].radius + (c[0].radius - c[1].radius);
        tail.position = add(c[n - 1].position, sub(c[n - 1].position, c[n - 2].position));
        tail.radius = c[n - 1].radius + (c[n - 1].radius - c[n - 2].radius);
        p.push_back(head);
        p.insert(p.end(), c.begin(), c.end());
        p.push_back(tail);
        std::vector<SweepPoint> out;
        out.push_back(p[1]);
        for (size_t i = 1; i + 2 < p.size(); i++) {
            for (int j = 1; j <= subdiv; j++) {
                const float u = (float)j / (float)subdiv;
                SweepPoint q;
                std::memset(&q, 0, sizeof(q));
                q.position = f3(cr(p[i - 1].position.x, p[i].position.x, p[i + 1].position.x, p[i + 2].position.x, u),
                                cr(p[i - 1].position.y, p[i].position.y, p[i + 1].position.y, p[i + 2].position.y, u),
                                cr(p[i - 1].position.z, p[i].position.z, p[i + 1].position.z, p[i + 2].position.z, u));
                q.radius = std::max(cr(p[i - 1].radius, p[i].radius, p[i + 1].radius, p[i + 2].radius, u), 1e-5f);
                out.push_back(q);
            }
        }
        sweep_linear(out, m);
    }
    Light& light(int type) {
        Light l;
        std::memset(&l, 0, sizeof(l));
        l.direction = f3(0, 0, -1);
        l.color = f3(1, 1, 1);
        l.intensity = 1.0f;
        l.type = type;
        l.flags = 1;  // seen by media
        l.texture_id = -1;
        lights.push_back(l);
        return lights.back();
    }
    void texture(int w, int h, std::vector<unsigned char> px) {
        tex.push_back(std::move(px));
        tex_w.push_back(w);
        tex_h.push_back(h);
    }

    void finalize() {
        // out-of-range material ids read as 0 (scene_io.cpp:945-962)
        const int nm = (int)mats.size();
        auto fix = [&](int& m) {
            if (nm > 0 && (m < 0 || m >= nm)) m = 0;
        };
        for (auto& s : strands) fix(s.material_id);
        for (auto& t : tris) fix(t.material_id);
        for (auto& o : objects) fix(o.material_id);
        for (auto& p : planes) fix(p.material_id);
        for (auto& h : hfs) fix(h.material_id);
        for (auto& s : ses) fix(s.material_id);
        for (auto& l : lathes) fix(l.material_id);
        for (auto& p : prisms) fix(p.material_id);
        tex_ptr.clear();
        for (auto& t : tex) tex_ptr.push_back(t.data());
        SheenScene& s = scene;
        std::memset(&s, 0, sizeof(s));
        s.camera = cam;
        auto set = [](auto& vec, auto*& ptr, int& count) {
            ptr = vec.empty() ? nullptr : vec.data();
            count = (int)vec.size();
        };
        set(objects, s.objects, s.num_objects);
        set(planes, s.planes, s.num_planes);
        set(tris, s.triangles, s.num_triangles);
        set(sweep_points, s.sweep_points, s.num_sweep_points);
        set(strands, s.sweep_strands, s.num_sweep_strands);
        set(mats, s.materials, s.num_materials);
        set(lights, s.lights, s.num_lights);
        set(media, s.media, s.num_media);
        set(hfs, s.heightfields, s.num_heightfields);
        set(ses, s.superellipsoids, s.num_superellipsoids);
        set(lathe_points, s.lathe_points, s.num_lathe_points);
        set(lathes, s.lathes, s.num_lathes);
        set(prism_points, s.prism_points, s.num_prism_points);
        set(prisms, s.prisms, s.num_prisms);
        set(csg, s.csg_nodes, s.num_csg_nodes);
        s.csg_root = csg_root;
        s.num_textures = (int)tex.size();
        s.texture_pixels = tex_ptr.empty() ? nullptr : tex_ptr.data();
        s.texture_widths = tex_w.empty() ? nullptr : tex_w.data();
        s.texture_heights = tex_h.empty() ? nullptr : tex_h.data();
        s.photons = photons;
        s.photon_cell = photon_cell;
        s.caustics = caustics;
    }
};

// Tokens of one line.  Numbers past the end read as 0 and set `short_`.
struct Line {
    std::vector<std::string> tok;
    size_t at = 1;
    bool short_ = false;
    bool bad = false;
    bool keyword() const { return at < tok.size() && tok[at].find('=') != std::string::npos; }
    bool more() const { return at < tok.size(); }
    float f() {
        if (at >= tok.size()) { short_ = true; return 0.0f; }
        char* end = nullptr;
        const float v = std::strtof(tok[at].c_str(), &end);
        if (end == tok[at].c_str()) bad = true;
        at++;
        return v;
    }
    int i() { return (int)f_int(); }
    long long f_int() {
        if (at >= tok.size()) { short_ = true; return 0; }
        char* end = nullptr;
        const long long v = std::strtoll(tok[at].c_str(), &end, 10);
        if (end == tok[at].c_str()) bad = true;
        at++;
        return v;
    }
    F3 v3() {
        const float x = f(), y = f(), z = f();
        return f3(x, y, z);
    }
    std::string s() {
        if (at >= tok.size()) { short_ = true; return ""; }
        return tok[at++];
    }
    // key=value tokens from the cursor on
    std::map<std::string, std::string> kv() {
        std::map<std::string, std::string> m;
        for (; at < tok.size(); at++) {
            const size_t eq = tok[at].find('=');
            if (eq != std::string::npos) m[tok[at].substr(0, eq)] = tok[at].substr(eq + 1);
        }
        return m;
    }
};

struct KV {
    std::map<std::string, std::string> m;
    bool bad = false;
    bool has(const char* k) const { return m.count(k) > 0; }
    float f(const char* k, float d) {
        auto it = m.find(k);
        if (it == m.end()) return d;
        char* end = nullptr;
        const float v = std::strtof(it->second.c_str(), &end);
        if (end == it->second.c_str()) bad = true;
        return v;
    }
    long long l(const char* k, long long d) {
        auto it = m.find(k);
        if (it == m.end()) return d;
        char* end = nullptr;
        const long long v = std::strtoll(it->second.c_str(), &end, 10);
        if (end == it->second.c_str()) bad = true;
        return v;
    }
    int i(const char* k, int d) { return (int)l(k, d); }
    F3 v3(const char* k, F3 d) {
        auto it = m.find(k);
        if (it == m.end()) return d;
        std::string v = it->second;
        std::replace(v.begin(), v.end(), ',', ' ');
        std::istringstream ss(v);
        float x, y, z;
        if (!(ss >> x >> y >> z)) return d;
        return f3(x, y, z);
    }
    std::string s(const char* k, const std::string& d) {
        auto it = m.find(k);
        return it == m.end() ? d : it->second;
    }
};

class Parser {
public:
    explicit Parser(Loaded& L) : L_(L) {}
    std::string error;

    bool line(const std::string& raw) {
        std::string text = raw.substr(0, raw.find('#'));
        Line ln;
        std::istringstream in(text);
        std::string t;
        while (in >> t) ln.tok.push_back(t);
        if (ln.tok.empty()) return true;
        const std::string& cmd = ln.tok[0];
        bool ok = dispatch(cmd, ln);
        if (ok && ln.bad) { error = "malformed number"; ok = false; }
        if (!ok && error.empty()) error = "bad directive";
        return ok;
    }

private:
    Loaded& L_;

    // a relative name as given, else next to the scene file, else next to
    // the directory holding it (scenes written relative to a project root)
    std::string path(const std::string& p) const {
        if (p.empty() || p[0] == '/' || L_.dir.empty()) return p;
        const std::string parent = L_.dir.find('/') == std::string::npos ? "." : L_.dir.substr(0, L_.dir.rfind('/'));
        for (const std::string& c : {p, L_.dir + "/" + p, parent + "/" + p}) {
            std::ifstream probe(c);
            if (probe) return c;
        }
        return p;
    }

    bool dispatch(const std::string& cmd, Line& ln) {
        if (cmd == "resolution") return resolution(ln);
        if (cmd == "camera") return camera(ln);
        if (cmd == "material") return material(ln);
        if (cmd == "sphere") { F3 c = ln.v3(); float r = ln.f(); L_.sphere(c, r, ln.i()); return true; }
        if (cmd == "plane") {
            Plane p;
            std::memset(&p, 0, sizeof(p));
            p.normal = ln.v3(); p.d = ln.f(); p.material_id = ln.i();
            L_.planes.push_back(p);
            return true;
        }
        if (cmd == "cube") { F3 c = ln.v3(); float s = ln.f(); L_.cube(c, s, ln.i()); return true; }
        if (cmd == "triangle") {
            F3 a = ln.v3(), b = ln.v3(), c = ln.v3();
            const F3 z = f3(0, 0, 0);
            L_.triangle(a, b, c, z, z, z, ln.i(), {0, 0}, {0, 0}, {0, 0});
            return true;
        }
        if (cmd == "grid") return grid(ln);
        if (cmd == "obj") return obj(ln);
        if (cmd == "mesh") return mesh(ln);
        if (cmd == "texture") return texture(ln);
        if (cmd == "strands") return strands(ln);
        if (cmd == "hair") return hair(ln);
        if (cmd == "hair_mesh") return hair_mesh(ln);
        if (cmd == "light" || cmd == "fill_light") {
            F3 d = ln.v3(), c = ln.v3();
            float inten = ln.f();
            Light& l = L_.light(0);
            l.direction = normalized(d); l.color = c; l.intensity = inten;
            if (cmd == "fill_light") l.flags = 0;  // not seen by media
            return true;
        }
        if (cmd == "point") {
            F3 p = ln.v3(), c = ln.v3();
            float pw = ln.f(), r = ln.f();
            Light& l = L_.light(1);
            l.position = p; l.color = c; l.intensity = pw; l.radius = r;
            return true;
        }
        if (cmd == "spot") {  // angle: the full cone angle
            F3 p = ln.v3(), d = ln.v3(), c = ln.v3();
            float pw = ln.f(), angle = ln.f(), blend = ln.f(), r = ln.f(), beam = ln.f();
            Light& l = L_.light(3);
            l.position = p; l.direction = normalized(d); l.color = c; l.intensity = pw;
            l.angle = angle * 0.5f; l.blend = blend; l.radius = r; l.beam = beam;
            return true;
        }
        if (cmd == "area") {  // a size_x x size_z rectangle in XZ facing -Y
            F3 p = ln.v3();
            float sx = ln.f(), sz = ln.f();
            F3 c = ln.v3();
            float pw = ln.f();
            Light& l = L_.light(2);
            l.position = p; l.u_axis = f3(sx * 0.5f, 0, 0); l.v_axis = f3(0, 0, sz * 0.5f); l.color = c; l.intensity = pw;
            return true;
        }
        if (cmd == "reflect") {
            int m = ln.i();
            float v = ln.f();
            if (m >= 0 && m < (int)L_.mats.size()) L_.mats[m].reflect = v;
            return true;
        }
        if (cmd == "emit") {
            int m = ln.i();
            F3 c = ln.v3();
            float s = ln.f();
            if (m >= 0 && m < (int)L_.mats.size()) { L_.mats[m].emission_color = c; L_.mats[m].emission = s; }
            return true;
        }
        if (cmd == "glass") {
            int m = ln.i();
            float t = ln.f(), ior = ln.f();
            if (m >= 0 && m < (int)L_.mats.size()) { L_.mats[m].transmission = t; L_.mats[m].ior = ior; }
            return true;
        }
        if (cmd == "sun_angle") {
            float a = ln.f();
            if (!L_.lights.empty()) L_.lights.back().angle = a;
            return true;
        }
        if (cmd == "gobo") {
            int li = ln.i(), ti = ln.i();
            if (li >= 0 && li < (int)L_.lights.size()) L_.lights[li].texture_id = ti;
            return true;
        }
        if (cmd == "photons") { L_.photons = ln.i(); return true; }
        if (cmd == "caustics") { L_.caustics = ln.f(); return true; }
        if (cmd == "photon_cell") { float v = ln.f(); if (v > 0.0f) L_.photon_cell = v; return true; }
        if (cmd == "media") {
            Medium m;
            std::memset(&m, 0, sizeof(m));
            m.center = ln.v3(); m.half = ln.v3(); m.sigma_s = ln.v3(); m.sigma_t = ln.f(); m.samples = ln.i();
            L_.media.push_back(m);
            return true;
        }
        if (cmd == "hf") {
            HeightField h;
            std::memset(&h, 0, sizeof(h));
            h.center = ln.v3(); h.size_x = ln.f(); h.size_z = ln.f(); h.height = ln.f(); h.freq = ln.f();
            h.octaves = ln.i(); h.seed = ln.i(); h.material_id = ln.i();
            L_.hfs.push_back(h);
            return true;
        }
        if (cmd == "hf_image") return hf_image(ln);
        if (cmd == "se" || cmd == "se_csg") {
            Superellipsoid s;
            std::memset(&s, 0, sizeof(s));
            s.center = ln.v3(); s.a = ln.f(); s.b = ln.f(); s.c = ln.f(); s.n = ln.f(); s.material_id = ln.i();
            s._pad0 = cmd == "se_csg" ? 1 : 0;  // CSG operand only: left out of the SE BVH
            L_.ses.push_back(s);
            return true;
        }
        if (cmd == "lathe" || cmd == "lathe_bezier") return lathe(ln, cmd == "lathe_bezier");
        if (cmd == "prism" || cmd == "prism_bezier") return prism(ln, cmd == "prism_bezier");
        if (cmd == "csg_begin") {
            CSGNode n;
            std::memset(&n, 0, sizeof(n));
            n.op = ln.i(); n.leaf_kind = -1; n.leaf_index = -1; n.left = n.right = -1;
            L_.csg.push_back(n);
            L_.csg_stack.push_back((int)L_.csg.size() - 1);
            return true;
        }
        if (cmd == "csg_leaf") {
            CSGNode n;
            std::memset(&n, 0, sizeof(n));
            n.leaf_kind = ln.i(); n.leaf_index = ln.i(); n.left = n.right = -1;
            L_.csg.push_back(n);
            attach((int)L_.csg.size() - 1);
            return true;
        }
        if (cmd == "csg_end") {
            if (!L_.csg_stack.empty()) {
                const int cur = L_.csg_stack.back();
                L_.csg_stack.pop_back();
                if (L_.csg_stack.empty()) L_.csg_root = cur;
                else attach(cur);
            }
            return true;
        }
        if (cmd == "sweep" || cmd == "sweep_smooth") {
            int m = ln.i(), subdiv = ln.i();
            std::vector<SweepPoint> pts;
            while (ln.more()) {
                SweepPoint p;
                std::memset(&p, 0, sizeof(p));
                p.position = ln.v3();
                p.radius = ln.f();
                if (ln.short_) break;
                pts.push_back(p);
            }
            ln.short_ = false;
            if (pts.size() < 2) { error = "sweep needs two points"; return false; }
            if (subdiv > 0) L_.sweep_smooth(pts, m, subdiv);
            else L_.sweep_linear(pts, m);
            return true;
        }
        error = "unknown directive '" + cmd + "'";
        return false;
    }

    // a node's first free child slot, left then right
    void attach(int node) {
        if (L_.csg_stack.empty()) return;
        CSGNode& p = L_.csg[L_.csg_stack.back()];
        if (p.left < 0) p.left = node;
        else p.right = node;
    }

    bool resolution(Line& ln) {
        if (ln.keyword()) {
            KV kv{ln.kv()};
            L_.width = kv.i("width", L_.width);
            L_.height = kv.i("height", L_.height);
            L_.aa = kv.i("aa", L_.aa);
            L_.output = kv.s("out", L_.output);
            return !kv.bad;
        }
        L_.width = ln.i(); L_.height = ln.i(); L_.aa = ln.i();
        if (ln.more()) L_.output = ln.s();
        return true;
    }

    bool camera(Line& ln) {
        CameraData& c = L_.cam;
        if (ln.keyword()) {
            KV kv{ln.kv()};
            c.position = kv.v3("loc", f3(0, 0, 3));
            c.right = kv.v3("right", f3(1, 0, 0));
            c.up = kv.v3("up", f3(0, 1, 0));
            c.forward = kv.v3("fwd", f3(0, 0, -1));
            c.focal_length = kv.f("lens", 50.0f);
            c.sensor_width = kv.f("sensor", 36.0f);
            c.shift_x = kv.f("shift_x", 0.0f);
            c.shift_y = kv.f("shift_y", 0.0f);
            return !kv.bad;
        }
        // position and forward; legacy field of view (focal length 0)
        c.position = ln.v3();
        c.forward = ln.v3();
        c.right = f3(1, 0, 0);
        c.up = f3(0, 1, 0);
        c.focal_length = 0.0f;
        return true;
    }

    bool material(Line& ln) {
        if (ln.keyword()) {
            KV kv{ln.kv()};
            L_.material(kv.v3("diffuse", f3(0.8f, 0.8f, 0.8f)), kv.f("spec", 0.0f), kv.f("shin", 32.0f), kv.f("amb", 0.1f),
                        kv.f("aniso", 0.0f), kv.f("shift", 0.1f), kv.f("spec_r", 0.0f), kv.f("exp_r", 0.0f),
                        kv.f("spec_t", 0.0f), kv.f("exp_t", 0.0f), kv.f("back", 0.0f));
            Material& m = L_.mats.back();
            if (kv.has("tex")) m.texture_id = kv.i("tex", -1);
            if (kv.has("mask")) m.mask_id = kv.i("mask", -1);
            if (kv.has("normal")) m.normal_id = kv.i("normal", -1);
            if (kv.has("nstr")) m.normal_strength = kv.f("nstr", 1.0f);
            return !kv.bad;
        }
        F3 c = ln.v3();
        float spec = ln.f(), shin = ln.f(), amb = ln.f();
        if (!ln.more()) { L_.material(c, spec, shin, amb); return true; }
        float aniso = ln.f(), shift = ln.f(), sr = ln.f(), er = ln.f(), st = ln.f(), et = ln.f(), back = ln.f();
        L_.material(c, spec, shin, amb, aniso, shift, sr, er, st, et, back);
        return true;
    }

    bool grid(Line& ln) {
        int nx = ln.i(), nz = ln.i();
        float step = ln.f(), y = ln.f(), r = ln.f();
        int modm = ln.i();
        if ((long long)nx * nz > 2000000) { error = "grid: too many spheres"; return false; }
        for (int i = 0; i < nx * nz; i++) {
            const float x = (i % nx) * step - nx * step * 0.5f;
            const float z = (i / nx) * step - 2.0f;
            L_.sphere(f3(x, y, z), r, modm > 0 ? i % modm : 0);
        }
        return true;
    }

    bool lathe(Line& ln, bool bezier) {
        F3 c = ln.v3();
        int m = ln.i();
        int subdiv = bezier ? ln.i() : 0;
        int n = ln.i();
        const int max_n = bezier ? 2048 : 512;
        if (n <= 0 || n > max_n || (bezier && (subdiv < 1 || subdiv > 64))) return true;
        std::vector<P2> prof;
        for (int k = 0; k < n; k++) {
            float y = ln.f(), r = ln.f();
            prof.push_back({y, r});
        }
        if (bezier) prof = bezier_run(prof, subdiv);
        Lathe l;
        std::memset(&l, 0, sizeof(l));
        l.center = c;
        l.point_offset = (int)L_.lathe_points.size();
        l.point_count = (int)prof.size();
        l.material_id = m;
        for (const P2& q : prof) {
            LathePoint p;
            std::memset(&p, 0, sizeof(p));
            p.y = q.a; p.radius = q.b;
            L_.lathe_points.push_back(p);
        }
        L_.lathes.push_back(l);
        return true;
    }

    bool prism(Line& ln, bool bezier) {
        F3 c = ln.v3();
        float y1 = ln.f(), y2 = ln.f(), sweep = ln.f();
        int m = ln.i();
        int subdiv = bezier ? ln.i() : 0;
        int n = ln.i();
        if (n <= 0 || n > 2048 || (bezier && (subdiv < 1 || subdiv > 64))) return true;
        std::vector<P2> prof;
        for (int k = 0; k < n; k++) {
            float x = ln.f(), z = ln.f();
            prof.push_back({x, z});
        }
        if (bezier) {
            prof = bezier_run(prof, subdiv);
            if (!prof.empty()) prof.push_back(prof.front());  // a Bezier contour is closed explicitly
        }
        Prism p;
        std::memset(&p, 0, sizeof(p));
        p.center = c; p.y_min = y1; p.y_max = y2; p.sweep = sweep;
        p.point_offset = (int)L_.prism_points.size();
        p.point_count = (int)prof.size();
        p.material_id = m;
        for (const P2& q : prof) {
            PrismPoint pp;
            std::memset(&pp, 0, sizeof(pp));
            pp.x = q.a; pp.z = q.b;
            L_.prism_points.push_back(pp);
        }
        L_.prisms.push_back(p);
        return true;
    }

    bool obj(Line& ln) {
        KV kv{ln.kv()};
        const std::string file = path(kv.s("file", ""));
        if (file.empty()) { error = "obj: file= missing"; return false; }
        const int m = kv.i("mat", 0);
        const F3 off = kv.v3("offset", f3(0, 0, 0));
        const float sc = kv.f("scale", 1.0f);
        const bool smooth = kv.i("smooth", 1) != 0;
        std::ifstream f(file);
        if (!f) { error = "obj: cannot open " + file; return false; }
        std::vector<F3> v, vn;
        std::vector<sheen_f2> vt;
        std::string line;
        int faces = 0;
        while (std::getline(f, line)) {
            std::istringstream ls(line);
            std::string tag;
            ls >> tag;
            if (tag == "v") { float x = 0, y = 0, z = 0; ls >> x >> y >> z; v.push_back(f3(x * sc + off.x, y * sc + off.y, z * sc + off.z)); }
            else if (tag == "vn") { float x = 0, y = 0, z = 0; ls >> x >> y >> z; vn.push_back(f3(x, y, z)); }
            else if (tag == "vt") { float a = 0, b = 0; ls >> a >> b; vt.push_back(sheen_f2{a, b}); }
            else if (tag == "f") {
                // v, v/t, v/t/n, v//n; negative indices count from the end
                std::vector<int> vi, ti, ni;
                std::string tok;
                bool ok = true;
                while (ls >> tok) {
                    int idx[3] = {0, 0, 0};
                    int field = 0;
                    std::string cur;
                    for (size_t k = 0; k <= tok.size(); k++) {
                        if (k == tok.size() || tok[k] == '/') {
                            if (field < 3) idx[field] = cur.empty() ? 0 : std::atoi(cur.c_str());
                            cur.clear();
                            field++;
                        } else {
                            cur += tok[k];
                        }
                    }
                    auto res = [](int i, size_t n) { return i == 0 ? -1 : (i < 0 ? (int)n + i : i - 1); };
                    const int a = res(idx[0], v.size());
                    if (a < 0 || a >= (int)v.size()) { ok = false; break; }
                    vi.push_back(a);
                    ti.push_back(res(idx[1], vt.size()));
                    ni.push_back(res(idx[2], vn.size()));
                }
                if (!ok || vi.size() < 3) continue;
                auto nrm = [&](int s) { return (smooth && ni[s] >= 0 && ni[s] < (int)vn.size()) ? vn[ni[s]] : f3(0, 0, 0); };
                auto uv = [&](int s) { return (ti[s] >= 0 && ti[s] < (int)vt.size()) ? vt[ti[s]] : sheen_f2{0, 0}; };
                for (size_t k = 2; k < vi.size(); k++) {  // fan
                    L_.triangle(v[vi[0]], v[vi[k - 1]], v[vi[k]], nrm(0), nrm((int)k - 1), nrm((int)k), m, uv(0),
                                uv((int)k - 1), uv((int)k));
                    faces++;
                }
            }
        }
        if (!faces) { error = "obj: no faces in " + file; return false; }
        return !kv.bad;
    }

    // SMESH1: "SMESH1" + 2 bytes, counts nv nn nu ntri (uint32), then
    // float3 vertices, float3 normals, float2 uvs, and per triangle int32
    // v[3] n[3] t[3] mat (scene_io.cpp:480-547)
    bool mesh(Line& ln) {
        KV kv{ln.kv()};
        const std::string file = path(kv.s("file", ""));
        if (file.empty()) { error = "mesh: file= missing"; return false; }
        const int mat0 = kv.i("mat0", 0);
        const float sc = kv.f("scale", 1.0f);
        std::vector<int> skip;
        if (kv.has("skip")) {
            std::string sk = kv.s("skip", "");
            std::replace(sk.begin(), sk.end(), ',', ' ');
            std::istringstream ss(sk);
            int x;
            while (ss >> x) skip.push_back(x);
        }
        std::ifstream f(file, std::ios::binary);
        if (!f) { error = "mesh: cannot open " + file; return false; }
        char magic[8] = {0};
        uint32_t cnt[4] = {0, 0, 0, 0};
        f.read(magic, 8);
        f.read(reinterpret_cast<char*>(cnt), 16);
        if (!f || std::memcmp(magic, "SMESH1", 6) != 0) { error = "mesh: not an SMESH1 file"; return false; }
        f.seekg(0, std::ios::end);
        const uint64_t size = (uint64_t)f.tellg();
        const uint64_t need = 24ull + cnt[0] * 12ull + cnt[1] * 12ull + cnt[2] * 8ull + cnt[3] * 40ull;
        if (size < need || cnt[0] > 50000000u || cnt[3] > 50000000u) { error = "mesh: header does not match the file"; return false; }
        f.seekg(24);
        std::vector<F3> v(cnt[0]), n(cnt[1]);
        std::vector<sheen_f2> t(cnt[2]);
        f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)v.size() * 12);
        f.read(reinterpret_cast<char*>(n.data()), (std::streamsize)n.size() * 12);
        f.read(reinterpret_cast<char*>(t.data()), (std::streamsize)t.size() * 8);
        for (uint32_t k = 0; k < cnt[3]; k++) {
            int32_t r[10];
            f.read(reinterpret_cast<char*>(r), 40);
            if (!f) { error = "mesh: truncated"; return false; }
            const int m = mat0 + r[9];
            if (std::find(skip.begin(), skip.end(), m) != skip.end()) continue;
            bool ok = true;
            for (int q = 0; q < 3; q++) ok = ok && r[q] >= 0 && r[q] < (int)v.size();
            if (!ok) { error = "mesh: vertex index out of range"; return false; }
            auto vv = [&](int q) { return mul(v[r[q]], sc); };
            auto nn = [&](int q) { return (r[3 + q] >= 0 && r[3 + q] < (int)n.size()) ? n[r[3 + q]] : f3(0, 0, 0); };
            auto tt = [&](int q) { return (r[6 + q] >= 0 && r[6 + q] < (int)t.size()) ? t[r[6 + q]] : sheen_f2{0, 0}; };
            L_.triangle(vv(0), vv(1), vv(2), nn(0), nn(1), nn(2), m, tt(0), tt(1), tt(2));
        }
        return !kv.bad;
    }

    // Binary PPM (P6), reduced by an integer nearest-neighbour step to at
    // most `max` texels a side.
    bool texture(Line& ln) {
        KV kv{ln.kv()};
        const std::string file = path(kv.s("file", ""));
        const int maxd = kv.i("max", 2048);
        if (file.empty()) { error = "texture: file= missing"; return false; }
        std::ifstream f(file, std::ios::binary);
        if (!f) { error = "texture: cannot open " + file; return false; }
        std::string magic;
        int w = 0, h = 0, maxv = 0;
        f >> magic >> w >> h >> maxv;
        if (magic != "P6" || w <= 0 || h <= 0 || w > 16384 || h > 16384) { error = "texture: needs a P6 PPM"; return false; }
        f.get();
        std::vector<unsigned char> px((size_t)w * h * 3);
        f.read(reinterpret_cast<char*>(px.data()), (std::streamsize)px.size());
        if (f.gcount() != (std::streamsize)px.size()) { error = "texture: file shorter than its header"; return false; }
        int step = 1;
        while (w / step > maxd || h / step > maxd) step++;
        if (step > 1) {
            const int nw = w / step, nh = h / step;
            std::vector<unsigned char> q((size_t)nw * nh * 3);
            for (int y = 0; y < nh; y++)
                for (int x = 0; x < nw; x++)
                    for (int c = 0; c < 3; c++)
                        q[((size_t)y * nw + x) * 3 + c] = px[((size_t)(y * step) * w + (size_t)(x * step)) * 3 + c];
            L_.texture(nw, nh, std::move(q));
        } else {
            L_.texture(w, h, std::move(px));
        }
        return !kv.bad;
    }

    // hf_image cx cy cz size_x size_z height file.pgm mat: binary PGM (P5)
    bool hf_image(Line& ln) {
        F3 c = ln.v3();
        float sx = ln.f(), sz = ln.f(), hgt = ln.f();
        const std::string file = path(ln.s());
        int m = ln.i();
        std::ifstream f(file, std::ios::binary);
        if (!f) { error = "hf_image: cannot open " + file; return false; }
        std::string magic;
        int w = 0, h = 0, maxv = 0;
        f >> magic >> w >> h >> maxv;
        if (magic != "P5" || w <= 0 || h <= 0 || w > 16384 || h > 16384) { error = "hf_image: needs a P5 PGM"; return false; }
        f.get();
        auto img = std::make_unique<std::vector<unsigned char>>((size_t)w * h);
        f.read(reinterpret_cast<char*>(img->data()), (std::streamsize)img->size());
        if (f.gcount() != (std::streamsize)img->size()) { error = "hf_image: file shorter than its header"; return false; }
        HeightField hf;
        std::memset(&hf, 0, sizeof(hf));
        hf.center = c; hf.size_x = sx; hf.size_z = sz; hf.height = hgt; hf.freq = 1.0f; hf.octaves = 1; hf.seed = -1;
        hf.material_id = m;
        hf.image = img->data();
        hf.img_w = w; hf.img_h = h;
        L_.hf_images.push_back(std::move(img));
        L_.hfs.push_back(hf);
        return true;
    }

    // strands file= : lines "strand r0 r1 count x y z ...", radii linear along
    // the strand (or radius=), long strands cut into 8-point pieces that share
    // their end points, each a Catmull-Rom sweep with one subdivision
    bool strands(Line& ln) {
        KV kv{ln.kv()};
        const std::string file = path(kv.s("file", ""));
        const int m = kv.i("mat", 1);
        const float sc = kv.f("scale", 1.0f);
        const float rad = kv.f("radius", 0.0f);
        if (file.empty()) { error = "strands: file= missing"; return false; }
        std::ifstream f(file);
        if (!f) { error = "strands: cannot open " + file; return false; }
        std::string line;
        while (std::getline(f, line)) {
            std::istringstream ls(line);
            std::string tag;
            ls >> tag;
            if (tag != "strand") continue;
            float r0 = 0, r1 = 0;
            int cnt = 0;
            ls >> r0 >> r1 >> cnt;
            if (cnt < 2 || cnt > 100000) continue;
            std::vector<SweepPoint> pts;
            for (int k = 0; k < cnt; k++) {
                float x, y, z;
                if (!(ls >> x >> y >> z)) break;
                const float t = (float)k / (float)(cnt - 1);
                SweepPoint p;
                std::memset(&p, 0, sizeof(p));
                p.position = f3(x * sc, y * sc, z * sc);
                p.radius = rad > 0.0f ? rad * sc : (r0 + (r1 - r0) * t) * sc;
                pts.push_back(p);
            }
            if (pts.size() < 2) continue;
            const size_t piece = 8;
            if (pts.size() <= piece) { L_.sweep_smooth(pts, m, 1); continue; }
            for (size_t a = 0; a + 1 < pts.size(); a += piece - 1) {
                const size_t b = std::min(a + piece, pts.size());
                std::vector<SweepPoint> part(pts.begin() + (long)a, pts.begin() + (long)b);
                if (part.size() >= 2) L_.sweep_smooth(part, m, 1);
            }
        }
        return !kv.bad;
    }

    // Procedural hair on a head sphere (scene_io.cpp:104-157): strands start
    // on the upper sphere and fall past its side; std::mt19937 streams as in
    // the reference, so the same seed gives the same strands.
    bool hair(Line& ln) {
        KV kv;
        if (ln.keyword()) {
            kv.m = ln.kv();
        } else {
            // cx cy cz head count segs seed mat, passed on as text with a
            // stream's default formatting (6 significant digits), as the
            // reference does
            F3 c = ln.v3();
            float head = ln.f();
            long long count = ln.f_int(), segs = ln.f_int(), seed = ln.f_int(), mat = ln.f_int();
            auto txt = [](auto v) {
                std::ostringstream o;
                o << v;
                return o.str();
            };
            kv.m["center"] = txt(c.x) + "," + txt(c.y) + "," + txt(c.z);
            kv.m["head"] = txt(head);
            kv.m["count"] = txt(count);
            kv.m["segs"] = txt(segs);
            kv.m["seed"] = txt(seed);
            kv.m["mat"] = txt(mat);
        }
        const F3 c = kv.v3("center", f3(0.0f, 1.6f, 0.0f));
        const float head = kv.f("head", 0.55f);
        const int count = kv.i("count", 500), segs = kv.i("segs", 10);
        const unsigned seed = (unsigned)kv.l("seed", 1);
        const int m = kv.i("mat", 0);
        const float rad = kv.f("radius", 0.009f), taper = kv.f("taper", 0.55f), len = kv.f("length", 0.9f);
        const float wob = kv.f("wobble", 0.012f), curl = kv.f("curl", 0.9f);
        const float phi0 = kv.f("phi0", 0.15f), phi1 = kv.f("phi1", 1.20f), spread = kv.f("spread", 1.02f);
        if (kv.bad) return false;
        std::mt19937 rng(seed);
        std::uniform_real_distribution<float> U(0.0f, 1.0f);
        for (int i = 0; i < count; i++) {
            const float theta = 6.2831853f * U(rng);
            const float ph = phi0 + (phi1 - phi0) * U(rng);
            const float r0 = rad * (0.75f + 0.5f * U(rng));
            const float fall = len * head * (0.6f + 0.8f * U(rng));
            std::vector<SweepPoint> pts;
            F3 pos = c;
            for (int k = 0; k < segs; k++) {
                const float t = (float)k / (float)(segs - 1);
                const float phi = ph + t * curl;
                F3 q;
                if (phi < 1.75f) {  // still on the scalp sphere
                    const float rr = head * (spread + 0.10f * t);
                    q = f3(c.x + rr * sinf(phi) * cosf(theta), c.y + rr * cosf(phi), c.z + rr * sinf(phi) * sinf(theta));
                } else {            // falling
                    q = pos;
                    q.y -= fall * t;
                    q.x += sinf(t * 6.0f + i) * wob;
                    q.z += cosf(t * 5.0f + i * 0.7f) * wob;
                }
                pos = q;
                SweepPoint p;
                std::memset(&p, 0, sizeof(p));
                p.position = q;
                p.radius = r0 * (1.0f - taper * t);
                pts.push_back(p);
            }
            L_.sweep_smooth(pts, m, 2);
        }
        return true;
    }

    // Fur grown from the vertices of the mesh loaded so far
    // (scene_io.cpp:158-214), roots within [y0, y1].
    bool hair_mesh(Line& ln) {
        KV kv{ln.kv()};
        const int m = kv.i("mat", 1), count = kv.i("count", 500), segs = kv.i("segs", 10);
        const unsigned seed = (unsigned)kv.l("seed", 7);
        const float rad = kv.f("radius", 0.004f), taper = kv.f("taper", 0.6f), len = kv.f("length", 0.25f);
        const float wob = kv.f("wobble", 0.01f), y0 = kv.f("y0", 0.0f), y1 = kv.f("y1", 1e9f), lift = kv.f("lift", 0.5f);
        if (kv.bad) return false;
        std::vector<F3> roots, norms;
        for (const Triangle& t : L_.tris) {
            const F3 n = normalized(cross3(sub(t.v1, t.v0), sub(t.v2, t.v0)));
            const F3 vs[3] = {t.v0, t.v1, t.v2};
            for (const F3& v : vs)
                if (v.y >= y0 && v.y <= y1) { roots.push_back(v); norms.push_back(n); }
        }
        if (roots.empty()) return true;
        std::mt19937 rng(seed);
        std::uniform_int_distribution<size_t> pick(0, roots.size() - 1);
        for (int i = 0; i < count; i++) {
            const size_t k = pick(rng);
            const F3 root = add(roots[k], mul(norms[k], 0.002f));
            F3 dir = normalized(add(mul(norms[k], lift), f3(0.0f, -1.0f, 0.0f)));
            const float r0 = rad * (0.7f + 0.6f * (float)(rng() % 1000) / 1000.0f);
            std::vector<SweepPoint> pts;
            F3 pos = root;
            const float step = len / (float)(segs - 1);
            for (int s = 0; s < segs; s++) {
                const float t = (float)s / (float)(segs - 1);
                SweepPoint p;
                std::memset(&p, 0, sizeof(p));
                p.position = pos;
                p.radius = r0 * (1.0f - taper * t);
                pts.push_back(p);
                const F3 w = f3(sinf(t * 9.0f + i * 1.3f) * wob, 0.0f, cosf(t * 7.0f + i * 0.9f) * wob);
                dir = normalized(add(add(dir, f3(0.0f, -0.6f * step, 0.0f)), w));
                pos = add(pos, mul(dir, step));
            }
            L_.sweep_smooth(pts, m, 1);
        }
        return true;
    }
};

}  // namespace

struct SheenFile {
    Loaded scene;
    std::string error;
};

extern "C" {

SheenFile* sheen_file_load(const char* path) {
    std::unique_ptr<SheenFile> h(new SheenFile());
    if (!path) { h->error = "null path"; return h.release(); }
    std::ifstream f(path);
    if (!f) { h->error = std::string("cannot open scene: ") + path; return h.release(); }
    const std::string p(path);
    const size_t slash = p.rfind('/');
    h->scene.dir = slash == std::string::npos ? "" : p.substr(0, slash);
    Parser parser(h->scene);
    std::string line;
    int n = 0;
    while (std::getline(f, line)) {
        n++;
        if (!parser.line(line)) {
            h->error = "line " + std::to_string(n) + ": " + parser.error;
            return h.release();
        }
    }
    h->scene.finalize();
    return h.release();
}

const char* sheen_file_error(const SheenFile* h) { return h ? h->error.c_str() : "null handle"; }

const SheenScene* sheen_file_scene(const SheenFile* h) {
    return (h && h->error.empty()) ? &h->scene.scene : nullptr;
}

int sheen_file_settings(const SheenFile* h, int* width, int* height, int* aa, const char** output) {
    if (!h || !h->error.empty()) return -1;
    if (width) *width = h->scene.width;
    if (height) *height = h->scene.height;
    if (aa) *aa = h->scene.aa;
    if (output) *output = h->scene.output.c_str();
    return 0;
}

void sheen_file_free(SheenFile* h) { delete h; }

}  // extern "C"
