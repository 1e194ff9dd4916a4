// host_scene.cpp — scene validation, per-category wide-BVH builds, triangle
// leaf copies and texture mip chains (host side of the HIP path).
#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "host_util.h"

void host_prim_bounds(const SheenScene& s, int cat, int i, float* mn, float* mx) {
    auto set = [&](float a0, float a1, float a2, float b0, float b1, float b2) {
        mn[0] = a0; mn[1] = a1; mn[2] = a2; mx[0] = b0; mx[1] = b1; mx[2] = b2;
    };
    switch (cat) {
        case 0: {  // sphere; the reference folds leaf bounds into +-1e10 (bvh.h:80-97)
            const Object& o = s.objects[i];
            set(std::fmin(1e10f, o.param1.x - o.param2), std::fmin(1e10f, o.param1.y - o.param2),
                std::fmin(1e10f, o.param1.z - o.param2), std::fmax(-1e10f, o.param1.x + o.param2),
                std::fmax(-1e10f, o.param1.y + o.param2), std::fmax(-1e10f, o.param1.z + o.param2));
            break;
        }
        case 1: {  // strand (bvh.h:40-62, folded as above in the strand builder)
            const SweepStrand& st = s.sweep_strands[i];
            float a[3] = {1e10f, 1e10f, 1e10f}, b[3] = {-1e10f, -1e10f, -1e10f};
            for (int k = 0; k < st.point_count; k++) {
                const SweepPoint& p = s.sweep_points[st.point_offset + k];
                const float pp[3] = {p.position.x, p.position.y, p.position.z};
                for (int d = 0; d < 3; d++) {
                    a[d] = std::fmin(a[d], pp[d] - p.radius);
                    b[d] = std::fmax(b[d], pp[d] + p.radius);
                }
            }
            set(a[0], a[1], a[2], b[0], b[1], b[2]);
            break;
        }
        case 2: {  // triangle (bvh.h:235-242)
            const Triangle& t = s.triangles[i];
            set(std::fmin(t.v0.x, std::fmin(t.v1.x, t.v2.x)), std::fmin(t.v0.y, std::fmin(t.v1.y, t.v2.y)),
                std::fmin(t.v0.z, std::fmin(t.v1.z, t.v2.z)), std::fmax(t.v0.x, std::fmax(t.v1.x, t.v2.x)),
                std::fmax(t.v0.y, std::fmax(t.v1.y, t.v2.y)), std::fmax(t.v0.z, std::fmax(t.v1.z, t.v2.z)));
            break;
        }
        case 3: {  // height field (bvh.h:304-309)
            const HeightField& h = s.heightfields[i];
            set(h.center.x - 0.5f * h.size_x, h.center.y, h.center.z - 0.5f * h.size_z, h.center.x + 0.5f * h.size_x,
                h.center.y + h.height, h.center.z + 0.5f * h.size_z);
            break;
        }
        case 4: {  // superellipsoid (superellipsoid.h:31-36)
            const Superellipsoid& e = s.superellipsoids[i];
            set(e.center.x - e.a, e.center.y - e.b, e.center.z - e.c, e.center.x + e.a, e.center.y + e.b,
                e.center.z + e.c);
            break;
        }
        case 5: {  // lathe (bvh.h:394-407): +-1e30 in x/z, y from the profile with the centre
            const Lathe& l = s.lathes[i];
            float a[3] = {l.center.x - 1e30f, 1e30f, l.center.z - 1e30f};
            float b[3] = {l.center.x + 1e30f, -1e30f, l.center.z + 1e30f};
            for (int k = 0; k < l.point_count; k++) {
                const LathePoint& p = s.lathe_points[l.point_offset + k];
                float r = p.radius;
                if (r < 0.0f) continue;
                a[0] = std::fmin(a[0], l.center.x - r); a[1] = std::fmin(a[1], l.center.y + p.y - r);
                a[2] = std::fmin(a[2], l.center.z - r);
                b[0] = std::fmax(b[0], l.center.x + r); b[1] = std::fmax(b[1], l.center.y + p.y + r);
                b[2] = std::fmax(b[2], l.center.z + r);
            }
            if (a[0] > b[0]) { a[0] = a[1] = a[2] = b[0] = b[1] = b[2] = 0.0f; }
            set(a[0], a[1], a[2], b[0], b[1], b[2]);
            break;
        }
        case 6: {  // prism (bvh.h:321-332)
            const Prism& p = s.prisms[i];
            float a[3] = {1e30f, p.center.y + p.y_min, 1e30f}, b[3] = {-1e30f, p.center.y + p.y_max, -1e30f};
            for (int k = 0; k < p.point_count; k++) {
                const PrismPoint& q = s.prism_points[p.point_offset + k];
                a[0] = std::fmin(a[0], p.center.x + q.x); b[0] = std::fmax(b[0], p.center.x + q.x);
                a[2] = std::fmin(a[2], p.center.z + q.z); b[2] = std::fmax(b[2], p.center.z + q.z);
            }
            if (a[0] > b[0]) { a[0] = 0; a[1] = p.center.y; a[2] = 0; b[0] = 0; b[1] = p.center.y; b[2] = 0; }
            set(a[0], a[1], a[2], b[0], b[1], b[2]);
            break;
        }
    }
}

void host_csg_boxes(const SheenScene& s, std::vector<float>& out) {
    const int n = (s.csg_nodes && s.num_csg_nodes > 0) ? s.num_csg_nodes : 0;
    out.assign(6 * (size_t)n, 0.0f);
    const float inf = INFINITY;
    auto set = [&](int i, const float* b) { std::copy(b, b + 6, out.begin() + 6 * (size_t)i); };
    const float empty[6] = {inf, inf, inf, -inf, -inf, -inf};
    const float all[6] = {-inf, -inf, -inf, inf, inf, inf};
    std::vector<char> state(n, 0);  // 0 new, 1 in progress, 2 done
    struct Item { int node; bool expanded; };
    std::vector<Item> stack;
    for (int r = 0; r < n; r++) {
        if (state[r]) continue;
        stack.push_back({r, false});
        while (!stack.empty()) {
            Item it = stack.back();
            stack.pop_back();
            const int i = it.node;
            const CSGNode& nd = s.csg_nodes[i];
            if (nd.leaf_kind >= 0) {
                float b[6];
                std::copy(empty, empty + 6, b);
                if (nd.leaf_kind == 0 && nd.leaf_index >= 0 && nd.leaf_index < s.num_objects && s.objects) {
                    const Object& o = s.objects[nd.leaf_index];
                    const float c[3] = {o.param1.x, o.param1.y, o.param1.z}, rr = std::fabs(o.param2);
                    for (int a = 0; a < 3; a++) { b[a] = c[a] - rr; b[3 + a] = c[a] + rr; }
                } else if (nd.leaf_kind == 1 && nd.leaf_index >= 0 && nd.leaf_index < s.num_superellipsoids &&
                           s.superellipsoids) {
                    const Superellipsoid& e = s.superellipsoids[nd.leaf_index];
                    const float c[3] = {e.center.x, e.center.y, e.center.z};
                    const float h[3] = {std::fabs(e.a), std::fabs(e.b), std::fabs(e.c)};
                    for (int a = 0; a < 3; a++) { b[a] = c[a] - h[a]; b[3 + a] = c[a] + h[a]; }
                }
                bool finite = true;
                for (int a = 0; a < 6; a++) finite = finite && std::isfinite(b[a]);
                set(i, (finite || nd.leaf_kind > 1 || b[0] == inf) ? b : all);
                state[i] = 2;
                continue;
            }
            const int ch[2] = {nd.left, nd.right};
            if (!it.expanded) {
                state[i] = 1;
                stack.push_back({i, true});
                for (int k = 0; k < 2; k++) {
                    const int c = ch[k];
                    if (c < 0 || c >= n) continue;
                    if (state[c] == 1) { set(i, all); continue; }  // cycle: no culling below it
                    if (state[c] == 0) stack.push_back({c, false});
                }
                continue;
            }
            float cb[2][6];
            for (int k = 0; k < 2; k++) {
                const int c = ch[k];
                if (c < 0 || c >= n) std::copy(empty, empty + 6, cb[k]);
                else if (state[c] == 2) std::copy(out.begin() + 6 * (size_t)c, out.begin() + 6 * (size_t)c + 6, cb[k]);
                else std::copy(all, all + 6, cb[k]);  // on a cycle
            }
            float b[6];
            for (int a = 0; a < 3; a++) {
                if (nd.op == 0) { b[a] = std::fmin(cb[0][a], cb[1][a]); b[3 + a] = std::fmax(cb[0][3 + a], cb[1][3 + a]); }
                else if (nd.op == 1 || nd.op == 2) { b[a] = cb[0][a]; b[3 + a] = cb[0][3 + a]; }
                else { b[a] = inf; b[3 + a] = -inf; }  // unknown op: no spans
            }
            set(i, b);
            state[i] = 2;
        }
    }
}

void host_csg_spine(const SheenScene& s, const std::vector<float>& boxes, HostCsgSpine& out) {
    out = HostCsgSpine();
    const int n = (s.csg_nodes && s.num_csg_nodes > 0) ? s.num_csg_nodes : 0;
    const int root = s.csg_root;
    if (root < 0 || root >= n) return;
    auto spine_union = [&](int i) { return s.csg_nodes[i].leaf_kind < 0 && s.csg_nodes[i].op == 0; };
    if (!spine_union(root)) return;
    out.spine = true;
    // in-order walk of the union spine; a node at frame depth >= CSG_DEPTH is
    // never evaluated (the evaluation stack holds CSG_DEPTH frames)
    struct Sub { int node, rank, depth; };
    std::vector<Sub> subs;
    std::vector<std::pair<int, int>> stack;  // (node, depth); right pushed first
    stack.push_back({root, 0});
    int rank = 0;
    while (!stack.empty()) {
        const auto [i, d] = stack.back();
        stack.pop_back();
        if (i < 0 || i >= n || d >= HOST_CSG_DEPTH) continue;
        if (spine_union(i)) {
            stack.push_back({s.csg_nodes[i].right, d + 1});
            stack.push_back({s.csg_nodes[i].left, d + 1});
            continue;
        }
        subs.push_back({i, rank++, d});
    }
    // unbounded boxes first (always evaluated), empty boxes dropped (no spans)
    std::vector<Sub> bounded;
    for (const Sub& u : subs) {
        const float* b = &boxes[6 * (size_t)u.node];
        bool finite = true, empty = false;
        for (int a = 0; a < 3; a++) {
            finite = finite && std::isfinite(b[a]) && std::isfinite(b[3 + a]);
            empty = empty || b[a] > b[3 + a];
        }
        if (finite && empty) continue;
        if (!finite) { out.sub.insert(out.sub.end(), {u.node, u.rank, u.depth, 0}); out.always++; }
        else bounded.push_back(u);
    }
    if (bounded.empty()) return;
    // median-split hierarchy, at most 2 subtrees per leaf
    struct Range { int node, s, e; };
    std::vector<Range> work;
    out.nodes.push_back(HostSpineNode());
    work.push_back({0, 0, (int)bounded.size()});
    auto cen = [&](const Sub& u, int a) {
        const float* b = &boxes[6 * (size_t)u.node];
        return 0.5 * ((double)b[a] + (double)b[3 + a]);
    };
    while (!work.empty()) {
        const Range w = work.back();
        work.pop_back();
        HostSpineNode nd;
        for (int a = 0; a < 3; a++) { nd.lo[a] = INFINITY; nd.hi[a] = -INFINITY; }
        for (int k = w.s; k < w.e; k++) {
            const float* b = &boxes[6 * (size_t)bounded[k].node];
            for (int a = 0; a < 3; a++) { nd.lo[a] = std::fmin(nd.lo[a], b[a]); nd.hi[a] = std::fmax(nd.hi[a], b[3 + a]); }
        }
        if (w.e - w.s <= 2) {
            nd.a = -((int)(out.sub.size() / 4) + 1);
            nd.b = w.e - w.s;
            for (int k = w.s; k < w.e; k++)
                out.sub.insert(out.sub.end(), {bounded[k].node, bounded[k].rank, bounded[k].depth, 0});
        } else {
            int axis = 0;
            for (int a = 1; a < 3; a++)
                if (nd.hi[a] - nd.lo[a] > nd.hi[axis] - nd.lo[axis]) axis = a;
            const int mid = (w.s + w.e) / 2;
            std::nth_element(bounded.begin() + w.s, bounded.begin() + mid, bounded.begin() + w.e,
                             [&](const Sub& x, const Sub& y) {
                                 const double cx = cen(x, axis), cy = cen(y, axis);
                                 return cx < cy || (cx == cy && x.rank < y.rank);
                             });
            nd.a = (int)out.nodes.size();
            out.nodes.push_back(HostSpineNode());
            nd.b = (int)out.nodes.size();
            out.nodes.push_back(HostSpineNode());
            work.push_back({nd.a, w.s, mid});
            work.push_back({nd.b, mid, w.e});
        }
        out.nodes[w.node] = nd;
    }
}

void host_build_mips(HostTexture& t) {
    t.mips = 1;
    if (t.w < 2 && t.h < 2) return;
    int sw = t.w, sh = t.h;
    size_t src = 0;
    while (sw > 1 || sh > 1) {
        const int nw = sw > 1 ? sw / 2 : 1, nh = sh > 1 ? sh / 2 : 1;
        const size_t dst = t.px.size();
        t.px.resize(dst + (size_t)nw * nh * 3);
        unsigned char* P = t.px.data();
        for (int y = 0; y < nh; y++) {
            const int ya = std::min(2 * y, sh - 1), yb = std::min(2 * y + 1, sh - 1);
            for (int x = 0; x < nw; x++) {
                const int xa = std::min(2 * x, sw - 1), xb = std::min(2 * x + 1, sw - 1);
                for (int c = 0; c < 3; c++) {
                    const int sum = P[src + ((size_t)ya * sw + xa) * 3 + c] + P[src + ((size_t)ya * sw + xb) * 3 + c] +
                                    P[src + ((size_t)yb * sw + xa) * 3 + c] + P[src + ((size_t)yb * sw + xb) * 3 + c];
                    P[dst + ((size_t)y * nw + x) * 3 + c] = (unsigned char)((sum + 2) / 4);
                }
            }
        }
        src = dst; sw = nw; sh = nh;
        t.mips++;
    }
}

static int prim_material(const SheenScene& s, int cat, int i) {
    switch (cat) {
        case 0: return s.objects[i].material_id;
        case 1: return s.sweep_strands[i].material_id;
        case 2: return s.triangles[i].material_id;
        case 3: return s.heightfields[i].material_id;
        case 4: return s.superellipsoids[i].material_id;
        case 5: return s.lathes[i].material_id;
        case 6: return s.prisms[i].material_id;
    }
    return 0;
}

bool host_prepare_scene(const SheenScene& s, HostScene& out, std::string& err, bool build_bvh) {
    auto t0 = std::chrono::steady_clock::now();
    const int counts[7] = {s.objects ? s.num_objects : 0, s.sweep_strands ? s.num_sweep_strands : 0,
                           s.triangles ? s.num_triangles : 0, s.heightfields ? s.num_heightfields : 0,
                           s.superellipsoids ? s.num_superellipsoids : 0, s.lathes ? s.num_lathes : 0,
                           s.prisms ? s.num_prisms : 0};
    auto bad_mat = [&](int m) { return m < 0 || m >= s.num_materials; };
    // reference accesses are unchecked; invalid references are errors here
    for (int c = 0; c < 7; c++)
        for (int i = 0; i < counts[c]; i++) {
            if (c == 0 && s.objects[i].type != 0) continue;
            if (bad_mat(prim_material(s, c, i))) { err = "material id out of range"; return false; }
        }
    for (int i = 0; i < (s.planes ? s.num_planes : 0); i++)
        if (bad_mat(s.planes[i].material_id)) { err = "material id out of range"; return false; }
    for (int i = 0; i < counts[1]; i++) {
        const SweepStrand& st = s.sweep_strands[i];
        if (st.point_count > 0 && (st.point_offset < 0 || st.point_offset + st.point_count > s.num_sweep_points || !s.sweep_points)) {
            err = "strand points out of range"; return false;
        }
    }
    for (int i = 0; i < counts[5]; i++) {
        const Lathe& l = s.lathes[i];
        if (l.point_count > 0 && (l.point_offset < 0 || l.point_offset + l.point_count > s.num_lathe_points || !s.lathe_points)) {
            err = "lathe points out of range"; return false;
        }
    }
    for (int i = 0; i < counts[6]; i++) {
        const Prism& p = s.prisms[i];
        if (p.point_count > 0 && (p.point_offset < 0 || p.point_offset + p.point_count > s.num_prism_points || !s.prism_points)) {
            err = "prism points out of range"; return false;
        }
    }
    if (s.csg_nodes && s.csg_root >= 0) {
        for (int i = 0; i < s.num_csg_nodes; i++) {
            const CSGNode& n = s.csg_nodes[i];
            if (n.leaf_kind == 0 && (n.leaf_index < 0 || n.leaf_index >= counts[0] || bad_mat(s.objects[n.leaf_index].material_id))) {
                err = "csg leaf out of range"; return false;
            }
            if (n.leaf_kind == 1 && (n.leaf_index < 0 || n.leaf_index >= counts[4])) { err = "csg leaf out of range"; return false; }
            if (n.leaf_kind < 0 && (n.left >= s.num_csg_nodes || n.right >= s.num_csg_nodes)) { err = "csg child out of range"; return false; }
        }
    }
    for (int c = 0; c < 7; c++) {
        HostBvh& hb = out.bvh[c];
        hb.count = counts[c];
        if (counts[c] == 0) continue;
        // triangle bounds are computed where the BVH is built
        const bool want_bounds = build_bvh || c != 2;
        if (want_bounds) hb.bounds.resize(6 * (size_t)counts[c]);
        hb.skip.assign(counts[c], 0);
        float maxtr = 0.0f;
        for (int i = 0; i < counts[c]; i++) {
            if (want_bounds) host_prim_bounds(s, c, i, &hb.bounds[6 * (size_t)i], &hb.bounds[6 * (size_t)i + 3]);
            if (c == 0 && s.objects[i].type != 0) { hb.skip[i] = 1; continue; }
            // a superellipsoid flagged _pad0 = 1 is a CSG operand only (.sheen se_csg)
            if (c == 4 && s.superellipsoids[i]._pad0 == 1) { hb.skip[i] = 1; continue; }
            maxtr = std::fmax(maxtr, s.materials[prim_material(s, c, i)].transmission);
        }
        hb.opaque = maxtr <= 0.5f ? 1 : 0;
        if (!build_bvh) continue;
        std::vector<float> mn(3 * (size_t)counts[c]), mx(3 * (size_t)counts[c]);
        for (int i = 0; i < counts[c]; i++)
            for (int a = 0; a < 3; a++) {
                mn[3 * (size_t)i + a] = hb.bounds[6 * (size_t)i + a];
                mx[3 * (size_t)i + a] = hb.bounds[6 * (size_t)i + 3 + a];
            }
        WBvhHost wb;
        try {
            wbvh_build(mn.data(), mx.data(), counts[c], hb.skip.data(), WBVH_MAX_LEAF, WBVH_BINS, wb);
        } catch (const std::exception& e) {
            err = e.what();
            return false;
        }
        if (wb.max_stack > 128) { err = "BVH too deep for the traversal stack"; return false; }
        hb.nodes.swap(wb.nodes);
        hb.prims.swap(wb.prims);
        hb.max_stack = wb.max_stack;
        hb.depth = wb.depth;
        if (c == 2) {
            out.tri_leaf.resize(12 * hb.prims.size());
            for (size_t k = 0; k < hb.prims.size(); k++) {
                const Triangle& t = s.triangles[hb.prims[k]];
                float* q = &out.tri_leaf[12 * k];
                int idx = (int)hb.prims[k];
                q[0] = t.v0.x; q[1] = t.v0.y; q[2] = t.v0.z; std::memcpy(&q[3], &idx, 4);
                q[4] = t.v1.x; q[5] = t.v1.y; q[6] = t.v1.z; q[7] = 0.0f;
                q[8] = t.v2.x; q[9] = t.v2.y; q[10] = t.v2.z; q[11] = 0.0f;
            }
        }
    }
    out.tex.resize(s.num_textures > 0 ? s.num_textures : 0);
    for (int i = 0; i < s.num_textures; i++) {
        HostTexture& t = out.tex[i];
        t.w = s.texture_widths ? s.texture_widths[i] : 0;
        t.h = s.texture_heights ? s.texture_heights[i] : 0;
        if (s.texture_pixels && s.texture_pixels[i] && t.w > 0 && t.h > 0) {
            t.px.assign(s.texture_pixels[i], s.texture_pixels[i] + (size_t)t.w * t.h * 3);
            host_build_mips(t);
        }
    }
    out.info.bvh_build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return true;
}
