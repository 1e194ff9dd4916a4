// wbvh_build.cpp — host builder of the 8-wide compressed BVH (see wbvh.h).
//
// Replaces the reference's per-category object-median binary builders
// (reference bvh.h:71-576, driven by scene.h:410-481): same input (one AABB
// per primitive, computed by the reference's bounds rules), different tree.
// Closest-hit results do not depend on the tree, only on which primitives a
// ray can reach; the traversal keeps that set a superset of the reference's.
#include "wbvh.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace {

struct Aabb {
    float mn[3], mx[3];
    void reset() { mn[0] = mn[1] = mn[2] = INFINITY; mx[0] = mx[1] = mx[2] = -INFINITY; }
    void grow(const Aabb& b) {
        for (int a = 0; a < 3; a++) { mn[a] = std::fmin(mn[a], b.mn[a]); mx[a] = std::fmax(mx[a], b.mx[a]); }
    }
    void grow_pt(const float* p) {
        for (int a = 0; a < 3; a++) { mn[a] = std::fmin(mn[a], p[a]); mx[a] = std::fmax(mx[a], p[a]); }
    }
    double area() const {
        double dx = (double)mx[0] - mn[0], dy = (double)mx[1] - mn[1], dz = (double)mx[2] - mn[2];
        if (!(dx >= 0) || !(dy >= 0) || !(dz >= 0)) return 0.0;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct BinNode {
    Aabb box;
    int left = -1, right = -1;  // children (internal)
    int start = 0, count = 0;   // primitive range (leaf when count > 0)
};

struct Builder {
    const float* bmin;
    const float* bmax;
    int max_leaf, bins;
    std::vector<int> idx;
    std::vector<float> cen;  // 3 per prim
    std::vector<Aabb> pbox;
    std::vector<BinNode> nodes;

    Aabb range_box(int s, int e) const {
        Aabb b; b.reset();
        for (int i = s; i < e; i++) b.grow(pbox[idx[i]]);
        return b;
    }

    // Returns node index.  Iterative top-down split with an explicit work list.
    void build(int n) {
        nodes.reserve(2 * (size_t)n / std::max(1, max_leaf) + 4);
        struct Work { int node, s, e; };
        std::vector<Work> work;
        nodes.push_back(BinNode());
        work.push_back({0, 0, n});
        std::vector<Aabb> bb(bins), lb(bins);
        std::vector<int> bc(bins), lc(bins);
        while (!work.empty()) {
            Work w = work.back();
            work.pop_back();
            int s = w.s, e = w.e, cnt = e - s;
            nodes[w.node].box = range_box(s, e);
            if (cnt <= max_leaf) { nodes[w.node].start = s; nodes[w.node].count = cnt; continue; }
            Aabb cb; cb.reset();
            for (int i = s; i < e; i++) cb.grow_pt(&cen[3 * (size_t)idx[i]]);
            int best_axis = -1, best_bin = -1;
            double best_cost = INFINITY;
            for (int a = 0; a < 3; a++) {
                float lo = cb.mn[a], ext = cb.mx[a] - cb.mn[a];
                if (!(ext > 0.0f)) continue;
                float k = (float)bins / ext;
                for (int b = 0; b < bins; b++) { bb[b].reset(); bc[b] = 0; }
                for (int i = s; i < e; i++) {
                    int p = idx[i];
                    int b = (int)((cen[3 * (size_t)p + a] - lo) * k);
                    b = std::min(std::max(b, 0), bins - 1);
                    bb[b].grow(pbox[p]);
                    bc[b]++;
                }
                Aabb acc; acc.reset();
                int ac = 0;
                for (int b = 0; b < bins - 1; b++) {
                    acc.grow(bb[b]); ac += bc[b];
                    lb[b] = acc; lc[b] = ac;
                }
                acc.reset(); ac = 0;
                for (int b = bins - 1; b >= 1; b--) {
                    acc.grow(bb[b]); ac += bc[b];
                    int l = lc[b - 1];
                    if (l == 0 || ac == 0) continue;
                    double c = lb[b - 1].area() * l + acc.area() * ac;
                    if (c < best_cost) { best_cost = c; best_axis = a; best_bin = b; }
                }
            }
            int mid;
            if (best_axis < 0) {
                mid = s + cnt / 2;  // all centroids coincide: split by index
            } else {
                float lo = cb.mn[best_axis], ext = cb.mx[best_axis] - cb.mn[best_axis];
                float k = (float)bins / ext;
                auto it = std::partition(idx.begin() + s, idx.begin() + e, [&](int p) {
                    int b = (int)((cen[3 * (size_t)p + best_axis] - lo) * k);
                    b = std::min(std::max(b, 0), bins - 1);
                    return b < best_bin;
                });
                mid = (int)(it - idx.begin());
                if (mid == s || mid == e) mid = s + cnt / 2;
            }
            int l = (int)nodes.size();
            nodes.push_back(BinNode());
            int r = (int)nodes.size();
            nodes.push_back(BinNode());
            nodes[w.node].left = l;
            nodes[w.node].right = r;
            work.push_back({r, mid, e});
            work.push_back({l, s, mid});
        }
    }
};

// Exponent byte e (scale 2^(e-127)) such that extent/scale <= 65000.
uint8_t pick_exp(float ext) {
    if (!(ext > 0.0f)) return 1;
    int k;
    std::frexp((double)ext / 65000.0, &k);  // ext/65000 < 2^k
    int e = k + 127;
    if (e < 1) e = 1;
    if (e > 254) e = 254;
    return (uint8_t)e;
}
inline float exp_scale(uint8_t e) {
    uint32_t bits = (uint32_t)e << 23;
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

void quantize(WNode& wn, const Aabb* cb, int nc) {
    Aabb nb; nb.reset();
    for (int i = 0; i < nc; i++) nb.grow(cb[i]);
    float org[3] = {nb.mn[0], nb.mn[1], nb.mn[2]};
    wn.ox = org[0]; wn.oy = org[1]; wn.oz = org[2];
    uint8_t ex[3];
    for (int a = 0; a < 3; a++) ex[a] = pick_exp(nb.mx[a] - nb.mn[a]);
    wn.ex = ex[0]; wn.ey = ex[1]; wn.ez = ex[2];
    for (int a = 0; a < 3; a++) {
        for (;;) {
            float sc = exp_scale(ex[a]);
            bool ok = true;
            for (int i = 0; i < WBVH_WIDTH; i++) {
                if (i >= nc) { wn.q[2 * a][i] = 65535; wn.q[2 * a + 1][i] = 0; continue; }
                double lo = std::floor(((double)cb[i].mn[a] - (double)org[a]) / (double)sc);
                double hi = std::ceil(((double)cb[i].mx[a] - (double)org[a]) / (double)sc);
                long ql = (long)std::max(0.0, lo), qh = (long)std::max(0.0, hi);
                while (ql > 0 && std::fmaf((float)ql, sc, org[a]) > cb[i].mn[a]) ql--;
                while (qh < 65535 && std::fmaf((float)qh, sc, org[a]) < cb[i].mx[a]) qh++;
                if (std::fmaf((float)ql, sc, org[a]) > cb[i].mn[a] || std::fmaf((float)qh, sc, org[a]) < cb[i].mx[a] ||
                    qh > 65535) {
                    ok = false;
                    break;
                }
                wn.q[2 * a][i] = (uint16_t)ql;
                wn.q[2 * a + 1][i] = (uint16_t)qh;
            }
            if (ok) break;
            if (ex[a] >= 254) throw std::runtime_error("wbvh: cannot quantise child box");
            ex[a]++;
        }
    }
    wn.ex = ex[0]; wn.ey = ex[1]; wn.ez = ex[2];
}

}  // namespace

void wbvh_build(const float* bmin, const float* bmax, int n, const unsigned char* skip, int max_leaf, int bins,
                WBvhHost& out) {
    out.nodes.clear();
    out.prims.clear();
    out.max_stack = 0;
    out.depth = 0;
    out.leaf_count = 0;
    if (max_leaf < 1) max_leaf = 1;
    if (max_leaf > 7) max_leaf = 7;
    Builder B;
    B.bmin = bmin; B.bmax = bmax; B.max_leaf = max_leaf; B.bins = bins < 2 ? 2 : bins;
    B.pbox.resize(n);
    B.cen.resize(3 * (size_t)n);
    for (int i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) {
            B.pbox[i].mn[a] = bmin[3 * (size_t)i + a];
            B.pbox[i].mx[a] = bmax[3 * (size_t)i + a];
            // centroid in double then rounded: robust for the +-1e30 bounds of lathes
            B.cen[3 * (size_t)i + a] = (float)(0.5 * ((double)bmin[3 * (size_t)i + a] + (double)bmax[3 * (size_t)i + a]));
        }
        if (!skip || !skip[i]) B.idx.push_back(i);
    }
    int m = (int)B.idx.size();
    if (m == 0) return;
    B.build(m);

    // Collapse to 8-wide nodes, breadth first.
    struct Item { int bin, wide; };
    std::vector<Item> queue;
    out.nodes.push_back(WNode());
    queue.push_back({0, 0});
    std::vector<int> need_parent;  // wide parent per node (for the stack bound)
    need_parent.push_back(-1);
    std::vector<int> n_internal;   // internal children per wide node
    std::vector<int> level;
    level.push_back(1);
    for (size_t qi = 0; qi < queue.size(); qi++) {
        Item it = queue[qi];
        std::vector<int> ch;
        const BinNode& bn = B.nodes[it.bin];
        if (bn.count > 0) ch.push_back(it.bin);
        else { ch.push_back(bn.left); ch.push_back(bn.right); }
        while ((int)ch.size() < WBVH_WIDTH) {
            int best = -1;
            double ba = -1.0;
            for (int k = 0; k < (int)ch.size(); k++) {
                const BinNode& c = B.nodes[ch[k]];
                if (c.count > 0) continue;
                double a = c.box.area();
                if (a > ba) { ba = a; best = k; }
            }
            if (best < 0) break;
            int b = ch[best];
            ch[best] = B.nodes[b].left;
            ch.push_back(B.nodes[b].right);
        }
        WNode wn;
        std::memset(&wn, 0, sizeof(wn));
        std::memset(wn.meta, 0xFF, sizeof(wn.meta));
        wn.nchild = (uint8_t)ch.size();
        int ni = 0;
        for (int c : ch) if (B.nodes[c].count == 0) ni++;
        wn.child_base = (uint32_t)out.nodes.size();
        wn.prim_base = (uint32_t)out.prims.size();
        Aabb cb[WBVH_WIDTH];
        int k_int = 0, off = 0;
        for (int k = 0; k < (int)ch.size(); k++) {
            const BinNode& c = B.nodes[ch[k]];
            cb[k] = c.box;
            if (c.count > 0) {
                wn.meta[k] = (uint8_t)((c.count << 5) | off);
                for (int p = 0; p < c.count; p++) out.prims.push_back((uint32_t)B.idx[c.start + p]);
                off += c.count;
                out.leaf_count++;
            } else {
                wn.meta[k] = (uint8_t)k_int;
                int wi = (int)out.nodes.size() + 0;  // allocated below in order
                (void)wi;
                k_int++;
            }
        }
        // allocate internal children contiguously in slot order
        for (int k = 0; k < (int)ch.size(); k++) {
            if (B.nodes[ch[k]].count > 0) continue;
            int wi = (int)out.nodes.size();
            out.nodes.push_back(WNode());
            need_parent.push_back(it.wide);
            level.push_back(level[it.wide] + 1);
            queue.push_back({ch[k], wi});
        }
        quantize(wn, cb, (int)ch.size());
        out.nodes[it.wide] = wn;
        if ((int)n_internal.size() <= it.wide) n_internal.resize(it.wide + 1, 0);
        n_internal[it.wide] = ni;
    }
    // Worst-case stack entries: a node with two or more internal children
    // keeps one entry (its unvisited siblings) while a child is explored:
    // need(v) = [internal(v) >= 2] + max need(child).
    int nn = (int)out.nodes.size();
    n_internal.resize(nn, 0);
    std::vector<int> need(nn, 0);
    for (int v = nn - 1; v >= 0; v--) {
        int best = 0;
        const WNode& w = out.nodes[v];
        for (int k = 0; k < WBVH_WIDTH; k++) {
            uint8_t mt = w.meta[k];
            if (mt == WBVH_EMPTY || (mt >> 5) != 0) continue;
            best = std::max(best, need[w.child_base + (mt & 31)]);
        }
        need[v] = (n_internal[v] >= 2 ? 1 : 0) + best;
    }
    out.max_stack = need[0];
    for (int l : level) out.depth = std::max(out.depth, l);
}
