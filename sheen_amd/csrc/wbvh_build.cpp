// wbvh_build.cpp — host builder of the 8-wide compressed BVH (see wbvh.h).
//
// Replaces the reference's per-category object-median binary builders
// (reference bvh.h:71-576, driven by scene.h:410-481): same input (one AABB
// per primitive, computed by the reference's bounds rules), different tree.
// Closest-hit results do not depend on the tree, only on which primitives a
// ray can reach; the traversal keeps that set a superset of the reference's.
#include "wbvh.h"
#include "wbvh_rules.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace {

struct BinNode {
    WbBox box;
    int left = -1, right = -1;  // children (internal)
    int start = 0, count = 0;   // primitive range (leaf when count > 0)
};

struct Builder {
    int max_leaf;
    std::vector<int> idx;
    std::vector<float> cen;  // 3 per prim
    std::vector<WbBox> pbox;
    std::vector<BinNode> nodes;

    WbBox range_box(int s, int e) const {
        WbBox b;
        wb_reset(b);
        for (int i = s; i < e; i++) wb_grow(b, pbox[idx[i]]);
        return b;
    }

    // Iterative top-down binned-SAH split with an explicit work list.  The
    // partition is stable, so primitives keep their input order within a
    // node (the device builder reproduces it with prefix sums).
    void build(int n) {
        nodes.reserve(2 * (size_t)n / std::max(1, max_leaf) + 4);
        struct Work { int node, s, e; };
        std::vector<Work> work;
        nodes.push_back(BinNode());
        work.push_back({0, 0, n});
        WbBox bb[3][WBVH_BINS];
        int bc[3][WBVH_BINS];
        while (!work.empty()) {
            Work w = work.back();
            work.pop_back();
            int s = w.s, e = w.e, cnt = e - s;
            nodes[w.node].box = range_box(s, e);
            if (cnt <= max_leaf) { nodes[w.node].start = s; nodes[w.node].count = cnt; continue; }
            WbBox cb;
            wb_reset(cb);
            for (int i = s; i < e; i++) {
                const float* c = &cen[3 * (size_t)idx[i]];
                const WbBox pt = {{c[0], c[1], c[2]}, {c[0], c[1], c[2]}};
                wb_grow(cb, pt);
            }
            for (int a = 0; a < 3; a++) {
                for (int b = 0; b < WBVH_BINS; b++) { wb_reset(bb[a][b]); bc[a][b] = 0; }
                const float ext = cb.mx[a] - cb.mn[a];
                if (!(ext > 0.0f)) continue;
                const float lo = cb.mn[a], k = wb_bin_scale(ext);
                for (int i = s; i < e; i++) {
                    const int p = idx[i];
                    const int b = wb_bin_of(cen[3 * (size_t)p + a], lo, k);
                    wb_grow(bb[a][b], pbox[p]);
                    bc[a][b]++;
                }
            }
            int best_axis, best_bin;
            wb_sah(cb.mn, cb.mx, [&](int a, int b) -> const WbBox& { return bb[a][b]; },
                   [&](int a, int b) { return bc[a][b]; }, best_axis, best_bin);
            int mid;
            if (best_axis < 0) {
                mid = s + cnt / 2;  // all centroids coincide: split by index
            } else {
                const float lo = cb.mn[best_axis], k = wb_bin_scale(cb.mx[best_axis] - cb.mn[best_axis]);
                auto it = std::stable_partition(idx.begin() + s, idx.begin() + e, [&](int p) {
                    return wb_bin_of(cen[3 * (size_t)p + best_axis], lo, k) < best_bin;
                });
                mid = (int)(it - idx.begin());  // both sides non-empty by the SAH choice
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

}  // namespace

void wbvh_build(const float* bmin, const float* bmax, int n, const unsigned char* skip, int max_leaf, int bins,
                WBvhHost& out) {
    out.nodes.clear();
    out.prims.clear();
    out.max_stack = 0;
    out.depth = 0;
    out.leaf_count = 0;
    if (max_leaf < 1) max_leaf = 1;
    if (max_leaf > 32 / WBVH_WIDTH) max_leaf = 32 / WBVH_WIDTH;  // leaf offsets fit 5 bits (wbvh.h)
    if (bins != WBVH_BINS) throw std::runtime_error("wbvh: the builder bins WBVH_BINS ways");
    Builder B;
    B.max_leaf = max_leaf;
    B.pbox.resize(n);
    B.cen.resize(3 * (size_t)n);
    for (int i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) {
            B.pbox[i].mn[a] = wb_canon(bmin[3 * (size_t)i + a]);
            B.pbox[i].mx[a] = wb_canon(bmax[3 * (size_t)i + a]);
            B.cen[3 * (size_t)i + a] = wb_centroid(B.pbox[i].mn[a], B.pbox[i].mx[a]);
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
        int chv[WBVH_WIDTH];
        const int nch = wb_collapse(it.bin, [&](int i, WbBox& box, int& l, int& r, int& cnt) {
            const BinNode& b = B.nodes[i];
            box = b.box; l = b.left; r = b.right; cnt = b.count;
        }, chv);
        // internal children take the first slots, leaves the next ones
        std::vector<int> ch;
        for (int k = 0; k < nch; k++) if (B.nodes[chv[k]].count == 0) ch.push_back(chv[k]);
        const int ni = (int)ch.size();
        for (int k = 0; k < nch; k++) if (B.nodes[chv[k]].count > 0) ch.push_back(chv[k]);
        WNode wn;
        std::memset(&wn, 0, sizeof(wn));
        std::memset(wn.meta, 0xFF, sizeof(wn.meta));
        wn.slots = (uint8_t)(ch.size() | ((size_t)ni << 4));
        wn.child_base = (uint32_t)out.nodes.size();
        wn.prim_base = (uint32_t)out.prims.size();
        WbBox cb[WBVH_WIDTH];
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
                wn.meta[k] = (uint8_t)k_int++;  // internal children are allocated below, in slot order
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
        if (!wb_quantize(wn, cb, (int)ch.size())) throw std::runtime_error("wbvh: cannot quantise child box");
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
