// host_api.cpp — host-only checks exported by libsheen_host.so for CPU tests.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "host_util.h"
#include "wbvh_rules.h"

static float decode(uint16_t q, uint8_t e, float o) { return std::fmaf(wb_q_value(q), wb_exp_scale(e), o); }

// Checks a wide BVH over `count` primitive boxes (bounds: 6 floats each,
// min xyz then max xyz): every node's slots hold its internal children
// first, then its leaves, then empty slots, as its slot counts say; every
// primitive not flagged in skip (nullable) is
// referenced exactly once, every decoded child box lies inside the decoded
// boxes of its ancestors, and every leaf's decoded box and all its
// ancestors' contain the bounds of the leaf's primitives.  The quantiser
// takes a node's origin at the exact lower corner of its children, so the
// lower sides nest exactly; an upper side may pass its parent's by the
// rounding of one step of its own node (scale s), allowed as 2s plus a few
// float spacings.  info[0..1] = leaves, violations.  Returns 0.
extern "C" int sheen_host_wbvh_check(const void* nodes_v, int n_nodes, const unsigned int* prims, int n_prims,
                                     const float* bounds, int count, const unsigned char* skip, int* info) {
    const WNode* nodes = static_cast<const WNode*>(nodes_v);
    info[0] = info[1] = 0;
    if (n_nodes <= 0) return 0;
    std::vector<int> seen(count, 0);
    int violations = 0, leaves = 0;
    struct Item { uint32_t node; float mn[3], mx[3]; };
    std::vector<Item> stack;
    Item root;
    root.node = 0;
    for (int a = 0; a < 3; a++) { root.mn[a] = -INFINITY; root.mx[a] = INFINITY; }
    stack.push_back(root);
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        if (it.node >= (uint32_t)n_nodes) { violations++; continue; }
        const WNode& n = nodes[it.node];
        const uint8_t ex[3] = {n.ex, n.ey, n.ez};
        const float org[3] = {n.ox, n.oy, n.oz};
        // slot layout: internal children (meta = slot) first, then leaves, then empty slots
        const int nc = n.slots & 15, ni = n.slots >> 4;
        if (nc > 8 || ni > nc) violations++;
        for (int i = 0; i < 8; i++) {
            const uint8_t mi = n.meta[i];
            const bool ok = i < ni ? mi == i : i < nc ? (mi != WBVH_EMPTY && (mi >> 5) != 0) : mi == WBVH_EMPTY;
            if (!ok) violations++;
        }
        for (int i = 0; i < 8; i++) {
            uint8_t m = n.meta[i];
            if (m == WBVH_EMPTY) continue;
            float mn[3], mx[3];
            for (int a = 0; a < 3; a++) {
                mn[a] = decode(n.q[2 * a][i], ex[a], org[a]);
                mx[a] = decode(n.q[2 * a + 1][i], ex[a], org[a]);
                const float slack = 2.0f * wb_exp_scale(ex[a]) + std::fabs(mx[a]) * 0x1p-21f;
                if (!(mn[a] >= it.mn[a]) || !(mx[a] <= it.mx[a] + slack)) violations++;
            }
            if (m >> 5) {
                leaves++;
                for (uint32_t k = 0; k < (uint32_t)(m >> 5); k++) {
                    const uint32_t slot = n.prim_base + (m & 31) + k;
                    if (slot >= (uint32_t)n_prims || prims[slot] >= (uint32_t)count) { violations++; continue; }
                    const uint32_t p = prims[slot];
                    seen[p]++;
                    const float* b = &bounds[6 * (size_t)p];
                    for (int a = 0; a < 3; a++)
                        if (!(mn[a] <= b[a]) || !(mx[a] >= b[3 + a]) || !(it.mn[a] <= b[a]) || !(it.mx[a] >= b[3 + a]))
                            violations++;
                }
            } else {
                Item c;
                c.node = n.child_base + (m & 31);
                for (int a = 0; a < 3; a++) { c.mn[a] = std::fmax(mn[a], it.mn[a]); c.mx[a] = std::fmin(mx[a], it.mx[a]); }
                stack.push_back(c);
            }
        }
    }
    for (int p = 0; p < count; p++) {
        const bool skipped = skip && skip[p];
        if (!skipped && seen[p] != 1) violations++;
    }
    info[0] = leaves;
    info[1] = violations;
    return 0;
}

// The host-built wide BVH of category `cat`: info[0..4] = nodes, leaf
// slots, primitives, max_stack, depth.  nodes (128 B each), prims, bounds
// (6 floats per primitive) and skip (1 byte per primitive) may be null to
// query the sizes; otherwise they must hold info[0], info[1], 6 info[2] and
// info[2] entries.  Returns 0, or -1 with an invalid scene.
extern "C" int sheen_host_bvh_nodes(const SheenScene* s, int cat, void* nodes, unsigned int* prims, float* bounds,
                                    unsigned char* skip, int* info) {
    if (!s || cat < 0 || cat > 6 || !info) return -1;
    HostScene hs;
    std::string err;
    if (!host_prepare_scene(*s, hs, err)) return -1;
    const HostBvh& hb = hs.bvh[cat];
    info[0] = (int)hb.nodes.size();
    info[1] = (int)hb.prims.size();
    info[2] = hb.count;
    info[3] = hb.max_stack;
    info[4] = hb.depth;
    if (nodes && !hb.nodes.empty()) std::memcpy(nodes, hb.nodes.data(), hb.nodes.size() * sizeof(WNode));
    if (prims && !hb.prims.empty()) std::memcpy(prims, hb.prims.data(), hb.prims.size() * sizeof(uint32_t));
    if (bounds && !hb.bounds.empty()) std::memcpy(bounds, hb.bounds.data(), hb.bounds.size() * sizeof(float));
    if (skip)
        for (int p = 0; p < hb.count; p++) skip[p] = (p < (int)hb.skip.size()) ? hb.skip[p] : 0;
    return 0;
}

// Builds category `cat` of a scene and checks its wide BVH with
// sheen_host_wbvh_check.  info[0..5] = nodes, max_stack, depth, leaves,
// violations, references.  Returns 0, or -1 with an invalid scene.
extern "C" int sheen_host_bvh_check(const SheenScene* s, int cat, int* info) {
    if (!s || cat < 0 || cat > 6 || !info) return -1;
    HostScene hs;
    std::string err;
    if (!host_prepare_scene(*s, hs, err)) return -1;
    const HostBvh& hb = hs.bvh[cat];
    for (int k = 0; k < 6; k++) info[k] = 0;
    info[0] = (int)hb.nodes.size();
    info[1] = hb.max_stack;
    info[2] = hb.depth;
    info[5] = (int)hb.prims.size();
    if (hb.nodes.empty()) return 0;
    std::vector<unsigned char> skip(hb.count, 0);
    for (int p = 0; p < hb.count && p < (int)hb.skip.size(); p++) skip[p] = hb.skip[p];
    int r[2];
    sheen_host_wbvh_check(hb.nodes.data(), (int)hb.nodes.size(), hb.prims.data(), (int)hb.prims.size(), hb.bounds.data(),
                          hb.count, skip.data(), r);
    info[3] = r[0];
    info[4] = r[1];
    return 0;
}

// CSG node boxes the render library uploads (6 floats per node).  Returns
// the node count.
extern "C" int sheen_host_csg_boxes(const SheenScene* s, float* out) {
    if (!s) return -1;
    std::vector<float> b;
    host_csg_boxes(*s, b);
    if (out && !b.empty()) std::memcpy(out, b.data(), b.size() * sizeof(float));
    return (int)(b.size() / 6);
}

// The CSG root's union spine as the render library builds it: info[0] = 1
// when the root is a union, info[1] = subtrees, info[2] = leading unbounded
// ones, info[3] = hierarchy nodes.  sub (4 ints per subtree: node, rank,
// depth, 0) and nodes (8 words per node) may be null to query the sizes.
extern "C" int sheen_host_csg_spine(const SheenScene* s, int* info, int* sub, void* nodes) {
    if (!s || !info) return -1;
    std::vector<float> b;
    host_csg_boxes(*s, b);
    HostCsgSpine sp;
    host_csg_spine(*s, b, sp);
    info[0] = sp.spine ? 1 : 0;
    info[1] = (int)(sp.sub.size() / 4);
    info[2] = sp.always;
    info[3] = (int)sp.nodes.size();
    if (sub && !sp.sub.empty()) std::memcpy(sub, sp.sub.data(), sp.sub.size() * sizeof(int));
    if (nodes && !sp.nodes.empty()) std::memcpy(nodes, sp.nodes.data(), sp.nodes.size() * sizeof(HostSpineNode));
    return 0;
}
