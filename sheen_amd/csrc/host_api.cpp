// host_api.cpp — host-only checks exported by libsheen_host.so for CPU tests.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "host_util.h"
#include "wbvh_rules.h"

static float decode(uint16_t q, uint8_t e, float o) { return std::fmaf(wb_q_value(q), wb_exp_scale(e), o); }

// Builds category `cat` of a scene and checks the wide BVH: every primitive
// is referenced exactly once and every decoded child box along its path
// contains its bounds.  info[0..5] = nodes, max_stack, depth, leaves,
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
    std::vector<int> seen(hb.count, 0);
    int violations = 0, leaves = 0;
    // walk with the chain of decoded ancestor boxes
    struct Item { uint32_t node; float mn[3], mx[3]; };
    std::vector<Item> stack;
    Item root;
    root.node = 0;
    for (int a = 0; a < 3; a++) { root.mn[a] = -INFINITY; root.mx[a] = INFINITY; }
    stack.push_back(root);
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        const WNode& n = hb.nodes[it.node];
        const uint8_t ex[3] = {n.ex, n.ey, n.ez};
        const float org[3] = {n.ox, n.oy, n.oz};
        for (int i = 0; i < 8; i++) {
            uint8_t m = n.meta[i];
            if (m == WBVH_EMPTY) continue;
            float mn[3], mx[3];
            for (int a = 0; a < 3; a++) {
                mn[a] = decode(n.q[2 * a][i], ex[a], org[a]);
                mx[a] = decode(n.q[2 * a + 1][i], ex[a], org[a]);
                // the child box must lie inside the parent's decoded box
                if (mn[a] < it.mn[a] && !(mn[a] <= it.mn[a])) violations++;
            }
            if (m >> 5) {
                leaves++;
                for (uint32_t k = 0; k < (uint32_t)(m >> 5); k++) {
                    uint32_t p = hb.prims[n.prim_base + (m & 31) + k];
                    seen[p]++;
                    const float* b = &hb.bounds[6 * (size_t)p];
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
    for (int p = 0; p < hb.count; p++) {
        bool skipped = (cat == 0 && s->objects[p].type != 0);
        if (!skipped && seen[p] != 1) violations++;
    }
    info[3] = leaves;
    info[4] = violations;
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
