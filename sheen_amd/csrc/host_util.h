// host_util.h — host-side scene preparation shared by the C ABI library and
// the host tools library: reference bounds rules, wide-BVH builds per
// category, triangle leaf copies, texture mip chains, validation.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "sheen_c_api.h"
#include "wbvh.h"

struct HostBvh {
    std::vector<WNode> nodes;
    std::vector<uint32_t> prims;
    std::vector<float> bounds;  // 6 per primitive (all primitives of the category); triangles: empty without a host build
    std::vector<unsigned char> skip;  // primitives left out of the BVH (non-sphere objects, CSG-only superellipsoids)
    int count = 0;
    int opaque = 1;
    int max_stack = 0;
    int depth = 0;
};

struct HostTexture {
    std::vector<unsigned char> px;  // base level + mip chain, RGB
    int w = 0, h = 0, mips = 1;
};

struct HostSceneInfo {
    uint64_t photon_rays = 0;
    int photons_emitted = 0;
    int photons_stored = 0;
    double bvh_build_ms = 0.0;
};

struct HostScene {
    HostBvh bvh[7];                // sphere, strand, triangle, hf, se, lathe, prism
    std::vector<float> tri_leaf;   // 12 floats per triangle leaf slot
    std::vector<HostTexture> tex;
    HostSceneInfo info;
};

// Bounds of primitive i of category cat by the reference's rules
// (bvh.h:24-66, 235-242, 304-331, 394-407).
void host_prim_bounds(const SheenScene& s, int cat, int i, float* mn, float* mx);

// Box of every CSG node (6 floats: min xyz, max xyz) containing every span
// its subtree can return: a sphere or superellipsoid leaf's own box, the
// union of the children's boxes for a union, the left operand's box for an
// intersection or a difference (their spans are parts of the left operand's).
// Missing children give an empty box; cycles and non-finite values an
// unbounded one.
void host_csg_boxes(const SheenScene& s, std::vector<float>& out);

// Frames of the device's CSG evaluation stack.
#define HOST_CSG_DEPTH 16

// The CSG root's union spine: the subtrees hanging off the unions reachable
// from the root through unions, in the order an in-order evaluation meets
// them (rank), with the frame depth of each, and a median-split hierarchy
// over their boxes.  Layout of a node as the device's CsgSpineNode.
struct HostSpineNode {
    float lo[3];
    int a;  // internal: first child; leaf: -(first subtree + 1)
    float hi[3];
    int b;  // internal: second child; leaf: subtree count
};
static_assert(sizeof(HostSpineNode) == 32, "spine node layout");
struct HostCsgSpine {
    bool spine = false;               // the root is a union
    std::vector<HostSpineNode> nodes;
    std::vector<int> sub;             // 4 ints per subtree: node, rank, depth, 0
    int always = 0;                   // leading subtrees with unbounded boxes
};
void host_csg_spine(const SheenScene& s, const std::vector<float>& boxes, HostCsgSpine& out);

// 2x2 box mip chain with rounding (s+2)/4 (reference scene.h:230-261).
void host_build_mips(HostTexture& t);

// Validates references, computes primitive bounds and, with build_bvh, builds
// every category BVH and the triangle leaf copies on the host (without it the
// device builder does both).  false + err on error.
bool host_prepare_scene(const SheenScene& s, HostScene& out, std::string& err, bool build_bvh = true);

// Linear -> sRGB byte (reference png_color.h:11-21).
unsigned char host_srgb_byte(float c);

struct SheenContext;
bool photon_pass(SheenContext* c, const SheenScene* in);
void sheen_set_error(const char* fmt, ...);
