// bvh8_build.hpp -- host flattening of the Go binary SAH BVH into bc_node8 / bc_prim arrays.
#pragma once
#include <cmath>
#include <string>
#include <vector>

#include "../../include/beamcast_hip.h"
#include "bc_layout.h"

// Leaf slots hold BC_LEAF_MAX primitives (see bc_leaf_run for larger Go leaves); BC_PRIM_COST is
// the cost of one primitive test relative to one wide-node visit in the collapse.  One primitive
// per slot gives every triangle its own quantised box: a box test costs about a third of a
// triangle test in VALU instructions, and closest-hit traversal is bound by VALU issue (measured
// on the 1M-triangle soup: incoherent closest hit 11% faster than with two per slot).
#ifndef BC_LEAF_MAX
#define BC_LEAF_MAX 1
#endif
#ifndef BC_PRIM_COST
#define BC_PRIM_COST 0.3
#endif

// Primitives per leaf slot when a binary leaf of `count` primitives is split into slots:
// BC_LEAF_MAX, or more when that would take more than the eight slots of one node (the Go build
// makes leaves of up to 16 primitives), so a leaf always fits one wide node.
BC_HD inline int bc_leaf_run(int count) {
    const int r = (count + 7) / 8;
    return r > BC_LEAF_MAX ? r : BC_LEAF_MAX;
}

namespace bcx {

struct Bvh8 {
    std::vector<bc_node8> nodes;
    std::vector<bc_prim> prims;  // leaf order
    std::vector<uint32_t> ids;   // by record: Scene.Primitives index | BC_SPHERE_BIT
    std::vector<int32_t> mats;   // by record: material
    int depth = 0;               // levels of bc_node8 (root = 1)
    std::vector<int32_t> node_depth;  // by node: its level (root = 1), for level-ordered refits
};

// Primitive record of Scene.Primitives[id] from a scene view (the fp32 operands of the tests).
bc_prim prim_record(const bc_scene_view& v, int id);

// validateGPUSceneEncoding analogue: counts, leaf sizes, child links and leaf ranges.
bool validate_view(const bc_scene_view& v, std::string& err);

// Returns false with a message when the view is not encodable.
bool build_bvh8(const bc_scene_view& v, Bvh8& out, std::string& err);

// Smallest e with 2^e >= x for x > 0, exactly (the quantisation grid exponent).
inline int ceil_log2(double x) {
    int k;
    const double m = std::frexp(x, &k);
    return m == 0.5 ? k - 1 : k;
}

}  // namespace bcx
