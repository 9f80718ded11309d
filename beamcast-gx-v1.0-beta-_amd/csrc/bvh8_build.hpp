// bvh8_build.hpp -- host flattening of the Go binary SAH BVH into bc_node8 / bc_prim arrays.
#pragma once
#include <cmath>
#include <string>
#include <vector>

#include "../../include/beamcast_hip.h"
#include "bc_layout.h"

// Leaf runs hold at most BC_LEAF_MAX primitives (8 slots x BC_LEAF_MAX <= 32); BC_PRIM_COST is
// the cost of one primitive test relative to one wide-node visit in the collapse.
#ifndef BC_LEAF_MAX
#define BC_LEAF_MAX 2
#endif
#ifndef BC_PRIM_COST
#define BC_PRIM_COST 0.3
#endif

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
