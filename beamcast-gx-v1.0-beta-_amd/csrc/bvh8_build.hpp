// bvh8_build.hpp -- host flattening of the Go binary SAH BVH into bc_node8 / bc_prim arrays.
#pragma once
#include <string>
#include <vector>

#include "../../include/beamcast_hip.h"
#include "bc_layout.h"

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

// Returns false with a message when the view is not encodable.
bool build_bvh8(const bc_scene_view& v, Bvh8& out, std::string& err);

}  // namespace bcx
