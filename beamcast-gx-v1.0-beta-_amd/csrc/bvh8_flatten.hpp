// bvh8_flatten.hpp -- the BVH2 -> BVH8 flattening on the device (flatten.hip).
#pragma once
#include <cstdint>
#include <vector>

#include "api_internal.hpp"
#include "bvh8_build.hpp"

namespace bcd {

struct FlattenResult {
    std::vector<uint32_t> ids;        // record ids (primitive | sphere bit), leaf order
    std::vector<uint32_t> level_off;  // refit level lists in d_levels, deepest level first
    int64_t nnode8 = 0;
    int depth = 0;
};

// Flattens the view's tree into c->d_nodes / d_prims / d_ids / d_pmat / d_levels with the same
// bytes as the host flattening (bcx::build_bvh8).  The view must pass bcx::validate_view.
int flatten_device(bc_ctx* c, const bc_scene_view& v, FlattenResult& res);

}  // namespace bcd
