// bvh8_build.hpp -- host flattening of the Go binary SAH BVH into bc_node8 / bc_prim arrays.
#pragma once
#include <cmath>
#include <string>
#include <vector>

#include "../../include/beamcast_hip.h"
#include "bc_layout.h"

// Every leaf slot holds one primitive, so every triangle has its own quantised box (a box test
// costs about a third of a triangle test in VALU instructions) and a leaf's record sits at a
// fixed position of its node (8 * node + slot); binary leaves of more than eight primitives are
// split in two first (prepare_view).  BC_PRIM_COST is the cost of one primitive test relative to
// one wide-node visit in the collapse.
#ifndef BC_PRIM_COST
#define BC_PRIM_COST 0.3
#endif
#define BC_BINARY_LEAF_MAX 8  // primitives per binary leaf after prepare_view

namespace bcx {

struct Bvh8 {
    std::vector<bc_node8> nodes;
    std::vector<bc_prim> prims;  // 8 record slots per node (holes: zero)
    std::vector<uint32_t> ids;   // by record slot: Scene.Primitives index | BC_SPHERE_BIT, or BC_REC_HOLE
    std::vector<int32_t> mats;   // by record slot: material (-1 in holes)
    int depth = 0;               // levels of bc_node8 (root = 1)
    std::vector<int32_t> node_depth;  // by node: its level (root = 1), for level-ordered refits
};

// Primitive record of Scene.Primitives[id] from a scene view (the fp32 operands of the tests).
bc_prim prim_record(const bc_scene_view& v, int id);

// validateGPUSceneEncoding analogue: counts, leaf sizes, child links and leaf ranges.
bool validate_view(const bc_scene_view& v, std::string& err);

// The view with every binary leaf of more than eight primitives turned into an internal node
// over two leaves (the first ceil(count / 2) primitives of its range, then the rest), appended
// after the original nodes with boxes of their primitives; the view itself when no leaf is that
// large.  Both flattenings (host and device) take the prepared view, so they emit the same bytes.
struct PreparedView {
    bc_scene_view view;
    std::vector<int32_t> link;
    std::vector<double> box;
};
const bc_scene_view& prepare_view(const bc_scene_view& v, PreparedView& pv);

// Returns false with a message when the view is not encodable.  Prepares the view itself.
bool build_bvh8(const bc_scene_view& v, Bvh8& out, std::string& err);

// Smallest e with 2^e >= x for x > 0, exactly (the quantisation grid exponent).
inline int ceil_log2(double x) {
    int k;
    const double m = std::frexp(x, &k);
    return m == 0.5 ? k - 1 : k;
}

}  // namespace bcx
