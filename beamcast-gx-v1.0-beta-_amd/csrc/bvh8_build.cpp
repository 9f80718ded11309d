// bvh8_build.cpp -- flattens the host-built binary SAH BVH into quantised 8-wide nodes.
//
// Input is the Go Scene's BVH exactly as Scene.Build leaves it (renderer.go:409-633:
// Nodes in DFS pre-order, leaves as [Start, Start+Count) ranges of Order).  The binary tree is
// collapsed into 8-wide nodes by a surface-area cost optimisation over the binary tree (the
// dynamic program of Ylitie, Karras and Laine, "Efficient Incoherent Ray Traversal on GPUs
// Through Compressed Wide BVHs", HPG 2017): for every binary node and every slot budget
// 1..8 it keeps the cheapest way to represent the subtree as at most that many children of
// one wide node -- a leaf of at most kLeafMax primitives, a wide internal node, or the budget
// split between the two binary children -- with cost = area x (kNodeCost per node visited,
// kPrimCost per primitive tested).  Leaf slots hold one primitive each; a binary leaf becomes
// several leaf slots, or a wide node of its own when it needs more than four (binary leaves of
// more than eight primitives are split in two beforehand, prepare_view).  A node's leaf records
// sit at its eight record slots, 8 * node + slot.
// Every primitive appears in exactly one leaf; which leaf only changes culling, never a hit
// result (equal distances resolve to the smaller primitive index).
#include "bvh8_build.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace bcx {

namespace {

struct Box {
    double mn[3], mx[3];
    void empty() {
        for (int a = 0; a < 3; a++) { mn[a] = INFINITY; mx[a] = -INFINITY; }
    }
    void grow(const Box& b) {
        for (int a = 0; a < 3; a++) {
            mn[a] = std::min(mn[a], b.mn[a]);
            mx[a] = std::max(mx[a], b.mx[a]);
        }
    }
    double area() const {
        double e0 = mx[0] - mn[0], e1 = mx[1] - mn[1], e2 = mx[2] - mn[2];
        if (e0 < 0 || e1 < 0 || e2 < 0) return 0;
        return 2 * (e0 * e1 + e1 * e2 + e2 * e0);
    }
};

struct Child {
    bool leaf;
    int node2;       // internal: binary node index
    int start, count;  // leaf: range of Order
    Box box;
};

Box node_box(const bc_scene_view& v, int i) {
    Box b;
    for (int a = 0; a < 3; a++) {
        b.mn[a] = v.node_box[6 * (size_t)i + a];
        b.mx[a] = v.node_box[6 * (size_t)i + 3 + a];
    }
    return b;
}

Box prim_box(const bc_scene_view& v, int p) {
    const double* g = v.prim_geom + 10 * (size_t)p;
    Box b;
    if (v.prim_kind[p] == 0) {
        double r = g[3];
        for (int a = 0; a < 3; a++) { b.mn[a] = g[a] - r; b.mx[a] = g[a] + r; }
    } else {
        for (int a = 0; a < 3; a++) {
            double A = g[a], B = g[a] + g[3 + a], C = g[a] + g[6 + a];
            b.mn[a] = std::min(A, std::min(B, C)) - 1e-6;
            b.mx[a] = std::max(A, std::max(B, C)) + 1e-6;
        }
    }
    return b;
}

// Record slot r of the output (grown to hold it, holes filled as empty slots).
void write_prim(const bc_scene_view& v, int id, size_t r, Bvh8& out) {
    if (out.prims.size() <= r) {
        const size_t n = (r / BC_REC_PER_NODE + 1) * BC_REC_PER_NODE;
        out.prims.resize(n, bc_prim{});
        out.ids.resize(n, BC_REC_HOLE);
        out.mats.resize(n, -1);
    }
    out.prims[r] = prim_record(v, id);
    out.ids[r] = (uint32_t)id | (v.prim_kind[id] == 0 ? BC_SPHERE_BIT : 0u);
    out.mats[r] = v.prim_mat[id];
}

constexpr int kLeafMax = 1;                 // primitives per leaf slot
constexpr double kNodeCost = 1.0;           // relative cost of visiting one wide node
constexpr double kPrimCost = BC_PRIM_COST;  // relative cost of one primitive test

// Per binary node: cost of representing its subtree with at most i (1..8) child slots of the
// enclosing wide node, and how.
struct Plan {
    float cost[8];
    uint8_t how[8];  // 0: use the plan for one slot fewer; 1: leaf; 2: internal wide node;
                     // 3..9: split k = how-2 slots to the left child, the rest to the right
    int32_t start, count;  // primitive range of the subtree in Order (count -1: not contiguous)
};

// Child items of wide nodes, produced top-down from the plans.
void expand(const bc_scene_view& v, const std::vector<Plan>& plan, int n, int slots, std::vector<Child>& out) {
    const Plan& P = plan[(size_t)n];
    int i = slots;
    while (i > 1 && P.how[i - 1] == 0) i--;
    const uint8_t h = P.how[i - 1];
    const int32_t* l = v.node_link + 4 * (size_t)n;
    if (h == 1) {  // leaf: the subtree's primitives, in runs of at most kLeafMax
        if (P.count <= kLeafMax) {
            Child c;
            c.leaf = true; c.node2 = -1; c.start = P.start; c.count = P.count;
            c.box = node_box(v, n);
            out.push_back(c);
        } else {
            for (int s = 0; s < P.count; s++) {
                Child c;
                c.leaf = true; c.node2 = -1; c.start = P.start + s; c.count = 1;
                c.box.empty();
                c.box.grow(prim_box(v, v.order[c.start]));
                out.push_back(c);
            }
        }
    } else if (h == 2) {  // a wide node of its own (an internal binary node, or a large leaf)
        Child c;
        c.leaf = false; c.node2 = n; c.start = 0; c.count = 0;
        c.box = node_box(v, n);
        out.push_back(c);
    } else {
        const int k = h - 2;
        expand(v, plan, l[0], k, out);
        expand(v, plan, l[1], i - k, out);
    }
}

// Bottom-up cost plans of every binary node (children have larger indices than parents).
void make_plans(const bc_scene_view& v, std::vector<Plan>& plan) {
    const float inf = INFINITY;
    plan.resize((size_t)v.nnode);
    for (int n = v.nnode - 1; n >= 0; n--) {
        Plan& P = plan[(size_t)n];
        const int32_t* l = v.node_link + 4 * (size_t)n;
        const double A = node_box(v, n).area();
        if (l[3] > 0) {  // binary leaf: one slot per primitive, at most eight
            P.start = l[2];
            P.count = l[3];
            const int need = l[3];
            const float c = (float)(A * kPrimCost * l[3]);
            // a leaf needing more than half a node may also be a wide node of its own (one slot)
            const float c_wide = need > 4 ? c + (float)(A * kNodeCost) : inf;
            for (int i = 1; i <= 8; i++) {
                P.cost[i - 1] = i >= need ? c : c_wide;
                P.how[i - 1] = i >= need ? 1 : 2;
            }
            continue;
        }
        const Plan& L = plan[(size_t)l[0]];
        const Plan& R = plan[(size_t)l[1]];
        const bool contiguous = L.count >= 0 && R.count >= 0 && R.start == L.start + L.count;
        P.start = L.start;
        P.count = contiguous ? L.count + R.count : -1;
        float dist[9];
        uint8_t dk[9];
        for (int j = 2; j <= 8; j++) {
            dist[j] = inf;
            dk[j] = 0;
            for (int k = 1; k < j; k++) {
                const float c = L.cost[k - 1] + R.cost[j - k - 1];
                if (c < dist[j]) { dist[j] = c; dk[j] = (uint8_t)k; }
            }
        }
        const float c_int = dist[8] + (float)(A * kNodeCost);
        const float c_leaf = (contiguous && P.count <= kLeafMax) ? (float)(A * kPrimCost * P.count) : inf;
        P.cost[0] = c_leaf <= c_int ? c_leaf : c_int;
        P.how[0] = c_leaf <= c_int ? 1 : 2;
        for (int i = 2; i <= 8; i++) {
            if (dist[i] < P.cost[i - 2]) {
                P.cost[i - 1] = dist[i];
                P.how[i - 1] = (uint8_t)(2 + dk[i]);
            } else {
                P.cost[i - 1] = P.cost[i - 2];
                P.how[i - 1] = 0;
            }
        }
    }
}

}  // namespace

bc_prim prim_record(const bc_scene_view& v, int id) {
    bc_prim r;
    std::memset(&r, 0, sizeof r);
    const double* g = v.prim_geom + 10 * (size_t)id;
    if (v.prim_kind[id] == 0) {
        const double rad = g[3];
        r.v[0] = (float)g[0]; r.v[1] = (float)g[1]; r.v[2] = (float)g[2];
        r.v[3] = (float)rad; r.v[4] = (float)(rad * rad); r.v[5] = (float)(1.0 / rad);
    } else {
        for (int k = 0; k < 9; k++) r.v[k] = (float)g[k];
    }
    return r;
}

bool validate_view(const bc_scene_view& v, std::string& err) {
    if (v.nnode <= 0 || v.nprim <= 0) { err = "scene has no BVH nodes or primitives"; return false; }
    // validateGPUSceneEncoding analogue (gpu_opencl_windows.go:470-482)
    if (v.nnode >= (1 << 27) || v.nprim >= (1 << 27)) { err = "BVH supports fewer than 2^27 nodes/primitives"; return false; }
    for (int i = 0; i < v.nnode; i++) {
        const int32_t* l = v.node_link + 4 * (size_t)i;
        if (l[3] < 0 || l[3] > 16) { err = "BVH leaf " + std::to_string(i) + " has unsupported count " + std::to_string(l[3]); return false; }
        if (l[3] == 0 && (l[0] <= i || l[1] <= i || l[0] >= v.nnode || l[1] >= v.nnode)) {
            err = "BVH node " + std::to_string(i) + " has invalid children"; return false;
        }
        if (l[3] > 0 && (l[2] < 0 || (int64_t)l[2] + l[3] > v.nprim)) { err = "BVH leaf range out of bounds"; return false; }
    }
    return true;
}

const bc_scene_view& prepare_view(const bc_scene_view& v, PreparedView& pv) {
    size_t big = 0;
    for (int i = 0; i < v.nnode; i++) big += v.node_link[4 * (size_t)i + 3] > BC_BINARY_LEAF_MAX ? 1 : 0;
    if (big == 0) return v;
    const size_t n0 = (size_t)v.nnode, n = n0 + 2 * big;
    pv.link.assign(v.node_link, v.node_link + 4 * n0);
    pv.link.resize(4 * n);
    pv.box.assign(v.node_box, v.node_box + 6 * n0);
    pv.box.resize(6 * n);
    size_t next = n0;
    for (size_t i = 0; i < n0; i++) {
        const int32_t start = pv.link[4 * i + 2], cnt = pv.link[4 * i + 3];
        if (cnt <= BC_BINARY_LEAF_MAX) continue;
        const int32_t first = (cnt + 1) / 2;
        const int32_t part[2][2] = {{start, first}, {start + first, cnt - first}};
        pv.link[4 * i + 0] = (int32_t)next;
        pv.link[4 * i + 1] = (int32_t)next + 1;
        pv.link[4 * i + 2] = 0;
        pv.link[4 * i + 3] = 0;
        for (int h = 0; h < 2; h++, next++) {
            int32_t* c = &pv.link[4 * next];
            c[0] = c[1] = -1;
            c[2] = part[h][0];
            c[3] = part[h][1];
            Box b;
            b.empty();
            for (int k = 0; k < part[h][1]; k++) b.grow(prim_box(v, v.order[part[h][0] + k]));
            for (int a = 0; a < 3; a++) {
                pv.box[6 * next + (size_t)a] = b.mn[a];
                pv.box[6 * next + 3 + (size_t)a] = b.mx[a];
            }
        }
    }
    pv.view = v;
    pv.view.nnode = (int32_t)n;
    pv.view.node_link = pv.link.data();
    pv.view.node_box = pv.box.data();
    return pv.view;
}

bool build_bvh8(const bc_scene_view& v0, Bvh8& out, std::string& err) {
    out.nodes.clear();
    out.prims.clear();
    out.ids.clear();
    out.mats.clear();
    out.node_depth.clear();
    out.depth = 0;
    if (!validate_view(v0, err)) return false;
    PreparedView pv;
    const bc_scene_view& v = prepare_view(v0, pv);
    out.nodes.reserve((size_t)v.nnode / 3 + 1);
    int64_t written = 0;
    std::vector<Plan> plan;
    make_plans(v, plan);
    if (!(plan[0].cost[7] < INFINITY)) { err = "internal: no BVH8 plan for the scene"; return false; }
    out.nodes.emplace_back();
    struct Work { int node2, i8, depth; };
    std::vector<Work> work;
    work.push_back({0, 0, 1});
    while (!work.empty()) {
        Work w = work.back();
        work.pop_back();
        out.depth = std::max(out.depth, w.depth);
        // the wide node's children: the root's (or an internal item's) best split of 8 slots
        std::vector<Child> fin;
        const int32_t* l0 = v.node_link + 4 * (size_t)w.node2;
        if (l0[3] > 0) {
            expand(v, plan, w.node2, 8, fin);  // a single binary leaf as the whole scene
        } else {
            // distribute 8 slots over the two binary children (the node itself is internal)
            const Plan& L = plan[(size_t)l0[0]];
            const Plan& R = plan[(size_t)l0[1]];
            float best = INFINITY;
            int bk = 1;
            for (int k = 1; k < 8; k++) {
                const float c = L.cost[k - 1] + R.cost[8 - k - 1];
                if (c < best) { best = c; bk = k; }
            }
            expand(v, plan, l0[0], bk, fin);
            expand(v, plan, l0[1], 8 - bk, fin);
        }
        if (fin.size() > 8) { err = "internal: BVH8 node overflow"; return false; }
        Box nb;
        nb.empty();
        for (auto& c : fin) nb.grow(c.box);
        // octant slot assignment: slot bit a set <=> child centre on the + side of axis a
        double pc[3];
        for (int a = 0; a < 3; a++) pc[a] = 0.5 * (nb.mn[a] + nb.mx[a]);
        int slot_of[8];
        bool slot_used[8] = {false, false, false, false, false, false, false, false};
        std::vector<bool> done(fin.size(), false);
        for (size_t it = 0; it < fin.size(); it++) {
            double bc = -INFINITY;
            int bi = -1, bs = -1;
            for (size_t i = 0; i < fin.size(); i++) {
                if (done[i]) continue;
                double cc[3];
                for (int a = 0; a < 3; a++) cc[a] = 0.5 * (fin[i].box.mn[a] + fin[i].box.mx[a]) - pc[a];
                for (int s = 0; s < 8; s++) {
                    if (slot_used[s]) continue;
                    double cost = 0;
                    for (int a = 0; a < 3; a++) cost += ((s >> a) & 1) ? cc[a] : -cc[a];
                    if (cost > bc) { bc = cost; bi = (int)i; bs = s; }
                }
            }
            done[bi] = true;
            slot_used[bs] = true;
            slot_of[bi] = bs;
        }
        int child_at_slot[8];
        for (int s = 0; s < 8; s++) child_at_slot[s] = -1;
        for (size_t i = 0; i < fin.size(); i++) child_at_slot[slot_of[i]] = (int)i;

        bc_node8 nd;
        std::memset(&nd, 0, sizeof nd);
        // quantisation grid (conservative: padded box, origin on the grid below it, power-of-two step)
        double pad[3], org[3], scale[3];
        for (int a = 0; a < 3; a++) {
            int32_t k;
            int e;
            if (!bc_grid_axis(nb.mn[a], nb.mx[a], pad[a], k, e, org[a], scale[a])) {
                err = "scene extent too large for quantised BVH8";
                return false;
            }
            bc_node_set_grid(nd, a, k, e);
        }
        int ninternal = 0;
        for (auto& c : fin) ninternal += c.leaf ? 0 : 1;
        const uint32_t child_base = (uint32_t)out.nodes.size();
        if ((uint64_t)child_base + (uint64_t)ninternal > (uint64_t)BC_NODE_MAX) {
            err = "scene needs more than 2^24 BVH8 nodes";
            return false;
        }
        out.nodes.resize(out.nodes.size() + (size_t)ninternal);
        uint32_t imask = 0, lmask = 0;
        bool spheres = false;
        int rank = 0;
        for (int s = 0; s < 8; s++) {
            int i = child_at_slot[s];
            if (i < 0) {
                nd.qlox[s] = nd.qloy[s] = nd.qloz[s] = 255;
                nd.qhix[s] = nd.qhiy[s] = nd.qhiz[s] = 0;
                continue;
            }
            const Child& c = fin[i];
            uint8_t* qlo[3] = {nd.qlox, nd.qloy, nd.qloz};
            uint8_t* qhi[3] = {nd.qhix, nd.qhiy, nd.qhiz};
            for (int a = 0; a < 3; a++) {
                double lo = std::floor(((c.box.mn[a] - pad[a]) - org[a]) / scale[a]);
                double hi = std::ceil(((c.box.mx[a] + pad[a]) - org[a]) / scale[a]);
                qlo[a][s] = (uint8_t)std::min(255.0, std::max(0.0, lo));
                qhi[a][s] = (uint8_t)std::min(255.0, std::max(0.0, hi));
            }
            if (c.leaf) {
                if (c.count != 1) { err = "internal: a BVH8 leaf slot holds more than one primitive"; return false; }
                const int id = v.order[c.start];
                write_prim(v, id, BC_REC_PER_NODE * (size_t)w.i8 + (size_t)s, out);
                written++;
                lmask |= 1u << s;
                spheres = spheres || v.prim_kind[id] == 0;
            } else {
                imask |= 1u << s;
                work.push_back({c.node2, (int)child_base + rank, w.depth + 1});
                rank++;
            }
        }
        bc_node_set_link(nd, child_base, imask);
        bc_node_set_flags(nd, lmask, spheres);
        out.nodes[(size_t)w.i8] = nd;
        if (out.node_depth.size() < out.nodes.size()) out.node_depth.resize(out.nodes.size(), 0);
        out.node_depth[(size_t)w.i8] = w.depth;
    }
    if (written != v.nprim) {
        err = "BVH leaves do not cover every primitive exactly once";
        return false;
    }
    const size_t nrec = BC_REC_PER_NODE * out.nodes.size();
    out.prims.resize(nrec, bc_prim{});
    out.ids.resize(nrec, BC_REC_HOLE);
    out.mats.resize(nrec, -1);
    return true;
}

}  // namespace bcx
