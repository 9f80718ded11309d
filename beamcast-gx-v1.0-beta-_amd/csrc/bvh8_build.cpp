// bvh8_build.cpp -- flattens the host-built binary SAH BVH into quantised 8-wide nodes.
//
// Input is the Go Scene's BVH exactly as Scene.Build leaves it (renderer.go:409-633:
// Nodes in DFS pre-order, leaves as [Start, Start+Count) ranges of Order).  The binary
// tree is collapsed top-down: a node's child set starts as its two children and the
// internal child with the largest surface area is repeatedly replaced by its own two
// children while the set still fits 8 slots and 32 primitives.  Binary leaves are kept
// whole (split into runs of at most 8 primitives when larger), so every primitive the
// reference would test in a leaf is tested here too; only culling boxes change.
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

float f_down(double v) {
    float f = (float)v;
    if ((double)f > v) f = std::nextafter(f, -INFINITY);
    return f;
}

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

Child make_child(const bc_scene_view& v, int k) {
    Child c;
    const int32_t* l = v.node_link + 4 * (size_t)k;
    c.box = node_box(v, k);
    c.node2 = k;
    if (l[3] > 0) { c.leaf = true; c.start = l[2]; c.count = l[3]; }
    else { c.leaf = false; c.start = 0; c.count = 0; }
    return c;
}

int slots_of(const Child& c) { return c.leaf ? (c.count + 7) / 8 : 1; }

void write_prim(const bc_scene_view& v, int id, Bvh8& out) {
    bc_prim r;
    std::memset(&r, 0, sizeof r);
    const double* g = v.prim_geom + 10 * (size_t)id;
    uint32_t idbits = (uint32_t)id;
    if (v.prim_kind[id] == 0) {
        const double rad = g[3];
        r.v[0] = (float)g[0]; r.v[1] = (float)g[1]; r.v[2] = (float)g[2];
        r.v[3] = (float)rad; r.v[4] = (float)(rad * rad); r.v[5] = (float)(1.0 / rad);
        idbits |= BC_SPHERE_BIT;
    } else {
        for (int k = 0; k < 9; k++) r.v[k] = (float)g[k];
    }
    out.prims.push_back(r);
    out.ids.push_back(idbits);
    out.mats.push_back(v.prim_mat[id]);
}

}  // namespace

bool build_bvh8(const bc_scene_view& v, Bvh8& out, std::string& err) {
    out.nodes.clear();
    out.prims.clear();
    out.ids.clear();
    out.mats.clear();
    out.depth = 0;
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
    out.prims.reserve((size_t)v.nprim);
    out.ids.reserve((size_t)v.nprim);
    out.mats.reserve((size_t)v.nprim);
    out.nodes.reserve((size_t)v.nnode / 3 + 1);
    out.nodes.emplace_back();
    struct Work { int node2, i8, depth; };
    std::vector<Work> work;
    work.push_back({0, 0, 1});
    while (!work.empty()) {
        Work w = work.back();
        work.pop_back();
        out.depth = std::max(out.depth, w.depth);
        std::vector<Child> ch;
        const int32_t* l0 = v.node_link + 4 * (size_t)w.node2;
        if (l0[3] > 0) {
            ch.push_back(make_child(v, w.node2));
        } else {
            ch.push_back(make_child(v, l0[0]));
            ch.push_back(make_child(v, l0[1]));
        }
        for (;;) {  // greedy expansion by surface area
            int slots = 0, prims = 0;
            for (auto& c : ch) { slots += slots_of(c); prims += c.leaf ? c.count : 0; }
            int best = -1;
            double best_area = -1;
            for (size_t i = 0; i < ch.size(); i++) {
                if (ch[i].leaf) continue;
                const int32_t* l = v.node_link + 4 * (size_t)ch[i].node2;
                Child a = make_child(v, l[0]), b = make_child(v, l[1]);
                int ns = slots - 1 + slots_of(a) + slots_of(b);
                int np = prims + (a.leaf ? a.count : 0) + (b.leaf ? b.count : 0);
                if (ns > 8 || np > 32) continue;
                double ar = ch[i].box.area();
                if (ar > best_area) { best_area = ar; best = (int)i; }
            }
            if (best < 0) break;
            const int32_t* l = v.node_link + 4 * (size_t)ch[best].node2;
            Child a = make_child(v, l[0]), b = make_child(v, l[1]);
            ch[best] = a;
            ch.push_back(b);
        }
        // split leaves larger than 8 primitives into runs with their own boxes
        std::vector<Child> fin;
        for (auto& c : ch) {
            if (!c.leaf || c.count <= 8) { fin.push_back(c); continue; }
            for (int s = 0; s < c.count; s += 8) {
                Child p;
                p.leaf = true; p.node2 = -1; p.start = c.start + s; p.count = std::min(8, c.count - s);
                p.box.empty();
                for (int k = 0; k < p.count; k++) p.box.grow(prim_box(v, v.order[p.start + k]));
                fin.push_back(p);
            }
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
        // quantisation grid (conservative: padded box, origin rounded down, power-of-two step)
        double pad[3], org[3], scale[3];
        for (int a = 0; a < 3; a++) {
            double mag = std::max(std::fabs(nb.mn[a]), std::fabs(nb.mx[a]));
            double ext = nb.mx[a] - nb.mn[a];
            pad[a] = std::max((mag + ext) * 0x1p-20, 1e-20);
            float o = f_down(nb.mn[a] - pad[a]);
            double ext2 = (nb.mx[a] + pad[a]) - (double)o;
            int e = (int)std::ceil(std::log2(ext2 / 255.0));
            while (std::ldexp(255.0, e) < ext2) e++;
            if (e < -126) e = -126;
            if (e > 127) { err = "scene extent too large for quantised BVH8"; return false; }
            org[a] = o;
            scale[a] = std::ldexp(1.0, e);
            (&nd.ox)[a] = o;
            (&nd.ex)[a] = (uint8_t)(e + 127);
        }
        int ninternal = 0;
        for (auto& c : fin) ninternal += c.leaf ? 0 : 1;
        uint32_t child_base = (uint32_t)out.nodes.size();
        nd.child_base = child_base;
        nd.prim_base = (uint32_t)out.prims.size();
        out.nodes.resize(out.nodes.size() + (size_t)ninternal);
        int rank = 0, poff = 0;
        for (int s = 0; s < 8; s++) {
            int i = child_at_slot[s];
            if (i < 0) {
                nd.meta[s] = BC_META_EMPTY;
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
                nd.meta[s] = (uint8_t)(((c.count - 1) << 5) | poff);
                for (int k = 0; k < c.count; k++) {
                    const int id = v.order[c.start + k];
                    write_prim(v, id, out);
                    if (v.prim_kind[id] == 0) nd.prim_base |= BC_NODE_HAS_SPHERES;
                }
                poff += c.count;
            } else {
                nd.imask |= (uint8_t)(1u << s);
                nd.meta[s] = 0;
                work.push_back({c.node2, (int)child_base + rank, w.depth + 1});
                rank++;
            }
        }
        out.nodes[(size_t)w.i8] = nd;
    }
    if ((int64_t)out.prims.size() != v.nprim) {
        err = "BVH leaves do not cover every primitive exactly once";
        return false;
    }
    return true;
}

}  // namespace bcx
