// flatten.hip -- the BVH2 -> quantised BVH8 flattening on the device, byte for byte the host's.
//
// Same algorithm and arithmetic as the host flattening (bvh8_build.cpp): the surface-area cost
// plans of every binary node for 1..8 slots (Ylitie, Karras and Laine, HPG 2017), their top-down
// expansion into wide nodes (one primitive per leaf slot), the octant slot
// assignment and the conservative 8-bit quantisation.  The host emits wide nodes from a LIFO
// work stack; here the wide tree is expanded level by level and the stack's numbering is
// recovered from subtree sizes: the stack processes a node, allocates its internal children
// contiguously, then processes the children's subtrees from the highest slot down, so
//     child_base(C) = child_base(W) + nint(W) + sum over internal siblings C' of C at higher
//                     slots of (wide nodes in subtree(C') - 1),
// with the root at index 0, child_base 1.  Records sit at 8 * node + slot.  Nodes, records, ids
// and materials are therefore identical to the host's and every traversal result is unchanged.
//
// Passes: binary levels (top-down), plans (bottom-up, one launch per level), wide levels
// (top-down: children, slots, quantised node), subtree sums (bottom-up), numbering (top-down),
// emission (nodes, records at their record slots, refit level lists).  The view must be
// prepared (bcx::prepare_view: binary leaves of at most eight primitives).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "api_internal.hpp"
#include "bvh8_flatten.hpp"

namespace bcd {
namespace {

constexpr int kLeafMax = 1;  // primitives per leaf slot
constexpr double kNodeCost = 1.0;
constexpr double kPrimCost = BC_PRIM_COST;
constexpr int kT = 256;

struct Plan {
    float cost[8];
    uint8_t how[8];
    int32_t start, count;
};

struct DBox {
    double mn[3], mx[3];
};

// std::min / std::max as the host flattening uses them
__device__ __forceinline__ double smin(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return a < b ? b : a; }

__device__ __forceinline__ void box_empty(DBox& b) {
    for (int a = 0; a < 3; a++) { b.mn[a] = INFINITY; b.mx[a] = -INFINITY; }
}
__device__ __forceinline__ void box_grow(DBox& b, const DBox& c) {
    for (int a = 0; a < 3; a++) { b.mn[a] = smin(b.mn[a], c.mn[a]); b.mx[a] = smax(b.mx[a], c.mx[a]); }
}
__device__ __forceinline__ double box_area(const DBox& b) {
    const double e0 = b.mx[0] - b.mn[0], e1 = b.mx[1] - b.mn[1], e2 = b.mx[2] - b.mn[2];
    if (e0 < 0 || e1 < 0 || e2 < 0) return 0;
    return 2 * (e0 * e1 + e1 * e2 + e2 * e0);
}

struct Inputs {
    const double* node_box;   // nnode x 6
    const int32_t* link;      // nnode x 4
    const int32_t* order;     // nprim
    const int32_t* kind;      // nprim
    const int32_t* mat;       // nprim
    const double* geom;       // nprim x 10
};

__device__ __forceinline__ DBox node_box(const Inputs& in, int i) {
    DBox b;
    for (int a = 0; a < 3; a++) {
        b.mn[a] = in.node_box[6 * (size_t)i + a];
        b.mx[a] = in.node_box[6 * (size_t)i + 3 + a];
    }
    return b;
}

__device__ __forceinline__ DBox prim_box(const Inputs& in, int p) {
    const double* g = in.geom + 10 * (size_t)p;
    DBox b;
    if (in.kind[p] == 0) {
        const double r = g[3];
        for (int a = 0; a < 3; a++) { b.mn[a] = g[a] - r; b.mx[a] = g[a] + r; }
    } else {
        for (int a = 0; a < 3; a++) {
            const double A = g[a], B = g[a] + g[3 + a], C = g[a] + g[6 + a];
            b.mn[a] = smin(A, smin(B, C)) - 1e-6;
            b.mx[a] = smax(A, smax(B, C)) + 1e-6;
        }
    }
    return b;
}

__device__ __forceinline__ float f_down(double v) {
    float f = (float)v;
    if ((double)f > v) f = nextafterf(f, -INFINITY);
    return f;
}

// ---------------------------------------------------------------------------------------------
// binary levels and plans

__global__ void k_bin_level(const int32_t* __restrict__ link, const int* __restrict__ cur, int n, int* __restrict__ next,
                            int* __restrict__ count, long long room, int* __restrict__ err) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const int32_t* l = link + 4 * (size_t)cur[i];
    if (l[3] > 0) return;
    const int k = atomicAdd(count, 2);
    if ((long long)k + 2 > room) { atomicOr(err, 8); return; }  // a node reached twice: not a tree
    next[k] = l[0];
    next[k + 1] = l[1];
}

__global__ void k_plans(Inputs in, const int* __restrict__ level, int n, Plan* __restrict__ plan) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const int nd = level[i];
    const float inf = INFINITY;
    Plan P;
    const int32_t* l = in.link + 4 * (size_t)nd;
    const double A = box_area(node_box(in, nd));
    if (l[3] > 0) {
        P.start = l[2];
        P.count = l[3];
        const int need = l[3];  // one slot per primitive (leaves of at most eight: prepare_view)
        const float c = (float)(A * kPrimCost * l[3]);
        const float c_wide = need > 4 ? c + (float)(A * kNodeCost) : inf;
        for (int s = 1; s <= 8; s++) {
            P.cost[s - 1] = s >= need ? c : c_wide;
            P.how[s - 1] = s >= need ? 1 : 2;
        }
        plan[nd] = P;
        return;
    }
    const Plan L = plan[l[0]], R = plan[l[1]];
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
    for (int s = 2; s <= 8; s++) {
        if (dist[s] < P.cost[s - 2]) {
            P.cost[s - 1] = dist[s];
            P.how[s - 1] = (uint8_t)(2 + dk[s]);
        } else {
            P.cost[s - 1] = P.cost[s - 2];
            P.how[s - 1] = 0;
        }
    }
    plan[nd] = P;
}

// ---------------------------------------------------------------------------------------------
// wide levels

struct WNode {
    bc_node8 q;            // quantised node; child_base filled by the numbering
    int32_t node2, depth;
    int32_t child[8];      // internal slot: temporary id of the child wide node; leaf slot: Order index
    int32_t nint;
    int32_t wsize;         // wide nodes in the subtree
    int32_t idx, alloc;
};

struct Child {
    DBox box;
    int32_t node2, start, count;  // internal: node2 >= 0; leaf: node2 = -1
};

// expand() of the host flattening, iteratively (left subtree first)
__device__ int expand(const Inputs& in, const Plan* __restrict__ plan, int n0, int slots0, Child* out, int nout, int* err) {
    int st_n[16], st_s[16], top = 0;
    st_n[0] = n0;
    st_s[0] = slots0;
    top = 1;
    while (top > 0) {
        top--;
        const int n = st_n[top], slots = st_s[top];
        const Plan P = plan[n];
        int i = slots;
        while (i > 1 && P.how[i - 1] == 0) i--;
        const int h = P.how[i - 1];
        const int32_t* l = in.link + 4 * (size_t)n;
        if (h == 1) {
            if (P.count <= kLeafMax) {
                if (nout >= 8) { atomicOr(err, 2); return nout; }
                Child& c = out[nout++];
                c.node2 = -1; c.start = P.start; c.count = P.count;
                c.box = node_box(in, n);
            } else {
                for (int s = 0; s < P.count; s++) {
                    if (nout >= 8) { atomicOr(err, 2); return nout; }
                    Child& c = out[nout++];
                    c.node2 = -1; c.start = P.start + s; c.count = 1;
                    box_empty(c.box);
                    box_grow(c.box, prim_box(in, in.order[c.start]));
                }
            }
        } else if (h == 2) {
            if (nout >= 8) { atomicOr(err, 2); return nout; }
            Child& c = out[nout++];
            c.box = node_box(in, n);
            c.node2 = n;
            c.start = 0;
            c.count = 0;
        } else {
            if (top + 2 > 16) { atomicOr(err, 2); return nout; }
            const int k = h - 2;
            st_n[top] = l[1]; st_s[top] = i - k; top++;  // right pushed first: left is expanded first
            st_n[top] = l[0]; st_s[top] = k; top++;
        }
    }
    return nout;
}

// Smallest e with 2^e >= x (x > 0), exactly.
__device__ __forceinline__ int ceil_log2(double x) {
    int k;
    const double m = frexp(x, &k);  // x = m 2^k, m in [0.5, 1)
    return m == 0.5 ? k - 1 : k;
}

__global__ void k_wide(Inputs in, const Plan* __restrict__ plan, WNode* __restrict__ w, int first, int n,
                       int* __restrict__ counter, int cap, int* __restrict__ err) {
    const int t = blockIdx.x * kT + threadIdx.x;
    if (t >= n) return;
    WNode& W = w[first + t];
    Child fin[8];
    int nf = 0;
    const int32_t* l0 = in.link + 4 * (size_t)W.node2;
    if (l0[3] > 0) {
        nf = expand(in, plan, W.node2, 8, fin, 0, err);
    } else {
        const Plan L = plan[l0[0]], R = plan[l0[1]];
        float best = INFINITY;
        int bk = 1;
        for (int k = 1; k < 8; k++) {
            const float c = L.cost[k - 1] + R.cost[8 - k - 1];
            if (c < best) { best = c; bk = k; }
        }
        nf = expand(in, plan, l0[0], bk, fin, 0, err);
        nf = expand(in, plan, l0[1], 8 - bk, fin, nf, err);
    }
    DBox nb;
    box_empty(nb);
    for (int i = 0; i < nf; i++) box_grow(nb, fin[i].box);
    // octant slot assignment
    double pc[3];
    for (int a = 0; a < 3; a++) pc[a] = 0.5 * (nb.mn[a] + nb.mx[a]);
    int slot_of[8];
    bool slot_used[8] = {false, false, false, false, false, false, false, false};
    bool done[8] = {false, false, false, false, false, false, false, false};
    for (int it = 0; it < nf; it++) {
        double bc = -INFINITY;
        int bi = -1, bs = -1;
        for (int i = 0; i < nf; i++) {
            if (done[i]) continue;
            double cc[3];
            for (int a = 0; a < 3; a++) cc[a] = 0.5 * (fin[i].box.mn[a] + fin[i].box.mx[a]) - pc[a];
            for (int s = 0; s < 8; s++) {
                if (slot_used[s]) continue;
                double cost = 0;
                for (int a = 0; a < 3; a++) cost += ((s >> a) & 1) ? cc[a] : -cc[a];
                if (cost > bc) { bc = cost; bi = i; bs = s; }
            }
        }
        if (bi < 0) { atomicOr(err, 2); return; }
        done[bi] = true;
        slot_used[bs] = true;
        slot_of[bi] = bs;
    }
    int at[8];
    for (int s = 0; s < 8; s++) at[s] = -1;
    for (int i = 0; i < nf; i++) at[slot_of[i]] = i;
    bc_node8 nd;
    memset(&nd, 0, sizeof nd);
    double pad[3], org[3], scale[3];
    for (int a = 0; a < 3; a++) {
        int32_t k = 0;
        int e = 127;
        if (!bc_grid_axis(nb.mn[a], nb.mx[a], pad[a], k, e, org[a], scale[a])) atomicOr(err, 4);
        bc_node_set_grid(nd, a, k, e);
    }
    uint8_t* qlo[3] = {nd.qlox, nd.qloy, nd.qloz};
    uint8_t* qhi[3] = {nd.qhix, nd.qhiy, nd.qhiz};
    uint32_t imask = 0, lmask = 0;
    int nint = 0;
    bool sph = false;
    for (int s = 0; s < 8; s++) {
        W.child[s] = -1;
        const int i = at[s];
        if (i < 0) {
            qlo[0][s] = qlo[1][s] = qlo[2][s] = 255;
            qhi[0][s] = qhi[1][s] = qhi[2][s] = 0;
            continue;
        }
        const Child& c = fin[i];
        for (int a = 0; a < 3; a++) {
            const double lo = floor(((c.box.mn[a] - pad[a]) - org[a]) / scale[a]);
            const double hi = ceil(((c.box.mx[a] + pad[a]) - org[a]) / scale[a]);
            qlo[a][s] = (uint8_t)fmin(255.0, fmax(0.0, lo));
            qhi[a][s] = (uint8_t)fmin(255.0, fmax(0.0, hi));
        }
        if (c.node2 < 0) {
            if (c.count != 1) atomicOr(err, 2);
            lmask |= 1u << s;
            W.child[s] = c.start;
            sph = sph || in.kind[in.order[c.start]] == 0;
        } else {
            imask |= 1u << s;
            const int cid = atomicAdd(counter, 1);
            if (cid >= cap) { atomicOr(err, 8); continue; }
            WNode& C = w[cid];
            C.node2 = c.node2;
            C.depth = W.depth + 1;
            W.child[s] = cid;
            nint++;
        }
    }
    bc_node_set_link(nd, 0u, imask);
    bc_node_set_flags(nd, lmask, sph);
    W.q = nd;
    W.nint = nint;
}

__global__ void k_wsums(WNode* __restrict__ w, int first, int n) {
    const int t = blockIdx.x * kT + threadIdx.x;
    if (t >= n) return;
    WNode& W = w[first + t];
    const uint32_t imask = bc_node_imask(W.q);
    int ws = 1;
    for (int s = 0; s < 8; s++)
        if ((imask >> s) & 1) ws += w[W.child[s]].wsize;
    W.wsize = ws;
}

__global__ void k_wnumber(WNode* __restrict__ w, int first, int n) {
    const int t = blockIdx.x * kT + threadIdx.x;
    if (t >= n) return;
    const WNode& W = w[first + t];
    const uint32_t imask = bc_node_imask(W.q);
    int rank = 0;
    for (int s = 0; s < 8; s++)
        if ((imask >> s) & 1) w[W.child[s]].idx = W.alloc + rank++;
    int alloc = W.alloc + W.nint;
    for (int s = 7; s >= 0; s--) {  // the LIFO stack visits the highest slot first
        if (!((imask >> s) & 1)) continue;
        WNode& C = w[W.child[s]];
        C.alloc = alloc;
        alloc += C.wsize - 1;
    }
}

// Nodes, and the records of their leaf slots at 8 * node + slot (the holes were filled before).
__global__ void k_wemit(Inputs in, const WNode* __restrict__ w, int n, bc_node8* __restrict__ nodes, bc_prim* __restrict__ prims,
                        uint32_t* __restrict__ ids, int32_t* __restrict__ mats, int nnode_out) {
    const int t = blockIdx.x * kT + threadIdx.x;
    if (t >= n) return;
    const WNode& W = w[t];
    if (W.idx < 0 || W.idx >= nnode_out) return;
    bc_node8 nd = W.q;
    bc_node_set_link(nd, (uint32_t)W.alloc, bc_node_imask(nd));
    nodes[W.idx] = nd;
    const uint32_t lmask = bc_node_lmask(nd);
    for (int s = 0; s < 8; s++) {
        if (!((lmask >> s) & 1u)) continue;
        const size_t r = (size_t)W.idx * BC_REC_PER_NODE + (size_t)s;
        const int id = in.order[W.child[s]];
        const double* g = in.geom + 10 * (size_t)id;
        bc_prim p;
        for (int q = 0; q < 9; q++) p.v[q] = 0.f;
        if (in.kind[id] == 0) {
            const double rad = g[3];
            p.v[0] = (float)g[0]; p.v[1] = (float)g[1]; p.v[2] = (float)g[2];
            p.v[3] = (float)rad; p.v[4] = (float)(rad * rad); p.v[5] = (float)(1.0 / rad);
        } else {
            for (int q = 0; q < 9; q++) p.v[q] = (float)g[q];
        }
        prims[r] = p;
        ids[r] = (uint32_t)id | (in.kind[id] == 0 ? BC_SPHERE_BIT : 0u);
        mats[r] = in.mat[id];
    }
}

// refit level lists: level l (depth l + 1) at off[l]
__global__ void k_wlevels(const WNode* __restrict__ w, int first, int n, uint32_t* __restrict__ out) {
    const int t = blockIdx.x * kT + threadIdx.x;
    if (t < n) out[t] = (uint32_t)w[first + t].idx;
}

__global__ void k_wroot(WNode* w) {
    WNode& r = w[0];
    r.node2 = 0;
    r.depth = 1;
    r.idx = 0;
    r.alloc = 1;
}

inline unsigned grid(size_t n) { return (unsigned)((n + kT - 1) / kT); }

struct Buf {
    void* p = nullptr;
    ~Buf() { if (p) (void)hipFree(p); }
    bool alloc(bc_ctx* c, size_t bytes, const char* what) {
        if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) {
            p = nullptr;
            c->err = std::string("out of device memory: ") + what;
            return false;
        }
        return true;
    }
    template <class T> T* as() const { return (T*)p; }
};

}  // namespace

// Flattens the view's binary tree into the context's device scene arrays (d_nodes, d_prims,
// d_ids, d_pmat, d_levels) and fills the host-side facts the upload records.
int flatten_device(bc_ctx* c, const bc_scene_view& v, FlattenResult& res) {
    hipStream_t st = c->stream;
    const size_t nn = (size_t)v.nnode, np = (size_t)v.nprim;
    Buf nbox, link, order, kind, mat, geom, plan, lv, wn, cnt;
    if (!nbox.alloc(c, 48 * nn, "flatten inputs") || !link.alloc(c, 16 * nn, "flatten inputs") ||
        !order.alloc(c, 4 * np, "flatten inputs") || !kind.alloc(c, 4 * np, "flatten inputs") ||
        !mat.alloc(c, 4 * np, "flatten inputs") || !geom.alloc(c, 80 * np, "flatten inputs") ||
        !plan.alloc(c, sizeof(Plan) * nn, "plans") || !lv.alloc(c, 4 * nn, "binary levels") || !cnt.alloc(c, 16, "counters"))
        return BC_ERR_NOMEM;
    if (!bc_h2d(c, nbox.p, v.node_box, 48 * nn, st, "upload nodes") ||
        !bc_h2d(c, link.p, v.node_link, 16 * nn, st, "upload links") ||
        !bc_h2d(c, order.p, v.order, 4 * np, st, "upload order") ||
        !bc_h2d(c, kind.p, v.prim_kind, 4 * np, st, "upload kinds") ||
        !bc_h2d(c, mat.p, v.prim_mat, 4 * np, st, "upload materials") ||
        !bc_h2d(c, geom.p, v.prim_geom, 80 * np, st, "upload geometry") ||
        !bc_check(c, hipMemsetAsync(c->d_err, 0, sizeof(int), st), "clear error"))
        return BC_ERR_DEVICE;
    const Inputs in{nbox.as<double>(), link.as<int32_t>(), order.as<int32_t>(), kind.as<int32_t>(), mat.as<int32_t>(),
                    geom.as<double>()};
    int* count = cnt.as<int>();
    // binary levels, top down
    std::vector<size_t> boff(1, 0);
    {
        const int zero = 0;
        if (!bc_check(c, hipMemcpyAsync(lv.p, &zero, 4, hipMemcpyHostToDevice, st), "levels")) return BC_ERR_DEVICE;
        boff.push_back(1);
        for (;;) {
            const size_t a = boff[boff.size() - 2], b = boff.back();
            if (!bc_check(c, hipMemsetAsync(count, 0, 4, st), "levels")) return BC_ERR_DEVICE;
            hipLaunchKernelGGL(k_bin_level, dim3(grid(b - a)), dim3(kT), 0, st, in.link, lv.as<int>() + a, (int)(b - a),
                               lv.as<int>() + b, count, (long long)(nn - b), c->d_err);
            int h = 0;
            if (!bc_check(c, hipMemcpyAsync(&h, count, 4, hipMemcpyDeviceToHost, st), "levels") ||
                !bc_check(c, hipStreamSynchronize(st), "levels"))
                return BC_ERR_DEVICE;
            if (h == 0) break;
            if (b + (size_t)h > nn || bc_check_kernel_error(c) != BC_OK) {
                c->err = "flatten: the binary nodes do not form a tree";
                return BC_ERR_SCENE;
            }
            boff.push_back(b + (size_t)h);
        }
    }
    for (size_t l = boff.size() - 1; l-- > 0;) {
        const size_t a = boff[l], b = boff[l + 1];
        hipLaunchKernelGGL(k_plans, dim3(grid(b - a)), dim3(kT), 0, st, in, lv.as<int>() + a, (int)(b - a), plan.as<Plan>());
    }
    // wide levels, top down (at most one wide node per binary node)
    const int wcap = (int)nn;
    if (!wn.alloc(c, sizeof(WNode) * (size_t)wcap, "wide nodes")) return BC_ERR_NOMEM;
    hipLaunchKernelGGL(k_wroot, dim3(1), dim3(1), 0, st, wn.as<WNode>());
    std::vector<int> woff(1, 0);
    woff.push_back(1);
    int total = 1;
    {
        if (!bc_check(c, hipMemcpyAsync(count, &total, 4, hipMemcpyHostToDevice, st), "wide levels")) return BC_ERR_DEVICE;
        for (;;) {
            const int a = woff[woff.size() - 2], b = woff.back();
            hipLaunchKernelGGL(k_wide, dim3(grid((size_t)(b - a))), dim3(kT), 0, st, in, plan.as<Plan>(), wn.as<WNode>(), a, b - a,
                               count, wcap, c->d_err);
            if (!bc_check(c, hipMemcpyAsync(&total, count, 4, hipMemcpyDeviceToHost, st), "wide levels") ||
                !bc_check(c, hipStreamSynchronize(st), "wide levels"))
                return BC_ERR_DEVICE;
            const int kerr = bc_check_kernel_error(c);
            if (kerr != BC_OK) { c->err = "flatten: BVH8 node overflow or extent too large"; return BC_ERR_SCENE; }
            if (total > wcap) { c->err = "flatten: wide node capacity exceeded"; return BC_ERR_SCENE; }
            if (total == b) break;
            woff.push_back(total);
        }
    }
    const int nlevels = (int)woff.size() - 1;
    for (int l = nlevels - 1; l >= 0; l--)
        hipLaunchKernelGGL(k_wsums, dim3(grid((size_t)(woff[l + 1] - woff[l]))), dim3(kT), 0, st, wn.as<WNode>(), woff[l],
                           woff[l + 1] - woff[l]);
    for (int l = 0; l < nlevels; l++)
        hipLaunchKernelGGL(k_wnumber, dim3(grid((size_t)(woff[l + 1] - woff[l]))), dim3(kT), 0, st, wn.as<WNode>(), woff[l],
                           woff[l + 1] - woff[l]);
    if ((size_t)total > (size_t)BC_NODE_MAX) {
        c->err = "scene needs more than 2^24 BVH8 nodes";
        return BC_ERR_SCENE;
    }
    // outputs: nodes, and eight record slots per node (holes: zero record, id BC_REC_HOLE, material -1)
    const size_t nrec = BC_REC_PER_NODE * (size_t)total;
    void* p;
    p = c->d_nodes;
    if (!bc_ensure(c, p, c->cap_nodes, sizeof(bc_node8) * (size_t)total, "upload nodes")) return BC_ERR_NOMEM;
    c->d_nodes = (bc_node8*)p;
    p = c->d_prims;
    if (!bc_ensure(c, p, c->cap_prims, sizeof(bc_prim) * nrec, "upload prims")) return BC_ERR_NOMEM;
    c->d_prims = (bc_prim*)p;
    p = c->d_ids;
    if (!bc_ensure(c, p, c->cap_ids, 4 * nrec, "upload prim ids")) return BC_ERR_NOMEM;
    c->d_ids = (uint32_t*)p;
    p = c->d_pmat;
    if (!bc_ensure(c, p, c->cap_pmat, 4 * nrec, "upload prim materials")) return BC_ERR_NOMEM;
    c->d_pmat = (int32_t*)p;
    p = c->d_levels;
    if (!bc_ensure(c, p, c->cap_levels, 4 * (size_t)total, "upload node levels")) return BC_ERR_NOMEM;
    c->d_levels = (uint32_t*)p;
    if (!bc_check(c, hipMemsetAsync(c->d_prims, 0, sizeof(bc_prim) * nrec, st), "record holes") ||
        !bc_check(c, hipMemsetAsync(c->d_ids, 0xFF, 4 * nrec, st), "record holes") ||
        !bc_check(c, hipMemsetAsync(c->d_pmat, 0xFF, 4 * nrec, st), "record holes"))
        return BC_ERR_DEVICE;
    hipLaunchKernelGGL(k_wemit, dim3(grid((size_t)total)), dim3(kT), 0, st, in, wn.as<WNode>(), total, c->d_nodes, c->d_prims,
                       c->d_ids, c->d_pmat, total);
    // refit lists: deepest level first
    res.level_off.assign(1, 0u);
    uint32_t off = 0;
    for (int l = nlevels - 1; l >= 0; l--) {
        const int n = woff[l + 1] - woff[l];
        hipLaunchKernelGGL(k_wlevels, dim3(grid((size_t)n)), dim3(kT), 0, st, wn.as<WNode>(), woff[l], n, c->d_levels + off);
        off += (uint32_t)n;
        res.level_off.push_back(off);
    }
    if (!bc_check(c, hipGetLastError(), "flatten launch")) return BC_ERR_DEVICE;
    // the root's subtree must be every wide node, and the leaf slots every primitive exactly once
    WNode root;
    if (!bc_check(c, hipMemcpyAsync(&root, wn.p, sizeof root, hipMemcpyDeviceToHost, st), "flatten") ||
        !bc_check(c, hipStreamSynchronize(st), "flatten"))
        return BC_ERR_DEVICE;
    res.ids.resize(nrec);
    if (root.wsize != total || !bc_d2h(c, res.ids.data(), c->d_ids, 4 * nrec, st, "read ids")) {
        if (root.wsize != total) c->err = "flatten: the wide tree does not cover its nodes";
        return root.wsize != total ? BC_ERR_SCENE : BC_ERR_DEVICE;
    }
    std::vector<char> seen(np, 0);
    size_t nseen = 0;
    for (uint32_t id : res.ids) {
        if (id == BC_REC_HOLE) continue;
        const uint32_t q = id & ~BC_SPHERE_BIT;
        if (q >= np || seen[q]) { nseen = np + 1; break; }
        seen[q] = 1;
        nseen++;
    }
    if (nseen != np) {
        c->err = "BVH leaves do not cover every primitive exactly once";
        return BC_ERR_SCENE;
    }
    res.nnode8 = total;
    res.depth = nlevels;
    return BC_OK;
}

}  // namespace bcd
