// refit.hip -- device refit of the resident BVH8 after primitives moved (same topology).
//
// Replaces the geometry-only update of the reference backend, which re-packs every node and
// primitive on the host and re-uploads them (updateGeometryLocked, gpu_opencl_windows.go:447-468,
// after Scene.Refit, renderer.go:451-473).  Here only the moved primitive records are uploaded;
// node boxes are recomputed bottom-up on the device, one launch per tree level (deepest first),
// one thread per node: the union of the node's child boxes (leaf slots: the box of their fp32
// primitive record; internal children: the union box the child's own thread stored), then the
// node's quantisation grid and child planes with the conservative rule of the host flattening
// (bc_grid_axis: outward padding, origin on the grid below the box, power-of-two step).
// Topology, slot layout and record slots are unchanged, so traversal results stay exact; only
// culling tightness can change.
#include <hip/hip_runtime.h>

#include <cmath>

#include "bc_layout.h"

namespace bcd {

struct DBox {
    double mn[3], mx[3];
};

__device__ __forceinline__ void dbox_empty(DBox& b) {
    for (int a = 0; a < 3; a++) { b.mn[a] = INFINITY; b.mx[a] = -INFINITY; }
}
__device__ __forceinline__ void dbox_grow(DBox& b, double x, double y, double z) {
    const double p[3] = {x, y, z};
    for (int a = 0; a < 3; a++) { b.mn[a] = fmin(b.mn[a], p[a]); b.mx[a] = fmax(b.mx[a], p[a]); }
}

// Box of one child slot s of node idx (slot s is not empty).
__device__ __forceinline__ DBox slot_box(const bc_node8& nd, uint32_t idx, int s, const bc_prim* __restrict__ prims,
                                         const uint32_t* __restrict__ ids, const float4* __restrict__ nbox) {
    DBox b;
    dbox_empty(b);
    const uint32_t imask = bc_node_imask(nd);
    if ((imask >> s) & 1u) {
        const uint32_t child = bc_node_child_base(nd) + (uint32_t)__popc(imask & ((1u << s) - 1u));
        const float4 lo = nbox[2 * (size_t)child], hi = nbox[2 * (size_t)child + 1];
        b.mn[0] = lo.x; b.mn[1] = lo.y; b.mn[2] = lo.z;
        b.mx[0] = hi.x; b.mx[1] = hi.y; b.mx[2] = hi.z;
        return b;
    }
    const size_t r = (size_t)idx * BC_REC_PER_NODE + (size_t)s;
    const float* v = prims[r].v;
    if (ids[r] & BC_SPHERE_BIT) {
        const double rad = v[3];
        dbox_grow(b, (double)v[0] - rad, (double)v[1] - rad, (double)v[2] - rad);
        dbox_grow(b, (double)v[0] + rad, (double)v[1] + rad, (double)v[2] + rad);
    } else {  // the fp32 triangle A, A+E1, A+E2 (the sums are exact in fp64)
        const double ax = v[0], ay = v[1], az = v[2];
        dbox_grow(b, ax, ay, az);
        dbox_grow(b, ax + (double)v[3], ay + (double)v[4], az + (double)v[5]);
        dbox_grow(b, ax + (double)v[6], ay + (double)v[7], az + (double)v[8]);
    }
    return b;
}

__device__ __forceinline__ float f_down(double v) {
    float f = (float)v;
    if ((double)f > v) f = nextafterf(f, -INFINITY);
    return f;
}
__device__ __forceinline__ float f_up(double v) {
    float f = (float)v;
    if ((double)f < v) f = nextafterf(f, INFINITY);
    return f;
}

__global__ __launch_bounds__(256) void k_refit_level(bc_node8* __restrict__ nodes, const bc_prim* __restrict__ prims,
                                                     const uint32_t* __restrict__ ids, float4* __restrict__ nbox,
                                                     const uint32_t* __restrict__ list, uint32_t n, int* __restrict__ err) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t idx = list[i];
    bc_node8 nd = nodes[idx];
    const uint32_t used = bc_node_imask(nd) | bc_node_lmask(nd);
    DBox nb;
    dbox_empty(nb);
    for (int s = 0; s < 8; s++) {
        if (!((used >> s) & 1u)) continue;
        const DBox c = slot_box(nd, idx, s, prims, ids, nbox);
        dbox_grow(nb, c.mn[0], c.mn[1], c.mn[2]);
        dbox_grow(nb, c.mx[0], c.mx[1], c.mx[2]);
    }
    double pad[3], org[3], scale[3];
    for (int a = 0; a < 3; a++) {
        int32_t k;
        int e;
        if (!bc_grid_axis(nb.mn[a], nb.mx[a], pad[a], k, e, org[a], scale[a])) atomicOr(err, 4);
        bc_node_set_grid(nd, a, k, e);
    }
    uint8_t* qlo[3] = {nd.qlox, nd.qloy, nd.qloz};
    uint8_t* qhi[3] = {nd.qhix, nd.qhiy, nd.qhiz};
    for (int s = 0; s < 8; s++) {
        if (!((used >> s) & 1u)) continue;
        const DBox c = slot_box(nd, idx, s, prims, ids, nbox);
        for (int a = 0; a < 3; a++) {
            const double lo = floor(((c.mn[a] - pad[a]) - org[a]) / scale[a]);
            const double hi = ceil(((c.mx[a] + pad[a]) - org[a]) / scale[a]);
            qlo[a][s] = (uint8_t)fmin(255.0, fmax(0.0, lo));
            qhi[a][s] = (uint8_t)fmin(255.0, fmax(0.0, hi));
        }
    }
    nodes[idx] = nd;
    nbox[2 * (size_t)idx] = make_float4(f_down(nb.mn[0]), f_down(nb.mn[1]), f_down(nb.mn[2]), 0.0f);
    nbox[2 * (size_t)idx + 1] = make_float4(f_up(nb.mx[0]), f_up(nb.mx[1]), f_up(nb.mx[2]), 0.0f);
}

// Records of the moved primitives at their record slots (prim_record of the host flattening) and
// their materials, over all n record slots (holes skipped); a primitive whose kind differs from
// the resident record's sets error bit 16.
__global__ __launch_bounds__(256) void k_refit_records(const double* __restrict__ geom, const int32_t* __restrict__ kind,
                                                       const int32_t* __restrict__ mat, const uint32_t* __restrict__ ids,
                                                       uint32_t n, bc_prim* __restrict__ prims, int32_t* __restrict__ pmat,
                                                       int* __restrict__ err) {
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    if (r >= n || ids[r] == BC_REC_HOLE) return;
    const uint32_t id = ids[r] & ~BC_SPHERE_BIT;
    const bool sphere = (ids[r] & BC_SPHERE_BIT) != 0;
    if ((kind[id] == 0) != sphere) { atomicOr(err, 16); return; }
    const double* g = geom + 10 * (size_t)id;
    bc_prim p;
    for (int q = 0; q < 9; q++) p.v[q] = 0.f;
    if (sphere) {
        const double rad = g[3];
        p.v[0] = (float)g[0]; p.v[1] = (float)g[1]; p.v[2] = (float)g[2];
        p.v[3] = (float)rad; p.v[4] = (float)(rad * rad); p.v[5] = (float)(1.0 / rad);
    } else {
        for (int q = 0; q < 9; q++) p.v[q] = (float)g[q];
    }
    prims[r] = p;
    pmat[r] = mat[id];
}

hipError_t launch_refit_records(const double* geom, const int32_t* kind, const int32_t* mat, const uint32_t* ids, uint32_t n,
                                bc_prim* prims, int32_t* pmat, int* err, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_refit_records, dim3((n + 255u) / 256u), dim3(256), 0, st, geom, kind, mat, ids, n, prims, pmat, err);
    return hipGetLastError();
}

hipError_t launch_refit_level(bc_node8* nodes, const bc_prim* prims, const uint32_t* ids, float4* nbox,
                              const uint32_t* list, uint32_t n, int* err, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_refit_level, dim3((n + 255u) / 256u), dim3(256), 0, st, nodes, prims, ids, nbox, list, n, err);
    return hipGetLastError();
}

}  // namespace bcd
