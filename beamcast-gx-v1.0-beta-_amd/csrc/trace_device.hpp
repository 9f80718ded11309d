// trace_device.hpp -- BVH8 closest-hit / any-hit traversal for CDNA4 (gfx950), wave64.
//
// Replaces closest()/boxhit()/phit() of the reference OpenCL kernel
// (gpu_opencl_windows.go:736-777) and Scene.Hit / Scene.Occluded (renderer.go:777-895).
//
// One ray per lane.  A node visit loads one 64-byte bc_node8 (four 16-byte loads), tests
// its eight quantised child boxes with FMA slab tests and produces
//   * a node group  : (child_base, imask, hit bits in near-to-far slot order)
//   * a prim group  : (record base 8 * node, mask of hit leaf slots, one primitive each)
// Groups that still hold children are pushed to a per-lane stack in LDS laid out
// [depth][lane] so that a wave's pushes and pops are conflict-free 8-byte accesses.
// A group consumes one stack entry per tree level above the deepest, so a BVH8 of depth D
// needs D - 1 entries per lane; kernels size their (dynamic) LDS stack to exactly that, so
// shallow trees leave room for more resident waves.
//
// Exactness: only culling uses approximate arithmetic (quantised, outward-padded boxes,
// hardware reciprocal).  Primitive tests are evaluated with the operation order of
// triHit (renderer.go:684-711) in IEEE single precision, no contraction, correctly
// rounded divide and square root, and equal distances are resolved to the smaller
// primitive index, so results do not depend on traversal order.
#pragma once
#include <hip/hip_runtime.h>

#include "bc_layout.h"

namespace bcd {

struct Ray {
    float ox, oy, oz, dx, dy, dz;
    float tmin;
};

struct HitRec {
    float t;
    int32_t id;    // Scene.Primitives index, -1 = none
    uint32_t rec;  // index of the bc_prim record
};

__device__ __forceinline__ float u8f(uint32_t w, uint32_t sh) { return (float)((w >> sh) & 0xffu); }

// Workgroups are dispatched to the eight XCDs of MI355X round-robin (workgroup b runs on XCD
// b mod 8), and each XCD has its own L2.  This maps workgroup b to a work chunk so that the
// eight XCDs take turns on runs of K consecutive chunks: the workgroups resident on one XCD at
// a time then work on neighbouring rays (pixels, path segments) and share that XCD's L2, while
// every region of the work is still spread over all XCDs (whole contiguous eighths per XCD
// would leave XCDs idle when cost varies across the image).  A bijection on [0, nb).
__device__ __forceinline__ uint32_t xcd_chunk(uint32_t b, uint32_t nb, uint32_t K) {
    const uint32_t full = nb - nb % (8u * K);
    if (b >= full) return b;
    return (b / (8u * K)) * (8u * K) + (b & 7u) * K + (b >> 3) % K;
}

// triHit in fp32 (renderer.go:684-711) on a record g = A.xyz E1.xyz E2.xyz; t through tout.
__device__ __forceinline__ bool tri_test(const float* __restrict__ g, const Ray& r, float tmin, float tmax,
                                         float& tout) {
    const float ax = g[0], ay = g[1], az = g[2];
    const float e1x = g[3], e1y = g[4], e1z = g[5], e2x = g[6], e2y = g[7], e2z = g[8];
    const float hx = r.dy * e2z - r.dz * e2y;
    const float hy = r.dz * e2x - r.dx * e2z;
    const float hz = r.dx * e2y - r.dy * e2x;
    const float det = e1x * hx + e1y * hy + e1z * hz;
    if (fabsf(det) < 1e-10f) return false;
    const float f = 1.0f / det;
    const float sx = r.ox - ax, sy = r.oy - ay, sz = r.oz - az;
    const float u = f * (sx * hx + sy * hy + sz * hz);
    if (u < 0.0f || u > 1.0f) return false;
    const float qx = sy * e1z - sz * e1y;
    const float qy = sz * e1x - sx * e1z;
    const float qz = sx * e1y - sy * e1x;
    const float v = f * (r.dx * qx + r.dy * qy + r.dz * qz);
    if (v < 0.0f || u + v > 1.0f) return false;
    const float t = f * (e2x * qx + e2y * qy + e2z * qz);
    if (t < tmin || t > tmax) return false;
    tout = t;
    return true;
}

// sphereHitAccel (renderer.go:650-678) with the line-distance discriminant of Haines et al.,
// "Precision Improvements for Ray/Sphere Intersection" (Ray Tracing Gems, 2019).
__device__ __forceinline__ bool sphere_test(const float* __restrict__ g, const Ray& r, float tmin, float tmax,
                                            float& tout) {
    const float a = r.dx * r.dx + r.dy * r.dy + r.dz * r.dz;
    if (!(a > 1e-18f)) return false;
    const float inva = 1.0f / a;
    const float ocx = r.ox - g[0], ocy = r.oy - g[1], ocz = r.oz - g[2];
    const float hb = ocx * r.dx + ocy * r.dy + ocz * r.dz;
    const float k = hb * inva;
    const float lx = ocx - r.dx * k, ly = ocy - r.dy * k, lz = ocz - r.dz * k;
    const float disc = a * (g[4] - (lx * lx + ly * ly + lz * lz));
    if (disc < 0.0f) return false;
    const float sq = sqrtf(disc);
    float root = (-hb - sq) * inva;
    if (root < tmin || root > tmax) {
        root = (-hb + sq) * inva;
        if (root < tmin || root > tmax) return false;
    }
    tout = root;
    return true;
}

__device__ __forceinline__ float safe_rcp(float d) {
    const float a = fabsf(d) > 1e-20f ? d : copysignf(1e-20f, d);
    return __builtin_amdgcn_rcpf(a);
}

// Tests the 8 children of node `idx`.  Outputs: the internal-hit bits, the node's imask and
// child_base, and the primitive group: its record base (8 * idx, bit 31 set when the node holds
// spheres) and the mask of hit leaf slots (record = base + slot).  Closest hit orders the internal
// hits near to far (bit k <=> slot k ^ oct); any hit keeps slot order (bit k <=> slot k), since an
// occlusion query gains nothing from the order (incoherent any hit 4.31 -> 4.05 ms without it).
template <bool ANY>
__device__ __forceinline__ void visit_node(const bc_node8* __restrict__ nodes, uint32_t idx, const Ray& r,
                                           float ix, float iy, float iz, uint32_t oct, float tmax,
                                           uint32_t& ng_base, uint32_t& ng_bits, uint32_t& pg_base,
                                           uint32_t& pg_mask) {
    const uint4* p = reinterpret_cast<const uint4*>(nodes + idx);
    const uint4 w0 = p[0], w1 = p[1], w2 = p[2], w3 = p[3];
    const uint32_t imask = w0.x >> 24;
    // per axis: grid step 2^e (the exponent field shifted into place) and origin k * 2^e, exact
    const float sxf = __uint_as_float(((w0.y >> 21) & 0xffu) << 23);
    const float syf = __uint_as_float(((w0.z >> 21) & 0xffu) << 23);
    const float szf = __uint_as_float(((w0.w >> 21) & 0xffu) << 23);
    const float ox = (float)(((int32_t)(w0.y << 11)) >> 11) * sxf;
    const float oy = (float)(((int32_t)(w0.z << 11)) >> 11) * syf;
    const float oz = (float)(((int32_t)(w0.w << 11)) >> 11) * szf;
    const float adjx = sxf * ix, adjy = syf * iy, adjz = szf * iz;
    const float offx = (ox - r.ox) * ix, offy = (oy - r.oy) * iy, offz = (oz - r.oz) * iz;
    // near/far planes per axis by direction sign: words (lo0-3, lo4-7) and (hi0-3, hi4-7)
    const uint32_t lox0 = w1.x, lox1 = w1.y, loy0 = w1.z, loy1 = w1.w, loz0 = w2.x, loz1 = w2.y;
    const uint32_t hix0 = w2.z, hix1 = w2.w, hiy0 = w3.x, hiy1 = w3.y, hiz0 = w3.z, hiz1 = w3.w;
    const bool nx = ix < 0.0f, ny = iy < 0.0f, nz = iz < 0.0f;
    const uint32_t nx0 = nx ? hix0 : lox0, nx1 = nx ? hix1 : lox1, fx0 = nx ? lox0 : hix0, fx1 = nx ? lox1 : hix1;
    const uint32_t ny0 = ny ? hiy0 : loy0, ny1 = ny ? hiy1 : loy1, fy0 = ny ? loy0 : hiy0, fy1 = ny ? loy1 : hiy1;
    const uint32_t nz0 = nz ? hiz0 : loz0, nz1 = nz ? hiz1 : loz1, fz0 = nz ? loz0 : hiz0, fz1 = nz ? loz1 : hiz1;
    uint32_t hits = 0;
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const uint32_t sh = (uint32_t)(s & 3) * 8u;
        const bool hiw = s >= 4;
        const float tnx = __builtin_fmaf(u8f(hiw ? nx1 : nx0, sh), adjx, offx);
        const float tny = __builtin_fmaf(u8f(hiw ? ny1 : ny0, sh), adjy, offy);
        const float tnz = __builtin_fmaf(u8f(hiw ? nz1 : nz0, sh), adjz, offz);
        const float tfx = __builtin_fmaf(u8f(hiw ? fx1 : fx0, sh), adjx, offx);
        const float tfy = __builtin_fmaf(u8f(hiw ? fy1 : fy0, sh), adjy, offy);
        const float tfz = __builtin_fmaf(u8f(hiw ? fz1 : fz0, sh), adjz, offz);
        const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, r.tmin));
        const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
        // an empty slot's box is inverted (255 / 0); rounding can still let it pass here, which
        // is harmless: it is in neither imask nor the leaf mask
        const bool h = tn <= tf;
        hits |= (h ? 1u : 0u) << s;
    }
    uint32_t ord = hits & imask;
    if (!ANY) {
        // near-to-far order, bit k = slot k ^ oct: the slot mask with its bit positions XORed by
        // the octant, one swap stage per octant bit
        ord = (oct & 1u) ? (((ord & 0x55u) << 1) | ((ord >> 1) & 0x55u)) : ord;
        ord = (oct & 2u) ? (((ord & 0x33u) << 2) | ((ord >> 2) & 0x33u)) : ord;
        ord = (oct & 4u) ? (((ord & 0x0fu) << 4) | ((ord >> 4) & 0x0fu)) : ord;
    }
    ng_base = w0.x & 0xffffffu;
    ng_bits = ord | (imask << 8);
    // leaf slots (one primitive each, record 8 * idx + slot): bits 29-31 of the three grid words
    const uint32_t lmask = (w0.y >> 29) | ((w0.z >> 29) << 3) | (((w0.w >> 29) & 3u) << 6);
    pg_base = (idx * BC_REC_PER_NODE) | (w0.w & BC_NODE_HAS_SPHERES);
    pg_mask = hits & lmask;
}

// Primitive records fetched together in a leaf (see the leaf loop of trav_step).
#ifndef BC_PRIM_BATCH
#define BC_PRIM_BATCH 2
#endif
constexpr int kPrimBatch = BC_PRIM_BATCH;

// Traversal state of one lane: the ray, its reciprocal direction and octant, the best hit so
// far, the node group being expanded and the primitive group being tested, and the depth of its
// LDS stack.  trav_step advances it by one node visit, the primitive group that visit produced,
// and the choice of the next node (a push, or pops), so a wave can run it until some lanes
// finish and then give those lanes new rays (see k_trace_refill) while the others continue.
struct TravState {
    Ray r;
    float ix, iy, iz, tmax;
    uint32_t oct;
    HitRec best;
    uint32_t ng_base, ng_bits, pg_base, pg_mask, node;
    int sp;
    bool have_node;
    // the counting instantiation (COUNT = true) only: node visits, record tests and id-table reads
    uint32_t n_node, n_rec, n_id;
};

// root: index of the scene's root node (0 for a single scene; scenes packed in one arena
// each have their own).
__device__ __forceinline__ void trav_init(TravState& st, const Ray& r, float tmax, uint32_t root) {
    st.r = r;
    st.ix = safe_rcp(r.dx);
    st.iy = safe_rcp(r.dy);
    st.iz = safe_rcp(r.dz);
    st.oct = (r.dx < 0.0f ? 1u : 0u) | (r.dy < 0.0f ? 2u : 0u) | (r.dz < 0.0f ? 4u : 0u);
    st.best.t = tmax;
    st.best.id = -1;
    st.best.rec = 0xffffffffu;
    st.ng_base = st.ng_bits = st.pg_base = st.pg_mask = 0;
    st.node = root;
    st.sp = 0;
    st.have_node = true;
    st.n_node = st.n_rec = st.n_id = 0;
}

// One step; returns true when the traversal of this lane's ray is complete (st.best final).
// ANY=true stops at the first primitive hit in [tmin, tmax].  stk points at this lane's column
// of the LDS stack (entry i at stk[i * stride], i < cap).  COUNT=true also tallies the memory
// work of the step in st.n_node / n_rec / n_id (the traversal itself is unchanged).
template <bool ANY, bool COUNT = false>
__device__ __forceinline__ bool trav_step(TravState& st, const bc_node8* __restrict__ nodes, const bc_prim* __restrict__ prims,
                                          const uint32_t* __restrict__ ids, uint2* stk, int stride, int cap,
                                          int* __restrict__ err) {
    const Ray& r = st.r;
    if (st.have_node) {
        visit_node<ANY>(nodes, st.node, r, st.ix, st.iy, st.iz, st.oct, st.best.t, st.ng_base, st.ng_bits, st.pg_base, st.pg_mask);
        st.have_node = false;
        if (COUNT) st.n_node++;
    }
    const bool sph_node = (st.pg_base & BC_NODE_HAS_SPHERES) != 0;
    const uint32_t pbase = st.pg_base & ~BC_NODE_HAS_SPHERES;
    // primitives kPrimBatch at a time: their records are loaded before any is tested, so the
    // memory round trips overlap (the loop is latency-bound); tests stay in mask order
    while (st.pg_mask) {
        uint32_t rr[kPrimBatch];
        int cnt = 0;
#pragma unroll
        for (int u = 0; u < kPrimBatch; u++) {
            const bool any = st.pg_mask != 0u;
            rr[u] = pbase + (any ? (uint32_t)__builtin_ctz(st.pg_mask) : 0u);
            if (!any) rr[u] = rr[0];
            cnt += any ? 1 : 0;
            st.pg_mask &= st.pg_mask - 1u;
        }
        // a lane with fewer records than the batch loads only its own (no repeated fetch of its
        // first record: every lane-load costs texture-data time)
        float g[kPrimBatch][9];
        bool sp[kPrimBatch];
#pragma unroll
        for (int u = 0; u < kPrimBatch; u++) {
            if (u == 0 || u < cnt) {
#pragma unroll
                for (int q = 0; q < 9; q++) g[u][q] = prims[rr[u]].v[q];
                sp[u] = sph_node && (ids[rr[u]] & BC_SPHERE_BIT);
            } else {
#pragma unroll
                for (int q = 0; q < 9; q++) g[u][q] = 0.0f;
                sp[u] = false;
            }
        }
        if (COUNT) {
            st.n_rec += (uint32_t)cnt;
            st.n_id += sph_node ? (uint32_t)cnt : 0u;
        }
#pragma unroll
        for (int u = 0; u < kPrimBatch; u++) {
            if (u >= cnt) break;
            const uint32_t ri = rr[u];
            float t;
            const bool h = sp[u] ? sphere_test(g[u], r, r.tmin, st.best.t, t) : tri_test(g[u], r, r.tmin, st.best.t, t);
            if (h) {
                if (COUNT && (ANY || !(st.best.rec == 0xffffffffu || t < st.best.t))) st.n_id += ANY ? 1u : 2u;
                if (ANY) {
                    st.best.t = t;
                    st.best.id = (int32_t)(ids[ri] & ~BC_SPHERE_BIT);
                    st.best.rec = ri;
                    return true;
                }
                // the test admits t <= best.t; only an exact tie needs the primitive indices
                // (smaller wins), so the id table is read in that rare case only
                if (st.best.rec == 0xffffffffu || t < st.best.t ||
                    (ids[ri] & ~BC_SPHERE_BIT) < (ids[st.best.rec] & ~BC_SPHERE_BIT)) {
                    st.best.t = t;
                    st.best.rec = ri;
                }
            }
        }
    }
    // next node: the next remaining child of the current group, else of the last pending group
    // on the stack, popped and descended into in the same step, so that every step of a lane
    // that still has work visits a node.  A popped group is not tested against the current hit
    // first: its children's boxes are (coherent closest hit 2.58 -> 2.43 ms without the test and
    // the entry distance it needed).
    if (!(st.ng_bits & 0xffu)) {
        if (st.sp == 0) return true;
        st.sp--;
        const uint2 e = stk[st.sp * stride];
        st.ng_base = e.x;
        st.ng_bits = e.y;
    }
    const uint32_t k = (uint32_t)__builtin_ctz(st.ng_bits & 0xffu);
    st.ng_bits &= ~(1u << k);
    const uint32_t slot = ANY ? k : k ^ st.oct;
    const uint32_t imask = (st.ng_bits >> 8) & 0xffu;
    st.node = st.ng_base + (uint32_t)__builtin_popcount(imask & ((1u << slot) - 1u));
    if (st.ng_bits & 0xffu) {
        if (st.sp >= cap) {  // the host sizes cap to the tree depth; report instead of overflowing
            atomicOr(err, 1);
            return true;
        }
        stk[st.sp * stride] = make_uint2(st.ng_base, st.ng_bits);
        st.sp++;
    }
    st.have_node = true;
    return false;
}

// The closest hit resolves its primitive index once, at the end (one more id-table read).
__device__ __forceinline__ HitRec trav_result(const TravState& st, const uint32_t* __restrict__ ids, bool any) {
    HitRec h = st.best;
    if (!any && h.rec != 0xffffffffu) h.id = (int32_t)(ids[h.rec] & ~BC_SPHERE_BIT);
    return h;
}

// Traverses the BVH8 for one ray to completion.
template <bool ANY>
__device__ __forceinline__ HitRec traverse(const bc_node8* __restrict__ nodes, const bc_prim* __restrict__ prims,
                                           const uint32_t* __restrict__ ids, uint32_t root, const Ray& r, float tmax,
                                           uint2* stk, int stride, int cap, int* __restrict__ err) {
    TravState st;
    trav_init(st, r, tmax, root);
    while (!trav_step<ANY>(st, nodes, prims, ids, stk, stride, cap, err)) {
    }
    return trav_result(st, ids, ANY);
}


}  // namespace bcd
