// trace_device.hpp -- BVH8 closest-hit / any-hit traversal for CDNA4 (gfx950), wave64.
//
// Replaces closest()/boxhit()/phit() of the reference OpenCL kernel
// (gpu_opencl_windows.go:736-777) and Scene.Hit / Scene.Occluded (renderer.go:777-895).
//
// One ray per lane.  A node visit loads one 80-byte bc_node8 (five 16-byte loads), tests
// its eight quantised child boxes with FMA slab tests and produces
//   * a node group  : (child_base, imask, hit bits in near-to-far slot order)
//   * a prim group  : (prim_base, 32-bit mask of primitives in hit leaves)
// Groups that still hold children are pushed to a per-lane stack in LDS laid out
// [depth][lane] so that a wave's pushes and pops are conflict-free 8-byte accesses.
// A group consumes one stack entry per tree level, so the stack depth needed equals
// the BVH8 depth, which the host checks before launch.
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

// triHit in fp32 (renderer.go:684-711); t returned through tout.
__device__ __forceinline__ bool tri_test(const float4 q0, const float4 q1, const float4 q2, const Ray& r, float tmin,
                                         float tmax, float& tout) {
    const float e1x = q0.w, e1y = q1.x, e1z = q1.y, e2x = q1.z, e2y = q1.w, e2z = q2.x;
    const float hx = r.dy * e2z - r.dz * e2y;
    const float hy = r.dz * e2x - r.dx * e2z;
    const float hz = r.dx * e2y - r.dy * e2x;
    const float det = e1x * hx + e1y * hy + e1z * hz;
    if (fabsf(det) < 1e-10f) return false;
    const float f = 1.0f / det;
    const float sx = r.ox - q0.x, sy = r.oy - q0.y, sz = r.oz - q0.z;
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
__device__ __forceinline__ bool sphere_test(const float4 q0, const float4 q1, const Ray& r, float tmin, float tmax,
                                            float& tout) {
    const float a = r.dx * r.dx + r.dy * r.dy + r.dz * r.dz;
    if (!(a > 1e-18f)) return false;
    const float inva = 1.0f / a;
    const float ocx = r.ox - q0.x, ocy = r.oy - q0.y, ocz = r.oz - q0.z;
    const float hb = ocx * r.dx + ocy * r.dy + ocz * r.dz;
    const float k = hb * inva;
    const float lx = ocx - r.dx * k, ly = ocy - r.dy * k, lz = ocz - r.dz * k;
    const float disc = a * (q1.x - (lx * lx + ly * ly + lz * lz));
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

// Per-axis quantised slab terms of one node for this ray.
struct NodeAxis {
    float adj, off;
};

// Tests the 8 children of node `nd`.  Outputs: ordered internal-hit bits (bit k <=> slot k^oct),
// the node's imask and child_base, and the primitive group.
__device__ __forceinline__ void visit_node(const bc_node8* __restrict__ nodes, uint32_t idx, const Ray& r,
                                           float ix, float iy, float iz, uint32_t oct, float tmax,
                                           uint32_t& ng_base, uint32_t& ng_bits, uint32_t& pg_base,
                                           uint32_t& pg_mask) {
    const uint4* p = reinterpret_cast<const uint4*>(nodes + idx);
    const uint4 w0 = p[0], w1 = p[1], w2 = p[2], w3 = p[3], w4 = p[4];
    const float ox = __uint_as_float(w0.x), oy = __uint_as_float(w0.y), oz = __uint_as_float(w0.z);
    const uint32_t ex = w0.w & 0xffu, ey = (w0.w >> 8) & 0xffu, ez = (w0.w >> 16) & 0xffu;
    const uint32_t imask = w0.w >> 24;
    const float sxf = __uint_as_float(ex << 23), syf = __uint_as_float(ey << 23), szf = __uint_as_float(ez << 23);
    const float adjx = sxf * ix, adjy = syf * iy, adjz = szf * iz;
    const float offx = (ox - r.ox) * ix, offy = (oy - r.oy) * iy, offz = (oz - r.oz) * iz;
    // near/far planes per axis by direction sign: words (lo0-3, lo4-7) and (hi0-3, hi4-7)
    const uint32_t lox0 = w2.x, lox1 = w2.y, loy0 = w2.z, loy1 = w2.w, loz0 = w3.x, loz1 = w3.y;
    const uint32_t hix0 = w3.z, hix1 = w3.w, hiy0 = w4.x, hiy1 = w4.y, hiz0 = w4.z, hiz1 = w4.w;
    const bool nx = ix < 0.0f, ny = iy < 0.0f, nz = iz < 0.0f;
    const uint32_t nx0 = nx ? hix0 : lox0, nx1 = nx ? hix1 : lox1, fx0 = nx ? lox0 : hix0, fx1 = nx ? lox1 : hix1;
    const uint32_t ny0 = ny ? hiy0 : loy0, ny1 = ny ? hiy1 : loy1, fy0 = ny ? loy0 : hiy0, fy1 = ny ? loy1 : hiy1;
    const uint32_t nz0 = nz ? hiz0 : loz0, nz1 = nz ? hiz1 : loz1, fz0 = nz ? loz0 : hiz0, fz1 = nz ? loz1 : hiz1;
    const uint32_t meta0 = w1.z, meta1 = w1.w;
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
        const uint32_t meta = ((hiw ? meta1 : meta0) >> sh) & 0xffu;
        const bool h = (tn <= tf) && (meta != BC_META_EMPTY);
        hits |= (h ? 1u : 0u) << s;
    }
    // internal children: reorder hit bits so that bit k is slot k ^ oct
    const uint32_t ih = hits & imask;
    uint32_t ord = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) ord |= ((ih >> ((uint32_t)k ^ oct)) & 1u) << k;
    ng_base = w1.x;
    ng_bits = ord | (imask << 8);
    // leaf children: union of their primitive runs
    uint32_t lh = hits & ~imask & 0xffu, pm = 0;
    while (lh) {
        const uint32_t s = (uint32_t)__builtin_ctz(lh);
        lh &= lh - 1u;
        const uint32_t meta = ((s >= 4 ? meta1 : meta0) >> ((s & 3u) * 8u)) & 0xffu;
        const uint32_t cnt = (meta >> 5) + 1u, offs = meta & 31u;
        pm |= (uint32_t)(((1ull << cnt) - 1ull) << offs);
    }
    pg_base = w1.y;
    pg_mask = pm;
}

// Traverses the BVH8.  ANY=true stops at the first primitive hit in [tmin, tmax].
// stk points at this lane's column of the LDS stack (entry i at stk[i * stride]).
template <bool ANY, int STACK>
__device__ __forceinline__ HitRec traverse(const bc_node8* __restrict__ nodes, const bc_prim* __restrict__ prims,
                                           const Ray& r, float tmax, uint2* stk, int stride, int* __restrict__ err) {
    const float ix = safe_rcp(r.dx), iy = safe_rcp(r.dy), iz = safe_rcp(r.dz);
    const uint32_t oct = (r.dx < 0.0f ? 1u : 0u) | (r.dy < 0.0f ? 2u : 0u) | (r.dz < 0.0f ? 4u : 0u);
    HitRec best;
    best.t = tmax;
    best.id = -1;
    best.rec = 0xffffffffu;
    uint32_t ng_base = 0, ng_bits = 0, pg_base = 0, pg_mask = 0;
    uint32_t node = 0;
    bool have_node = true;
    int sp = 0;
    for (;;) {
        if (have_node) {
            visit_node(nodes, node, r, ix, iy, iz, oct, best.t, ng_base, ng_bits, pg_base, pg_mask);
            have_node = false;
        }
        while (pg_mask) {
            const uint32_t k = (uint32_t)__builtin_ctz(pg_mask);
            pg_mask &= pg_mask - 1u;
            const uint32_t ri = pg_base + k;
            const float4* pq = reinterpret_cast<const float4*>(prims + ri);
            const float4 q0 = pq[0], q1 = pq[1], q2 = pq[2];
            const uint32_t idbits = __float_as_uint(q2.y);
            float t;
            bool h;
            if (idbits & BC_SPHERE_BIT) h = sphere_test(q0, q1, r, r.tmin, best.t, t);
            else h = tri_test(q0, q1, q2, r, r.tmin, best.t, t);
            if (h) {
                const int32_t id = (int32_t)(idbits & ~BC_SPHERE_BIT);
                if (ANY) {
                    best.t = t;
                    best.id = id;
                    best.rec = ri;
                    return best;
                }
                if (best.id < 0 || t < best.t || id < best.id) {
                    best.t = t;
                    best.id = id;
                    best.rec = ri;
                }
            }
        }
        if (ng_bits & 0xffu) {
            const uint32_t k = (uint32_t)__builtin_ctz(ng_bits & 0xffu);
            ng_bits &= ~(1u << k);
            const uint32_t slot = k ^ oct;
            const uint32_t imask = ng_bits >> 8;
            node = ng_base + (uint32_t)__builtin_popcount(imask & ((1u << slot) - 1u));
            if (ng_bits & 0xffu) {
                if (sp >= STACK) {  // host guarantees depth <= STACK; report instead of overflowing
                    atomicOr(err, 1);
                    return best;
                }
                stk[sp * stride] = make_uint2(ng_base, ng_bits);
                sp++;
            }
            have_node = true;
        } else if (sp > 0) {
            sp--;
            const uint2 e = stk[sp * stride];
            ng_base = e.x;
            ng_bits = e.y;
        } else {
            break;
        }
    }
    return best;
}


// ------------------------------------------------------------------------------ stepwise form
// The same traversal as a resumable state machine, for persistent kernels that refill a
// lane with a new ray as soon as its ray finishes (instead of waiting for the wave's
// slowest ray).  PPS = primitives tested per step (0 = all pending primitives).
struct TravState {
    Ray r;
    float ix, iy, iz, tmax;
    uint32_t oct;
    HitRec best;
    uint32_t ng_base, ng_bits, pg_base, pg_mask, node;
    int sp;
    bool have_node;
};

__device__ __forceinline__ void trav_init(TravState& s, const Ray& r, float tmax) {
    s.r = r;
    s.ix = safe_rcp(r.dx);
    s.iy = safe_rcp(r.dy);
    s.iz = safe_rcp(r.dz);
    s.oct = (r.dx < 0.0f ? 1u : 0u) | (r.dy < 0.0f ? 2u : 0u) | (r.dz < 0.0f ? 4u : 0u);
    s.tmax = tmax;
    s.best.t = tmax;
    s.best.id = -1;
    s.best.rec = 0xffffffffu;
    s.ng_base = 0;
    s.ng_bits = 0;
    s.pg_base = 0;
    s.pg_mask = 0;
    s.node = 0;
    s.sp = 0;
    s.have_node = true;
}

// One step; returns true when the ray is finished (result in s.best).
template <bool ANY, int STACK, int PPS>
__device__ __forceinline__ bool trav_step(TravState& s, const bc_node8* __restrict__ nodes,
                                          const bc_prim* __restrict__ prims, uint2* stk, int stride,
                                          int* __restrict__ err) {
    if (s.have_node) {
        visit_node(nodes, s.node, s.r, s.ix, s.iy, s.iz, s.oct, s.best.t, s.ng_base, s.ng_bits, s.pg_base, s.pg_mask);
        s.have_node = false;
    }
    int tested = 0;
    while (s.pg_mask && (PPS == 0 || tested < PPS)) {
        tested++;
        const uint32_t k = (uint32_t)__builtin_ctz(s.pg_mask);
        s.pg_mask &= s.pg_mask - 1u;
        const uint32_t ri = s.pg_base + k;
        const float4* pq = reinterpret_cast<const float4*>(prims + ri);
        const float4 q0 = pq[0], q1 = pq[1], q2 = pq[2];
        const uint32_t idbits = __float_as_uint(q2.y);
        float t;
        bool h;
        if (idbits & BC_SPHERE_BIT) h = sphere_test(q0, q1, s.r, s.r.tmin, s.best.t, t);
        else h = tri_test(q0, q1, q2, s.r, s.r.tmin, s.best.t, t);
        if (h) {
            const int32_t id = (int32_t)(idbits & ~BC_SPHERE_BIT);
            if (ANY) {
                s.best.t = t;
                s.best.id = id;
                s.best.rec = ri;
                return true;
            }
            if (s.best.id < 0 || t < s.best.t || id < s.best.id) {
                s.best.t = t;
                s.best.id = id;
                s.best.rec = ri;
            }
        }
    }
    if (s.pg_mask) return false;
    if (!(s.ng_bits & 0xffu)) {
        if (s.sp == 0) return true;
        s.sp--;
        const uint2 e = stk[s.sp * stride];
        s.ng_base = e.x;
        s.ng_bits = e.y;
    }
    const uint32_t k = (uint32_t)__builtin_ctz(s.ng_bits & 0xffu);
    s.ng_bits &= ~(1u << k);
    const uint32_t slot = k ^ s.oct;
    s.node = s.ng_base + (uint32_t)__builtin_popcount((s.ng_bits >> 8) & ((1u << slot) - 1u));
    if (s.ng_bits & 0xffu) {
        if (s.sp >= STACK) {
            atomicOr(err, 1);
            return true;
        }
        stk[s.sp * stride] = make_uint2(s.ng_base, s.ng_bits);
        s.sp++;
    }
    s.have_node = true;
    return false;
}

}  // namespace bcd
