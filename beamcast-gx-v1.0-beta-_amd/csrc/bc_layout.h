// bc_layout.h -- records resident in HBM, shared by the host packer and the HIP kernels.
//
// The reference GPU path uploads a threaded binary BVH of 32-byte nodes and 40-byte
// triangles (gpu_opencl_windows.go:54-77, packed at :337-377).  This backend replaces
// that with an 8-wide BVH whose child boxes are quantised to 8 bits per plane relative
// to a per-node power-of-two grid, one primitive per leaf slot, and primitive records at fixed
// positions of their node (eight record slots per node), so a node needs no record base.
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define BC_HD __host__ __device__
#else
#define BC_HD
#endif

// One 8-wide node: 64 bytes = four 16-byte loads.
//   link           : bits 0-23 child_base, bits 24-31 imask
//   grid[a]        : bits 0-20 origin k_a (two's complement), bits 21-28 e_a + 127, bits 29-31
//                    three bits of lmask and the sphere flag (bc_node_lmask, bc_node_spheres)
//   child box (slot i, axis a) = [o_a + qlo_a[i] * 2^e_a, o_a + qhi_a[i] * 2^e_a], o_a = k_a * 2^e_a
//   imask bit i    : slot i is an internal child, node index = child_base + popcount(imask & ((1<<i)-1))
//   lmask bit i    : slot i is a leaf holding one primitive, whose record is 8 * node + i
//   neither        : empty slot
// Slot i holds the child lying most towards (+-x,+-y,+-z) = (i&1, i&2, i&4): a ray whose
// direction sign bits are oct visits slots in the order k ^ oct, k = 0..7 (near to far).
struct alignas(16) bc_node8 {
    uint32_t link;
    uint32_t grid[3];
    uint8_t qlox[8], qloy[8], qloz[8];
    uint8_t qhix[8], qhiy[8], qhiz[8];
};
static_assert(sizeof(bc_node8) == 64, "bc_node8 must be 64 bytes");

#define BC_NODE_MAX (1u << 24)     // child_base field: nodes per scene (or batch arena)
#define BC_GRID_K_BITS 21          // origin grid integer: |k| < 2^20
#define BC_REC_PER_NODE 8u         // record slots per node: record of leaf slot s of node i = 8 i + s
#define BC_REC_HOLE 0xFFFFFFFFu    // id of a record slot that holds no primitive

BC_HD static inline uint32_t bc_node_child_base(const bc_node8& n) { return n.link & 0xFFFFFFu; }
BC_HD static inline uint32_t bc_node_imask(const bc_node8& n) { return n.link >> 24; }
BC_HD static inline uint32_t bc_node_lmask(const bc_node8& n) {
    return (n.grid[0] >> 29) | ((n.grid[1] >> 29) << 3) | (((n.grid[2] >> 29) & 3u) << 6);
}
BC_HD static inline bool bc_node_spheres(const bc_node8& n) { return (n.grid[2] >> 31) != 0u; }
BC_HD static inline int32_t bc_node_k(const bc_node8& n, int a) { return ((int32_t)(n.grid[a] << 11)) >> 11; }
BC_HD static inline int bc_node_e(const bc_node8& n, int a) { return (int)((n.grid[a] >> 21) & 0xFFu) - 127; }
BC_HD static inline void bc_node_set_link(bc_node8& n, uint32_t child_base, uint32_t imask) {
    n.link = (child_base & 0xFFFFFFu) | (imask << 24);
}
BC_HD static inline void bc_node_set_grid(bc_node8& n, int a, int32_t k, int e) {
    n.grid[a] = ((uint32_t)k & 0x1FFFFFu) | ((uint32_t)(e + 127) << 21) | (n.grid[a] & 0xE0000000u);
}
BC_HD static inline void bc_node_set_flags(bc_node8& n, uint32_t lmask, bool spheres) {
    n.grid[0] = (n.grid[0] & 0x1FFFFFFFu) | ((lmask & 7u) << 29);
    n.grid[1] = (n.grid[1] & 0x1FFFFFFFu) | (((lmask >> 3) & 7u) << 29);
    n.grid[2] = (n.grid[2] & 0x1FFFFFFFu) | (((lmask >> 6) & 3u) << 29) | (spheres ? 0x80000000u : 0u);
}

// The quantisation grid of one axis of a node whose children's union box is [mn, mx] (fp64):
// outward padding pad (relative 2^-20), the smallest power-of-two step 2^e (e >= -126) for which
// the origin k 2^e = floor((mn - pad) / 2^e) 2^e has |k| < 2^20 and 255 steps from it reach
// mx + pad.  The same arithmetic runs on the host and the device (exact fp64 operations only),
// so both produce the same bytes.  Returns false when the step would exceed 2^127 or the origin
// leaves the fp32 range.
BC_HD static inline bool bc_grid_axis(double mn, double mx, double& pad, int32_t& k, int& e, double& org, double& scale) {
    const double mag = fmax(fabs(mn), fabs(mx));
    const double ext = mx - mn;
    pad = fmax((mag + ext) * 0x1p-20, 1e-20);
    const double lo = mn - pad, hi = mx + pad;
    int k2;
    const double m2 = frexp((hi - lo) / 255.0, &k2);  // smallest e with 2^e >= (hi - lo) / 255, exactly
    e = m2 == 0.5 ? k2 - 1 : k2;
    if (e < -126) e = -126;
    for (;;) {
        if (e > 127) return false;
        const double step = ldexp(1.0, e);
        const double kd = floor(lo / step);
        if (fabs(kd) >= 1048576.0 || (kd + 255.0) * step < hi) {
            e++;
            continue;
        }
        k = (int32_t)kd;
        org = kd * step;
        scale = step;
        return fabs(org) < 0x1p127 && fabs(org + 255.0 * step) < 0x1p127;
    }
}

#define BC_SPHERE_BIT 0x80000000u

// Primitive record (36 bytes, nine floats), at 8 * node + slot of its leaf slot (record slots
// of internal and empty slots are holes: zero record, id BC_REC_HOLE): exactly the operands of the
// intersection test.
//   triangle: A.xyz, E1.xyz, E2.xyz
//   sphere  : C.xyz, r, r*r, 1/r, 0, 0, 0
// Parallel arrays by record: id = Scene.Primitives index (renderer.go:233) | SPHERE_BIT for
// spheres, and material.  Traversal reads the id only to resolve an accepted hit, and the
// kind only in nodes flagged as holding spheres (bc_node_spheres).
struct bc_prim {
    float v[9];
};
static_assert(sizeof(bc_prim) == 36, "bc_prim must be 36 bytes");
// bit 31 of a traversal's record base: the node holds spheres
#define BC_NODE_HAS_SPHERES 0x80000000u

// Material record: kinds as renderer.go:86-91.
struct alignas(16) bc_mat {
    float color[3], rough;
    float emission[3], ior;
    int32_t kind, pad0, pad1, pad2;
};
static_assert(sizeof(bc_mat) == 48, "bc_mat must be 48 bytes");

// NEE light record (emissive spheres, and emissive triangles when enabled).
struct alignas(16) bc_light {
    float g0[4];   // sphere: C.xyz, r        triangle: A.xyz, 0
    float g1[4];   // triangle: E1.xyz, 0
    float g2[4];   // triangle: E2.xyz, 0
    float le[4];   // emission of the light's material (rgb), 0
    float area;    // surface area rounded from fp64
    float cdf;     // normalised CDF value
    int32_t prim;  // Scene.Primitives index
    int32_t kind;  // 0 sphere, 1 triangle
};
static_assert(sizeof(bc_light) == 80, "bc_light must be 80 bytes");

// Ray record used by the trace entry points: 32 bytes.
struct alignas(16) bc_ray {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
};
// Closest-hit record: 8 bytes (t, primitive id; id = -1 on miss).
struct bc_hit {
    float t;
    int32_t prim;
};

// Random-number dimension map of the integrator (counter-based; identical in the oracle).
#define BC_DIM_CAM_X 0u
#define BC_DIM_CAM_Y 1u
#define BC_DIM_BOUNCE0 16u
#define BC_DIMS_PER_BOUNCE 64u
#define BC_D_LIGHT 0u
#define BC_D_GLASS 1u
#define BC_D_RR 2u
#define BC_D_TRI 3u
#define BC_D_SPH 5u
#define BC_D_COS 21u
#define BC_D_METAL 37u
#define BC_NTRY_SPH 8
#define BC_NTRY_COS 8
#define BC_NTRY_METAL 9
/* resampled direct lighting at a vertex of bounce b: the vertex key is bc_rng_u32(path key,
 * BC_DIM_RIS0 + b); candidate k draws dims k * BC_RIS_DIMS + o of the vertex key, o = 0 component
 * (fewer candidates than components only), 1 light pick, 2-3 triangle point (2: the sky sample's
 * squared disk radius), 4-19 sphere point / sky angle tries, 20 reservoir update */
#define BC_DIM_RIS0 0x100000u
#define BC_RIS_DIMS 32u
#define BC_RIS_MAX 64
/* spatial reuse: neighbour i draws dims BC_DIM_SPATIAL0 + 4 i + o, o = 0 x offset, 1 y offset
 * (integer hashes), 2 reservoir update (i = 0: the pixel's own reservoir) */
#define BC_DIM_SPATIAL0 0x200000u

BC_HD static inline uint32_t bc_mix32(uint32_t x) {
    // "lowbias32" integer hash (C. Wellons, hash-prospector; public domain)
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
BC_HD static inline uint32_t bc_path_key(uint32_t seed, uint32_t pixel, uint32_t sample) {
    return bc_mix32(seed ^ bc_mix32(pixel ^ bc_mix32(sample + 0x9e3779b9u)));
}
BC_HD static inline uint32_t bc_rng_u32(uint32_t key, uint32_t dim) {
    return bc_mix32(key ^ bc_mix32(dim * 0x85ebca6bu + 0x632be5abu));
}
BC_HD static inline float bc_rng(uint32_t key, uint32_t dim) {
    return (float)(bc_rng_u32(key, dim) >> 8) * (1.0f / 16777216.0f);
}
