// integrator_device.hpp -- device-side pieces of the wavefront integrator shared by the plain
// (pass per sample batch) and persistent (in-place regeneration) schedules.
//
// Semantics follow the Go CPU integrator: trace (renderer.go:1484-1563), directLight
// (:1427-1465), chooseLight (:1371-1393), lightDirectionalPDF (:1404-1425), the samplers
// (:1262-1311) and the camera jitter of adaptivePixel (:1940-1950), evaluated in fp32 with the
// operation order of the oracle so that per-sample results match it.
#pragma once
#include <hip/hip_runtime.h>

#include "bc_layout.h"
#include "trace_device.hpp"

namespace bcd {

constexpr int kBlock = 256;

struct RenderParams {
    int32_t width, height, spp, bounces, rr_depth, sampler, mis, power, sky_mode, want_guides;
    float firefly, sky[3];
    uint32_t seed;
    float origin[3], lower[3], horiz[3], vert[3];
    int32_t y0, npix;          // chunk: first row, pixels in chunk
    int32_t s0, ns;            // pass: first sample, samples in this pass
    int32_t nmat, nlight;
    uint32_t root, lbase;      // scene root node and first light of the scene (packed scenes)
    int32_t restir;            // candidates of the resampled direct lighting (0 = one-sample NEE)
    float sun_dir[3], sun[3];  // directional light (restir only)
    int32_t spatial, radius;   // spatial reuse at primary vertices: neighbours, pixel radius (0 = off)
    int32_t band;              // persistent window: pixels per band of its sample order (window_sample)
};

// Order of a persistent window's samples (g < npix * ns of one job): bands of P.band consecutive
// pixels (the last one shorter), each holding all of the window's samples of its pixels, sample
// after sample.  The path pool restarts samples in g order, so the paths in flight come from one
// or two bands of the image and share the scene's nodes in cache.
__device__ __forceinline__ void window_sample(const RenderParams& P, uint32_t g, uint32_t& sl, uint32_t& pl) {
    const uint32_t band = (uint32_t)P.band, per = band * (uint32_t)P.ns;
    const uint32_t b = g / per, b0 = b * band;
    const uint32_t bs = min(band, (uint32_t)P.npix - b0);
    const uint32_t r = g - b * per;
    sl = r / bs;
    pl = b0 + (r - sl * bs);
}
// The inverse: the window index g of sample slot sl of pixel pl.
__device__ __forceinline__ uint32_t window_index(const RenderParams& P, uint32_t sl, uint32_t pl) {
    const uint32_t band = (uint32_t)P.band;
    const uint32_t b = pl / band, b0 = b * band;
    const uint32_t bs = min(band, (uint32_t)P.npix - b0);
    return b * band * (uint32_t)P.ns + sl * bs + (pl - b0);
}

// Path slots are grouped in segments of B.seg consecutive slots (a multiple of the workgroup
// size).  One workgroup owns a segment in every kernel: it walks the segment's queues, and
// compacts the paths that continue (and the shadow rays, and the terminated paths) into the
// same segment of the output queues, writing the segment's lengths itself.  No queue needs a
// device-wide counter: a same-address atomic from all eight XCDs costs ~10 ns on MI355X, which
// made one-atomic-per-wave appends the dominant cost of shading.  Larger segments pack the
// survivors of late bounces into fewer, fuller waves; smaller ones give more workgroups.
constexpr int kWaves = kBlock / 64;
constexpr uint32_t kSegPlain = 1024;     // segment of the plain schedule
constexpr uint32_t kSegPersist = 256;    // default segment of the persistent schedule
constexpr uint32_t kSegMax = 4096;
constexpr int kMaxGrid = 4096;  // largest grid of the per-pixel kernels (one tally word each)

struct PathBufs {
    uint32_t seg;   // slots per segment
    uint32_t xg;    // segments per XCD turn (see xcd_chunk)
    uint32_t* bnc;  // per-path bounce (persistent schedule)
    uint32_t* gid;  // per-path window sample index, 0xffffffff = idle (persistent schedule)
    uint32_t* sid;  // per-path scene index (packed scenes)
    float4* res;    // per-sample result: rgb, path segments (persistent schedule)
    uint32_t* dq;   // terminated paths (persistent schedule)
    float4* o;      // ray origin
    float4* d;      // ray direction
    float4* T;      // throughput xyz, prevPdf w
    float4* L;      // radiance
    float4* pp;     // prevPoint xyz, w = prevDiffuse flag
    uint4* meta;    // pixel, sample, key, packed jitter
    uint32_t* nseg; // path segments
    float2* hit;    // t, record index bits (0xffffffff: miss)
    float4* so;     // shadow origin xyz, tmax w
    float4* sd;     // shadow dir
    float4* sc;     // shadow contribution
    uint32_t* q[2];  // path queues (segment s holds entries [s*seg, s*seg + qc[s]))
    uint32_t* sq;    // shadow-ray queue
    uint32_t* qc[2]; // per-segment queue lengths
    uint32_t* sqc;   // per-segment shadow queue length
    uint32_t* dqc;   // per-segment terminated-path count (persistent schedule)
    uint32_t* snext; // per-segment next sample ordinal (persistent schedule)
    uint32_t* flag;  // [0] != 0: some segment had work (persistent drain test)
    float4* rv[5];   // spatial reuse: primary-vertex reservoirs (p, t | N, code | f, W | y, M | light normal)
    unsigned long long* tally;   // path segments per pixel-kernel block [0, kMaxGrid), then shadow rays per segment
    unsigned long long* totals;  // [0] path segments, [1] shadow rays
};

struct SceneBufs {
    const bc_node8* nodes;
    const bc_prim* prims;
    const uint32_t* ids;
    const int32_t* pmat;
    const bc_mat* mats;
    const bc_light* lights;
    const float* pdf;
    const int32_t* light_of;
    int* err;
    int stack;  // traversal stack entries per lane (BVH8 depth - 1); kernels take that much LDS
    int shadow_refill;  // shadow rays: refill a wave's lanes once this many are idle (0: one ray per thread)
    int extend_refill;  // the plain schedule's second-segment extension rays: the same for closest hit (0: off)
};

__device__ __forceinline__ float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 add(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 sub(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 mul(float3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 had(float3 a, float3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross(float3 a, float3 b) {
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float3 unitv(float3 a) {  // Vec3.Unit, renderer.go:32-38
    const float l = sqrtf(dot(a, a));
    if (l < (float)1e-12) return f3(0.0f, 0.0f, 0.0f);
    return mul(a, 1.0f / l);
}
__device__ __forceinline__ float rmax(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float rmin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float rclamp(float x, float a, float b) { return x < a ? a : (x > b ? b : x); }
__device__ __forceinline__ float3 xyz(float4 v) { return f3(v.x, v.y, v.z); }

#define PI_F ((float)3.141592653589793)
#define INVPI_F ((float)0.3183098861837907)

// Block-wide ranks of NQ predicates for one round of a segment walk: rank[q] is the number
// of lanes of the workgroup below this lane whose pred[q] holds, total[q] the round's count.
// lds holds 2 x NQ x kWaves words; consecutive rounds alternate halves (parity), so a single
// barrier per round suffices.  Every thread of the workgroup must call it the same number of times.
template <int NQ>
__device__ __forceinline__ void block_rank(const bool (&pred)[NQ], uint32_t (&rank)[NQ], uint32_t (&total)[NQ],
                                           uint32_t* lds, uint32_t& parity) {
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    uint32_t* L = lds + parity * (NQ * kWaves);
    uint32_t below[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const uint64_t m = __ballot(pred[q]);
        below[q] = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) L[q * kWaves + w] = (uint32_t)__popcll(m);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        uint32_t before = 0, tot = 0;
#pragma unroll
        for (int k = 0; k < kWaves; k++) {
            const uint32_t c = L[q * kWaves + k];
            before += (uint32_t)k < w ? c : 0u;
            tot += c;
        }
        rank[q] = before + below[q];
        total[q] = tot;
    }
    parity ^= 1u;
}

__device__ __forceinline__ double radical_inverse(uint64_t n, uint32_t base) {  // renderer.go:1844-1858
    if (base == 2) return (double)__brevll(n) * 5.421010862427522e-20;
    const double invBase = 1.0 / (double)base;
    double inv = invBase, result = 0.0;
    while (n > 0) {
        const uint64_t digit = n % base;
        result += (double)digit * inv;
        n /= base;
        inv *= invBase;
    }
    return result;
}

__device__ __forceinline__ void pixel_rotation(int x, int y, double& rx, double& ry) {  // renderer.go:1906-1915
    uint32_t h1 = (uint32_t)x * 0x9e3779b1u ^ (uint32_t)y * 0x85ebca77u;
    h1 ^= h1 >> 16;
    h1 *= 0x7feb352du;
    h1 ^= h1 >> 15;
    uint32_t h2 = h1 ^ 0x68bc21ebu;
    h2 ^= h2 >> 13;
    const double inv = 1.0 / 4294967296.0;
    rx = (double)h1 * inv;
    ry = (double)h2 * inv;
}

__device__ __forceinline__ uint32_t pack_guide(float3 n, int mat) {  // gpu_opencl_windows.go:787-792
    const float inv = 1.0f / (fabsf(n.x) + fabsf(n.y) + fabsf(n.z) + 1e-20f);
    const float nx = n.x * inv, ny = n.y * inv, nz = n.z * inv;
    float px = nx, py = ny;
    if (nz < 0.0f) {
        const float ax = 1.0f - fabsf(py), ay = 1.0f - fabsf(px);
        px = copysignf(ax, px);
        py = copysignf(ay, py);
    }
    px = px * 0.5f + 0.5f;
    py = py * 0.5f + 0.5f;
    px = px < 0.0f ? 0.0f : (px > 1.0f ? 1.0f : px);
    py = py < 0.0f ? 0.0f : (py > 1.0f ? 1.0f : py);
    const uint32_t qx = (uint32_t)rintf(px * 1023.0f), qy = (uint32_t)rintf(py * 1023.0f);
    const uint32_t m = (uint32_t)(mat < 0 ? 0 : (mat > 2047 ? 2047 : mat));
    return 0x80000000u | (m << 20) | (qy << 10) | qx;
}
__device__ __forceinline__ uint32_t pack_jitter(float ju, float jv) {  // gpu_opencl_windows.go:793
    ju = ju < 0.0f ? 0.0f : (ju > 1.0f ? 1.0f : ju);
    jv = jv < 0.0f ? 0.0f : (jv > 1.0f ? 1.0f : jv);
    return ((uint32_t)rintf(jv * 65535.0f) << 16) | (uint32_t)rintf(ju * 65535.0f);
}


// Camera ray of adaptivePixel (renderer.go:1940-1950) with adaptive sampling off, written to
// path slot j for chunk pixel pl and sample `sample`.
__device__ __forceinline__ void init_camera_path(const RenderParams& P, const PathBufs& B, uint32_t j, int pl,
                                                 int sample) {
    const double g = 1.324717957244746, a1 = 1.0 / g, a2 = 1.0 / (g * g);
    const float3 origin = f3(P.origin[0], P.origin[1], P.origin[2]);
    const float3 lower = f3(P.lower[0], P.lower[1], P.lower[2]);
    const float3 horiz = f3(P.horiz[0], P.horiz[1], P.horiz[2]);
    const float3 vert = f3(P.vert[0], P.vert[1], P.vert[2]);
    const int x = pl % P.width, y = P.y0 + pl / P.width;
    const uint32_t pixel = (uint32_t)y * (uint32_t)P.width + (uint32_t)x;
    const uint32_t key = bc_path_key(P.seed, pixel, (uint32_t)sample);
    double jx, jy;
    if (P.sampler == 1) {
        double rx, ry;
        pixel_rotation(x, y, rx, ry);
        jx = radical_inverse((uint64_t)sample + 1, 2) + rx;
        jx = jx - floor(jx);
        jy = radical_inverse((uint64_t)sample + 1, 3) + ry;
        jy = jy - floor(jy);
    } else if (P.sampler == 2) {
        double rx, ry;
        pixel_rotation(x, y, rx, ry);
        double sx = 0.5 + (double)(sample + 1) * a1, sy = 0.5 + (double)(sample + 1) * a2;
        sx = sx - floor(sx);
        sy = sy - floor(sy);
        jx = sx + rx;
        jx = jx - floor(jx);
        jy = sy + ry;
        jy = jy - floor(jy);
    } else {
        jx = (double)(bc_rng_u32(key, BC_DIM_CAM_X) >> 8) * (1.0 / 16777216.0);
        jy = (double)(bc_rng_u32(key, BC_DIM_CAM_Y) >> 8) * (1.0 / 16777216.0);
    }
    const double uu = ((double)x + jx) / (double)(P.width - 1 > 1 ? P.width - 1 : 1);
    const double vv = ((double)(P.height - 1 - y) + jy) / (double)(P.height - 1 > 1 ? P.height - 1 : 1);
    const float u = (float)uu, v = (float)vv;
    const float3 dir = sub(add(add(lower, mul(horiz, u)), mul(vert, v)), origin);
    B.o[j] = make_float4(origin.x, origin.y, origin.z, 0.0f);
    B.d[j] = make_float4(dir.x, dir.y, dir.z, 0.0f);
    B.T[j] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    B.L[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    B.pp[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    B.meta[j] = make_uint4(pixel, (uint32_t)sample, key, pack_jitter((float)jx, (float)jy));
    B.nseg[j] = 0;
}

__device__ __forceinline__ float light_dir_pdf(const SceneBufs& S, const RenderParams& P, int prim, float3 from,
                                               float3 lp, float3 ln) {  // renderer.go:1404-1425
    const int li = S.light_of[prim];
    if (li < 0) return 0.0f;
    const float3 to = sub(lp, from);
    const float dist2 = dot(to, to);
    if (dist2 <= (float)1e-12) return 0.0f;
    const float3 ld = unitv(to);
    const float cosL = rmax(0.0f, dot(ln, mul(ld, -1.0f)));
    if (cosL <= (float)1e-12) return 0.0f;
    const float area = S.lights[li].area;
    const float sel = P.power ? S.pdf[prim] : 1.0f / (float)P.nlight;
    return sel * dist2 / (cosL * area);
}


// Background radiance seen along d (renderer.go:1326-1331): constant sky or the gradient.
__device__ __forceinline__ float3 sky_radiance(const RenderParams& P, float3 d) {
    if (P.sky_mode == 1) return f3(P.sky[0], P.sky[1], P.sky[2]);
    const float3 u = unitv(d);
    const float tt = (float)0.5 * (u.y + 1.0f);
    return add(mul(f3((float)0.06, (float)0.08, (float)0.12), 1.0f - tt), mul(f3((float)0.32, (float)0.48, (float)0.72), tt));
}

__device__ __forceinline__ float lum(float3 v) { return (float)0.2126 * v.x + (float)0.7152 * v.y + (float)0.0722 * v.z; }

// Direction of the cosine-weighted hemisphere sample (x, y, z) about N (the frame of
// randCosHemisphere, renderer.go:1290-1311).
__device__ __forceinline__ float3 hemi_dir(float3 N, float x, float y, float z) {
    const float sign = N.z < 0.0f ? -1.0f : 1.0f;
    const float a = -1.0f / (sign + N.z);
    const float b = N.x * N.y * a;
    const float3 tt = f3(1.0f + sign * N.x * N.x * a, sign * b, -sign * N.x);
    const float3 bt = f3(b, sign + N.y * N.y * a, -N.y);
    return add(add(mul(tt, x), mul(bt, y)), mul(N, z));
}

// Resampled direct lighting at a diffuse vertex (an extension; the Go integrator has
// one-sample NEE): RIS over P.restir candidates from the components {sun (delta), emitters
// (chooseLight + point sampling of directLight), sky (cosine hemisphere)}, target = luminance of
// the unshadowed contribution, one streaming reservoir.  With at least one candidate per
// component (M >= nc) the candidates are split into contiguous strata, candidate k taking
// component k * nc / M, so the lanes of a wave evaluate the same component at each k; component
// c's M_c candidates each weigh target x (M / M_c) / pdf, since only those M_c candidates can
// produce a sample in c's domain (the domains are disjoint: a delta direction, emitter points,
// sky directions).  With fewer candidates than components each draws its component with
// probability 1 / nc and weighs target x nc / pdf.  Either way the survivor y carries
// W = sum(w) / (M target(y)), unbiased.  ris_direct returns false when nothing survives;
// otherwise the unshadowed contribution F(y) W (before throughput) and the visibility ray:
// towards the point y when `point`, else along the direction y.
// The reservoir of resampled direct lighting after its candidates: the survivor, its target and
// weight sum (ris_sample), from which ris_direct forms F(y) W and spatial reuse starts.
struct Reservoir {
    int comp;         // -1 none, 0 sun, 1 emitter point, 2 sky direction
    int li;           // emitter: light index in the scene's table
    float3 y, lnrm;   // point or direction; emitter normal at the point
    float3 F;         // unshadowed contribution at the vertex (solid-angle integrand)
    float ph, wsum, cosL, dist2;
    int M;
};

// (x, y) in the tangent plane of a cosine-weighted hemisphere sample whose squared disk radius is
// u1 (Malley's method: cos(theta) = sqrt(1 - u1)): the radius sqrt(u1) along a uniform angle,
// the angle of a point drawn uniformly in the unit disk by rejection from dims 4-19 of the
// candidate (after BC_NTRY_COS misses, probability 5e-6, the angle is 0).
__device__ __forceinline__ void sky_disk(uint32_t rk, uint32_t base, float u1, float& x, float& y) {
    const float r = sqrtf(u1);
    x = r;
    y = 0.0f;
    for (int t = 0; t < BC_NTRY_COS; t++) {
        const float a = bc_rng(rk, base + 4u + 2u * (uint32_t)t) * 2.0f - 1.0f;
        const float b = bc_rng(rk, base + 5u + 2u * (uint32_t)t) * 2.0f - 1.0f;
        const float ss = a * a + b * b;
        if (ss < 1.0f && ss > (float)1e-18) {
            const float s = r / sqrtf(ss);
            x = a * s;
            y = b * s;
            break;
        }
    }
}

__device__ __forceinline__ void ris_sample(const SceneBufs& S, const RenderParams& P, float3 p, float3 N, float3 color,
                                           uint32_t key, int bounce, Reservoir& rs) {
    const int M = P.restir < BC_RIS_MAX ? P.restir : BC_RIS_MAX;
    const float3 sd = f3(P.sun_dir[0], P.sun_dir[1], P.sun_dir[2]);
    const float3 E = f3(P.sun[0], P.sun[1], P.sun[2]);
    const bool sun_on = (E.x > 0.0f || E.y > 0.0f || E.z > 0.0f) && (sd.x != 0.0f || sd.y != 0.0f || sd.z != 0.0f);
    const bool area_on = P.nlight > 0;
    const int nc = 1 + (sun_on ? 1 : 0) + (area_on ? 1 : 0);
    const bool strat = M >= nc;
    const float3 f = mul(color, INVPI_F);
    // the sun's candidates are all the same sample (a delta direction)
    float3 F_sun = f3(0.0f, 0.0f, 0.0f);
    float ph_sun = 0.0f;
    if (sun_on) {
        const float cs = dot(N, sd);
        if (cs > 0.0f) {
            F_sun = mul(had(f, E), cs);
            ph_sun = lum(F_sun);
        }
    }
    float wsum = 0.0f, ph_s = 0.0f, cosL_s = 0.0f, dist2_s = 0.0f;
    float3 F_s = f3(0.0f, 0.0f, 0.0f), y_s = F_s, ln_s = F_s;
    bool have = false;
    int comp_s = -1, li_s = -1, k_s = 0;
    // The candidates draw from a key of their own per vertex, so the dimension half of every hash
    // depends only on (k, offset): the same for all lanes of a wave, whatever their bounce.
    const uint32_t rk = bc_rng_u32(key, BC_DIM_RIS0 + (uint32_t)bounce);
    for (int k = 0; k < M; k++) {
        const uint32_t base = (uint32_t)k * BC_RIS_DIMS;
        int ci;
        float mult;  // M / M_c (strata) or nc (random components)
        if (strat) {
            ci = k * nc / M;
            const int c0 = (ci * M + nc - 1) / nc, c1 = ((ci + 1) * M + nc - 1) / nc;
            mult = (float)M / (float)(c1 - c0);
        } else {
            ci = (int)(bc_rng(rk, base) * (float)nc);
            if (ci >= nc) ci = nc - 1;
            mult = (float)nc;
        }
        int comp;  // 0 sun, 1 emitters, 2 sky
        if (sun_on) comp = ci == 0 ? 0 : ((area_on && ci == 1) ? 1 : 2);
        else comp = (area_on && ci == 0) ? 1 : 2;
        float w = 0.0f, ph = 0.0f, cosL_c = 0.0f, dist2_c = 0.0f;
        float3 F = f3(0.0f, 0.0f, 0.0f), y = F, ln_c = F;
        int li_c = -1;
        if (comp == 0) {
            if (ph_sun > 0.0f) {
                F = F_sun;
                ph = ph_sun;
                w = ph_sun * mult;
                y = sd;
            }
        } else if (comp == 1) {
            const float u = bc_rng(rk, base + 1u);
            int li;
            float sel;
            if (!P.power) {
                li = (int)(u * (float)P.nlight);
                if (li >= P.nlight) li = P.nlight - 1;
                sel = 1.0f / (float)P.nlight;
            } else {
                int lo = 0, hi = P.nlight;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (S.lights[P.lbase + mid].cdf >= u) hi = mid; else lo = mid + 1;
                }
                li = lo >= P.nlight ? P.nlight - 1 : lo;
                sel = S.pdf[S.lights[P.lbase + li].prim];
                if (sel <= 0.0f) sel = 1.0f / (float)P.nlight;
            }
            const bc_light& Lr = S.lights[P.lbase + li];
            float3 qpt, lnrm;
            const bool sph = Lr.kind == 0;
            if (sph) {
                float3 ln = f3(0.0f, 0.0f, 1.0f);
                for (int t = 0; t < BC_NTRY_SPH; t++) {
                    const float a = bc_rng(rk, base + 4u + 2u * (uint32_t)t) * 2.0f - 1.0f;
                    const float b = bc_rng(rk, base + 5u + 2u * (uint32_t)t) * 2.0f - 1.0f;
                    const float ss = a * a + b * b;
                    if (ss >= 1.0f || ss <= (float)1e-18) continue;
                    const float ff = 2.0f * sqrtf(1.0f - ss);
                    ln = f3(a * ff, b * ff, 1.0f - 2.0f * ss);
                    break;
                }
                qpt = add(f3(Lr.g0[0], Lr.g0[1], Lr.g0[2]), mul(ln, Lr.g0[3]));
                lnrm = ln;
            } else {
                const float u1 = bc_rng(rk, base + 2u), u2 = bc_rng(rk, base + 3u);
                const float su = sqrtf(u1);
                const float b1 = su * (1.0f - u2), b2 = su * u2;
                const float3 e1 = f3(Lr.g1[0], Lr.g1[1], Lr.g1[2]), e2 = f3(Lr.g2[0], Lr.g2[1], Lr.g2[2]);
                qpt = add(add(f3(Lr.g0[0], Lr.g0[1], Lr.g0[2]), mul(e1, b1)), mul(e2, b2));
                lnrm = unitv(cross(e1, e2));
            }
            const float3 to = sub(qpt, p);
            const float dist2 = dot(to, to);
            if (!(dist2 < (float)1e-8)) {
                const float3 ld = unitv(to);
                const float nd = rmax(0.0f, dot(N, ld));
                const float cl = dot(mul(lnrm, -1.0f), ld);
                const float cosL = sph ? rmax(0.0f, cl) : fabsf(cl);
                if (!(nd <= 0.0f || cosL <= 0.0f)) {
                    const float pin = sel * dist2 / (cosL * Lr.area);
                    if (pin > 0.0f) {
                        F = mul(had(f, f3(Lr.le[0], Lr.le[1], Lr.le[2])), nd);
                        ph = lum(F);
                        w = (ph * mult) / pin;
                        y = qpt;
                        ln_c = lnrm;
                        li_c = li;
                        cosL_c = cosL;
                        dist2_c = dist2;
                    }
                }
            }
        } else {
            // cos(theta) of a cosine-weighted sample from its squared disk radius u1; the angle
            // about N is drawn only where the direction is needed: a sky that varies with
            // direction, and the survivor (after the loop)
            const float u1 = bc_rng(rk, base + 2u);
            const float z = sqrtf(rmax(0.0f, 1.0f - u1));
            if (z > 0.0f) {
                const float pin = z / PI_F;
                float3 Ls = f3(P.sky[0], P.sky[1], P.sky[2]);
                if (P.sky_mode != 1) {
                    float x, yy;
                    sky_disk(rk, base, u1, x, yy);
                    Ls = sky_radiance(P, hemi_dir(N, x, yy, z));
                }
                F = mul(had(f, Ls), z);
                ph = lum(F);
                w = (ph * mult) / pin;
                y = f3(u1, 0.0f, z);
            }
        }
        if (w > 0.0f) {
            wsum = wsum + w;
            if (bc_rng(rk, base + 20u) * wsum < w) {
                have = true;
                F_s = F;
                ph_s = ph;
                y_s = y;
                comp_s = comp;
                li_s = li_c;
                ln_s = ln_c;
                cosL_s = cosL_c;
                dist2_s = dist2_c;
                k_s = k;
            }
        }
    }
    rs.comp = have ? comp_s : -1;
    rs.li = li_s;
    if (comp_s == 2) {  // sky: the survivor's local sample -> direction
        float x, yy;
        sky_disk(rk, (uint32_t)k_s * BC_RIS_DIMS, y_s.x, x, yy);
        rs.y = hemi_dir(N, x, yy, y_s.z);
    } else {
        rs.y = y_s;
    }
    rs.lnrm = ln_s;
    rs.F = F_s;
    rs.ph = ph_s;
    rs.wsum = wsum;
    rs.cosL = cosL_s;
    rs.dist2 = dist2_s;
    rs.M = M;
}

__device__ __forceinline__ bool ris_direct(const SceneBufs& S, const RenderParams& P, float3 p, float3 N,
                                           float3 color, uint32_t key, int bounce, float3& contrib, float3& y_out,
                                           bool& point_out) {
    Reservoir rs;
    ris_sample(S, P, p, N, color, key, bounce, rs);
    if (rs.comp < 0) return false;
    const float W = rs.wsum / ((float)rs.M * rs.ph);
    contrib = mul(rs.F, W);
    y_out = rs.y;
    point_out = rs.comp == 1;
    return true;
}

// Spatial reuse of primary-vertex reservoirs (ReSTIR DI, Bitterli et al., SIGGRAPH 2020, the
// biased combination with 1/M normalisation).  Samples live in the measure of their kind:
// emitter points by area, sky directions by solid angle, the sun as a discrete direction; a
// reservoir's W is stored in that measure (an emitter sample's solid-angle W times dist^2 /
// cosL).  spatial_eval is the integrand of a sample at receiver (p, N, f) in the same measure.
__device__ __forceinline__ float3 spatial_eval(const SceneBufs& S, const RenderParams& P, float3 p, float3 N, float3 f,
                                               int comp, int li, float3 y, float3 lnrm) {
    const float3 zero = f3(0.0f, 0.0f, 0.0f);
    if (comp == 0) {
        const float3 sd = f3(P.sun_dir[0], P.sun_dir[1], P.sun_dir[2]);
        const float cs = dot(N, sd);
        return cs > 0.0f ? mul(had(f, f3(P.sun[0], P.sun[1], P.sun[2])), cs) : zero;
    }
    if (comp == 2) {
        const float cv = rmax(0.0f, dot(N, y));
        return cv > 0.0f ? mul(had(f, sky_radiance(P, y)), cv) : zero;
    }
    const bc_light& Lr = S.lights[P.lbase + li];
    const float3 to = sub(y, p);
    const float dist2 = dot(to, to);
    if (dist2 < (float)1e-8) return zero;
    const float3 ld = unitv(to);
    const float nd = rmax(0.0f, dot(N, ld));
    const float cl = dot(mul(lnrm, -1.0f), ld);
    const float cosL = Lr.kind == 0 ? rmax(0.0f, cl) : fabsf(cl);
    if (nd <= 0.0f || cosL <= 0.0f) return zero;
    return mul(had(f, f3(Lr.le[0], Lr.le[1], Lr.le[2])), (nd * cosL) / dist2);
}

// Adjoint state of the path-replay backward pass: per-pixel d loss / d image, the replayed
// path's total radiance from the forward pass, and per-block gradient accumulators in LDS.
struct Adj {
    const float* dl;      // rows*W*3 d loss / d image (chunk pixels)
    float inv_spp;
    const float4* Lt;     // per path slot: total radiance of this sample (forward pass)
    float4* sg;           // per path slot: d NEE contribution / d emission (rgb), light material
    float* gcol;          // LDS: nmat*3 d loss / d Color
    float* gle;           // LDS: nmat*3 d loss / d Emission
};

// d loss / d color at a scattering vertex: every later contribution (this vertex's NEE term
// included) carries the factor color, so its derivative is the remaining radiance / color.
__device__ __forceinline__ void adj_color(const Adj& A, int mat, float3 g, float3 Lt, float3 L, float3 color) {
    const float r[3] = {Lt.x - L.x, Lt.y - L.y, Lt.z - L.z};
    const float c[3] = {color.x, color.y, color.z};
    const float gg[3] = {g.x, g.y, g.z};
#pragma unroll
    for (int k = 0; k < 3; k++)
        if (c[k] != 0.0f && gg[k] != 0.0f && r[k] != 0.0f) atomicAdd(&A.gcol[mat * 3 + k], gg[k] * (r[k] / c[k]));
}
__device__ __forceinline__ void adj_emission(const Adj& A, int mat, float3 g, float3 w) {
    if (g.x != 0.0f && w.x != 0.0f) atomicAdd(&A.gle[mat * 3 + 0], g.x * w.x);
    if (g.y != 0.0f && w.y != 0.0f) atomicAdd(&A.gle[mat * 3 + 1], g.y * w.y);
    if (g.z != 0.0f && w.z != 0.0f) atomicAdd(&A.gle[mat * 3 + 2], g.z * w.z);
}

// One shading step of trace for path j at `bounce` (renderer.go:1490-1560).  Returns true when
// the path continues (a new ray is stored); an NEE shadow ray, if any, is returned in so/sd/sc.
// ADJ adds the path-replay gradient of the step (Russian roulette and light selection are
// treated as constants of the parameters).
template <bool ADJ>
__device__ __forceinline__ bool shade_path(const SceneBufs& S, const PathBufs& B, const RenderParams& P, uint32_t j,
                                           int bounce, uint32_t* __restrict__ guides, bool& shadow,
                                           float4& shadow_o, float4& shadow_d, float4& shadow_c, const Adj& A) {
    const uint32_t dim = BC_DIM_BOUNCE0 + BC_DIMS_PER_BOUNCE * (uint32_t)bounce;
    const int rr = P.rr_depth < 1 ? 4 : P.rr_depth;
    bool alive_out = false;
    const float2 hv = B.hit[j];
    const uint32_t rec = __float_as_uint(hv.y);
    const float t = hv.x;
    float3 ro = xyz(B.o[j]), rd = xyz(B.d[j]);
    float4 Tv = B.T[j];
    float3 T = xyz(Tv);
    float3 L = xyz(B.L[j]);
    const uint4 meta = B.meta[j];
    const uint32_t key = meta.z;
    float3 g = f3(0.0f, 0.0f, 0.0f), Lt = g;
    if (ADJ) {
        const uint32_t pl = meta.x - (uint32_t)P.y0 * (uint32_t)P.width;
        g = mul(f3(A.dl[3 * pl], A.dl[3 * pl + 1], A.dl[3 * pl + 2]), A.inv_spp);
        Lt = xyz(A.Lt[j]);
    }
    if (rec == 0xffffffffu) {  // background, renderer.go:1492-1495 / :1326-1331
        // with resampled direct lighting the sky seen from a diffuse vertex is a candidate there
        if (!(P.restir > 0 && B.pp[j].w != 0.0f)) {
            L = add(L, had(T, sky_radiance(P, rd)));
            B.L[j] = make_float4(L.x, L.y, L.z, 0.0f);
        }
    } else {
        const float* gq = S.prims[rec].v;
        const uint32_t idbits = S.ids[rec];
        const int id = (int)(idbits & ~BC_SPHERE_BIT);
        const int mat = S.pmat[rec];
        float3 nrm;
        if (idbits & BC_SPHERE_BIT) {
            const float3 pos = add(ro, mul(rd, t));
            nrm = mul(sub(pos, f3(gq[0], gq[1], gq[2])), gq[5]);
        } else {
            nrm = unitv(cross(f3(gq[3], gq[4], gq[5]), f3(gq[6], gq[7], gq[8])));
        }
        const bool front = dot(rd, nrm) < 0.0f;
        const float3 N = front ? nrm : mul(nrm, -1.0f);
        const float3 p = add(ro, mul(rd, t));
        if (bounce == 0 && P.want_guides && meta.y == 0) {  // gpu_opencl_windows.go:811
            uint32_t* gg = guides + 3 * (size_t)(meta.x - (uint32_t)P.y0 * (uint32_t)P.width);
            gg[0] = __float_as_uint(t * sqrtf(dot(rd, rd)));
            gg[1] = pack_guide(N, mat);
            gg[2] = meta.w;
        }
        const bc_mat M = S.mats[mat];
        const float3 color = f3(M.color[0], M.color[1], M.color[2]);
        bool alive = true;
        if (M.kind == 3) {  // emissive: MIS weight on BSDF-sampled hits, renderer.go:1504-1511
            const float4 pp = B.pp[j];
            float w = 1.0f;
            if (P.restir > 0 && pp.w != 0.0f) {
                // registered emitters seen from a diffuse vertex are resampling candidates there
                w = (id >= 0 && S.light_of[id] >= 0) ? 0.0f : 1.0f;
            } else if (P.mis && pp.w != 0.0f && id >= 0) {
                const float lpdf = light_dir_pdf(S, P, id, xyz(pp), p, N);
                const float a2 = Tv.w * Tv.w, b2 = lpdf * lpdf;
                w = (a2 + b2 <= (float)1e-20) ? 0.0f : a2 / (a2 + b2);
            }
            L = add(L, mul(had(T, f3(M.emission[0], M.emission[1], M.emission[2])), w));
            B.L[j] = make_float4(L.x, L.y, L.z, 0.0f);
            if (ADJ) adj_emission(A, mat, g, mul(T, w));
            alive = false;
        } else if (M.kind == 0) {  // diffuse: NEE then cosine sampling
            if (ADJ) adj_color(A, mat, g, Lt, L, color);
            if (!ADJ && P.restir > 0 && P.spatial > 0 && bounce == 0) {
                // primary vertex: keep the reservoir; k_spatial combines it with its neighbours'
                Reservoir rs;
                ris_sample(S, P, p, N, color, key, bounce, rs);
                float W = 0.0f;
                if (rs.comp >= 0) {
                    W = rs.wsum / ((float)rs.M * rs.ph);
                    if (rs.comp == 1) W = W * (rs.dist2 / rs.cosL);
                }
                const float3 f = mul(color, INVPI_F);
                const uint32_t code = (uint32_t)(rs.comp + 1) | ((uint32_t)(rs.li < 0 ? 0 : rs.li) << 2);
                B.rv[0][j] = make_float4(p.x, p.y, p.z, t);
                B.rv[1][j] = make_float4(N.x, N.y, N.z, __uint_as_float(code));
                B.rv[2][j] = make_float4(f.x, f.y, f.z, W);
                B.rv[3][j] = make_float4(rs.y.x, rs.y.y, rs.y.z, __int_as_float(rs.M));
                B.rv[4][j] = make_float4(rs.lnrm.x, rs.lnrm.y, rs.lnrm.z, 0.0f);
            } else if (!ADJ && P.restir > 0) {
                float3 dl, y;
                bool point;
                if (ris_direct(S, P, p, N, color, key, bounce, dl, y, point)) {
                    const float3 c = had(T, dl);
                    const float3 from = add(p, mul(N, (float)2e-4));
                    if (point) {  // visible(), renderer.go:1333-1341
                        const float3 dd = sub(y, from);
                        const float dd2 = dot(dd, dd);
                        if (dd2 < (float)1e-12) {
                            L = add(L, c);
                            B.L[j] = make_float4(L.x, L.y, L.z, 0.0f);
                        } else {
                            const float dist = sqrtf(dd2);
                            const float3 dir = mul(dd, 1.0f / dist);
                            shadow_o = make_float4(from.x, from.y, from.z, dist - (float)1e-4);
                            shadow_d = make_float4(dir.x, dir.y, dir.z, 0.0f);
                            shadow_c = make_float4(c.x, c.y, c.z, 0.0f);
                            shadow = true;
                        }
                    } else {
                        shadow_o = make_float4(from.x, from.y, from.z, (float)1e30);
                        shadow_d = make_float4(y.x, y.y, y.z, 0.0f);
                        shadow_c = make_float4(c.x, c.y, c.z, 0.0f);
                        shadow = true;
                    }
                }
            } else if (P.nlight > 0) {   // directLight, renderer.go:1427-1465
                const float u = bc_rng(key, dim + BC_D_LIGHT);
                int li;
                float sel;
                if (!P.power) {  // chooseLight, renderer.go:1371-1393 (Intn -> floor(u*n))
                    li = (int)(u * (float)P.nlight);
                    if (li >= P.nlight) li = P.nlight - 1;
                    sel = 1.0f / (float)P.nlight;
                } else {
                    int lo = 0, hi = P.nlight;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (S.lights[P.lbase + mid].cdf >= u) hi = mid; else lo = mid + 1;
                    }
                    li = lo >= P.nlight ? P.nlight - 1 : lo;
                    sel = S.pdf[S.lights[P.lbase + li].prim];
                    if (sel <= 0.0f) sel = 1.0f / (float)P.nlight;
                }
                const bc_light& Lr = S.lights[P.lbase + li];
                float3 qpt, lnrm;
                const bool sph = Lr.kind == 0;
                if (sph) {  // randUnitVector, renderer.go:1273-1285
                    float3 ln = f3(0.0f, 0.0f, 1.0f);
                    for (int k = 0; k < BC_NTRY_SPH; k++) {
                        const float a = bc_rng(key, dim + BC_D_SPH + 2u * (uint32_t)k) * 2.0f - 1.0f;
                        const float b = bc_rng(key, dim + BC_D_SPH + 2u * (uint32_t)k + 1u) * 2.0f - 1.0f;
                        const float ss = a * a + b * b;
                        if (ss >= 1.0f || ss <= (float)1e-18) continue;
                        const float f = 2.0f * sqrtf(1.0f - ss);
                        ln = f3(a * f, b * f, 1.0f - 2.0f * ss);
                        break;
                    }
                    qpt = add(f3(Lr.g0[0], Lr.g0[1], Lr.g0[2]), mul(ln, Lr.g0[3]));
                    lnrm = ln;
                } else {  // uniform point on an emissive triangle
                    const float u1 = bc_rng(key, dim + BC_D_TRI), u2 = bc_rng(key, dim + BC_D_TRI + 1u);
                    const float su = sqrtf(u1);
                    const float b1 = su * (1.0f - u2), b2 = su * u2;
                    const float3 e1 = f3(Lr.g1[0], Lr.g1[1], Lr.g1[2]), e2 = f3(Lr.g2[0], Lr.g2[1], Lr.g2[2]);
                    qpt = add(add(f3(Lr.g0[0], Lr.g0[1], Lr.g0[2]), mul(e1, b1)), mul(e2, b2));
                    lnrm = unitv(cross(e1, e2));
                }
                const float3 to = sub(qpt, p);
                const float dist2 = dot(to, to);
                if (!(dist2 < (float)1e-8)) {
                    const float3 ld = unitv(to);
                    const float nd = rmax(0.0f, dot(N, ld));
                    const float cl = dot(mul(lnrm, -1.0f), ld);
                    const float cosL = sph ? rmax(0.0f, cl) : fabsf(cl);
                    if (!(nd <= 0.0f || cosL <= 0.0f)) {
                        const float lightPDF = sel * dist2 / (cosL * Lr.area);
                        if (lightPDF > 0.0f) {
                            const float bsdfPDF = nd / PI_F;
                            float w = 1.0f;
                            if (P.mis) {
                                const float a2 = lightPDF * lightPDF, b2 = bsdfPDF * bsdfPDF;
                                w = (a2 + b2 <= (float)1e-20) ? 0.0f : a2 / (a2 + b2);
                            }
                            const float3 f = mul(color, INVPI_F);
                            const float3 dl = mul(had(f, f3(Lr.le[0], Lr.le[1], Lr.le[2])), nd * w / lightPDF);
                            const float3 c = had(T, dl);  // throughput before the albedo update
                            const float3 wle = mul(had(T, f), nd * w / lightPDF);  // d c / d emission
                            // visible(), renderer.go:1333-1341
                            const float3 from = add(p, mul(N, (float)2e-4));
                            const float3 dd = sub(qpt, from);
                            const float dd2 = dot(dd, dd);
                            if (dd2 < (float)1e-12) {
                                L = add(L, c);
                                B.L[j] = make_float4(L.x, L.y, L.z, 0.0f);
                                if (ADJ) adj_emission(A, (int)__float_as_uint(Lr.g1[3]), g, wle);
                            } else {
                                if (ADJ) A.sg[j] = make_float4(wle.x, wle.y, wle.z, Lr.g1[3]);
                                const float dist = sqrtf(dd2);
                                const float3 dir = mul(dd, 1.0f / dist);
                                shadow_o = make_float4(from.x, from.y, from.z, dist - (float)1e-4);
                                shadow_d = make_float4(dir.x, dir.y, dir.z, 0.0f);
                                shadow_c = make_float4(c.x, c.y, c.z, 0.0f);
                                shadow = true;
                            }
                        }
                    }
                }
            }
            T = had(T, color);
            float x = 0.0f, y = 0.0f, ss = 0.0f;  // randCosHemisphere, renderer.go:1290-1311
            bool ok = false;
            for (int k = 0; k < BC_NTRY_COS && !ok; k++) {
                x = bc_rng(key, dim + BC_D_COS + 2u * (uint32_t)k) * 2.0f - 1.0f;
                y = bc_rng(key, dim + BC_D_COS + 2u * (uint32_t)k + 1u) * 2.0f - 1.0f;
                ss = x * x + y * y;
                ok = ss < 1.0f;
            }
            if (!ok) { x = 0.0f; y = 0.0f; ss = 0.0f; }
            const float z = sqrtf(rmax(0.0f, 1.0f - ss));
            const float3 nd = hemi_dir(N, x, y, z);
            Tv.w = rmax(0.0f, dot(N, nd)) / PI_F;
            B.pp[j] = make_float4(p.x, p.y, p.z, 1.0f);
            ro = add(p, mul(N, (float)2e-4));
            rd = nd;
        } else if (M.kind == 1) {  // metal, renderer.go:1520-1528
            float3 ru = f3(0.0f, 0.0f, 0.0f);
            for (int k = 0; k < BC_NTRY_METAL; k++) {
                const float a = bc_rng(key, dim + BC_D_METAL + 3u * (uint32_t)k) * 2.0f - 1.0f;
                const float b = bc_rng(key, dim + BC_D_METAL + 3u * (uint32_t)k + 1u) * 2.0f - 1.0f;
                const float c = bc_rng(key, dim + BC_D_METAL + 3u * (uint32_t)k + 2u) * 2.0f - 1.0f;
                const float l2 = a * a + b * b + c * c;
                if (l2 < 1.0f && l2 > (float)1e-18) { ru = f3(a, b, c); break; }
            }
            const float3 u = unitv(rd);
            const float3 refl = sub(u, mul(N, 2.0f * dot(u, N)));
            const float3 nd = unitv(add(refl, mul(ru, rclamp(M.rough, 0.0f, 1.0f))));
            if (dot(nd, N) <= 0.0f) {
                alive = false;
            } else {
                if (ADJ) adj_color(A, mat, g, Lt, L, color);
                B.pp[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                T = had(T, color);
                ro = add(p, mul(N, (float)2e-4));
                rd = nd;
            }
        } else if (M.kind == 2) {  // glass, renderer.go:1529-1549
            float eta = M.ior;
            if (eta <= 1.0f) eta = (float)1.5;
            const float ratio = front ? 1.0f / eta : eta;
            const float3 u = unitv(rd);
            const float cosT = rmin(dot(mul(u, -1.0f), N), 1.0f);
            const float sinT = sqrtf(rmax(0.0f, 1.0f - cosT * cosT));
            float r0 = (1.0f - ratio) / (1.0f + ratio);
            r0 *= r0;
            const float xx = 1.0f - cosT, x2 = xx * xx, x5 = x2 * x2 * xx;
            const float sch = r0 + (1.0f - r0) * x5;
            float3 nd;
            if (ratio * sinT > 1.0f || sch > bc_rng(key, dim + BC_D_GLASS)) {
                nd = sub(u, mul(N, 2.0f * dot(u, N)));
            } else {
                const float ct = rmin(dot(mul(u, -1.0f), N), 1.0f);
                const float3 perp = mul(add(u, mul(N, ct)), ratio);
                const float3 par = mul(N, -sqrtf(fabsf(1.0f - dot(perp, perp))));
                nd = add(perp, par);
            }
            if (ADJ) adj_color(A, mat, g, Lt, L, color);
            B.pp[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            T = had(T, color);
            ro = add(p, mul(nd, (float)2e-4));
            rd = nd;
        } else {
            alive = false;
        }
        if (alive && bounce >= rr) {  // Russian roulette, renderer.go:1550-1560
            const float mx = rmax(T.x, rmax(T.y, T.z));
            const float qq = rclamp(mx, (float)0.08, (float)0.95);
            if (bc_rng(key, dim + BC_D_RR) > qq) alive = false;
            else T = mul(T, 1.0f / qq);
        }
        if (alive) {
            B.o[j] = make_float4(ro.x, ro.y, ro.z, 0.0f);
            B.d[j] = make_float4(rd.x, rd.y, rd.z, 0.0f);
            B.T[j] = make_float4(T.x, T.y, T.z, Tv.w);
            alive_out = true;
        }
    }
    return alive_out;
}

}  // namespace bcd
