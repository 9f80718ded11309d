// integrator.hip -- wavefront path tracer (raygen / extend / shade / shadow / accumulate).
//
// Replaces the per-pixel megakernel renderPixel of the reference OpenCL path
// (gpu_opencl_windows.go:800-836) with separate kernels over SoA path state in HBM, and
// implements the CPU integrator's semantics (trace, renderer.go:1484-1563): next-event
// estimation with power-weighted light selection, power-heuristic MIS, Russian roulette
// from RRDepth, Halton + Cranley-Patterson camera jitter.
//
// Schedule of one chunk of rows: the pixels x samples of a pass are one batch of paths.
// For each bounce: extend (closest hit over the active queue) -> shade (emission, NEE
// shadow-ray generation, BSDF sampling, roulette; survivors are appended to the next
// queue and shadow rays to the shadow queue with one atomic per wave) -> shadow (any
// hit; unoccluded contributions are added to the path radiance).  After the last bounce
// the accumulate kernel adds the pass's samples to the pixel sums in sample order, so the
// result is independent of scheduling.  Queue lengths live in device memory and every
// kernel walks its queue with a grid-stride loop: no host round trip between bounces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "api_internal.hpp"
#include "kernels.hpp"
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
};

struct PathBufs {
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
    uint32_t* q[2];
    uint32_t* sq;
    uint32_t* cnt;  // [0],[1] queue lengths, [2] shadow queue length
    unsigned long long* totals;  // [0] path segments, [1] shadow rays
};

struct SceneBufs {
    const bc_node8* nodes;
    const bc_prim* prims;
    const bc_mat* mats;
    const bc_light* lights;
    const float* pdf;
    const int32_t* light_of;
    int* err;
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

// Appends the calling lanes whose pred is true to a queue with one atomic per wave.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred, uint32_t& slot) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return 0;
    const uint32_t lane = __lane_id();
    const uint32_t leader = (uint32_t)__builtin_ctzll(m);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, (int)leader);
    slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    return (uint32_t)__popcll(m);
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

// ------------------------------------------------------------------------------ raygen
// Camera ray of adaptivePixel (renderer.go:1940-1950) with adaptive sampling off.
__global__ __launch_bounds__(kBlock) void k_raygen(RenderParams P, PathBufs B, float* __restrict__ acc,
                                                   uint32_t* __restrict__ guides) {
    const int n = P.ns * P.npix;
    const int stride = gridDim.x * kBlock;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        B.cnt[0] = (uint32_t)n;
        B.cnt[1] = 0;
        B.cnt[2] = 0;
    }
    const double g = 1.324717957244746, a1 = 1.0 / g, a2 = 1.0 / (g * g);
    const float3 origin = f3(P.origin[0], P.origin[1], P.origin[2]);
    const float3 lower = f3(P.lower[0], P.lower[1], P.lower[2]);
    const float3 horiz = f3(P.horiz[0], P.horiz[1], P.horiz[2]);
    const float3 vert = f3(P.vert[0], P.vert[1], P.vert[2]);
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += stride) {
        const int sl = j / P.npix, pl = j - sl * P.npix;
        const int sample = P.s0 + sl;
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
        B.q[0][j] = (uint32_t)j;
        if (P.s0 == 0 && sl == 0) {
            float* a = acc + 4 * (size_t)pl;
            a[0] = 0.0f; a[1] = 0.0f; a[2] = 0.0f; a[3] = 0.0f;
            if (P.want_guides) {
                uint32_t* gg = guides + 3 * (size_t)pl;
                gg[0] = 0u; gg[1] = 0u; gg[2] = 0u;
            }
        }
    }
}

// ------------------------------------------------------------------------------ extend
template <int STACK>
__global__ __launch_bounds__(kBlock) void k_extend(SceneBufs S, PathBufs B, int cur) {
    __shared__ uint2 stk[STACK * kBlock];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        B.cnt[cur ^ 1] = 0;  // next queue, filled by shade of this bounce
        B.cnt[2] = 0;        // shadow queue of this bounce
    }
    const uint32_t n = B.cnt[cur];
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t* q = B.q[cur];
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        const uint32_t j = q[i];
        const float4 o = B.o[j], d = B.d[j];
        Ray r;
        r.ox = o.x; r.oy = o.y; r.oz = o.z;
        r.dx = d.x; r.dy = d.y; r.dz = d.z;
        r.tmin = 1e-4f;
        const HitRec h = traverse<false, STACK>(S.nodes, S.prims, r, 1e30f, stk + threadIdx.x, kBlock, S.err);
        B.hit[j] = make_float2(h.t, __uint_as_float(h.id >= 0 ? h.rec : 0xffffffffu));
        B.nseg[j] += 1u;
    }
}

// ------------------------------------------------------------------------------ shade
struct LightPick {
    bool ok;
    float3 q, lnrm;
    bool sphere;
    float sel, area;
    int mat;
};

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

__global__ __launch_bounds__(kBlock) void k_shade(SceneBufs S, PathBufs B, RenderParams P, int cur, int bounce,
                                                  uint32_t* __restrict__ guides) {
    const uint32_t n = B.cnt[cur];
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t* q = B.q[cur];
    uint32_t* qn = B.q[cur ^ 1];
    const uint32_t dim = BC_DIM_BOUNCE0 + BC_DIMS_PER_BOUNCE * (uint32_t)bounce;
    const int rr = P.rr_depth < 1 ? 4 : P.rr_depth;
    const uint32_t nloop = (n + stride - 1) / stride * stride;  // uniform trip count for wave ops
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < nloop; i += stride) {
        bool cont = false, shadow = false;
        uint32_t j = 0;
        float4 shadow_o = make_float4(0, 0, 0, 0), shadow_d = make_float4(0, 0, 0, 0), shadow_c = make_float4(0, 0, 0, 0);
        if (i < n) {
            j = q[i];
            const float2 hv = B.hit[j];
            const uint32_t rec = __float_as_uint(hv.y);
            const float t = hv.x;
            float3 ro = xyz(B.o[j]), rd = xyz(B.d[j]);
            float4 Tv = B.T[j];
            float3 T = xyz(Tv);
            float3 L = xyz(B.L[j]);
            const uint4 meta = B.meta[j];
            const uint32_t key = meta.z;
            if (rec == 0xffffffffu) {  // background, renderer.go:1492-1495 / :1326-1331
                float3 bg;
                if (P.sky_mode == 1) {
                    bg = f3(P.sky[0], P.sky[1], P.sky[2]);
                } else {
                    const float3 u = unitv(rd);
                    const float tt = (float)0.5 * (u.y + 1.0f);
                    bg = add(mul(f3((float)0.06, (float)0.08, (float)0.12), 1.0f - tt),
                             mul(f3((float)0.32, (float)0.48, (float)0.72), tt));
                }
                L = add(L, had(T, bg));
                B.L[j] = make_float4(L.x, L.y, L.z, 0.0f);
            } else {
                const float4* pr = reinterpret_cast<const float4*>(S.prims + rec);
                const float4 q0 = pr[0], q1 = pr[1], q2 = pr[2];
                const uint32_t idbits = __float_as_uint(q2.y);
                const int id = (int)(idbits & ~BC_SPHERE_BIT);
                const int mat = (int)__float_as_uint(q2.z);
                float3 nrm;
                if (idbits & BC_SPHERE_BIT) {
                    const float3 pos = add(ro, mul(rd, t));
                    nrm = mul(sub(pos, f3(q0.x, q0.y, q0.z)), q1.y);
                } else {
                    nrm = unitv(cross(f3(q0.w, q1.x, q1.y), f3(q1.z, q1.w, q2.x)));
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
                    if (P.mis && pp.w != 0.0f && id >= 0) {
                        const float lpdf = light_dir_pdf(S, P, id, xyz(pp), p, N);
                        const float a2 = Tv.w * Tv.w, b2 = lpdf * lpdf;
                        w = (a2 + b2 <= (float)1e-20) ? 0.0f : a2 / (a2 + b2);
                    }
                    L = add(L, mul(had(T, f3(M.emission[0], M.emission[1], M.emission[2])), w));
                    B.L[j] = make_float4(L.x, L.y, L.z, 0.0f);
                    alive = false;
                } else if (M.kind == 0) {  // diffuse: NEE then cosine sampling
                    if (P.nlight > 0) {   // directLight, renderer.go:1427-1465
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
                                if (S.lights[mid].cdf >= u) hi = mid; else lo = mid + 1;
                            }
                            li = lo >= P.nlight ? P.nlight - 1 : lo;
                            sel = S.pdf[S.lights[li].prim];
                            if (sel <= 0.0f) sel = 1.0f / (float)P.nlight;
                        }
                        const bc_light& Lr = S.lights[li];
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
                                    // visible(), renderer.go:1333-1341
                                    const float3 from = add(p, mul(N, (float)2e-4));
                                    const float3 dd = sub(qpt, from);
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
                    const float sign = N.z < 0.0f ? -1.0f : 1.0f;
                    const float a = -1.0f / (sign + N.z);
                    const float b = N.x * N.y * a;
                    const float3 tt = f3(1.0f + sign * N.x * N.x * a, sign * b, -sign * N.x);
                    const float3 bt = f3(b, sign + N.y * N.y * a, -N.y);
                    const float3 nd = add(add(mul(tt, x), mul(bt, y)), mul(N, z));
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
                    cont = bounce + 1 < P.bounces;
                }
            }
        }
        uint32_t slot = 0;
        wave_append(&B.cnt[cur ^ 1], cont, slot);
        if (cont) qn[slot] = j;
        uint32_t sslot = 0;
        wave_append(&B.cnt[2], shadow, sslot);
        if (shadow) {
            B.sq[sslot] = j;
            B.so[j] = shadow_o;
            B.sd[j] = shadow_d;
            B.sc[j] = shadow_c;
        }
    }
}

// ------------------------------------------------------------------------------ shadow
// Occluded(p + N*2e-4 -> q, 1e-4, dist - 1e-4) of visible() (renderer.go:1333-1341, 851-895).
template <int STACK>
__global__ __launch_bounds__(kBlock) void k_shadow(SceneBufs S, PathBufs B) {
    __shared__ uint2 stk[STACK * kBlock];
    const uint32_t n = B.cnt[2];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&B.totals[1], (unsigned long long)n);
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        const uint32_t j = B.sq[i];
        const float4 o = B.so[j], d = B.sd[j];
        Ray r;
        r.ox = o.x; r.oy = o.y; r.oz = o.z;
        r.dx = d.x; r.dy = d.y; r.dz = d.z;
        r.tmin = (float)1e-4;
        const HitRec h = traverse<true, STACK>(S.nodes, S.prims, r, o.w, stk + threadIdx.x, kBlock, S.err);
        if (h.id < 0) {
            const float4 c = B.sc[j];
            const float4 l = B.L[j];
            B.L[j] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.0f);
        }
    }
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off);
    return v;
}

// ------------------------------------------------------------------------------ accumulate
// Per-pixel sum in sample order with clampFirefly (renderer.go:1917-1926, 1951-1953).
__global__ __launch_bounds__(kBlock) void k_accumulate(RenderParams P, PathBufs B, float4* __restrict__ acc) {
    const int stride = gridDim.x * kBlock;
    const int nloop = (P.npix + stride - 1) / stride * stride;
    for (int pl = blockIdx.x * kBlock + threadIdx.x; pl < nloop; pl += stride) {
        uint32_t segs = 0;
        if (pl < P.npix) {
            float4 a = acc[pl];
            for (int sl = 0; sl < P.ns; sl++) {
                const int j = sl * P.npix + pl;
                const float4 l = B.L[j];
                float cx = l.x, cy = l.y, cz = l.z;
                if (P.firefly > 0.0f) {
                    const float lim = P.firefly;
                    const float m = rmax(cx, rmax(cy, cz));
                    if (!(m <= lim || m <= 0.0f)) {
                        const float k = lim / m;
                        cx = cx * k; cy = cy * k; cz = cz * k;
                    }
                }
                a.x = a.x + cx; a.y = a.y + cy; a.z = a.z + cz;
                const uint32_t ns = B.nseg[j];
                a.w = a.w + (float)ns;
                segs += ns;
            }
            acc[pl] = a;
        }
        const uint32_t tot = wave_sum(segs);
        if (__lane_id() == 0 && tot) atomicAdd(&B.totals[0], (unsigned long long)tot);
    }
}

__global__ __launch_bounds__(kBlock) void k_finalize(RenderParams P, const float4* __restrict__ acc,
                                                     float4* __restrict__ rgba) {
    const float inv = 1.0f / (float)P.spp;
    for (int pl = blockIdx.x * kBlock + threadIdx.x; pl < P.npix; pl += gridDim.x * kBlock) {
        const float4 a = acc[pl];
        rgba[pl] = make_float4(a.x * inv, a.y * inv, a.z * inv, a.w);
    }
}

}  // namespace bcd

// ------------------------------------------------------------------------------ host side
struct RenderState {
    size_t cap_paths = 0, cap_pix = 0;
    void* buf[16] = {};
    size_t bufcap[16] = {};
    bcd::PathBufs B{};
    float4* acc = nullptr;
    float4* rgba_tmp = nullptr;
    uint32_t* guide_tmp = nullptr;
    unsigned long long* totals = nullptr;
};

namespace {

bool ensure_state(bc_ctx* c, size_t paths, size_t npix, bool guides) {
    if (!c->rs) c->rs = new RenderState();
    RenderState& R = *c->rs;
    struct Want { int slot; size_t bytes; };
    const Want w[] = {
        {0, paths * 16}, {1, paths * 16}, {2, paths * 16}, {3, paths * 16}, {4, paths * 16}, {5, paths * 16},
        {6, paths * 4},  {7, paths * 8},  {8, paths * 16}, {9, paths * 16}, {10, paths * 16}, {11, paths * 4},
        {12, paths * 4}, {13, paths * 4}, {14, 64},        {15, npix * 16 * 2 + (guides ? npix * 12 : 0) + 64},
    };
    for (const Want& x : w)
        if (!bc_ensure(c, R.buf[x.slot], R.bufcap[x.slot], std::max<size_t>(x.bytes, 64), "render buffers")) return false;
    R.B.o = (float4*)R.buf[0];
    R.B.d = (float4*)R.buf[1];
    R.B.T = (float4*)R.buf[2];
    R.B.L = (float4*)R.buf[3];
    R.B.pp = (float4*)R.buf[4];
    R.B.meta = (uint4*)R.buf[5];
    R.B.nseg = (uint32_t*)R.buf[6];
    R.B.hit = (float2*)R.buf[7];
    R.B.so = (float4*)R.buf[8];
    R.B.sd = (float4*)R.buf[9];
    R.B.sc = (float4*)R.buf[10];
    R.B.q[0] = (uint32_t*)R.buf[11];
    R.B.q[1] = (uint32_t*)R.buf[12];
    R.B.sq = (uint32_t*)R.buf[13];
    R.B.cnt = (uint32_t*)R.buf[14];
    R.B.totals = (unsigned long long*)((char*)R.buf[14] + 32);
    R.acc = (float4*)R.buf[15];
    R.rgba_tmp = R.acc + npix;
    R.guide_tmp = (uint32_t*)(R.rgba_tmp + npix);
    R.cap_paths = paths;
    R.cap_pix = npix;
    return true;
}

template <int STACK>
void launch_bounce(const bcd::SceneBufs& S, const bcd::PathBufs& B, const bcd::RenderParams& P, int cur, int bounce,
                   uint32_t* guides, int grid, hipStream_t st) {
    hipLaunchKernelGGL((bcd::k_extend<STACK>), dim3(grid), dim3(bcd::kBlock), 0, st, S, B, cur);
    hipLaunchKernelGGL(bcd::k_shade, dim3(grid), dim3(bcd::kBlock), 0, st, S, B, P, cur, bounce, guides);
    hipLaunchKernelGGL((bcd::k_shadow<STACK>), dim3(grid), dim3(bcd::kBlock), 0, st, S, B);
}

}  // namespace

void bc_render_release(bc_ctx* c) {
    if (!c || !c->rs) return;
    for (void* p : c->rs->buf)
        if (p) (void)hipFree(p);
    delete c->rs;
    c->rs = nullptr;
}

extern "C" int bc_render_device(bc_ctx* c, const bc_render_desc* d, int y0, int rows, float* d_rgba,
                                uint32_t* d_guides, uint64_t* rays_out, uint64_t* shadow_out) {
    if (!c || !d) return BC_ERR_ARG;
    if (!c->has_scene) { c->err = "no scene uploaded"; return BC_ERR_SCENE; }
    if (d->width <= 0 || d->height <= 0 || d->spp <= 0 || d->bounces < 0 || d->bounces > 1024 || y0 < 0 || rows <= 0 ||
        y0 + rows > d->height || !d_rgba || d->sampler < 0 || d->sampler > 2) {
        c->err = "invalid render description or row range";
        return BC_ERR_ARG;
    }
    (void)hipSetDevice(c->device);
    const size_t npix = (size_t)rows * (size_t)d->width;
    const size_t max_paths = std::max<size_t>(npix, (size_t)1 << 22);
    int spp_pass = d->samples_per_pass > 0 ? d->samples_per_pass : (int)std::max<size_t>(1, max_paths / npix);
    spp_pass = std::min(spp_pass, d->spp);
    const size_t paths = npix * (size_t)spp_pass;
    if (!ensure_state(c, paths, npix, d_guides != nullptr)) return BC_ERR_NOMEM;
    RenderState& R = *c->rs;
    bcd::SceneBufs S{c->d_nodes, c->d_prims, c->d_mats, c->d_lights, c->d_pdf, c->d_light_of, c->d_err};
    bcd::RenderParams P{};
    P.width = d->width; P.height = d->height; P.spp = d->spp; P.bounces = d->bounces; P.rr_depth = d->rr_depth;
    P.sampler = d->sampler; P.mis = d->mis != 0; P.power = d->power_lights != 0; P.sky_mode = d->sky_mode;
    P.want_guides = d_guides != nullptr;
    P.firefly = d->firefly_clamp;
    for (int k = 0; k < 3; k++) {
        P.sky[k] = d->sky_color[k];
        P.origin[k] = (float)d->origin[k]; P.lower[k] = (float)d->lower[k];
        P.horiz[k] = (float)d->horiz[k]; P.vert[k] = (float)d->vert[k];
    }
    P.seed = d->seed;
    P.y0 = y0;
    P.npix = (int32_t)npix;
    P.nmat = c->nmat;
    P.nlight = c->nlight;
    const int grid_paths = (int)std::min<size_t>((paths + bcd::kBlock - 1) / bcd::kBlock, (size_t)c->num_cus * 8);
    const int grid_pix = (int)std::min<size_t>((npix + bcd::kBlock - 1) / bcd::kBlock, (size_t)c->num_cus * 8);
    hipStream_t st = c->stream;
    if (!bc_check(c, hipMemsetAsync(R.B.totals, 0, 16, st), "clear totals")) return BC_ERR_DEVICE;
    uint32_t* guides = d_guides ? d_guides : R.guide_tmp;
    for (int s0 = 0; s0 < d->spp; s0 += spp_pass) {
        P.s0 = s0;
        P.ns = std::min(spp_pass, d->spp - s0);
        hipLaunchKernelGGL(bcd::k_raygen, dim3(grid_paths), dim3(bcd::kBlock), 0, st, P, R.B, (float*)R.acc, guides);
        for (int b = 0; b < d->bounces; b++) {
            const int cur = b & 1;
            switch (c->stack) {
            case 8: launch_bounce<8>(S, R.B, P, cur, b, guides, grid_paths, st); break;
            case 16: launch_bounce<16>(S, R.B, P, cur, b, guides, grid_paths, st); break;
            case 24: launch_bounce<24>(S, R.B, P, cur, b, guides, grid_paths, st); break;
            default: launch_bounce<32>(S, R.B, P, cur, b, guides, grid_paths, st); break;
            }
        }
        hipLaunchKernelGGL(bcd::k_accumulate, dim3(grid_pix), dim3(bcd::kBlock), 0, st, P, R.B, R.acc);
        if (!bc_check(c, hipGetLastError(), "render launch")) return BC_ERR_DEVICE;
    }
    hipLaunchKernelGGL(bcd::k_finalize, dim3(grid_pix), dim3(bcd::kBlock), 0, st, P, (const float4*)R.acc,
                       (float4*)d_rgba);
    unsigned long long tot[2] = {0, 0};
    if (!bc_check(c, hipGetLastError(), "finalize launch") ||
        !bc_check(c, hipMemcpyAsync(tot, R.B.totals, 16, hipMemcpyDeviceToHost, st), "read totals") ||
        !bc_check(c, hipStreamSynchronize(st), "render"))
        return BC_ERR_DEVICE;
    if (rays_out) *rays_out = tot[0];
    if (shadow_out) *shadow_out = tot[1];
    return bc_check_kernel_error(c);
}

extern "C" int bc_render(bc_ctx* c, const bc_render_desc* d, int y0, int rows, float* rgba, uint32_t* guides,
                         uint64_t* rays_out, const volatile int32_t* cancel) {
    if (!c || !d || !rgba) return BC_ERR_ARG;
    if (cancel && *cancel) { c->err = "render cancelled"; return BC_ERR_CANCELLED; }
    if (d->width <= 0 || rows <= 0) { c->err = "invalid render description or row range"; return BC_ERR_ARG; }
    const size_t npix = (size_t)rows * (size_t)d->width;
    void* out = nullptr;
    size_t cap = 0;
    if (!bc_ensure(c, out, cap, npix * 16 + (guides ? npix * 12 : 0) + 64, "render output")) return BC_ERR_NOMEM;
    uint32_t* dg = guides ? (uint32_t*)((char*)out + npix * 16) : nullptr;
    int rc = bc_render_device(c, d, y0, rows, (float*)out, dg, rays_out, nullptr);
    if (rc == BC_OK) {
        if (!bc_check(c, hipMemcpy(rgba, out, npix * 16, hipMemcpyDeviceToHost), "copy image") ||
            (guides && !bc_check(c, hipMemcpy(guides, dg, npix * 12, hipMemcpyDeviceToHost), "copy guides")))
            rc = BC_ERR_DEVICE;
    }
    (void)hipFree(out);
    return rc;
}
