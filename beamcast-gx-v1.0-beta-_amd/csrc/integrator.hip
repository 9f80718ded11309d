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
#include <cstddef>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "api_internal.hpp"
#include "kernels.hpp"
#include "integrator_device.hpp"
#include "trace_device.hpp"

namespace bcd {

// Resident waves per SIMD that a traversal kernel's register allocation is held to (7: at most
// 72 VGPRs, 6: at most 80).  Traversal is latency-bound, so more resident waves hide more of the
// node and primitive fetch latency; the persistent extension kernel, which also regenerates
// paths in place, spills a few registers outside its traversal loop at 6 and is still faster.
#define BC_WAVES(N) __attribute__((amdgpu_waves_per_eu(N, N)))
#ifndef BC_EXTEND_WAVES
#define BC_EXTEND_WAVES 7
#endif
#ifndef BC_P_EXTEND_WAVES
#define BC_P_EXTEND_WAVES 6
#endif

constexpr uint32_t kIdle = 0xffffffffu;

// Slots of segment `seg` (of `ss` slots each) among the first `total` slots.
__device__ __forceinline__ uint32_t seg_slots(uint32_t seg, uint32_t ss, uint32_t total) {
    const uint32_t first = seg * ss;
    return first >= total ? 0u : (total - first < ss ? total - first : ss);
}

// Sum over a workgroup of NW waves; the result is valid in thread 0.
template <int NW = kWaves>
__device__ __forceinline__ unsigned long long block_sum(unsigned long long v, unsigned long long* lds) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += (unsigned long long)__shfl_xor((long long)v, off);
    if (__lane_id() == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    unsigned long long t = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < NW; k++) t += lds[k];
    return t;
}

// ------------------------------------------------------------------------------ raygen
// Slot j = (sample - s0) * npix + pixel; the pass's segments start full.
__global__ __launch_bounds__(kBlock) void k_raygen(RenderParams P, PathBufs B, float* __restrict__ acc,
                                                   uint32_t* __restrict__ guides, uint32_t nseg) {
    const uint32_t n = (uint32_t)P.ns * (uint32_t)P.npix;
    for (uint32_t seg = xcd_chunk(blockIdx.x, gridDim.x, B.xg); seg < nseg; seg += gridDim.x) {
        const uint32_t cnt = seg_slots(seg, B.seg, n), base = seg * B.seg;
        for (uint32_t k = threadIdx.x; k < cnt; k += kBlock) {
            const uint32_t j = base + k;
            const uint32_t sl = j / (uint32_t)P.npix, pl = j - sl * (uint32_t)P.npix;
            init_camera_path(P, B, j, (int)pl, P.s0 + (int)sl);
            B.q[0][j] = j;
            if (P.spatial > 0) B.rv[0][j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // no reservoir yet
            if (P.s0 == 0 && sl == 0) {
                float* a = acc + 4 * (size_t)pl;
                a[0] = 0.0f; a[1] = 0.0f; a[2] = 0.0f; a[3] = 0.0f;
                if (P.want_guides) {
                    uint32_t* gg = guides + 3 * (size_t)pl;
                    gg[0] = 0u; gg[1] = 0u; gg[2] = 0u;
                }
            }
        }
        if (threadIdx.x == 0) B.qc[0][seg] = cnt;
    }
}

// The run [pos, end) of a queue of n rays that wave wib of a workgroup takes in the lane-refill
// kernels: consecutive runs of a whole number of wave widths (n / 4 rounded up to 64), so a short
// queue fills as few waves as it can, as one ray per thread would, instead of spreading a few
// rays over every wave.
__device__ __forceinline__ void queue_run(uint32_t n, uint32_t wib, uint32_t& pos, uint32_t& end) {
    const uint32_t chunk = (n + 64u * kWaves - 1u) / (64u * kWaves) * 64u;
    pos = min(n, wib * chunk);
    end = min(n, (wib + 1u) * chunk);
}

// ------------------------------------------------------------------------------ extend
// Closest hit of every queued path; the hit keeps the record index (the primitive index is
// looked up by shading only where it needs it).
__device__ __forceinline__ void store_extension(const PathBufs& B, uint32_t j, const TravState& st) {
    B.hit[j] = make_float2(st.best.t, __uint_as_float(st.best.rec));
    B.nseg[j] += 1u;
}

__device__ __forceinline__ void extension_ray(const PathBufs& B, uint32_t j, Ray& r) {
    const float4 o = B.o[j], d = B.d[j];
    r.ox = o.x; r.oy = o.y; r.oz = o.z;
    r.dx = d.x; r.dy = d.y; r.dz = d.z;
    r.tmin = 1e-4f;
}

__global__ __launch_bounds__(kBlock) BC_WAVES(BC_EXTEND_WAVES) void k_extend(SceneBufs S, PathBufs B, int cur, uint32_t nseg, uint32_t root) {
    extern __shared__ uint2 stk[];  // S.stack entries per lane
    const uint32_t* q = B.q[cur];
    for (uint32_t seg = xcd_chunk(blockIdx.x, gridDim.x, B.xg); seg < nseg; seg += gridDim.x) {
        const uint32_t n = B.qc[cur][seg], base = seg * B.seg;
        for (uint32_t k = threadIdx.x; k < n; k += kBlock) {
            const uint32_t j = q[base + k];
            Ray r;
            extension_ray(B, j, r);
            TravState st;
            trav_init(st, r, 1e30f, root);
            while (!trav_step<false>(st, S.nodes, S.prims, S.ids, stk + threadIdx.x, kBlock, S.stack, S.err)) {
            }
            store_extension(B, j, st);
        }
    }
}

// Extension rays with lane refill, for the second segment (whose rays leave scattering vertices in
// all directions; see launch_extend): wave w of a workgroup takes its run (queue_run) of the path queue
// of each segment the workgroup owns, and whenever at least S.extend_refill of its lanes are idle
// hands them the next paths of the run (ballot + prefix count), as k_shadow_refill does for
// shadow rays.  Each path is traced by one lane; results equal k_extend's.
__global__ __launch_bounds__(kBlock) BC_WAVES(BC_EXTEND_WAVES) void k_extend_refill(SceneBufs S, PathBufs B, int cur, uint32_t nseg,
                                                                                   uint32_t root) {
    extern __shared__ uint2 stk[];  // S.stack entries per lane
    const int lane = (int)(threadIdx.x & 63u);
    const uint32_t wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int refill = S.extend_refill;
    const uint32_t* q = B.q[cur];
    uint2* my = stk + threadIdx.x;
    // wave-uniform cursor: segment, this wave's range [pos, end) of its queue
    uint32_t seg = xcd_chunk(blockIdx.x, gridDim.x, B.xg), pos = 0, end = 0, base = 0;
    bool started = false;
    TravState st;
    bool active = false;
    uint32_t j = 0;
    for (;;) {
        if (active && trav_step<false>(st, S.nodes, S.prims, S.ids, my, kBlock, S.stack, S.err)) {
            store_extension(B, j, st);
            active = false;
        }
        while (pos >= end && seg < nseg) {
            if (started) {
                seg += gridDim.x;
                if (seg >= nseg) break;
            }
            started = true;
            const uint32_t n = B.qc[cur][seg];
            base = seg * B.seg;
            queue_run(n, wib, pos, end);
        }
        const uint64_t idle = __ballot(!active);
        const int nidle = __popcll(idle);
        if (pos >= end) {
            if (nidle == 64) break;
            continue;
        }
        if (nidle >= refill) {
            if (!active) {
                const uint32_t k = pos + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                if (k < end) {
                    j = q[base + k];
                    Ray r;
                    extension_ray(B, j, r);
                    trav_init(st, r, 1e30f, root);
                    active = true;
                }
            }
            pos = min(end, pos + (uint32_t)nidle);
        }
    }
}

// ------------------------------------------------------------------------------ shade
__global__ __launch_bounds__(kBlock) void k_shade(SceneBufs S, PathBufs B, RenderParams P, int cur, int bounce,
                                                  uint32_t* __restrict__ guides, uint32_t nseg) {
    __shared__ uint32_t lds[2 * 2 * kWaves];
    uint32_t parity = 0;
    const uint32_t* q = B.q[cur];
    uint32_t* qn = B.q[cur ^ 1];
    for (uint32_t seg = xcd_chunk(blockIdx.x, gridDim.x, B.xg); seg < nseg; seg += gridDim.x) {
        const uint32_t n = B.qc[cur][seg], base = seg * B.seg;
        uint32_t bc = 0, bs = 0;
        for (uint32_t r0 = 0; r0 < n; r0 += kBlock) {
            const uint32_t k = r0 + threadIdx.x;
            bool cont = false, shadow = false;
            uint32_t j = 0;
            float4 so = make_float4(0, 0, 0, 0), sd = so, sc = so;
            if (k < n) {
                j = q[base + k];
                cont = shade_path<false>(S, B, P, j, bounce, guides, shadow, so, sd, sc, Adj{}) && bounce + 1 < P.bounces;
            }
            const bool pr[2] = {cont, shadow};
            uint32_t rk[2], tot[2];
            block_rank<2>(pr, rk, tot, lds, parity);
            if (cont) qn[base + bc + rk[0]] = j;
            if (shadow) {
                B.sq[base + bs + rk[1]] = j;
                B.so[j] = so;
                B.sd[j] = sd;
                B.sc[j] = sc;
            }
            bc += tot[0];
            bs += tot[1];
        }
        if (threadIdx.x == 0) {
            B.qc[cur ^ 1][seg] = bc;
            B.sqc[seg] = bs;
        }
    }
}

// ------------------------------------------------------------------------------ spatial reuse
// After the primary-vertex shading of a pass: every path with a reservoir combines it with up to
// P.spatial neighbours' (same sample index, pixel offsets of at most P.radius within the rendered
// rows, normals within 25 degrees and hit distances within 10%), picks one sample with the
// streaming rule of the reservoir, and queues its shadow ray with contribution F(y) W.  The
// segment's shadow queue (empty after the primary shading in this mode) is written here.
__global__ __launch_bounds__(kBlock) void k_spatial(SceneBufs S, PathBufs B, RenderParams P, uint32_t nseg) {
    __shared__ uint32_t lds[2 * 2 * kWaves];
    uint32_t parity = 0;
    const uint32_t n = (uint32_t)P.ns * (uint32_t)P.npix;
    const int R = P.radius, span = 2 * P.radius + 1;
    const int rows = P.npix / P.width;
    for (uint32_t seg = xcd_chunk(blockIdx.x, gridDim.x, B.xg); seg < nseg; seg += gridDim.x) {
        const uint32_t cnt = seg_slots(seg, B.seg, n), base = seg * B.seg;
        uint32_t bs = 0;
        for (uint32_t r0 = 0; r0 < cnt; r0 += kBlock) {
            const uint32_t k = r0 + threadIdx.x;
            bool shadow = false;
            uint32_t j = 0;
            float4 so = make_float4(0, 0, 0, 0), sdv = so, sc = so;
            if (k < cnt) {
                j = base + k;
                const float4 a0 = B.rv[0][j];
                if (a0.w > 0.0f) {
                    const float4 a1 = B.rv[1][j], a2 = B.rv[2][j];
                    const float3 p = xyz(a0), N = xyz(a1), f = xyz(a2);
                    const uint32_t key = B.meta[j].z;
                    const uint32_t sl = j / (uint32_t)P.npix, pl = j - sl * (uint32_t)P.npix;
                    const int px = (int)(pl % (uint32_t)P.width), py = (int)(pl / (uint32_t)P.width);
                    float wsum = 0.0f, ph_s = 0.0f;
                    int Mtot = 0, comp_s = -1;
                    float3 F_s = f3(0.0f, 0.0f, 0.0f), y_s = F_s;
                    for (int i = 0; i <= P.spatial; i++) {
                        uint32_t jn = j;
                        if (i > 0) {
                            const uint32_t hx = bc_rng_u32(key, BC_DIM_SPATIAL0 + 4u * (uint32_t)i);
                            const uint32_t hy = bc_rng_u32(key, BC_DIM_SPATIAL0 + 4u * (uint32_t)i + 1u);
                            const int ox = (int)(hx % (uint32_t)span) - R, oy = (int)(hy % (uint32_t)span) - R;
                            const int nx = px + ox, ny = py + oy;
                            if ((ox == 0 && oy == 0) || nx < 0 || nx >= P.width || ny < 0 || ny >= rows) continue;
                            jn = sl * (uint32_t)P.npix + (uint32_t)(ny * P.width + nx);
                            const float4 b0 = B.rv[0][jn];
                            if (!(b0.w > 0.0f)) continue;
                            const float3 Nn = xyz(B.rv[1][jn]);
                            if (dot(N, Nn) < (float)0.9063 || fabsf(b0.w - a0.w) > (float)0.1 * a0.w) continue;
                        }
                        const uint32_t code = __float_as_uint(B.rv[1][jn].w);
                        const float4 c2 = B.rv[2][jn], c3 = B.rv[3][jn];
                        const int Mi = __float_as_int(c3.w);
                        Mtot += Mi;
                        const int comp = (int)(code & 3u) - 1;
                        if (comp < 0) continue;
                        const int li = (int)(code >> 2);
                        const float3 y = xyz(c3);
                        const float3 F = spatial_eval(S, P, p, N, f, comp, li, y, xyz(B.rv[4][jn]));
                        const float ph = lum(F);
                        const float w = ph * c2.w * (float)Mi;
                        if (w > 0.0f) {
                            wsum = wsum + w;
                            if (bc_rng(key, BC_DIM_SPATIAL0 + 4u * (uint32_t)i + 2u) * wsum < w) {
                                comp_s = comp;
                                F_s = F;
                                ph_s = ph;
                                y_s = y;
                            }
                        }
                    }
                    if (comp_s >= 0) {
                        const float W = wsum / ((float)Mtot * ph_s);
                        const float3 c = mul(F_s, W);  // the primary vertex has unit throughput
                        const float3 from = add(p, mul(N, (float)2e-4));
                        if (comp_s == 1) {  // visible(), renderer.go:1333-1341
                            const float3 dd = sub(y_s, from);
                            const float dd2 = dot(dd, dd);
                            if (dd2 < (float)1e-12) {
                                const float4 l = B.L[j];
                                B.L[j] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.0f);
                            } else {
                                const float dist = sqrtf(dd2);
                                const float3 dir = mul(dd, 1.0f / dist);
                                so = make_float4(from.x, from.y, from.z, dist - (float)1e-4);
                                sdv = make_float4(dir.x, dir.y, dir.z, 0.0f);
                                sc = make_float4(c.x, c.y, c.z, 0.0f);
                                shadow = true;
                            }
                        } else {
                            so = make_float4(from.x, from.y, from.z, (float)1e30);
                            sdv = make_float4(y_s.x, y_s.y, y_s.z, 0.0f);
                            sc = make_float4(c.x, c.y, c.z, 0.0f);
                            shadow = true;
                        }
                    }
                }
            }
            const bool pr[2] = {shadow, false};
            uint32_t rk[2], tot[2];
            block_rank<2>(pr, rk, tot, lds, parity);
            if (shadow) {
                B.sq[base + bs + rk[0]] = j;
                B.so[j] = so;
                B.sd[j] = sdv;
                B.sc[j] = sc;
            }
            bs += tot[0];
        }
        if (threadIdx.x == 0) B.sqc[seg] = bs;
    }
}

// ------------------------------------------------------------------------------ shadow
// Occluded(p + N*2e-4 -> q, 1e-4, dist - 1e-4) of visible() (renderer.go:1333-1341, 851-895).
// MULTI: packed scenes, the root of each path's scene from its scene index.
template <bool MULTI>
__global__ __launch_bounds__(kBlock) BC_WAVES(7) void k_shadow(SceneBufs S, PathBufs B, uint32_t nseg, uint32_t root,
                                                   const RenderParams* __restrict__ Ps) {
    extern __shared__ uint2 stk[];  // S.stack entries per lane
    for (uint32_t seg = xcd_chunk(blockIdx.x, gridDim.x, B.xg); seg < nseg; seg += gridDim.x) {
        const uint32_t n = B.sqc[seg], base = seg * B.seg;
        if (threadIdx.x == 0) B.tally[kMaxGrid + seg] += n;
        for (uint32_t k = threadIdx.x; k < n; k += kBlock) {
            const uint32_t j = B.sq[base + k];
            const float4 o = B.so[j], d = B.sd[j];
            Ray r;
            r.ox = o.x; r.oy = o.y; r.oz = o.z;
            r.dx = d.x; r.dy = d.y; r.dz = d.z;
            r.tmin = (float)1e-4;
            const uint32_t rt = MULTI ? Ps[B.sid[j]].root : root;
            const HitRec h = traverse<true>(S.nodes, S.prims, S.ids, rt, r, o.w, stk + threadIdx.x, kBlock, S.stack, S.err);
            if (h.rec == 0xffffffffu) {
                const float4 c = B.sc[j];
                const float4 l = B.L[j];
                B.L[j] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.0f);
            }
        }
    }
}

// Shadow rays with lane refill.  Wave w of a workgroup takes its run (queue_run) of the shadow
// queue of each segment the workgroup owns (segments seg, seg + grid, ...): whenever at least
// `refill` of its lanes are idle, it hands them the next rays of the run (ballot + prefix
// count), moving on to the next segment when the current run is used up, so lanes do
// not wait for the wave's slowest ray and the waves of a workgroup never wait for each other.
// Each path has at most one shadow ray in the queue, so its radiance is updated by one lane, and
// results equal k_shadow's.
template <bool MULTI>
__global__ __launch_bounds__(kBlock) BC_WAVES(8) void k_shadow_refill(SceneBufs S, PathBufs B, uint32_t nseg, uint32_t root,
                                                                      const RenderParams* __restrict__ Ps) {
    extern __shared__ uint2 stk[];  // S.stack entries per lane
    const int lane = (int)(threadIdx.x & 63u);
    const uint32_t wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int refill = S.shadow_refill;
    uint2* my = stk + threadIdx.x;
    // wave-uniform cursor: segment, this wave's range [pos, end) of its queue
    uint32_t seg = xcd_chunk(blockIdx.x, gridDim.x, B.xg), pos = 0, end = 0, base = 0;
    bool started = false;
    TravState st;
    bool active = false;
    uint32_t j = 0;
    for (;;) {
        if (active && trav_step<true>(st, S.nodes, S.prims, S.ids, my, kBlock, S.stack, S.err)) {
            if (st.best.rec == 0xffffffffu) {
                const float4 c = B.sc[j];
                const float4 l = B.L[j];
                B.L[j] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.0f);
            }
            active = false;
        }
        while (pos >= end && seg < nseg) {
            if (started) {
                seg += gridDim.x;
                if (seg >= nseg) break;
            }
            started = true;
            const uint32_t n = B.sqc[seg];
            if (threadIdx.x == 0) B.tally[kMaxGrid + seg] += n;
            base = seg * B.seg;
            queue_run(n, wib, pos, end);
        }
        const uint64_t idle = __ballot(!active);
        const int nidle = __popcll(idle);
        if (pos >= end) {
            if (nidle == 64) break;
            continue;
        }
        if (nidle >= refill) {
            if (!active) {
                const uint32_t k = pos + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                if (k < end) {
                    j = B.sq[base + k];
                    const float4 o = B.so[j], d = B.sd[j];
                    Ray r;
                    r.ox = o.x; r.oy = o.y; r.oz = o.z;
                    r.dx = d.x; r.dy = d.y; r.dz = d.z;
                    r.tmin = (float)1e-4;
                    trav_init(st, r, o.w, MULTI ? Ps[B.sid[j]].root : root);
                    active = true;
                }
            }
            pos = min(end, pos + (uint32_t)nidle);
        }
    }
}

// ------------------------------------------------------------------------------ accumulate
// Per-pixel sum in sample order with clampFirefly (renderer.go:1917-1926, 1951-1953).
__device__ __forceinline__ float4 clamp_sample(const RenderParams& P, float4 l) {
    if (P.firefly > 0.0f) {
        const float lim = P.firefly;
        const float m = rmax(l.x, rmax(l.y, l.z));
        if (!(m <= lim || m <= 0.0f)) {
            const float k = lim / m;
            l.x = l.x * k; l.y = l.y * k; l.z = l.z * k;
        }
    }
    return l;
}

__global__ __launch_bounds__(kBlock) void k_accumulate(RenderParams P, PathBufs B, float4* __restrict__ acc,
                                                       float4* __restrict__ keep) {
    __shared__ unsigned long long lds[kWaves];
    unsigned long long segs = 0;
    for (int pl = blockIdx.x * kBlock + threadIdx.x; pl < P.npix; pl += gridDim.x * kBlock) {
        float4 a = acc[pl];
        for (int sl = 0; sl < P.ns; sl++) {
            const int j = sl * P.npix + pl;
            const float4 l = B.L[j];
            if (keep) keep[(size_t)(P.s0 + sl) * (size_t)P.npix + (size_t)pl] = l;
            const float4 c = clamp_sample(P, l);
            a.x = a.x + c.x; a.y = a.y + c.y; a.z = a.z + c.z;
            const uint32_t ns = B.nseg[j];
            a.w = a.w + (float)ns;
            segs += ns;
        }
        acc[pl] = a;
    }
    const unsigned long long t = block_sum(segs, lds);
    if (threadIdx.x == 0) B.tally[blockIdx.x] += t;
}

__global__ __launch_bounds__(kBlock) void k_finalize(RenderParams P, const float4* __restrict__ acc,
                                                     float4* __restrict__ rgba) {
    const float inv = 1.0f / (float)P.spp;
    for (int pl = blockIdx.x * kBlock + threadIdx.x; pl < P.npix; pl += gridDim.x * kBlock) {
        const float4 a = acc[pl];
        rgba[pl] = make_float4(a.x * inv, a.y * inv, a.z * inv, a.w);
    }
}

// Reduces the per-block tallies into totals[0] (path segments) and totals[1] (shadow rays).
constexpr int kTallyBlock = 1024;
__global__ __launch_bounds__(kTallyBlock) void k_tally(PathBufs B, uint32_t nseg) {
    __shared__ unsigned long long lds[kTallyBlock / 64];
    unsigned long long a = 0, b = 0;
    for (uint32_t i = threadIdx.x; i < (uint32_t)kMaxGrid; i += kTallyBlock) a += B.tally[i];
    for (uint32_t i = threadIdx.x; i < nseg; i += kTallyBlock) b += B.tally[kMaxGrid + i];
    const unsigned long long ta = block_sum<kTallyBlock / 64>(a, lds);
    __syncthreads();
    const unsigned long long tb = block_sum<kTallyBlock / 64>(b, lds);
    if (threadIdx.x == 0) {
        B.totals[0] = ta;
        B.totals[1] = tb;
    }
}


// ------------------------------------------------------------------------------ persistent schedule
// A fixed pool of path slots in segments.  Each iteration: extend -> shade -> shadow.  A path
// that ends is listed in its segment's terminated list; the next extend of that segment stores
// its radiance in the per-sample result slot and restarts the path slot in place with the
// segment's next sample, then traces it with the segment's continuing paths.  The window's
// samples are cut into chunks of B.seg consecutive (pixel, sample) indices and segment s owns
// chunks s, s + nseg, s + 2 nseg, ..., so a segment restarts runs of neighbouring pixels and
// its camera rays stay coherent.  Samples are keyed by (pixel, sample), never by
// slot, so the image does not depend on the schedule; per-sample results are reduced in sample
// order afterwards.  Parameters live in device memory so that a run of iterations can be
// captured once as a hipGraph and replayed.

__global__ __launch_bounds__(kBlock) void k_p_start(const RenderParams* __restrict__ Pd, PathBufs B,
                                                    float4* __restrict__ acc, uint32_t* __restrict__ guides,
                                                    uint32_t nseg, int clear_pixels) {
    const RenderParams P = *Pd;
    if (clear_pixels) {
        for (int pl = blockIdx.x * kBlock + threadIdx.x; pl < P.npix; pl += gridDim.x * kBlock) {
            acc[pl] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (P.want_guides) {
                uint32_t* gg = guides + 3 * (size_t)pl;
                gg[0] = 0u; gg[1] = 0u; gg[2] = 0u;
            }
        }
    }
    for (uint32_t seg = xcd_chunk(blockIdx.x, gridDim.x, B.xg); seg < nseg; seg += gridDim.x) {
        const uint32_t base = seg * B.seg;
        for (uint32_t k = threadIdx.x; k < B.seg; k += kBlock) {
            B.dq[base + k] = base + k;
            B.gid[base + k] = kIdle;
        }
        if (threadIdx.x == 0) {
            B.qc[0][seg] = 0u;
            B.qc[1][seg] = 0u;
            B.sqc[seg] = 0u;
            B.dqc[seg] = B.seg;
            B.snext[seg] = 0u;
        }
    }
}

// Samples of the window (W of them, in chunks of ss) owned by segment `seg` of `nseg`.
__device__ __forceinline__ uint32_t seg_samples(uint32_t seg, uint32_t nseg, uint32_t ss, uint32_t W) {
    const uint32_t chunks = (W + ss - 1) / ss;
    if (seg >= chunks) return 0u;
    const uint32_t last = chunks - 1u;
    uint32_t n = ((last - seg) / nseg + 1u) * ss;
    if ((last - seg) % nseg == 0u) n -= chunks * ss - W;  // the short final chunk is this segment's
    return n;
}

__global__ void k_p_clear_flag(PathBufs B) {
    if (threadIdx.x == 0) B.flag[0] = 0u;
}

// Scene of window sample g among packed scenes: the last s with gpre[s] <= g (gpre: prefix
// sums of the scenes' window samples, nscene + 1 entries, g < gpre[nscene]).
__device__ __forceinline__ uint32_t scene_of(const uint32_t* __restrict__ gpre, uint32_t nscene, uint32_t g) {
    uint32_t lo = 0, hi = nscene;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (gpre[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

// MULTI: the window's samples are those of several packed scenes (Pd[s], s < nscene); each
// restarted path records its scene and is traced from that scene's root.
template <bool MULTI>
__global__ __launch_bounds__(kBlock) BC_WAVES(BC_P_EXTEND_WAVES) void k_p_extend(SceneBufs S, PathBufs B, const RenderParams* __restrict__ Pd,
                                                     int cur, uint32_t nseg, const uint32_t* __restrict__ gpre,
                                                     uint32_t nscene) {
    extern __shared__ uint2 stk[];  // S.stack entries per lane
    const RenderParams P = *Pd;
    const uint32_t W = MULTI ? gpre[nscene] : (uint32_t)P.ns * (uint32_t)P.npix;
    uint32_t* q = B.q[cur];
    bool work = false;
    for (uint32_t seg = xcd_chunk(blockIdx.x, gridDim.x, B.xg); seg < nseg; seg += gridDim.x) {
        const uint32_t n = B.qc[cur][seg], nd = B.dqc[seg], m0 = B.snext[seg], base = seg * B.seg;
        const uint32_t owned = seg_samples(seg, nseg, B.seg, W);
        const uint32_t avail = owned > m0 ? owned - m0 : 0u;
        const uint32_t nr = nd < avail ? nd : avail;
        for (uint32_t k = threadIdx.x; k < n + nd; k += kBlock) {
            uint32_t j;
            bool live = true;
            if (k < n) {
                j = q[base + k];
            } else {
                const uint32_t kk = k - n;
                j = B.dq[base + kk];
                const uint32_t g = B.gid[j];
                if (g != kIdle) {
                    const float4 l = B.L[j];
                    B.res[g] = make_float4(l.x, l.y, l.z, (float)B.nseg[j]);
                }
                if (kk < nr) {
                    const uint32_t m = m0 + kk;
                    const uint32_t gn = ((m / B.seg) * nseg + seg) * B.seg + m % B.seg;
                    if constexpr (MULTI) {
                        const uint32_t sc = scene_of(gpre, nscene, gn);
                        const RenderParams& Q = Pd[sc];
                        uint32_t sl, pl;
                        window_sample(Q, gn - gpre[sc], sl, pl);
                        init_camera_path(Q, B, j, (int)pl, Q.s0 + (int)sl);
                        B.sid[j] = sc;
                    } else {
                        uint32_t sl, pl;
                        window_sample(P, gn, sl, pl);
                        init_camera_path(P, B, j, (int)pl, P.s0 + (int)sl);
                    }
                    B.gid[j] = gn;
                    B.bnc[j] = 0u;
                    q[base + n + kk] = j;
                } else {
                    B.gid[j] = kIdle;
                    live = false;
                }
            }
            if (live) {
                const float4 o = B.o[j], d = B.d[j];
                Ray r;
                r.ox = o.x; r.oy = o.y; r.oz = o.z;
                r.dx = d.x; r.dy = d.y; r.dz = d.z;
                r.tmin = 1e-4f;
                const uint32_t rt = MULTI ? Pd[B.sid[j]].root : P.root;
                const HitRec h = traverse<false>(S.nodes, S.prims, S.ids, rt, r, 1e30f, stk + threadIdx.x, kBlock, S.stack,
                                                 S.err);
                B.hit[j] = make_float2(h.t, __uint_as_float(h.rec));
                B.nseg[j] += 1u;
            }
        }
        work = work || (n + nd) > 0u;
        __syncthreads();  // every lane has read this segment's lengths
        if (threadIdx.x == 0) {
            B.qc[cur][seg] = n + nr;
            B.dqc[seg] = 0u;
            B.snext[seg] = m0 + nr;
        }
    }
    if (work && threadIdx.x == 0) B.flag[0] = 1u;
}

// Material class of a path's hit: 0 escaped, 1 + material kind (diffuse, metal, glass, emissive).
__device__ __forceinline__ uint32_t hit_class(const SceneBufs& S, const PathBufs& B, uint32_t j) {
    const uint32_t rec = __float_as_uint(B.hit[j].y);
    if (rec == 0xffffffffu) return 0u;
    return 1u + (uint32_t)S.mats[S.pmat[rec]].kind;
}

// SORT: the segment's paths are first ordered by material class (a counting sort in LDS), so
// that the waves of the shading loop take one BSDF branch each.  The order does not change
// any result: paths are independent and results are keyed by sample.
// The sorted single-scene instance runs five waves per SIMD (96 VGPRs, a few spilled; 2% faster
// on the terrain than four); the others keep the compiler's choice (the batch's sorted instance
// would spill 56 registers at five).
template <bool MULTI, bool SORT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu((SORT && !MULTI) ? 5 : 1, (SORT && !MULTI) ? 5 : 8)))
void k_p_shade(SceneBufs S, PathBufs B, const RenderParams* __restrict__ Pd,
                                                    int cur, uint32_t* __restrict__ guides, uint32_t nseg) {
    // the sort's ranks keep their own buffer and parity: block_rank alternates the halves of a
    // buffer between consecutive calls, which is race-free only for calls of one width
    __shared__ uint32_t lds[2 * 3 * kWaves];
    __shared__ uint32_t lds_sort[SORT ? 2 * 5 * kWaves : 1];
    // the sort's arrays (SORT only): B.seg entries each, in the launch's dynamic LDS
    extern __shared__ uint2 stk[];
    uint32_t* const order = reinterpret_cast<uint32_t*>(stk);
    uint32_t* const place = order + B.seg;
    uint32_t parity = 0, parity_sort = 0;
    const RenderParams P = *Pd;
    const uint32_t* q = B.q[cur];
    uint32_t* qn = B.q[cur ^ 1];
    for (uint32_t seg = xcd_chunk(blockIdx.x, gridDim.x, B.xg); seg < nseg; seg += gridDim.x) {
        const uint32_t n = B.qc[cur][seg], base = seg * B.seg;
        if constexpr (SORT) {
            // place[k] = (class << 24) | rank of item k within its class
            uint32_t run[5] = {0u, 0u, 0u, 0u, 0u};
            for (uint32_t r0 = 0; r0 < n; r0 += kBlock) {
                const uint32_t k = r0 + threadIdx.x;
                const uint32_t c = k < n ? hit_class(S, B, q[base + k]) : 5u;
                const bool pr[5] = {c == 0u, c == 1u, c == 2u, c == 3u, c == 4u};
                uint32_t rk[5], tot[5];
                block_rank<5>(pr, rk, tot, lds_sort, parity_sort);
                if (k < n) {
                    uint32_t pos = 0;
#pragma unroll
                    for (int cc = 0; cc < 5; cc++)
                        if ((uint32_t)cc == c) pos = run[cc] + rk[cc];
                    place[k] = (c << 24) | pos;
                }
#pragma unroll
                for (int cc = 0; cc < 5; cc++) run[cc] += tot[cc];
            }
            __syncthreads();
            uint32_t off[5], acc = 0;
#pragma unroll
            for (int cc = 0; cc < 5; cc++) { off[cc] = acc; acc += run[cc]; }
            for (uint32_t k = threadIdx.x; k < n; k += kBlock) {
                const uint32_t pc = place[k], c = pc >> 24;
                uint32_t o = 0;
#pragma unroll
                for (int cc = 0; cc < 5; cc++)
                    if ((uint32_t)cc == c) o = off[cc];
                order[o + (pc & 0xffffffu)] = q[base + k];
            }
            __syncthreads();
        }
        uint32_t bc = 0, bd = 0, bs = 0;
        for (uint32_t r0 = 0; r0 < n; r0 += kBlock) {
            const uint32_t k = r0 + threadIdx.x;
            bool cont = false, shadow = false, done = false;
            uint32_t j = 0;
            float4 so = make_float4(0, 0, 0, 0), sd = so, sc = so;
            if (k < n) {
                j = SORT ? order[k] : q[base + k];
                const int bounce = (int)B.bnc[j];
                if constexpr (MULTI) {
                    const RenderParams& Q = Pd[B.sid[j]];
                    cont = shade_path<false>(S, B, Q, j, bounce, guides, shadow, so, sd, sc, Adj{}) && bounce + 1 < Q.bounces;
                } else {
                    cont = shade_path<false>(S, B, P, j, bounce, guides, shadow, so, sd, sc, Adj{}) && bounce + 1 < P.bounces;
                }
                if (cont) B.bnc[j] = (uint32_t)(bounce + 1);
                done = !cont;
            }
            const bool pr[3] = {cont, done, shadow};
            uint32_t rk[3], tot[3];
            block_rank<3>(pr, rk, tot, lds, parity);
            if (cont) qn[base + bc + rk[0]] = j;
            if (done) B.dq[base + bd + rk[1]] = j;
            if (shadow) {
                B.sq[base + bs + rk[2]] = j;
                B.so[j] = so;
                B.sd[j] = sd;
                B.sc[j] = sc;
            }
            bc += tot[0];
            bd += tot[1];
            bs += tot[2];
        }
        if (threadIdx.x == 0) {
            B.qc[cur ^ 1][seg] = bc;
            B.dqc[seg] = bd;
            B.sqc[seg] = bs;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_p_reduce(const RenderParams* __restrict__ Pd, PathBufs B,
                                                     float4* __restrict__ acc) {
    __shared__ unsigned long long lds[kWaves];
    const RenderParams P = *Pd;
    unsigned long long segs = 0;
    for (int pl = blockIdx.x * kBlock + threadIdx.x; pl < P.npix; pl += gridDim.x * kBlock) {
        float4 a = acc[pl];
        for (int sl = 0; sl < P.ns; sl++) {
            const float4 r = B.res[window_index(P, (uint32_t)sl, (uint32_t)pl)];
            const float4 c = clamp_sample(P, r);
            a.x = a.x + c.x; a.y = a.y + c.y; a.z = a.z + c.z;
            a.w = a.w + r.w;
            segs += (uint32_t)r.w;
        }
        acc[pl] = a;
    }
    const unsigned long long t = block_sum(segs, lds);
    if (threadIdx.x == 0) B.tally[blockIdx.x] += t;
}


// ------------------------------------------------------------------------------ adjoint (path replay)
// Backward pass of the differentiable render: every sample is replayed from its key with the
// forward pass's total radiance, and d loss / d (Color, Emission) of each material is summed
// per workgroup in LDS, then once per workgroup into one of kGradCopies copies of the global
// gradient (chosen by workgroup index, i.e. by XCD), which the host adds up: eight copies
// divide the same-address atomic traffic of the flush by eight.
constexpr int kAdjMaxMat = 1024;
constexpr int kGradCopies = 8;

__device__ __forceinline__ void adj_lds_init(float* lds, int n) {
    for (int i = threadIdx.x; i < n; i += kBlock) lds[i] = 0.0f;
    __syncthreads();
}
// grad: copy 0 of the gradient (nmat x 6 floats: colour rgb then emission rgb); `off`
// selects the part that lds holds.
__device__ __forceinline__ void adj_lds_flush(const float* lds, int n, float* __restrict__ grad, int nmat, int off) {
    __syncthreads();
    float* g = grad + (size_t)(blockIdx.x % kGradCopies) * (size_t)(nmat * 6) + off;
    for (int i = threadIdx.x; i < n; i += kBlock)
        if (lds[i] != 0.0f) atomicAdd(&g[i], lds[i]);
}

__global__ __launch_bounds__(kBlock) void k_raygen_adj(RenderParams P, PathBufs B, const float4* __restrict__ keep,
                                                       float4* __restrict__ Lt, uint32_t nseg) {
    const uint32_t n = (uint32_t)P.ns * (uint32_t)P.npix;
    for (uint32_t seg = xcd_chunk(blockIdx.x, gridDim.x, B.xg); seg < nseg; seg += gridDim.x) {
        const uint32_t cnt = seg_slots(seg, B.seg, n), base = seg * B.seg;
        for (uint32_t k = threadIdx.x; k < cnt; k += kBlock) {
            const uint32_t j = base + k;
            const uint32_t sl = j / (uint32_t)P.npix, pl = j - sl * (uint32_t)P.npix;
            init_camera_path(P, B, j, (int)pl, P.s0 + (int)sl);
            B.q[0][j] = j;
            Lt[j] = keep[(size_t)(P.s0 + (int)sl) * (size_t)P.npix + (size_t)pl];
        }
        if (threadIdx.x == 0) B.qc[0][seg] = cnt;
    }
}

__global__ __launch_bounds__(kBlock) void k_shade_adj(SceneBufs S, PathBufs B, RenderParams P, int cur, int bounce,
                                                      uint32_t* __restrict__ guides, Adj A, float* __restrict__ grad,
                                                      uint32_t nseg) {
    __shared__ float lds[2 * kAdjMaxMat * 3];
    __shared__ uint32_t rl[2 * 2 * kWaves];
    uint32_t parity = 0;
    const int ng = P.nmat * 6;
    adj_lds_init(lds, ng);
    A.gcol = lds;
    A.gle = lds + P.nmat * 3;
    const uint32_t* q = B.q[cur];
    uint32_t* qn = B.q[cur ^ 1];
    for (uint32_t seg = xcd_chunk(blockIdx.x, gridDim.x, B.xg); seg < nseg; seg += gridDim.x) {
        const uint32_t n = B.qc[cur][seg], base = seg * B.seg;
        uint32_t bc = 0, bs = 0;
        for (uint32_t r0 = 0; r0 < n; r0 += kBlock) {
            const uint32_t k = r0 + threadIdx.x;
            bool cont = false, shadow = false;
            uint32_t j = 0;
            float4 so = make_float4(0, 0, 0, 0), sd = so, sc = so;
            if (k < n) {
                j = q[base + k];
                cont = shade_path<true>(S, B, P, j, bounce, guides, shadow, so, sd, sc, A) && bounce + 1 < P.bounces;
            }
            const bool pr[2] = {cont, shadow};
            uint32_t rk[2], tot[2];
            block_rank<2>(pr, rk, tot, rl, parity);
            if (cont) qn[base + bc + rk[0]] = j;
            if (shadow) {
                B.sq[base + bs + rk[1]] = j;
                B.so[j] = so;
                B.sd[j] = sd;
                B.sc[j] = sc;
            }
            bc += tot[0];
            bs += tot[1];
        }
        if (threadIdx.x == 0) {
            B.qc[cur ^ 1][seg] = bc;
            B.sqc[seg] = bs;
        }
    }
    adj_lds_flush(lds, ng, grad, P.nmat, 0);
}

__global__ __launch_bounds__(kBlock) BC_WAVES(6) void k_shadow_adj(SceneBufs S, PathBufs B, RenderParams P, Adj A,
                                                       float* __restrict__ grad, uint32_t nseg) {
    extern __shared__ uint2 stk[];  // S.stack entries per lane
    __shared__ float lds[kAdjMaxMat * 3];
    const int ng = P.nmat * 3;
    adj_lds_init(lds, ng);
    A.gle = lds;
    for (uint32_t seg = xcd_chunk(blockIdx.x, gridDim.x, B.xg); seg < nseg; seg += gridDim.x) {
        const uint32_t n = B.sqc[seg], base = seg * B.seg;
        for (uint32_t k = threadIdx.x; k < n; k += kBlock) {
            const uint32_t j = B.sq[base + k];
            const float4 o = B.so[j], d = B.sd[j];
            Ray r;
            r.ox = o.x; r.oy = o.y; r.oz = o.z;
            r.dx = d.x; r.dy = d.y; r.dz = d.z;
            r.tmin = (float)1e-4;
            const HitRec h = traverse<true>(S.nodes, S.prims, S.ids, P.root, r, o.w, stk + threadIdx.x, kBlock, S.stack,
                                            S.err);
            if (h.rec == 0xffffffffu) {
                const float4 c = B.sc[j];
                const float4 l = B.L[j];
                B.L[j] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.0f);
                const uint32_t pl = B.meta[j].x - (uint32_t)P.y0 * (uint32_t)P.width;
                const float3 g = mul(f3(A.dl[3 * pl], A.dl[3 * pl + 1], A.dl[3 * pl + 2]), A.inv_spp);
                const float4 w = A.sg[j];
                adj_emission(A, (int)__float_as_uint(w.w), g, xyz(w));
            }
        }
    }
    adj_lds_flush(lds, ng, grad, P.nmat, P.nmat * 3);
}

// Per-row sums of the per-pixel path-segment counts (rgba.w, integers held in fp32) of a block of
// rows rendered ahead, one workgroup of 256 threads per row, so that a chunk's segment count is a
// sum over its rows on the host.
__global__ __launch_bounds__(256) void k_row_segments(const float4* __restrict__ rgba, int width,
                                                      unsigned long long* __restrict__ out) {
    const size_t base = (size_t)blockIdx.x * (size_t)width;
    unsigned long long sum = 0;
    for (int x = (int)threadIdx.x; x < width; x += 256) sum += (unsigned long long)(int)(rgba[base + x].w + 0.5f);
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_down(sum, o);
    __shared__ unsigned long long part[4];
    if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

}  // namespace bcd

// ------------------------------------------------------------------------------ host side
struct RenderState {
    void* buf[30] = {};
    size_t bufcap[30] = {};
    float4* sg = nullptr;
    // differentiable render: per-sample radiance of the last bc_forward, its shape
    void* keep = nullptr;
    size_t keep_cap = 0;
    bc_render_desc keep_desc{};
    int keep_y0 = -1, keep_rows = 0;
    uint64_t keep_topo = 0;
    bcd::PathBufs B{};
    uint32_t nseg = 0;  // segments of the current allocation
    float4* acc = nullptr;
    uint32_t* guide_tmp = nullptr;
    bcd::RenderParams* Pd = nullptr;        // per-job parameters (device)
    uint32_t* gpre = nullptr;               // per-job window sample prefix sums (device)
    bcd::RenderParams* h_params = nullptr;  // pinned staging for Pd, then gpre
    uint32_t* h_cnt = nullptr;              // pinned
    // captured iteration graph of the persistent schedule and what its kernel arguments hold
    hipGraphExec_t exec = nullptr;
    std::vector<const void*> g_key;
    // autotuner cache: (shape) -> (pool, iterations per graph)
    struct Tune {
        int w, rows, spp, bounces;
        uint64_t topo, geom, shape;  // shape: hash of a batch's frame sizes (0: one scene)
        int pool, iters, seg, sort, plain;
    };
    std::vector<Tune> tunes;
    int last_pool = 0, last_iters = 0, last_windows = 0, last_seg = 0, last_sort = 0, last_plain = 0;
};

namespace {

constexpr size_t kResultCap = (size_t)1 << 27;   // per-sample results held per window (2 GB)
constexpr size_t kBatchCap = (size_t)1 << 30;    // the same for a batch of scenes (16 GB)
constexpr int kMaxJobs = 256;                    // scenes of one batch render

size_t segments_for(size_t paths, uint32_t ss) { return (paths + ss - 1) / ss; }

// Allocates (or keeps) path state for `paths` slots in whole segments of `ss` slots.
bool ensure_state(bc_ctx* c, size_t paths, size_t npix, bool guides, size_t window_samples, uint32_t ss,
                  int njobs = 1, bool reservoirs = false) {
    if (!c->rs) c->rs = new RenderState();
    RenderState& R = *c->rs;
    if (!R.h_cnt) {
        const size_t pbytes = (size_t)kMaxJobs * sizeof(bcd::RenderParams) + (size_t)(kMaxJobs + 1) * 4;
        void* h = nullptr;
        if (!bc_check(c, hipHostMalloc(&h, 256 + pbytes, hipHostMallocDefault), "pinned staging")) return false;
        R.h_cnt = (uint32_t*)h;
        R.h_params = (bcd::RenderParams*)((char*)h + 256);
    }
    const size_t nseg = segments_for(paths, ss), slots = nseg * ss;
    struct Want { int slot; size_t bytes; };
    const Want w[] = {
        {0, slots * 16},  {1, slots * 16},  {2, slots * 16}, {3, slots * 16},  {4, slots * 16},
        {5, slots * 16},  {6, slots * 4},   {7, slots * 8},  {8, slots * 16},  {9, slots * 16},
        {10, slots * 16}, {11, slots * 4},  {12, slots * 4}, {13, slots * 4},  {14, 64},
        {15, npix * 16 + (guides ? npix * 12 : 0) + 64},     {16, slots * 4},  {17, slots * 4},
        {18, window_samples * 16},
        {19, (size_t)njobs * sizeof(bcd::RenderParams) + (size_t)(njobs + 1) * 4 + 16},
        {20, slots * 16}, {21, slots * 4}, {22, nseg * 5 * 4}, {23, ((size_t)bcd::kMaxGrid + nseg) * 8},
        {24, slots * 4},
    };
    for (const Want& x : w)
        if (!bc_ensure(c, R.buf[x.slot], R.bufcap[x.slot], std::max<size_t>(x.bytes, 64), "render buffers")) return false;
    for (int k = 0; k < 5; k++) {  // primary-vertex reservoirs of spatial reuse
        if (reservoirs && !bc_ensure(c, R.buf[25 + k], R.bufcap[25 + k], slots * 16, "reservoir buffers")) return false;
        R.B.rv[k] = reservoirs ? (float4*)R.buf[25 + k] : nullptr;
    }
    R.nseg = (uint32_t)nseg;
    R.B.seg = ss;
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
    R.B.totals = (unsigned long long*)R.buf[14];
    R.B.flag = (uint32_t*)((char*)R.buf[14] + 32);
    R.B.bnc = (uint32_t*)R.buf[16];
    R.B.dq = (uint32_t*)R.buf[17];
    R.B.res = (float4*)R.buf[18];
    R.acc = (float4*)R.buf[15];
    R.guide_tmp = (uint32_t*)(R.acc + npix);
    R.Pd = (bcd::RenderParams*)R.buf[19];
    R.gpre = (uint32_t*)(R.Pd + njobs);
    R.B.sid = (uint32_t*)R.buf[24];
    R.sg = (float4*)R.buf[20];
    R.B.gid = (uint32_t*)R.buf[21];
    uint32_t* sc = (uint32_t*)R.buf[22];
    R.B.qc[0] = sc;
    R.B.qc[1] = sc + nseg;
    R.B.sqc = sc + 2 * nseg;
    R.B.dqc = sc + 3 * nseg;
    R.B.snext = sc + 4 * nseg;
    R.B.tally = (unsigned long long*)R.buf[23];
    return true;
}

// One workgroup per segment: a segment's live paths are packed at its front, so a sparse
// segment leaves the later waves of its workgroup idle, and they exit at once.
int seg_grid(const bc_ctx*, size_t nseg) { return (int)nseg; }

// Segments per XCD turn of the workgroup mapping (reserved[3]; 0 = 1, plain round-robin).
uint32_t xcd_group(const bc_render_desc* d) { return d->reserved[3] > 0 ? (uint32_t)d->reserved[3] : 1u; }

// Plain schedule: the largest segment (up to kSegPlain) that still gives every CU eight
// workgroups, so late, sparse bounces are packed into few waves without starving the chip.
uint32_t plain_segment(const bc_ctx* c, size_t paths) {
    uint32_t ss = bcd::kSegPlain;
    while (ss > (uint32_t)bcd::kBlock && paths / ss < (size_t)c->num_cus * 8) ss >>= 1;
    return ss;
}
int pix_grid(const bc_ctx* c, size_t npix) {
    return (int)std::min<size_t>({(npix + bcd::kBlock - 1) / bcd::kBlock, (size_t)c->num_cus * 8, (size_t)bcd::kMaxGrid});
}

// scene: index into a resident batch, or -1 for the resident single scene.
bcd::RenderParams make_params(const bc_ctx* c, const bc_render_desc* d, int y0, int rows, bool guides, int scene = -1) {
    bcd::RenderParams P{};
    P.width = d->width; P.height = d->height; P.spp = d->spp; P.bounces = d->bounces; P.rr_depth = d->rr_depth;
    P.sampler = d->sampler; P.mis = d->mis != 0; P.power = d->power_lights != 0; P.sky_mode = d->sky_mode;
    P.want_guides = guides;
    P.firefly = d->firefly_clamp;
    for (int k = 0; k < 3; k++) {
        P.sky[k] = d->sky_color[k];
        P.origin[k] = (float)d->origin[k]; P.lower[k] = (float)d->lower[k];
        P.horiz[k] = (float)d->horiz[k]; P.vert[k] = (float)d->vert[k];
    }
    P.seed = d->seed;
    P.y0 = y0;
    P.npix = rows * d->width;
    P.nmat = c->nmat;
    P.nlight = c->nlight;
    P.root = 0;
    P.lbase = 0;
    if (scene >= 0) {
        const bc_ctx::BatchScene& b = c->batch[(size_t)scene];
        P.nlight = b.nlight;
        P.root = b.root;
        P.lbase = b.lbase;
    }
    P.restir = d->restir_candidates;
    // sun direction normalised in double precision, then rounded
    const double sx = d->sun_dir[0], sy = d->sun_dir[1], sz = d->sun_dir[2];
    const double sl = std::sqrt(sx * sx + sy * sy + sz * sz);
    P.sun_dir[0] = sl > 0 ? (float)(sx / sl) : 0.0f;
    P.sun_dir[1] = sl > 0 ? (float)(sy / sl) : 0.0f;
    P.sun_dir[2] = sl > 0 ? (float)(sz / sl) : 0.0f;
    for (int k = 0; k < 3; k++) P.sun[k] = d->sun_irradiance[k];
    P.spatial = d->restir_candidates > 0 ? d->restir_spatial : 0;
    P.radius = d->restir_radius > 0 ? d->restir_radius : 16;
    P.band = P.npix;
    return P;
}

// The resident scene as the kernels see it.  Refill thresholds: shadow rays take the any-hit
// setting; the plain schedule's second-segment extension rays the closest-hit one (automatic: 8
// lanes).
bcd::SceneBufs scene_bufs(const bc_ctx* c) {
    const int closest = c->trace_refill[0] < 0 ? 8 : c->trace_refill[0];
    return bcd::SceneBufs{c->d_nodes, c->d_prims, c->d_ids, c->d_pmat, c->d_mats, c->d_lights, c->d_pdf,
                          c->d_light_of, c->d_err, c->stack, c->trace_refill[1], closest};
}

// Dynamic LDS of the traversal kernels: S.stack stack entries per lane.
size_t stack_lds(const bcd::SceneBufs& S) { return (size_t)S.stack * bcd::kBlock * sizeof(uint2); }

// Shadow queue of every segment: lane refill unless S.shadow_refill is 0.
template <bool MULTI>
void launch_shadow(const bcd::SceneBufs& S, const bcd::PathBufs& B, int grid, uint32_t nseg, uint32_t root,
                   const bcd::RenderParams* Pd, hipStream_t st) {
    if (S.shadow_refill > 0)
        hipLaunchKernelGGL((bcd::k_shadow_refill<MULTI>), dim3(grid), dim3(bcd::kBlock), stack_lds(S), st, S, B, nseg, root, Pd);
    else
        hipLaunchKernelGGL((bcd::k_shadow<MULTI>), dim3(grid), dim3(bcd::kBlock), stack_lds(S), st, S, B, nseg, root, Pd);
}

// Extension rays of one segment: the second segment's (the first bounce off the scene, whose
// queues are still full and whose rays scatter) with lane refill unless S.extend_refill is 0, the
// others one per thread.  Measured on the terrain: refill takes the second segment from 26.7 to
// 20.3 ms per frame but the later ones (short queues) 5-19% slower, also when it only refills
// whole waves; camera rays are coherent, where refill loses on ray queries as well.
void launch_extend(const bcd::SceneBufs& S, const bcd::PathBufs& B, int cur, int bounce, int grid, uint32_t nseg,
                   uint32_t root, hipStream_t st) {
    if (bounce == 1 && S.extend_refill > 0)
        hipLaunchKernelGGL(bcd::k_extend_refill, dim3(grid), dim3(bcd::kBlock), stack_lds(S), st, S, B, cur, nseg, root);
    else
        hipLaunchKernelGGL(bcd::k_extend, dim3(grid), dim3(bcd::kBlock), stack_lds(S), st, S, B, cur, nseg, root);
}

void launch_bounce(const bcd::SceneBufs& S, const bcd::PathBufs& B, const bcd::RenderParams& P, int cur, int bounce,
                   uint32_t* guides, int grid, uint32_t nseg, hipStream_t st) {
    launch_extend(S, B, cur, bounce, grid, nseg, P.root, st);
    hipLaunchKernelGGL(bcd::k_shade, dim3(grid), dim3(bcd::kBlock), 0, st, S, B, P, cur, bounce, guides, nseg);
    if (bounce == 0 && P.spatial > 0 && P.restir > 0)
        hipLaunchKernelGGL(bcd::k_spatial, dim3(grid), dim3(bcd::kBlock), 0, st, S, B, P, nseg);
    launch_shadow<false>(S, B, grid, nseg, P.root, nullptr, st);
}

// One iteration of the persistent schedule; MULTI: the jobs are packed scenes of a batch.
template <bool MULTI>
void launch_iteration(const bcd::SceneBufs& S, const bcd::PathBufs& B, const bcd::RenderParams* Pd,
                      const uint32_t* gpre, uint32_t njobs, int cur, uint32_t* guides, int grid, uint32_t nseg,
                      bool last, bool sort, hipStream_t st) {
    // the drain test reads the flag written by the last iteration of a replay only
    if (last) hipLaunchKernelGGL(bcd::k_p_clear_flag, dim3(1), dim3(64), 0, st, B);
    hipLaunchKernelGGL((bcd::k_p_extend<MULTI>), dim3(grid), dim3(bcd::kBlock), stack_lds(S), st, S, B, Pd, cur, nseg,
                       gpre, njobs);
    if (sort)
        hipLaunchKernelGGL((bcd::k_p_shade<MULTI, true>), dim3(grid), dim3(bcd::kBlock), (size_t)B.seg * 8, st, S, B, Pd, cur,
                           guides, nseg);
    else
        hipLaunchKernelGGL((bcd::k_p_shade<MULTI, false>), dim3(grid), dim3(bcd::kBlock), 0, st, S, B, Pd, cur, guides, nseg);
    launch_shadow<MULTI>(S, B, grid, nseg, 0u, Pd, st);
}

bool clear_tally(bc_ctx* c) {
    RenderState& R = *c->rs;
    return bc_check(c, hipMemsetAsync(R.B.tally, 0, ((size_t)bcd::kMaxGrid + R.nseg) * 8, c->stream), "clear tally");
}

// Finalise: reduced pixel sums -> rgba (already done per job by the caller), then the tallies.
int finish_tally(bc_ctx* c, uint64_t* rays_out, uint64_t* shadow_out) {
    RenderState& R = *c->rs;
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(bcd::k_tally, dim3(1), dim3(bcd::kTallyBlock), 0, st, R.B, R.nseg);
    if (!bc_check(c, hipGetLastError(), "finalize launch") ||
        !bc_check(c, hipMemcpyAsync(R.h_cnt + 16, R.B.totals, 16, hipMemcpyDeviceToHost, st), "read totals") ||
        !bc_check(c, hipStreamSynchronize(st), "render"))
        return BC_ERR_DEVICE;
    unsigned long long tot[2];
    std::memcpy(tot, R.h_cnt + 16, 16);
    if (rays_out) *rays_out = tot[0];
    if (shadow_out) *shadow_out = tot[1];
    return bc_check_kernel_error(c);
}

int finish(bc_ctx* c, const bcd::RenderParams& P, float* d_rgba, int grid_pix, uint64_t* rays_out, uint64_t* shadow_out) {
    RenderState& R = *c->rs;
    hipLaunchKernelGGL(bcd::k_finalize, dim3(grid_pix), dim3(bcd::kBlock), 0, c->stream, P, (const float4*)R.acc,
                       (float4*)d_rgba);
    return finish_tally(c, rays_out, shadow_out);
}

// Plain wavefront schedule: passes of pixels x samples, three kernels per bounce.  scene: batch
// scene index (-1: the single resident scene).  A pass holds up to 2^22 paths.  When the rows'
// samples need several passes, they run band by band: bands of whole rows whose every sample
// fits one pass (one row per band, in passes of its samples, when a row's samples do not), so
// that the rays in flight come from one region of the image and share the scene's nodes in
// cache, instead of passes of a few samples over the whole image.  Each pixel still adds its
// samples in sample order, so the result does not depend on the split.  Spatial reuse (which
// reads neighbouring pixels among the rendered rows) and the forward pass of the adjoint (which
// keeps every sample by index) run whole-image passes.
int render_plain(bc_ctx* c, const bc_render_desc* d, int y0, int rows, float* d_rgba, uint32_t* d_guides,
                 uint64_t* rays_out, uint64_t* shadow_out, float4* keep = nullptr, int scene = -1) {
    const size_t W = (size_t)d->width;
    const size_t npix = (size_t)rows * W;
    const size_t max_paths = std::max<size_t>(npix, (size_t)1 << 22);
    const bool spatial = d->restir_candidates > 0 && d->restir_spatial > 0;
    int band = rows;  // rows per band
    int spp_pass;
    if (d->samples_per_pass > 0) {
        spp_pass = std::min(d->samples_per_pass, d->spp);
    } else if (npix * (size_t)d->spp <= max_paths || spatial || keep) {
        spp_pass = (int)std::min<size_t>((size_t)d->spp, std::max<size_t>(1, max_paths / npix));
    } else {
        const size_t per_row = W * (size_t)d->spp;
        band = (int)std::max<size_t>(1, std::min<size_t>((size_t)rows, ((size_t)1 << 22) / per_row));
        spp_pass = (int)std::min<size_t>((size_t)d->spp, std::max<size_t>(1, ((size_t)1 << 22) / ((size_t)band * W)));
    }
    const size_t paths = (size_t)band * W * (size_t)spp_pass;
    if (!ensure_state(c, paths, npix, d_guides != nullptr, 1, plain_segment(c, paths), 1, spatial)) return BC_ERR_NOMEM;
    RenderState& R = *c->rs;
    R.B.xg = xcd_group(d);
    const bcd::SceneBufs S = scene_bufs(c);
    const bcd::RenderParams Pall = make_params(c, d, y0, rows, d_guides != nullptr, scene);
    hipStream_t st = c->stream;
    if (!clear_tally(c)) return BC_ERR_DEVICE;
    uint32_t* guides = d_guides ? d_guides : R.guide_tmp;
    for (int yb = 0; yb < rows; yb += band) {
        const int hb = std::min(band, rows - yb);
        const size_t off = (size_t)yb * W, bpix = (size_t)hb * W;
        bcd::RenderParams P = make_params(c, d, y0 + yb, hb, d_guides != nullptr, scene);
        const uint32_t nseg = (uint32_t)segments_for(bpix * (size_t)spp_pass, R.B.seg);
        const int grid = seg_grid(c, nseg);
        const int grid_pix = pix_grid(c, bpix);
        for (int s0 = 0; s0 < d->spp; s0 += spp_pass) {
            P.s0 = s0;
            P.ns = std::min(spp_pass, d->spp - s0);
            hipLaunchKernelGGL(bcd::k_raygen, dim3(grid), dim3(bcd::kBlock), 0, st, P, R.B, (float*)(R.acc + off),
                               guides + 3 * off, nseg);
            for (int b = 0; b < d->bounces; b++) {
                const int cur = b & 1;
                launch_bounce(S, R.B, P, cur, b, guides + 3 * off, grid, nseg, st);
            }
            hipLaunchKernelGGL(bcd::k_accumulate, dim3(grid_pix), dim3(bcd::kBlock), 0, st, P, R.B, R.acc + off, keep);
            if (!bc_check(c, hipGetLastError(), "render launch")) return BC_ERR_DEVICE;
        }
    }
    return finish(c, Pall, d_rgba, pix_grid(c, npix), rays_out, shadow_out);
}

// One frame (or row chunk) of the persistent schedule: a single scene, or one scene of a batch.
struct Job {
    const bc_render_desc* d;
    int y0, rows, scene;
    float* d_rgba;
    uint32_t* d_guides;
};

void capture_iterations(const bcd::SceneBufs& S, const bcd::PathBufs& B, const bcd::RenderParams* Pd,
                        const uint32_t* gpre, uint32_t njobs, bool multi, uint32_t* guides, int grid, uint32_t nseg,
                        int iters, bool sort, hipStream_t st) {
    for (int it = 0; it < iters; it++) {
        if (multi) launch_iteration<true>(S, B, Pd, gpre, njobs, it & 1, guides, grid, nseg, it == iters - 1, sort, st);
        else launch_iteration<false>(S, B, Pd, gpre, njobs, it & 1, guides, grid, nseg, it == iters - 1, sort, st);
    }
}

// Persistent schedule with in-place regeneration over the samples of all jobs (one job: a
// scene; several: the packed scenes of a batch, one launch sequence for all of them); `iters`
// iterations are one captured graph.  Jobs are rendered in windows of samples [s0, s0 + ns)
// of every job whose per-sample results fit the result buffer.
int render_persistent_jobs(bc_ctx* c, const std::vector<Job>& jobs, uint64_t* rays_out, uint64_t* shadow_out, int pool,
                           int iters, uint32_t ss, bool sort) {
    const bool multi = !(jobs.size() == 1 && jobs[0].scene < 0);
    const uint32_t njobs = (uint32_t)jobs.size();
    size_t npix_all = 0;
    int max_spp = 0;
    std::vector<size_t> accoff;
    for (const Job& jb : jobs) {
        accoff.push_back(npix_all);
        npix_all += (size_t)jb.rows * (size_t)jb.d->width;
        max_spp = std::max(max_spp, jb.d->spp);
    }
    const size_t cap = multi ? kBatchCap : kResultCap;
    int window_ns = (int)std::min<size_t>((size_t)max_spp, std::max<size_t>(1, cap / npix_all));
    // reserved[5] > 0 caps the samples per window below what the result buffer holds
    if (jobs[0].d->reserved[5] > 0) window_ns = std::min(window_ns, jobs[0].d->reserved[5]);
    size_t total_first = 0;
    for (const Job& jb : jobs)
        total_first += (size_t)jb.rows * (size_t)jb.d->width * (size_t)std::min(window_ns, jb.d->spp);
    const size_t slots = segments_for(std::min<size_t>((size_t)std::max(pool, 1), total_first), ss) * ss;
    const bool guides_on = !multi && jobs[0].d_guides != nullptr;
    if (!ensure_state(c, slots, npix_all, guides_on, total_first, ss, (int)njobs)) return BC_ERR_NOMEM;
    RenderState& R = *c->rs;
    R.B.xg = xcd_group(jobs[0].d);
    const bcd::SceneBufs S = scene_bufs(c);
    std::vector<bcd::RenderParams> P(njobs);
    for (uint32_t k = 0; k < njobs; k++)
        P[k] = make_params(c, jobs[k].d, jobs[k].y0, jobs[k].rows, guides_on, jobs[k].scene);
    const uint32_t nseg = R.nseg;
    const int grid = seg_grid(c, nseg);
    hipStream_t st = c->stream;
    uint32_t* guides = guides_on ? jobs[0].d_guides : R.guide_tmp;
    iters = std::max(2, iters & ~1);
    std::vector<const void*> key = {c->d_nodes, c->d_prims, c->d_ids, c->d_pmat, c->d_mats, c->d_lights, c->d_pdf,
                                    c->d_light_of, c->d_err, R.Pd, guides,
                                    (const void*)(uintptr_t)nseg, (const void*)(uintptr_t)grid,
                                    (const void*)(uintptr_t)iters, (const void*)(uintptr_t)c->stack, (const void*)(uintptr_t)ss,
                                    (const void*)(uintptr_t)R.B.xg, (const void*)(uintptr_t)multi, (const void*)(uintptr_t)sort,
                                    (const void*)(uintptr_t)njobs};
    for (void* p : R.buf) key.push_back(p);
    if (!R.exec || key != R.g_key) {
        if (R.exec) { (void)hipGraphExecDestroy(R.exec); R.exec = nullptr; }
        hipGraph_t graph = nullptr;
        if (!bc_check(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture")) return BC_ERR_DEVICE;
        capture_iterations(S, R.B, R.Pd, R.gpre, njobs, multi, guides, grid, nseg, iters, sort, st);
        if (!bc_check(c, hipStreamEndCapture(st, &graph), "end capture")) return BC_ERR_DEVICE;
        hipError_t e = hipGraphInstantiate(&R.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (!bc_check(c, e, "graph instantiate")) { R.exec = nullptr; return BC_ERR_DEVICE; }
        R.g_key = key;
    }
    if (!clear_tally(c)) return BC_ERR_DEVICE;
    if (multi && !bc_check(c, hipMemsetAsync(R.acc, 0, npix_all * 16, st), "clear pixel sums")) return BC_ERR_DEVICE;
    int windows = 0;
    std::vector<uint32_t> gpre(njobs + 1);
    for (int s0 = 0; s0 < max_spp; s0 += window_ns) {
        gpre[0] = 0;
        for (uint32_t k = 0; k < njobs; k++) {
            P[k].s0 = s0;
            P[k].ns = std::max(0, std::min(window_ns, jobs[k].d->spp - s0));
            gpre[k + 1] = gpre[k] + (uint32_t)P[k].npix * (uint32_t)P[k].ns;
            // bands of whole rows holding about one pool of samples each (window_sample)
            const size_t row = (size_t)P[k].width * (size_t)std::max(1, P[k].ns);
            const size_t rows_b = std::max<size_t>(1, slots / row);
            P[k].band = (int32_t)std::min<size_t>((size_t)P[k].npix, rows_b * (size_t)P[k].width);
        }
        const size_t window_total = gpre[njobs];
        std::memcpy(R.h_params, P.data(), njobs * sizeof(bcd::RenderParams));
        std::memcpy((char*)R.h_params + njobs * sizeof(bcd::RenderParams), gpre.data(), (njobs + 1) * 4);
        if (!bc_check(c, hipMemcpyAsync(R.Pd, R.h_params, njobs * sizeof(bcd::RenderParams) + (njobs + 1) * 4,
                                        hipMemcpyHostToDevice, st), "params"))
            return BC_ERR_DEVICE;
        const int gstart = std::max(grid, multi ? 1 : pix_grid(c, (size_t)P[0].npix));
        hipLaunchKernelGGL(bcd::k_p_start, dim3(gstart), dim3(bcd::kBlock), 0, st, (const bcd::RenderParams*)R.Pd, R.B,
                           R.acc, guides, nseg, (!multi && s0 == 0) ? 1 : 0);
        // a slot holds one path for at most bounces + 1 iterations (the last one records it and
        // restarts the slot), and every segment restarts all of its free slots while it has
        // samples: bound the replays accordingly
        int max_b = 0;
        for (const Job& jb : jobs) max_b = std::max(max_b, jb.d->bounces);
        const size_t max_iters = (size_t)(max_b + 2) * (window_total / slots + 2) + 2;
        const size_t max_graphs = max_iters / (size_t)iters + 4;
        for (size_t g = 0;; g++) {
            if (g >= max_graphs) { c->err = "persistent schedule did not drain"; return BC_ERR_DEVICE; }
            if (!bc_check(c, hipGraphLaunch(R.exec, st), "graph launch") ||
                !bc_check(c, hipMemcpyAsync(R.h_cnt, R.B.flag, 4, hipMemcpyDeviceToHost, st), "drain flag") ||
                !bc_check(c, hipStreamSynchronize(st), "graph"))
                return BC_ERR_DEVICE;
            if (R.h_cnt[0] == 0) break;
        }
        for (uint32_t k = 0; k < njobs; k++) {
            if (P[k].ns == 0) continue;
            bcd::PathBufs Bk = R.B;
            Bk.res = R.B.res + gpre[k];
            hipLaunchKernelGGL(bcd::k_p_reduce, dim3(pix_grid(c, (size_t)P[k].npix)), dim3(bcd::kBlock), 0, st,
                               (const bcd::RenderParams*)(R.Pd + k), Bk, R.acc + accoff[k]);
        }
        if (!bc_check(c, hipGetLastError(), "reduce launch")) return BC_ERR_DEVICE;
        windows++;
    }
    for (uint32_t k = 0; k < njobs; k++) {
        hipLaunchKernelGGL(bcd::k_finalize, dim3(pix_grid(c, (size_t)P[k].npix)), dim3(bcd::kBlock), 0, st, P[k],
                           (const float4*)(R.acc + accoff[k]), (float4*)jobs[k].d_rgba);
    }
    R.last_pool = (int)slots;
    R.last_iters = iters;
    R.last_seg = (int)ss;
    R.last_sort = sort ? 1 : 0;
    R.last_windows = windows;
    R.last_plain = 0;
    return finish_tally(c, rays_out, shadow_out);
}

uint64_t jobs_shape(const std::vector<Job>& jobs) {
    if (jobs.size() == 1 && jobs[0].scene < 0) return 0;
    uint64_t h = 1469598103934665603ull;
    for (const Job& jb : jobs) {
        const uint64_t v[4] = {(uint64_t)jb.d->width, (uint64_t)jb.rows, (uint64_t)jb.d->spp, (uint64_t)jb.d->bounces};
        for (uint64_t x : v) h = (h ^ x) * 1099511628211ull;
    }
    return h;
}

// Renders the jobs with the plain schedule: each scene (or the single scene's rows) in turn.
int render_jobs_plain(bc_ctx* c, const std::vector<Job>& jobs, uint64_t* rays_out, uint64_t* shadow_out) {
    uint64_t rays = 0, shadow = 0;
    for (const Job& jb : jobs) {
        uint64_t r = 0, sh = 0;
        const int rc = render_plain(c, jb.d, jb.y0, jb.rows, jb.d_rgba, jb.d_guides, &r, &sh, nullptr, jb.scene);
        if (rc != BC_OK) return rc;
        rays += r;
        shadow += sh;
    }
    if (rays_out) *rays_out = rays;
    if (shadow_out) *shadow_out = shadow;
    return BC_OK;
}

// Autotuner: times short renders of the same jobs, with enough samples to refill the pool about
// three times (at least 1 spp), and keeps the fastest per sample for this image shape, scene and
// transport depth.  Pool size and segment size are swept at 8 iterations per graph, then the
// graph length for the best pair, then material-sorted shading for the best triple, and last the
// plain schedule on the same samples (on a small scene its passes already fill the chip and the
// persistent bookkeeping only costs).  Each trial is timed after one untimed run (which captures
// the graph) unless the trial itself holds more than 2^24 path samples.
bool autotune(bc_ctx* c, const std::vector<Job>& jobs, int& pool, int& iters, uint32_t& seg, bool& sort, bool& plain) {
    RenderState& R = *c->rs;
    const bc_render_desc* d = jobs[0].d;
    const uint64_t shape = jobs_shape(jobs);
    for (const auto& t : R.tunes)
        if (t.w == d->width && t.rows == jobs[0].rows && t.spp == d->spp && t.bounces == d->bounces && t.topo == c->topo &&
            t.geom == c->geom && t.shape == shape) {
            pool = t.pool;
            iters = t.iters;
            seg = (uint32_t)t.seg;
            sort = t.sort != 0;
            plain = t.plain != 0;
            return true;
        }
    size_t npix = 0;
    int max_spp = 0;
    for (const Job& jb : jobs) {
        npix += (size_t)jb.rows * (size_t)jb.d->width;
        max_spp = std::max(max_spp, jb.d->spp);
    }
    std::vector<bc_render_desc> trial(jobs.size());
    std::vector<Job> tj = jobs;
    const int pools[] = {1 << 18, 1 << 19, 1 << 20, 1 << 21, 1 << 22, 1 << 23, 1 << 24};
    const uint32_t segs[] = {256, 1024};
    float best = 1e30f;
    pool = 1 << 20;
    iters = 8;
    seg = bcd::kSegPersist;
    sort = false;
    plain = false;
    // p: pool slots (also the trial's sample count); it < 0 times the plain schedule
    auto time_trial = [&](int p, int it, uint32_t ss, bool so, float& per_sample) -> bool {
        const int tspp = (int)std::min<size_t>((size_t)max_spp, std::max<size_t>(1, (3 * (size_t)p + npix - 1) / npix));
        for (size_t k = 0; k < jobs.size(); k++) {
            trial[k] = *jobs[k].d;
            trial[k].spp = std::min(trial[k].spp, tspp);
            tj[k].d = &trial[k];
        }
        auto run = [&]() {
            return it < 0 ? render_jobs_plain(c, tj, nullptr, nullptr)
                          : render_persistent_jobs(c, tj, nullptr, nullptr, p, it, ss, so);
        };
        if (npix * (size_t)tspp <= ((size_t)1 << 24) && run() != BC_OK) return false;
        float ms = 0;
        if (!bc_check(c, hipEventRecord(c->ev0, c->stream), "tune") || run() != BC_OK ||
            !bc_check(c, hipEventRecord(c->ev1, c->stream), "tune") ||
            !bc_check(c, hipEventSynchronize(c->ev1), "tune") ||
            !bc_check(c, hipEventElapsedTime(&ms, c->ev0, c->ev1), "tune"))
            return false;
        per_sample = ms / (float)tspp;
        return true;
    };
    for (int p : pools) {
        if ((size_t)p > npix * (size_t)max_spp && p != pools[0]) continue;
        for (uint32_t ss : segs) {
            float ms;
            if (!time_trial(p, 8, ss, false, ms)) return false;
            if (ms < best) { best = ms; pool = p; seg = ss; }
        }
    }
    for (int it : {4, 16}) {
        float ms;
        if (!time_trial(pool, it, seg, false, ms)) return false;
        if (ms < best) { best = ms; iters = it; }
    }
    {
        float ms;
        if (!time_trial(pool, iters, seg, true, ms)) return false;
        if (ms < best) { best = ms; sort = true; }
    }
    {
        float ms;
        if (!time_trial(pool, -1, seg, false, ms)) return false;
        if (ms < best) { best = ms; plain = true; }
    }
    R.tunes.push_back({d->width, jobs[0].rows, d->spp, d->bounces, c->topo, c->geom, shape, pool, iters, (int)seg,
                       sort ? 1 : 0, plain ? 1 : 0});
    return true;
}

// Schedule shape of a mode-1 render from the first description: explicit or autotuned (which may
// choose the plain schedule).
int persistent_shape(bc_ctx* c, const std::vector<Job>& jobs, int& pool, int& iters, uint32_t& seg, bool& sort,
                     bool& plain) {
    const bc_render_desc* d = jobs[0].d;
    pool = d->reserved[0];
    iters = d->reserved[1];
    seg = d->reserved[2] > 0 ? (uint32_t)d->reserved[2] : bcd::kSegPersist;
    sort = d->reserved[4] != 0;
    plain = false;
    if (d->reserved[5] < 0) {
        c->err = "samples per window (reserved[5]) must not be negative";
        return BC_ERR_ARG;
    }
    if (seg % bcd::kBlock != 0 || seg > bcd::kSegMax) {
        c->err = "segment size (reserved[2]) must be a multiple of 256 up to 4096";
        return BC_ERR_ARG;
    }
    if (pool <= 0 || iters <= 0) {
        if (!autotune(c, jobs, pool, iters, seg, sort, plain)) return BC_ERR_DEVICE;
    }
    return BC_OK;
}

// Runs the jobs of a mode-1 call with the schedule persistent_shape chose.
int render_mode1(bc_ctx* c, const std::vector<Job>& jobs, uint64_t* rays_out, uint64_t* shadow_out) {
    int pool, iters;
    uint32_t seg;
    bool sort, plain;
    int rc = persistent_shape(c, jobs, pool, iters, seg, sort, plain);
    if (rc != BC_OK) return rc;
    if (!plain) return render_persistent_jobs(c, jobs, rays_out, shadow_out, pool, iters, seg, sort);
    rc = render_jobs_plain(c, jobs, rays_out, shadow_out);
    RenderState& R = *c->rs;
    R.last_pool = pool;
    R.last_iters = iters;
    R.last_seg = (int)seg;
    R.last_sort = sort ? 1 : 0;
    R.last_windows = 0;
    R.last_plain = 1;
    return rc;
}

bool valid_desc(bc_ctx* c, const bc_render_desc* d, int y0, int rows) {
    if (d->width <= 0 || d->height <= 0 || d->spp <= 0 || d->spp > 65536 || d->bounces < 0 || d->bounces > 1024 ||
        y0 < 0 || rows <= 0 || y0 + rows > d->height || d->sampler < 0 || d->sampler > 2 ||
        (int64_t)d->width * d->height > ((int64_t)1 << 28) || d->mode < 0 || d->mode > 1) {
        c->err = "invalid render description or row range";
        return false;
    }
    if (d->restir_spatial < 0 || d->restir_spatial > 16 || d->restir_radius < 0 || d->restir_radius > 64) {
        c->err = "restir_spatial must be in [0, 16] and restir_radius in [0, 64]";
        return false;
    }
    if (d->restir_spatial > 0 && (d->restir_candidates == 0 || d->mode != 0)) {
        c->err = "spatial reuse (restir_spatial > 0) needs restir_candidates > 0 and the plain schedule (mode 0)";
        return false;
    }
    if (d->restir_candidates < 0 || d->restir_candidates > BC_RIS_MAX) {
        c->err = "restir_candidates must be in [0, 64]";
        return false;
    }
    if (d->restir_candidates == 0 && (d->sun_irradiance[0] != 0.0f || d->sun_irradiance[1] != 0.0f ||
                                      d->sun_irradiance[2] != 0.0f)) {
        c->err = "a sun (sun_irradiance != 0) needs restir_candidates > 0";
        return false;
    }
    return true;
}

}  // namespace

void bc_render_release(bc_ctx* c) {
    if (!c || !c->rs) return;
    if (c->rs->exec) (void)hipGraphExecDestroy(c->rs->exec);
    if (c->rs->keep) (void)hipFree(c->rs->keep);
    for (void* p : c->rs->buf)
        if (p) (void)hipFree(p);
    if (c->rs->h_cnt) (void)hipHostFree(c->rs->h_cnt);
    delete c->rs;
    c->rs = nullptr;
}

extern "C" int bc_render_device(bc_ctx* c, const bc_render_desc* d, int y0, int rows, float* d_rgba,
                                uint32_t* d_guides, uint64_t* rays_out, uint64_t* shadow_out) {
    if (!c || !d) return BC_ERR_ARG;
    if (!c->has_scene) { c->err = "no scene uploaded"; return BC_ERR_SCENE; }
    if (!d_rgba) { c->err = "invalid render description or row range"; return BC_ERR_ARG; }
    if (!valid_desc(c, d, y0, rows)) return BC_ERR_ARG;
    (void)hipSetDevice(c->device);
    // zero transport segments render black on either schedule; the plain one does it without tracing
    if (d->mode == 0 || d->bounces == 0) return render_plain(c, d, y0, rows, d_rgba, d_guides, rays_out, shadow_out);
    if (!ensure_state(c, 1, (size_t)rows * (size_t)d->width, d_guides != nullptr, 1, bcd::kSegPersist)) return BC_ERR_NOMEM;
    const std::vector<Job> jobs = {Job{d, y0, rows, -1, d_rgba, d_guides}};
    return render_mode1(c, jobs, rays_out, shadow_out);
}

extern "C" int bc_batch_render_device(bc_ctx* c, int n, const bc_render_desc* descs, float* const* d_rgba,
                                      uint64_t* rays_out, uint64_t* shadow_out) {
    if (!c || n <= 0 || !descs || !d_rgba) return BC_ERR_ARG;
    if (c->batch.empty()) { c->err = "no batch uploaded"; return BC_ERR_SCENE; }
    if ((size_t)n != c->batch.size() || n > kMaxJobs) {
        c->err = "one description per batch scene is needed (at most 256 scenes)";
        return BC_ERR_ARG;
    }
    for (int k = 0; k < n; k++) {
        if (!d_rgba[k]) { c->err = "invalid render description or row range"; return BC_ERR_ARG; }
        if (!valid_desc(c, &descs[k], 0, descs[k].height)) {
            c->err = "batch scene " + std::to_string(k) + ": " + c->err;
            return BC_ERR_ARG;
        }
    }
    (void)hipSetDevice(c->device);
    uint64_t rays = 0, shadow = 0;
    if (descs[0].mode == 0) {  // the scenes one after another, plain schedule
        for (int k = 0; k < n; k++) {
            uint64_t r = 0, sh = 0;
            const int rc = render_plain(c, &descs[k], 0, descs[k].height, d_rgba[k], nullptr, &r, &sh, nullptr, k);
            if (rc != BC_OK) return rc;
            rays += r;
            shadow += sh;
        }
    } else {  // one persistent schedule over the samples of every scene
        std::vector<Job> jobs;
        int max_b = 0;
        for (int k = 0; k < n; k++) {
            jobs.push_back(Job{&descs[k], 0, descs[k].height, k, d_rgba[k], nullptr});
            max_b = std::max(max_b, descs[k].bounces);
        }
        if (max_b == 0) {
            c->err = "a batch render needs at least one transport segment";
            return BC_ERR_ARG;
        }
        if (!ensure_state(c, 1, 1, false, 1, bcd::kSegPersist, n)) return BC_ERR_NOMEM;
        const int rc = render_mode1(c, jobs, &rays, &shadow);
        if (rc != BC_OK) return rc;
    }
    if (rays_out) *rays_out = rays;
    if (shadow_out) *shadow_out = shadow;
    return BC_OK;
}

// ------------------------------------------------------------------------------ no-allocation render
// bc_render_noalloc runs the plain schedule with its path state, pixel sums, queue lengths and
// tallies in one fixed device array, `scratch`, instead of buffers the library allocates.  Each GPU
// holds its own copy of the array from the time the library's kernels are loaded on it until the
// process ends.  Host code never passes the array's name to a kernel (on the host the name is a
// stand-in, not a device address); it uses the address hipGetSymbolAddress returns for the current
// device, resolved once per device by bc_scratch_init and kept in every context on that device.
namespace {
constexpr size_t kNaSlots = (size_t)1 << 18;   // path slots of one pass
constexpr size_t kNaPixels = (size_t)1 << 18;  // pixels (rows x width) of one call
constexpr size_t kNaMaxSeg = kNaSlots / (size_t)bcd::kBlock;  // plain segments are >= kBlock slots
constexpr size_t na_align(size_t x) { return (x + 255) & ~(size_t)255; }
// Byte offsets of the arrays carved from the scratch array.
constexpr size_t kNaVec16 = 9;  // o d T L pp meta so sd sc: 16 bytes per slot each
constexpr size_t kNaOffHit = kNaVec16 * kNaSlots * 16;
constexpr size_t kNaOffU32 = kNaOffHit + kNaSlots * 8;   // nseg q0 q1 sq: 4 bytes per slot each
constexpr size_t kNaOffAcc = kNaOffU32 + 4 * kNaSlots * 4;
constexpr size_t kNaOffSeg = kNaOffAcc + kNaPixels * 16;  // qc0 qc1 sqc: 4 bytes per segment each
constexpr size_t kNaOffTally = na_align(kNaOffSeg + 3 * kNaMaxSeg * 4);
constexpr size_t kNaTallyWords = (size_t)bcd::kMaxGrid + kNaMaxSeg;
constexpr size_t kNaOffTotals = na_align(kNaOffTally + kNaTallyWords * 8);
constexpr size_t kNaBytes = na_align(kNaOffTotals + 64);
static_assert(kNaBytes < ((size_t)49 << 20), "scratch of the no-allocation render is stated as under 49 MiB");
constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
void* g_na_ptr[kNaMaxDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace bcd {
__global__ __launch_bounds__(kBlock) void k_na_clear(unsigned long long* __restrict__ w, uint32_t n) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) w[i] = 0ull;
}
}  // namespace bcd

void bc_scratch_init(bc_ctx* c) {
    if (!c) return;
    if (c->device < 0 || c->device >= kNaMaxDevices) {
        c->scratch_err = "device ordinal beyond the scratch table";
        return;
    }
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (!g_na_ptr[c->device]) {
        // bc_create made c->device current; the address is that device's copy of the array
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            c->scratch_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return;
        }
        g_na_ptr[c->device] = p;
    }
    c->scratch = g_na_ptr[c->device];
}

extern "C" int bc_render_noalloc(bc_ctx* c, const bc_render_desc* d, int y0, int rows, float* d_rgba,
                                 uint32_t* d_guides, uint64_t* rays_out, uint64_t* shadow_out) {
    if (!c || !d) return BC_ERR_ARG;
    if (!c->scratch) {
        c->err = "no scratch array on this device: " + c->scratch_err;
        return BC_ERR_DEVICE;
    }
    if (!c->has_scene) { c->err = "no scene uploaded"; return BC_ERR_SCENE; }
    if (!d_rgba) { c->err = "invalid render description or row range"; return BC_ERR_ARG; }
    if (!valid_desc(c, d, y0, rows)) return BC_ERR_ARG;
    if (d->mode != 0 || d->restir_spatial != 0) {
        c->err = "the no-allocation render runs the plain schedule (mode 0) without spatial reuse";
        return BC_ERR_ARG;
    }
    const size_t npix = (size_t)rows * (size_t)d->width;
    if (npix > kNaPixels) {
        c->err = "the no-allocation render takes at most 262144 pixels (rows x width) per call";
        return BC_ERR_ARG;
    }
    (void)hipSetDevice(c->device);
    int spp_pass = (int)std::max<size_t>(1, kNaSlots / npix);
    if (d->samples_per_pass > 0) spp_pass = std::min(spp_pass, d->samples_per_pass);
    spp_pass = std::min(spp_pass, d->spp);
    const size_t paths = npix * (size_t)spp_pass;
    const uint32_t ss = plain_segment(c, paths);
    const uint32_t nseg = (uint32_t)segments_for(paths, ss);
    char* base = static_cast<char*>(c->scratch);
    bcd::PathBufs B{};
    B.seg = ss;
    B.xg = xcd_group(d);
    float4** v16[kNaVec16 - 1] = {&B.o, &B.d, &B.T, &B.L, &B.pp, &B.so, &B.sd, &B.sc};
    for (size_t k = 0; k < kNaVec16 - 1; k++) *v16[k] = (float4*)(base + k * kNaSlots * 16);
    B.meta = (uint4*)(base + (kNaVec16 - 1) * kNaSlots * 16);
    B.hit = (float2*)(base + kNaOffHit);
    uint32_t* u32 = (uint32_t*)(base + kNaOffU32);
    B.nseg = u32;
    B.q[0] = u32 + kNaSlots;
    B.q[1] = u32 + 2 * kNaSlots;
    B.sq = u32 + 3 * kNaSlots;
    float4* acc = (float4*)(base + kNaOffAcc);
    uint32_t* segw = (uint32_t*)(base + kNaOffSeg);
    B.qc[0] = segw;
    B.qc[1] = segw + kNaMaxSeg;
    B.sqc = segw + 2 * kNaMaxSeg;
    B.tally = (unsigned long long*)(base + kNaOffTally);
    B.totals = (unsigned long long*)(base + kNaOffTotals);
    B.flag = (uint32_t*)(base + kNaOffTotals + 32);
    const bcd::SceneBufs S = scene_bufs(c);
    bcd::RenderParams P = make_params(c, d, y0, rows, d_guides != nullptr);
    const int grid = seg_grid(c, nseg);
    const int grid_pix = pix_grid(c, npix);
    hipStream_t st = c->stream;
    // the array keeps the previous call's contents: the tallies start at zero by a fill kernel
    // (the pixel sums are zeroed by the first pass's raygen, queue lengths written before use)
    const uint32_t ntally = (uint32_t)((size_t)bcd::kMaxGrid + nseg);
    hipLaunchKernelGGL(bcd::k_na_clear, dim3((ntally + bcd::kBlock - 1) / bcd::kBlock), dim3(bcd::kBlock), 0, st, B.tally,
                       ntally);
    for (int s0 = 0; s0 < d->spp; s0 += spp_pass) {
        P.s0 = s0;
        P.ns = std::min(spp_pass, d->spp - s0);
        hipLaunchKernelGGL(bcd::k_raygen, dim3(grid), dim3(bcd::kBlock), 0, st, P, B, (float*)acc, d_guides, nseg);
        for (int b = 0; b < d->bounces; b++) launch_bounce(S, B, P, b & 1, b, d_guides, grid, nseg, st);
        hipLaunchKernelGGL(bcd::k_accumulate, dim3(grid_pix), dim3(bcd::kBlock), 0, st, P, B, acc, (float4*)nullptr);
        if (!bc_check(c, hipGetLastError(), "render launch")) return BC_ERR_DEVICE;
    }
    hipLaunchKernelGGL(bcd::k_finalize, dim3(grid_pix), dim3(bcd::kBlock), 0, st, P, (const float4*)acc, (float4*)d_rgba);
    hipLaunchKernelGGL(bcd::k_tally, dim3(1), dim3(bcd::kTallyBlock), 0, st, B, nseg);
    unsigned long long tot[2] = {0, 0};
    if (!bc_check(c, hipGetLastError(), "finalize launch") ||
        !bc_check(c, hipMemcpyAsync(tot, B.totals, 16, hipMemcpyDeviceToHost, st), "read totals") ||
        !bc_check(c, hipStreamSynchronize(st), "render"))
        return BC_ERR_DEVICE;
    if (rays_out) *rays_out = tot[0];
    if (shadow_out) *shadow_out = tot[1];
    return bc_check_kernel_error(c);
}

extern "C" int bc_render_tuning(const bc_ctx* c, int32_t* out8) {
    if (!c || !out8) return BC_ERR_ARG;
    out8[0] = c->rs ? c->rs->last_pool : 0;
    out8[1] = c->rs ? c->rs->last_iters : 0;
    out8[2] = c->stack;
    out8[3] = c->rs ? c->rs->last_windows : 0;
    out8[4] = c->rs ? c->rs->last_seg : 0;
    out8[5] = c->rs ? c->rs->last_sort : 0;
    out8[6] = c->rs ? c->rs->last_plain : 0;
    out8[7] = 0;
    return BC_OK;
}

namespace {

// The description bytes that identify a render: every field (the struct has no inner padding),
// without the trailing padding.
constexpr size_t kDescBytes = offsetof(bc_render_desc, restir_radius) + sizeof(int32_t);
static_assert(offsetof(bc_render_desc, origin) == 14 * 4 && offsetof(bc_render_desc, mode) == 14 * 4 + 12 * 8 &&
                  offsetof(bc_render_desc, sun_dir) == offsetof(bc_render_desc, restir_candidates) + 4,
              "bc_render_desc is compared bytewise: no padding between its fields");
constexpr int kNotAhead = 1 << 30;

// Render-ahead for bc_render.  A frame rendered in row chunks pays one drain of the path pool
// (persistent schedule) or one short run of passes (plain schedule) per call, plus a host round
// trip (terrain 1080p, 64 spp, persistent: 147 ms in 64-row chunks against 111 ms in one call);
// so the first call of a run of chunks renders the rest of the frame, or as many further rows (a
// multiple of its chunk height) as one window of per-sample results holds (persistent) or as
// kAheadPixels allows (plain), into a buffer of the context, and the calls for those rows copy
// them out.  Every pixel
// depends only on the description, the scene and its own index (spatial reuse, which reads
// neighbouring pixels' reservoirs among the rendered rows, never renders ahead), so the rows are
// what rendering them alone gives.  Rows are served once, in increasing order: a request that
// starts before the end of the last rows served (a new frame, a repeated chunk) renders anew, so
// the block is a prefetch of the frame being read, never a cache of earlier frames.  Returns
// kNotAhead when the call should render its rows alone.
constexpr size_t kAheadPixels = size_t(1) << 23;
int render_ahead(bc_ctx* c, const bc_render_desc* d, int y0, int rows, float* rgba, uint32_t* guides,
                 uint64_t* rays_out) {
    const bool have = c->ahead_valid && c->ahead_epoch == c->scene_epoch && y0 >= c->ahead_next &&
                      y0 + rows <= c->ahead_y0 + c->ahead_rows && (!guides || c->ahead_guides) &&
                      std::memcmp(&c->ahead_desc, d, kDescBytes) == 0;
    if (!have) {
        if (!c->has_scene) { c->err = "no scene uploaded"; return BC_ERR_SCENE; }
        if (!valid_desc(c, d, y0, rows)) return BC_ERR_ARG;
        const size_t per_row = d->mode == 1 ? (size_t)d->width * (size_t)d->spp : (size_t)d->width;
        const size_t cap = d->mode == 1 ? kResultCap : kAheadPixels;
        const int fit = (int)std::min<size_t>((size_t)d->height, cap / per_row);
        // the rest of the frame when it fits, else whole chunks
        const int block = d->height - y0 <= fit ? d->height - y0 : std::max(rows, fit / rows * rows);
        if (block <= rows) return kNotAhead;
        const size_t bpix = (size_t)block * (size_t)d->width;
        c->ahead_valid = false;
        c->pin_ahead = false;
        c->pin_pre_y0 = -1;
        for (int i = 0; i < 2; i++)  // a copy still reading the old block ends before it can be freed
            if (c->pin_busy[i] && !bc_check(c, hipEventSynchronize(c->pin_ev[i]), "rows rendered ahead"))
                return BC_ERR_DEVICE;
        // image, guides, then the per-row segment counts (8-byte aligned)
        const size_t row_off = (bpix * 28 + 7) & ~size_t(7);
        if (!bc_ensure(c, c->ahead_buf, c->ahead_cap, row_off + (size_t)block * 8 + 64, "rows rendered ahead"))
            return BC_ERR_NOMEM;
        uint32_t* dg = guides ? (uint32_t*)((char*)c->ahead_buf + bpix * 16) : nullptr;
        const int rc = bc_render_device(c, d, y0, block, (float*)c->ahead_buf, dg, nullptr, nullptr);
        if (rc != BC_OK) return rc;
        unsigned long long* d_rows = (unsigned long long*)((char*)c->ahead_buf + row_off);
        c->ahead_row_segs.resize((size_t)block);
        hipLaunchKernelGGL(bcd::k_row_segments, dim3((unsigned)block), dim3(256), 0, c->stream,
                           (const float4*)c->ahead_buf, d->width, d_rows);
        if (!bc_check(c, hipGetLastError(), "row segment counts") ||
            !bc_check(c, hipMemcpyAsync(c->ahead_row_segs.data(), d_rows, (size_t)block * 8, hipMemcpyDeviceToHost,
                                        c->stream), "row segment counts") ||
            !bc_check(c, hipStreamSynchronize(c->stream), "row segment counts"))
            return BC_ERR_DEVICE;
        // a block that fits the pinned buffers is read back once, and its chunks are then copied
        // on the host (closed box 256x256 in 64-row chunks: one transfer instead of four pairs)
        if (bc_d2h_hold(c, c->ahead_buf, bpix * 16, dg, guides ? bpix * 12 : 0, c->stream, "read rows rendered ahead") < 0)
            return BC_ERR_DEVICE;
        c->ahead_valid = true;
        c->ahead_desc = *d;
        c->ahead_y0 = y0;
        c->ahead_next = y0;
        c->ahead_rows = block;
        c->ahead_guides = guides != nullptr;
        c->ahead_epoch = c->scene_epoch;
    }
    const size_t off = (size_t)(y0 - c->ahead_y0) * (size_t)d->width, npix = (size_t)rows * (size_t)d->width;
    const size_t bpix = (size_t)c->ahead_rows * (size_t)d->width;
    const uint32_t* dg = (const uint32_t*)((const char*)c->ahead_buf + bpix * 16);
    // starts copying the next chunk of the same height into a pinned buffer, so that its transfer
    // runs while this call copies its own rows out and the caller works on them (terrain 1080p in
    // 64-row chunks: 3.4 MB per chunk); false on a failed call
    auto start_next = [&]() {
        const int ny0 = y0 + rows, nrows = std::min(rows, c->ahead_y0 + c->ahead_rows - ny0);
        if (nrows <= 0) return true;
        const size_t noff = (size_t)(ny0 - c->ahead_y0) * (size_t)d->width, nn = (size_t)nrows * (size_t)d->width;
        const int b = bc_d2h_prefetch(c, (const float*)c->ahead_buf + noff * 4, nn * 16, guides ? dg + noff * 3 : nullptr,
                                      guides ? nn * 12 : 0, c->stream, "copy image and guides");
        if (b == -2) return false;
        if (b >= 0) {
            c->pin_pre_y0 = ny0;
            c->pin_pre_rows = nrows;
            c->pin_pre_buf = b;
            c->pin_pre_guides = guides != nullptr;
        }
        return true;
    };
    if (c->pin_ahead) {  // the block is still in the pinned buffers
        bc_host_copy(c, rgba, (const char*)c->h_pin[0] + off * 16, npix * 16);
        if (guides) bc_host_copy(c, guides, (const char*)c->h_pin[1] + off * 12, npix * 12);
    } else if (c->pin_pre_y0 == y0 && c->pin_pre_rows == rows && (c->pin_pre_guides || !guides)) {
        // these rows were copied into pinned buffer b after the last chunk was served; the next
        // chunk's copy goes to the other buffer
        const int b = c->pin_pre_buf;
        c->pin_pre_y0 = -1;
        if (!bc_check(c, hipEventSynchronize(c->pin_ev[b]), "copy image and guides")) return BC_ERR_DEVICE;
        c->pin_busy[b] = false;
        c->pin_next = b ^ 1;
        if (!start_next()) return BC_ERR_DEVICE;
        const char* h = static_cast<const char*>(c->h_pin[b]);
        bc_host_copy(c, rgba, h, npix * 16);
        if (guides) bc_host_copy(c, guides, h + npix * 16, npix * 12);
    } else if (!bc_d2h_pair(c, rgba, (const float*)c->ahead_buf + off * 4, npix * 16, guides, dg + off * 3,
                            guides ? npix * 12 : 0, c->stream, "copy image and guides") ||
               !start_next()) {
        return BC_ERR_DEVICE;
    }
    c->ahead_next = y0 + rows;
    // the rows' path segments, from the per-pixel counts as renderOpenCL sums them
    // (gpu_opencl_windows.go:681-685), summed per row on the device when the block was rendered
    if (rays_out) {
        uint64_t r = 0;
        for (int k = y0 - c->ahead_y0; k < y0 - c->ahead_y0 + rows; k++) r += c->ahead_row_segs[(size_t)k];
        *rays_out = r;
    }
    return BC_OK;
}

}  // namespace

extern "C" int bc_render(bc_ctx* c, const bc_render_desc* d, int y0, int rows, float* rgba, uint32_t* guides,
                         uint64_t* rays_out, const volatile int32_t* cancel) {
    if (!c || !d || !rgba) return BC_ERR_ARG;
    if (cancel && *cancel) { c->err = "render cancelled"; return BC_ERR_CANCELLED; }
    if (d->width <= 0 || rows <= 0) { c->err = "invalid render description or row range"; return BC_ERR_ARG; }
    if (c->ahead_on && d->bounces > 0 && !(d->restir_candidates > 0 && d->restir_spatial > 0)) {
        const int rc = render_ahead(c, d, y0, rows, rgba, guides, rays_out);
        if (rc != kNotAhead) return rc;
    }
    const size_t npix = (size_t)rows * (size_t)d->width;
    // output staging kept in the context, so repeated chunk renders reuse one buffer (and a
    // captured graph keyed on it)
    if (!bc_ensure(c, c->out_buf, c->out_cap, npix * 28 + 64, "render output")) return BC_ERR_NOMEM;
    uint32_t* dg = guides ? (uint32_t*)((char*)c->out_buf + npix * 16) : nullptr;
    int rc = bc_render_device(c, d, y0, rows, (float*)c->out_buf, dg, rays_out, nullptr);
    if (rc == BC_OK) {
        if (!bc_d2h_pair(c, rgba, c->out_buf, npix * 16, guides, dg, guides ? npix * 12 : 0, c->stream,
                         "copy image and guides"))
            rc = BC_ERR_DEVICE;
    }
    return rc;
}

namespace {
void launch_bounce_adj(const bcd::SceneBufs& S, const bcd::PathBufs& B, const bcd::RenderParams& P, int cur,
                       int bounce, uint32_t* guides, const bcd::Adj& A, float* grad, int grid, int grid_adj,
                       uint32_t nseg, hipStream_t st) {
    hipLaunchKernelGGL(bcd::k_extend, dim3(grid), dim3(bcd::kBlock), stack_lds(S), st, S, B, cur, nseg, P.root);
    hipLaunchKernelGGL(bcd::k_shade_adj, dim3(grid_adj), dim3(bcd::kBlock), 0, st, S, B, P, cur, bounce, guides, A, grad,
                       nseg);
    hipLaunchKernelGGL(bcd::k_shadow_adj, dim3(grid_adj), dim3(bcd::kBlock), stack_lds(S), st, S, B, P, A, grad, nseg);
}
bool same_shape(const bc_render_desc& a, const bc_render_desc& b) {
    return a.width == b.width && a.height == b.height && a.spp == b.spp && a.bounces == b.bounces &&
           a.rr_depth == b.rr_depth && a.sampler == b.sampler && a.mis == b.mis && a.power_lights == b.power_lights &&
           a.sky_mode == b.sky_mode && a.seed == b.seed && std::memcmp(a.origin, b.origin, sizeof a.origin) == 0 &&
           std::memcmp(a.lower, b.lower, sizeof a.lower) == 0 && std::memcmp(a.horiz, b.horiz, sizeof a.horiz) == 0 &&
           std::memcmp(a.vert, b.vert, sizeof a.vert) == 0 && a.restir_candidates == b.restir_candidates;
}
}  // namespace

extern "C" int bc_forward(bc_ctx* c, const bc_render_desc* d, int y0, int rows, float* rgba, uint64_t* rays_out) {
    if (!c || !d || !rgba) return BC_ERR_ARG;
    if (!c->has_scene) { c->err = "no scene uploaded"; return BC_ERR_SCENE; }
    if (d->firefly_clamp > 0.0f) { c->err = "the differentiable render needs firefly_clamp <= 0"; return BC_ERR_ARG; }
    if (d->width <= 0 || d->height <= 0 || d->spp <= 0 || d->spp > 65536 || d->bounces < 0 || d->bounces > 1024 ||
        y0 < 0 || rows <= 0 || y0 + rows > d->height || d->sampler < 0 || d->sampler > 2) {
        c->err = "invalid render description or row range";
        return BC_ERR_ARG;
    }
    if (c->nmat > bcd::kAdjMaxMat) { c->err = "the differentiable render supports at most 1024 materials"; return BC_ERR_ARG; }
    if (d->restir_candidates != 0 || d->sun_irradiance[0] != 0.0f || d->sun_irradiance[1] != 0.0f ||
        d->sun_irradiance[2] != 0.0f) {
        c->err = "the differentiable render supports one-sample light sampling only (restir_candidates = 0, no sun)";
        return BC_ERR_ARG;
    }
    (void)hipSetDevice(c->device);
    const size_t npix = (size_t)rows * (size_t)d->width;
    if (!ensure_state(c, 1, npix, false, 1, bcd::kSegPlain)) return BC_ERR_NOMEM;
    RenderState& R = *c->rs;
    if (!bc_ensure(c, R.keep, R.keep_cap, npix * (size_t)d->spp * 16, "forward sample buffer")) return BC_ERR_NOMEM;
    if (!bc_ensure(c, c->out_buf, c->out_cap, npix * 28 + 64, "render output")) return BC_ERR_NOMEM;
    int rc = render_plain(c, d, y0, rows, (float*)c->out_buf, nullptr, rays_out, nullptr, (float4*)R.keep);
    if (rc != BC_OK) return rc;
    if (!bc_d2h(c, rgba, c->out_buf, npix * 16, c->stream, "copy image")) return BC_ERR_DEVICE;
    R.keep_desc = *d;
    R.keep_y0 = y0;
    R.keep_rows = rows;
    R.keep_topo = c->topo;
    return BC_OK;
}

extern "C" int bc_backward(bc_ctx* c, const bc_render_desc* d, int y0, int rows, const float* dl, float* dcolor,
                           float* demission) {
    if (!c || !d || !dl || !dcolor || !demission) return BC_ERR_ARG;
    RenderState* Rp = c->rs;
    if (!Rp || !Rp->keep || Rp->keep_y0 != y0 || Rp->keep_rows != rows || Rp->keep_topo != c->topo ||
        !same_shape(Rp->keep_desc, *d)) {
        c->err = "bc_backward needs a bc_forward of the same description and rows first";
        return BC_ERR_ARG;
    }
    (void)hipSetDevice(c->device);
    const size_t npix = (size_t)rows * (size_t)d->width;
    const size_t max_paths = std::max<size_t>(npix, (size_t)1 << 22);
    const int spp_pass = (int)std::min<size_t>((size_t)d->spp, std::max<size_t>(1, max_paths / npix));
    const size_t paths = npix * (size_t)spp_pass;
    if (!ensure_state(c, paths, npix, false, paths, plain_segment(c, paths))) return BC_ERR_NOMEM;
    RenderState& R = *c->rs;
    R.B.xg = xcd_group(d);
    const size_t ng = (size_t)c->nmat * 6;
    const size_t ngc = ng * bcd::kGradCopies;
    if (!bc_ensure(c, c->out_buf, c->out_cap, std::max(npix * 28 + 64, npix * 12 + ngc * 4 + 64), "adjoint buffers"))
        return BC_ERR_NOMEM;
    float* d_dl = (float*)c->out_buf;
    float* d_grad = d_dl + npix * 3;
    hipStream_t st = c->stream;
    if (!bc_h2d(c, d_dl, dl, npix * 12, st, "upload d loss") ||
        !bc_check(c, hipMemsetAsync(d_grad, 0, ngc * 4, st), "clear gradient"))
        return BC_ERR_DEVICE;
    const bcd::SceneBufs S = scene_bufs(c);
    bcd::RenderParams P = make_params(c, d, y0, rows, false);
    bcd::Adj A{};
    A.dl = d_dl;
    A.inv_spp = 1.0f / (float)d->spp;
    A.Lt = R.B.res;
    A.sg = R.sg;
    const uint32_t nseg = R.nseg;
    const int grid = seg_grid(c, nseg);
    // the kernels that flush per-workgroup gradients walk several segments per workgroup
    const int grid_adj = (int)std::min<size_t>(nseg, (size_t)c->num_cus * 16);
    for (int s0 = 0; s0 < d->spp; s0 += spp_pass) {
        P.s0 = s0;
        P.ns = std::min(spp_pass, d->spp - s0);
        hipLaunchKernelGGL(bcd::k_raygen_adj, dim3(grid), dim3(bcd::kBlock), 0, st, P, R.B, (const float4*)R.keep,
                           R.B.res, nseg);
        for (int b = 0; b < d->bounces; b++) {
            const int cur = b & 1;
            launch_bounce_adj(S, R.B, P, cur, b, R.guide_tmp, A, d_grad, grid, grid_adj, nseg, st);
        }
        if (!bc_check(c, hipGetLastError(), "adjoint launch")) return BC_ERR_DEVICE;
    }
    std::vector<float> g(ngc);
    if (!bc_check(c, hipMemcpyAsync(g.data(), d_grad, ngc * 4, hipMemcpyDeviceToHost, st), "read gradient") ||
        !bc_check(c, hipStreamSynchronize(st), "backward"))
        return BC_ERR_DEVICE;
    const size_t n3 = (size_t)c->nmat * 3;
    for (int k = 0; k < bcd::kGradCopies; k++) {
        const float* gk = g.data() + (size_t)k * ng;
        for (size_t i = 0; i < n3; i++) {
            dcolor[i] += gk[i];
            demission[i] += gk[n3 + i];
        }
    }
    return bc_check_kernel_error(c);
}
