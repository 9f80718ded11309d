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
#include "integrator_device.hpp"
#include "trace_device.hpp"

namespace bcd {

// ------------------------------------------------------------------------------ raygen
__global__ __launch_bounds__(kBlock) void k_raygen(RenderParams P, PathBufs B, float* __restrict__ acc,
                                                   uint32_t* __restrict__ guides) {
    const int n = P.ns * P.npix;
    const int stride = gridDim.x * kBlock;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        B.cnt[0] = (uint32_t)n;
        B.cnt[1] = 0;
        B.cnt[2] = 0;
    }
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += stride) {
        const int sl = j / P.npix, pl = j - sl * P.npix;
        init_camera_path(P, B, (uint32_t)j, pl, P.s0 + sl);
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
        B.cnt[3] = 0;        // terminated paths (persistent schedule)
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
        const HitRec h = traverse<false, STACK>(S.nodes, S.prims, S.ids, r, 1e30f, stk + threadIdx.x, kBlock, S.err);
        B.hit[j] = make_float2(h.t, __uint_as_float(h.id >= 0 ? h.rec : 0xffffffffu));
        B.nseg[j] += 1u;
    }
}

// ------------------------------------------------------------------------------ shade
__global__ __launch_bounds__(kBlock) void k_shade(SceneBufs S, PathBufs B, RenderParams P, int cur, int bounce,
                                                  uint32_t* __restrict__ guides) {
    const uint32_t n = B.cnt[cur];
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t* q = B.q[cur];
    uint32_t* qn = B.q[cur ^ 1];
    const uint32_t nloop = (n + stride - 1) / stride * stride;  // uniform trip count for wave ops
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < nloop; i += stride) {
        bool cont = false, shadow = false;
        uint32_t j = 0;
        float4 so = make_float4(0, 0, 0, 0), sd = so, sc = so;
        if (i < n) {
            j = q[i];
            cont = shade_path<false>(S, B, P, j, bounce, guides, shadow, so, sd, sc, Adj{}) && bounce + 1 < P.bounces;
        }
        uint32_t slot = 0;
        wave_append(&B.cnt[cur ^ 1], cont, slot);
        if (cont) qn[slot] = j;
        uint32_t sslot = 0;
        wave_append(&B.cnt[2], shadow, sslot);
        if (shadow) {
            B.sq[sslot] = j;
            B.so[j] = so;
            B.sd[j] = sd;
            B.sc[j] = sc;
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
        const HitRec h = traverse<true, STACK>(S.nodes, S.prims, S.ids, r, o.w, stk + threadIdx.x, kBlock, S.err);
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
__global__ __launch_bounds__(kBlock) void k_accumulate(RenderParams P, PathBufs B, float4* __restrict__ acc,
                                                       float4* __restrict__ keep) {
    const int stride = gridDim.x * kBlock;
    const int nloop = (P.npix + stride - 1) / stride * stride;
    for (int pl = blockIdx.x * kBlock + threadIdx.x; pl < nloop; pl += stride) {
        uint32_t segs = 0;
        if (pl < P.npix) {
            float4 a = acc[pl];
            for (int sl = 0; sl < P.ns; sl++) {
                const int j = sl * P.npix + pl;
                const float4 l = B.L[j];
                if (keep) keep[(size_t)(P.s0 + sl) * (size_t)P.npix + (size_t)pl] = l;
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


// ------------------------------------------------------------------------------ persistent schedule
// A fixed pool of path slots.  Each iteration: extend -> shade -> shadow -> regen.  Paths that
// end leave their radiance in a per-sample slot and their path slot is refilled in place with
// the next (pixel, sample) of the window, claimed from a device counter.  Samples are keyed by
// (pixel, sample), never by slot, so the image does not depend on the schedule; the per-sample
// results are reduced in sample order afterwards.  Parameters live in device memory so the
// iteration sequence can be captured once as a hipGraph and replayed.

__global__ void k_p_reset(PathBufs B) {
    if (threadIdx.x < 5) B.cnt[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kBlock) void k_p_fill(const RenderParams* __restrict__ Pd, PathBufs B,
                                                   float4* __restrict__ acc, uint32_t* __restrict__ guides,
                                                   uint32_t slots, int clear_pixels) {
    const RenderParams P = *Pd;
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t total = (uint32_t)P.ns * (uint32_t)P.npix;
    if (clear_pixels) {
        for (uint32_t pl = blockIdx.x * kBlock + threadIdx.x; pl < (uint32_t)P.npix; pl += stride) {
            acc[pl] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (P.want_guides) {
                uint32_t* gg = guides + 3 * (size_t)pl;
                gg[0] = 0u; gg[1] = 0u; gg[2] = 0u;
            }
        }
    }
    const uint32_t nloop = (slots + stride - 1) / stride * stride;
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < nloop; j += stride) {
        uint32_t idx = 0xffffffffu;
        wave_append(&B.cnt[4], j < slots, idx);
        const bool start = j < slots && idx < total;
        if (start) {
            const uint32_t sl = idx / (uint32_t)P.npix, pl = idx - sl * (uint32_t)P.npix;
            init_camera_path(P, B, j, (int)pl, P.s0 + (int)sl);
            B.bnc[j] = 0;
        }
        uint32_t slot = 0;
        wave_append(&B.cnt[0], start, slot);
        if (start) B.q[0][slot] = j;
    }
}

__global__ __launch_bounds__(kBlock) void k_p_shade(SceneBufs S, PathBufs B, const RenderParams* __restrict__ Pd,
                                                    int cur, uint32_t* __restrict__ guides) {
    const RenderParams P = *Pd;
    const uint32_t n = B.cnt[cur];
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t* q = B.q[cur];
    uint32_t* qn = B.q[cur ^ 1];
    const uint32_t nloop = (n + stride - 1) / stride * stride;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < nloop; i += stride) {
        bool cont = false, shadow = false, done = false;
        uint32_t j = 0;
        float4 so = make_float4(0, 0, 0, 0), sd = so, sc = so;
        if (i < n) {
            j = q[i];
            const int bounce = (int)B.bnc[j];
            cont = shade_path<false>(S, B, P, j, bounce, guides, shadow, so, sd, sc, Adj{}) && bounce + 1 < P.bounces;
            if (cont) B.bnc[j] = (uint32_t)(bounce + 1);
            done = !cont;
        }
        uint32_t slot = 0;
        wave_append(&B.cnt[cur ^ 1], cont, slot);
        if (cont) qn[slot] = j;
        uint32_t dslot = 0;
        wave_append(&B.cnt[3], done, dslot);
        if (done) B.dq[dslot] = j;
        uint32_t sslot = 0;
        wave_append(&B.cnt[2], shadow, sslot);
        if (shadow) {
            B.sq[sslot] = j;
            B.so[j] = so;
            B.sd[j] = sd;
            B.sc[j] = sc;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_p_regen(PathBufs B, const RenderParams* __restrict__ Pd, int cur) {
    const RenderParams P = *Pd;
    const uint32_t n = B.cnt[3];
    const uint32_t total = (uint32_t)P.ns * (uint32_t)P.npix;
    const uint32_t stride = gridDim.x * kBlock;
    uint32_t* qn = B.q[cur ^ 1];
    const uint32_t nloop = (n + stride - 1) / stride * stride;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < nloop; i += stride) {
        uint32_t j = 0;
        if (i < n) {
            j = B.dq[i];
            const uint4 meta = B.meta[j];
            const uint32_t pl = meta.x - (uint32_t)P.y0 * (uint32_t)P.width;
            const uint32_t g = (meta.y - (uint32_t)P.s0) * (uint32_t)P.npix + pl;
            const float4 l = B.L[j];
            B.res[g] = make_float4(l.x, l.y, l.z, (float)B.nseg[j]);
        }
        uint32_t idx = 0xffffffffu;
        wave_append(&B.cnt[4], i < n, idx);
        const bool start = i < n && idx < total;
        if (start) {
            const uint32_t sl = idx / (uint32_t)P.npix, pl = idx - sl * (uint32_t)P.npix;
            init_camera_path(P, B, j, (int)pl, P.s0 + (int)sl);
            B.bnc[j] = 0;
        }
        uint32_t slot = 0;
        wave_append(&B.cnt[cur ^ 1], start, slot);
        if (start) qn[slot] = j;
    }
}

__global__ __launch_bounds__(kBlock) void k_p_reduce(const RenderParams* __restrict__ Pd, PathBufs B,
                                                     float4* __restrict__ acc) {
    const RenderParams P = *Pd;
    const int stride = gridDim.x * kBlock;
    const int nloop = (P.npix + stride - 1) / stride * stride;
    for (int pl = blockIdx.x * kBlock + threadIdx.x; pl < nloop; pl += stride) {
        uint32_t segs = 0;
        if (pl < P.npix) {
            float4 a = acc[pl];
            for (int sl = 0; sl < P.ns; sl++) {
                const float4 r = B.res[(size_t)sl * (size_t)P.npix + (size_t)pl];
                float cx = r.x, cy = r.y, cz = r.z;
                if (P.firefly > 0.0f) {
                    const float lim = P.firefly;
                    const float m = rmax(cx, rmax(cy, cz));
                    if (!(m <= lim || m <= 0.0f)) {
                        const float k = lim / m;
                        cx = cx * k; cy = cy * k; cz = cz * k;
                    }
                }
                a.x = a.x + cx; a.y = a.y + cy; a.z = a.z + cz;
                a.w = a.w + r.w;
                segs += (uint32_t)r.w;
            }
            acc[pl] = a;
        }
        const uint32_t tot = wave_sum(segs);
        if (__lane_id() == 0 && tot) atomicAdd(&B.totals[0], (unsigned long long)tot);
    }
}


// ------------------------------------------------------------------------------ adjoint (path replay)
// Backward pass of the differentiable render: every sample is replayed from its key with the
// forward pass's total radiance, and d loss / d (Color, Emission) of each material is summed
// per workgroup in LDS, then once per workgroup into the global gradient.
constexpr int kAdjMaxMat = 1024;

__device__ __forceinline__ void adj_lds_init(float* lds, int n) {
    for (int i = threadIdx.x; i < n; i += kBlock) lds[i] = 0.0f;
    __syncthreads();
}
__device__ __forceinline__ void adj_lds_flush(const float* lds, int n, float* __restrict__ grad) {
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kBlock)
        if (lds[i] != 0.0f) atomicAdd(&grad[i], lds[i]);
}

__global__ __launch_bounds__(kBlock) void k_raygen_adj(RenderParams P, PathBufs B, const float4* __restrict__ keep,
                                                       float4* __restrict__ Lt) {
    const int n = P.ns * P.npix;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        B.cnt[0] = (uint32_t)n;
        B.cnt[1] = 0;
        B.cnt[2] = 0;
    }
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        const int sl = j / P.npix, pl = j - sl * P.npix;
        init_camera_path(P, B, (uint32_t)j, pl, P.s0 + sl);
        B.q[0][j] = (uint32_t)j;
        Lt[j] = keep[(size_t)(P.s0 + sl) * (size_t)P.npix + (size_t)pl];
    }
}

__global__ __launch_bounds__(kBlock) void k_shade_adj(SceneBufs S, PathBufs B, RenderParams P, int cur, int bounce,
                                                      uint32_t* __restrict__ guides, Adj A, float* __restrict__ grad) {
    __shared__ float lds[2 * kAdjMaxMat * 3];
    const int ng = P.nmat * 6;
    adj_lds_init(lds, ng);
    A.gcol = lds;
    A.gle = lds + P.nmat * 3;
    const uint32_t n = B.cnt[cur];
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t nloop = (n + stride - 1) / stride * stride;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < nloop; i += stride) {
        bool cont = false, shadow = false;
        uint32_t j = 0;
        float4 so = make_float4(0, 0, 0, 0), sd = so, sc = so;
        if (i < n) {
            j = B.q[cur][i];
            cont = shade_path<true>(S, B, P, j, bounce, guides, shadow, so, sd, sc, A) && bounce + 1 < P.bounces;
        }
        uint32_t slot = 0;
        wave_append(&B.cnt[cur ^ 1], cont, slot);
        if (cont) B.q[cur ^ 1][slot] = j;
        uint32_t sslot = 0;
        wave_append(&B.cnt[2], shadow, sslot);
        if (shadow) {
            B.sq[sslot] = j;
            B.so[j] = so;
            B.sd[j] = sd;
            B.sc[j] = sc;
        }
    }
    adj_lds_flush(lds, ng, grad);
}

template <int STACK>
__global__ __launch_bounds__(kBlock) void k_shadow_adj(SceneBufs S, PathBufs B, RenderParams P, Adj A,
                                                       float* __restrict__ grad) {
    __shared__ uint2 stk[STACK * kBlock];
    __shared__ float lds[kAdjMaxMat * 3];
    const int ng = P.nmat * 3;
    adj_lds_init(lds, ng);
    A.gle = lds;
    const uint32_t n = B.cnt[2];
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        const uint32_t j = B.sq[i];
        const float4 o = B.so[j], d = B.sd[j];
        Ray r;
        r.ox = o.x; r.oy = o.y; r.oz = o.z;
        r.dx = d.x; r.dy = d.y; r.dz = d.z;
        r.tmin = (float)1e-4;
        const HitRec h = traverse<true, STACK>(S.nodes, S.prims, S.ids, r, o.w, stk + threadIdx.x, kBlock, S.err);
        if (h.id < 0) {
            const float4 c = B.sc[j];
            const float4 l = B.L[j];
            B.L[j] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.0f);
            const uint32_t pl = B.meta[j].x - (uint32_t)P.y0 * (uint32_t)P.width;
            const float3 g = mul(f3(A.dl[3 * pl], A.dl[3 * pl + 1], A.dl[3 * pl + 2]), A.inv_spp);
            const float4 w = A.sg[j];
            adj_emission(A, (int)__float_as_uint(w.w), g, xyz(w));
        }
    }
    adj_lds_flush(lds, ng, grad + P.nmat * 3);
}

}  // namespace bcd

// ------------------------------------------------------------------------------ host side
struct RenderState {
    void* buf[24] = {};
    size_t bufcap[24] = {};
    float4* sg = nullptr;
    // differentiable render: per-sample radiance of the last bc_forward, its shape
    void* keep = nullptr;
    size_t keep_cap = 0;
    bc_render_desc keep_desc{};
    int keep_y0 = -1, keep_rows = 0;
    uint64_t keep_topo = 0;
    bcd::PathBufs B{};
    float4* acc = nullptr;
    uint32_t* guide_tmp = nullptr;
    bcd::RenderParams* Pd = nullptr;
    bcd::RenderParams* h_params = nullptr;  // pinned staging for Pd
    uint32_t* h_cnt = nullptr;              // pinned
    // captured iteration graph of the persistent schedule
    hipGraphExec_t exec = nullptr;
    const void* g_key[13] = {};
    int g_iters = 0, g_stack = 0, g_grid = 0;
    // autotuner cache: (shape) -> (pool, iterations per graph)
    struct Tune {
        int w, rows, spp, bounces;
        uint64_t topo, geom;
        int pool, iters;
    };
    std::vector<Tune> tunes;
    int last_pool = 0, last_iters = 0, last_windows = 0;
};

namespace {

constexpr size_t kResultCap = (size_t)1 << 25;  // per-sample results held per window (512 MB)

bool ensure_state(bc_ctx* c, size_t paths, size_t npix, bool guides, size_t window_samples) {
    if (!c->rs) c->rs = new RenderState();
    RenderState& R = *c->rs;
    if (!R.h_cnt) {
        void* h = nullptr;
        if (!bc_check(c, hipHostMalloc(&h, 4096, hipHostMallocDefault), "pinned staging")) return false;
        R.h_cnt = (uint32_t*)h;
        R.h_params = (bcd::RenderParams*)((char*)h + 256);
    }
    struct Want { int slot; size_t bytes; };
    const Want w[] = {
        {0, paths * 16},  {1, paths * 16},  {2, paths * 16}, {3, paths * 16},  {4, paths * 16},
        {5, paths * 16},  {6, paths * 4},   {7, paths * 8},  {8, paths * 16},  {9, paths * 16},
        {10, paths * 16}, {11, paths * 4},  {12, paths * 4}, {13, paths * 4},  {14, 64},
        {15, npix * 16 + (guides ? npix * 12 : 0) + 64},     {16, paths * 4},  {17, paths * 4},
        {18, window_samples * 16},                           {19, sizeof(bcd::RenderParams)},
        {20, paths * 16},
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
    R.B.bnc = (uint32_t*)R.buf[16];
    R.B.dq = (uint32_t*)R.buf[17];
    R.B.res = (float4*)R.buf[18];
    R.acc = (float4*)R.buf[15];
    R.guide_tmp = (uint32_t*)(R.acc + npix);
    R.Pd = (bcd::RenderParams*)R.buf[19];
    R.sg = (float4*)R.buf[20];
    return true;
}

bcd::RenderParams make_params(const bc_ctx* c, const bc_render_desc* d, int y0, int rows, bool guides) {
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
    return P;
}

template <int STACK>
void launch_bounce(const bcd::SceneBufs& S, const bcd::PathBufs& B, const bcd::RenderParams& P, int cur, int bounce,
                   uint32_t* guides, int grid, hipStream_t st) {
    hipLaunchKernelGGL((bcd::k_extend<STACK>), dim3(grid), dim3(bcd::kBlock), 0, st, S, B, cur);
    hipLaunchKernelGGL(bcd::k_shade, dim3(grid), dim3(bcd::kBlock), 0, st, S, B, P, cur, bounce, guides);
    hipLaunchKernelGGL((bcd::k_shadow<STACK>), dim3(grid), dim3(bcd::kBlock), 0, st, S, B);
}

template <int STACK>
void launch_iteration(const bcd::SceneBufs& S, const bcd::PathBufs& B, const bcd::RenderParams* Pd, int cur,
                      uint32_t* guides, int grid, hipStream_t st) {
    hipLaunchKernelGGL((bcd::k_extend<STACK>), dim3(grid), dim3(bcd::kBlock), 0, st, S, B, cur);
    hipLaunchKernelGGL(bcd::k_p_shade, dim3(grid), dim3(bcd::kBlock), 0, st, S, B, Pd, cur, guides);
    hipLaunchKernelGGL((bcd::k_shadow<STACK>), dim3(grid), dim3(bcd::kBlock), 0, st, S, B);
    hipLaunchKernelGGL(bcd::k_p_regen, dim3(grid), dim3(bcd::kBlock), 0, st, B, Pd, cur);
}

int finish(bc_ctx* c, const bcd::RenderParams& P, float* d_rgba, int grid_pix, uint64_t* rays_out, uint64_t* shadow_out) {
    RenderState& R = *c->rs;
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(bcd::k_finalize, dim3(grid_pix), dim3(bcd::kBlock), 0, st, P, (const float4*)R.acc,
                       (float4*)d_rgba);
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

// Plain wavefront schedule: passes of pixels x samples, three kernels per bounce.
int render_plain(bc_ctx* c, const bc_render_desc* d, int y0, int rows, float* d_rgba, uint32_t* d_guides,
                 uint64_t* rays_out, uint64_t* shadow_out, float4* keep = nullptr) {
    const size_t npix = (size_t)rows * (size_t)d->width;
    const size_t max_paths = std::max<size_t>(npix, (size_t)1 << 22);
    int spp_pass = d->samples_per_pass > 0 ? d->samples_per_pass : (int)std::max<size_t>(1, max_paths / npix);
    spp_pass = std::min(spp_pass, d->spp);
    const size_t paths = npix * (size_t)spp_pass;
    if (!ensure_state(c, paths, npix, d_guides != nullptr, 1)) return BC_ERR_NOMEM;
    RenderState& R = *c->rs;
    bcd::SceneBufs S{c->d_nodes, c->d_prims, c->d_ids, c->d_pmat, c->d_mats, c->d_lights, c->d_pdf, c->d_light_of, c->d_err};
    bcd::RenderParams P = make_params(c, d, y0, rows, d_guides != nullptr);
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
        hipLaunchKernelGGL(bcd::k_accumulate, dim3(grid_pix), dim3(bcd::kBlock), 0, st, P, R.B, R.acc, keep);
        if (!bc_check(c, hipGetLastError(), "render launch")) return BC_ERR_DEVICE;
    }
    return finish(c, P, d_rgba, grid_pix, rays_out, shadow_out);
}

// Persistent schedule with in-place regeneration; `iters` iterations are one captured graph.
int render_persistent(bc_ctx* c, const bc_render_desc* d, int y0, int rows, float* d_rgba, uint32_t* d_guides,
                      uint64_t* rays_out, uint64_t* shadow_out, int pool, int iters) {
    const size_t npix = (size_t)rows * (size_t)d->width;
    const int window_ns = (int)std::min<size_t>((size_t)d->spp, std::max<size_t>(1, kResultCap / npix));
    const size_t total_first = npix * (size_t)window_ns;
    const size_t slots = std::min<size_t>((size_t)pool, total_first);
    if (!ensure_state(c, slots, npix, d_guides != nullptr, total_first)) return BC_ERR_NOMEM;
    RenderState& R = *c->rs;
    bcd::SceneBufs S{c->d_nodes, c->d_prims, c->d_ids, c->d_pmat, c->d_mats, c->d_lights, c->d_pdf, c->d_light_of, c->d_err};
    bcd::RenderParams P = make_params(c, d, y0, rows, d_guides != nullptr);
    const int grid = (int)std::min<size_t>((slots + bcd::kBlock - 1) / bcd::kBlock, (size_t)c->num_cus * 8);
    const int grid_pix = (int)std::min<size_t>((npix + bcd::kBlock - 1) / bcd::kBlock, (size_t)c->num_cus * 8);
    hipStream_t st = c->stream;
    uint32_t* guides = d_guides ? d_guides : R.guide_tmp;
    iters = std::max(2, iters & ~1);
    const void* key[13] = {c->d_nodes, c->d_prims, c->d_ids, c->d_pmat, c->d_mats, c->d_lights, c->d_pdf, c->d_light_of, c->d_err,
                           R.B.o, R.B.res, R.Pd, guides};
    if (!R.exec || R.g_iters != iters || R.g_stack != c->stack || R.g_grid != grid ||
        std::memcmp(key, R.g_key, sizeof key) != 0) {
        if (R.exec) { (void)hipGraphExecDestroy(R.exec); R.exec = nullptr; }
        hipGraph_t graph = nullptr;
        if (!bc_check(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture")) return BC_ERR_DEVICE;
        for (int it = 0; it < iters; it++) {
            const int cur = it & 1;
            switch (c->stack) {
            case 8: launch_iteration<8>(S, R.B, R.Pd, cur, guides, grid, st); break;
            case 16: launch_iteration<16>(S, R.B, R.Pd, cur, guides, grid, st); break;
            case 24: launch_iteration<24>(S, R.B, R.Pd, cur, guides, grid, st); break;
            default: launch_iteration<32>(S, R.B, R.Pd, cur, guides, grid, st); break;
            }
        }
        if (!bc_check(c, hipStreamEndCapture(st, &graph), "end capture")) return BC_ERR_DEVICE;
        hipError_t e = hipGraphInstantiate(&R.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (!bc_check(c, e, "graph instantiate")) { R.exec = nullptr; return BC_ERR_DEVICE; }
        std::memcpy(R.g_key, key, sizeof key);
        R.g_iters = iters;
        R.g_stack = c->stack;
        R.g_grid = grid;
    }
    if (!bc_check(c, hipMemsetAsync(R.B.totals, 0, 16, st), "clear totals")) return BC_ERR_DEVICE;
    int windows = 0;
    for (int s0 = 0; s0 < d->spp; s0 += window_ns) {
        P.s0 = s0;
        P.ns = std::min(window_ns, d->spp - s0);
        *R.h_params = P;
        if (!bc_check(c, hipMemcpyAsync(R.Pd, R.h_params, sizeof P, hipMemcpyHostToDevice, st), "params")) return BC_ERR_DEVICE;
        hipLaunchKernelGGL(bcd::k_p_reset, dim3(1), dim3(64), 0, st, R.B);
        const int gfill = (int)std::min<size_t>((std::max(slots, npix) + bcd::kBlock - 1) / bcd::kBlock, (size_t)c->num_cus * 8);
        hipLaunchKernelGGL(bcd::k_p_fill, dim3(gfill), dim3(bcd::kBlock), 0, st, (const bcd::RenderParams*)R.Pd, R.B,
                           R.acc, guides, (uint32_t)slots, s0 == 0 ? 1 : 0);
        // every path ends within `bounces` iterations of its start, and each iteration starts
        // at least min(slots, remaining) new samples: bound the replays accordingly
        const size_t window_total = npix * (size_t)P.ns;
        const size_t max_graphs = ((size_t)(d->bounces + 1) * (window_total / slots + 2)) / (size_t)iters + 4;
        size_t g = 0;
        for (;; g++) {
            if (g >= max_graphs) { c->err = "persistent schedule did not drain"; return BC_ERR_DEVICE; }
            if (!bc_check(c, hipGraphLaunch(R.exec, st), "graph launch") ||
                !bc_check(c, hipMemcpyAsync(R.h_cnt, R.B.cnt, 4, hipMemcpyDeviceToHost, st), "queue length") ||
                !bc_check(c, hipStreamSynchronize(st), "graph"))
                return BC_ERR_DEVICE;
            if (R.h_cnt[0] == 0) break;
        }
        hipLaunchKernelGGL(bcd::k_p_reduce, dim3(grid_pix), dim3(bcd::kBlock), 0, st, (const bcd::RenderParams*)R.Pd,
                           R.B, R.acc);
        if (!bc_check(c, hipGetLastError(), "reduce launch")) return BC_ERR_DEVICE;
        windows++;
    }
    R.last_pool = (int)slots;
    R.last_iters = iters;
    R.last_windows = windows;
    return finish(c, P, d_rgba, grid_pix, rays_out, shadow_out);
}

// Autotuner: times a short render (2 spp of the same chunk) for each pool size and graph
// length and keeps the fastest for this image shape, scene and transport depth.
bool autotune(bc_ctx* c, const bc_render_desc* d, int y0, int rows, float* d_rgba, uint32_t* d_guides, int& pool,
              int& iters) {
    RenderState& R = *c->rs;
    for (const auto& t : R.tunes)
        if (t.w == d->width && t.rows == rows && t.spp == d->spp && t.bounces == d->bounces && t.topo == c->topo &&
            t.geom == c->geom) {
            pool = t.pool;
            iters = t.iters;
            return true;
        }
    bc_render_desc trial = *d;
    trial.spp = std::min(d->spp, 2);
    const size_t npix = (size_t)rows * (size_t)d->width;
    const int pools[] = {1 << 18, 1 << 19, 1 << 20, 1 << 21};
    const int its[] = {4, 8, 16};
    float best = 1e30f;
    pool = 1 << 20;
    iters = 8;
    for (int p : pools) {
        if ((size_t)p > npix * (size_t)trial.spp * 2 && p != pools[0]) continue;
        for (int it : its) {
            if (render_persistent(c, &trial, y0, rows, d_rgba, d_guides, nullptr, nullptr, p, it) != BC_OK) return false;
            float ms = 0;
            if (!bc_check(c, hipEventRecord(c->ev0, c->stream), "tune") ||
                render_persistent(c, &trial, y0, rows, d_rgba, d_guides, nullptr, nullptr, p, it) != BC_OK ||
                !bc_check(c, hipEventRecord(c->ev1, c->stream), "tune") ||
                !bc_check(c, hipEventSynchronize(c->ev1), "tune") ||
                !bc_check(c, hipEventElapsedTime(&ms, c->ev0, c->ev1), "tune"))
                return false;
            if (ms < best) { best = ms; pool = p; iters = it; }
        }
    }
    R.tunes.push_back({d->width, rows, d->spp, d->bounces, c->topo, c->geom, pool, iters});
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
    if (d->width <= 0 || d->height <= 0 || d->spp <= 0 || d->spp > 65536 || d->bounces < 0 || d->bounces > 1024 ||
        y0 < 0 || rows <= 0 || y0 + rows > d->height || !d_rgba || d->sampler < 0 || d->sampler > 2 ||
        (int64_t)d->width * d->height > ((int64_t)1 << 28) || d->mode < 0 || d->mode > 1) {
        c->err = "invalid render description or row range";
        return BC_ERR_ARG;
    }
    (void)hipSetDevice(c->device);
    if (d->mode == 0) return render_plain(c, d, y0, rows, d_rgba, d_guides, rays_out, shadow_out);
    if (!ensure_state(c, 1, (size_t)rows * (size_t)d->width, d_guides != nullptr, 1)) return BC_ERR_NOMEM;
    int pool = d->reserved[0], iters = d->reserved[1];
    if (pool <= 0 || iters <= 0) {
        if (!autotune(c, d, y0, rows, d_rgba, d_guides, pool, iters)) return BC_ERR_DEVICE;
    }
    return render_persistent(c, d, y0, rows, d_rgba, d_guides, rays_out, shadow_out, pool, iters);
}

extern "C" int bc_render_tuning(const bc_ctx* c, int32_t* out4) {
    if (!c || !out4) return BC_ERR_ARG;
    out4[0] = c->rs ? c->rs->last_pool : 0;
    out4[1] = c->rs ? c->rs->last_iters : 0;
    out4[2] = c->stack;
    out4[3] = c->rs ? c->rs->last_windows : 0;
    return BC_OK;
}

extern "C" int bc_render(bc_ctx* c, const bc_render_desc* d, int y0, int rows, float* rgba, uint32_t* guides,
                         uint64_t* rays_out, const volatile int32_t* cancel) {
    if (!c || !d || !rgba) return BC_ERR_ARG;
    if (cancel && *cancel) { c->err = "render cancelled"; return BC_ERR_CANCELLED; }
    if (d->width <= 0 || rows <= 0) { c->err = "invalid render description or row range"; return BC_ERR_ARG; }
    const size_t npix = (size_t)rows * (size_t)d->width;
    // output staging kept in the context, so repeated chunk renders reuse one buffer (and a
    // captured graph keyed on it)
    if (!bc_ensure(c, c->out_buf, c->out_cap, npix * 28 + 64, "render output")) return BC_ERR_NOMEM;
    uint32_t* dg = guides ? (uint32_t*)((char*)c->out_buf + npix * 16) : nullptr;
    int rc = bc_render_device(c, d, y0, rows, (float*)c->out_buf, dg, rays_out, nullptr);
    if (rc == BC_OK) {
        if (!bc_check(c, hipMemcpy(rgba, c->out_buf, npix * 16, hipMemcpyDeviceToHost), "copy image") ||
            (guides && !bc_check(c, hipMemcpy(guides, dg, npix * 12, hipMemcpyDeviceToHost), "copy guides")))
            rc = BC_ERR_DEVICE;
    }
    return rc;
}

namespace {
template <int STACK>
void launch_bounce_adj(const bcd::SceneBufs& S, const bcd::PathBufs& B, const bcd::RenderParams& P, int cur,
                       int bounce, uint32_t* guides, const bcd::Adj& A, float* grad, int grid, hipStream_t st) {
    hipLaunchKernelGGL((bcd::k_extend<STACK>), dim3(grid), dim3(bcd::kBlock), 0, st, S, B, cur);
    hipLaunchKernelGGL(bcd::k_shade_adj, dim3(grid), dim3(bcd::kBlock), 0, st, S, B, P, cur, bounce, guides, A, grad);
    hipLaunchKernelGGL((bcd::k_shadow_adj<STACK>), dim3(grid), dim3(bcd::kBlock), 0, st, S, B, P, A, grad);
}
bool same_shape(const bc_render_desc& a, const bc_render_desc& b) {
    return a.width == b.width && a.height == b.height && a.spp == b.spp && a.bounces == b.bounces &&
           a.rr_depth == b.rr_depth && a.sampler == b.sampler && a.mis == b.mis && a.power_lights == b.power_lights &&
           a.sky_mode == b.sky_mode && a.seed == b.seed && std::memcmp(a.origin, b.origin, sizeof a.origin) == 0 &&
           std::memcmp(a.lower, b.lower, sizeof a.lower) == 0 && std::memcmp(a.horiz, b.horiz, sizeof a.horiz) == 0 &&
           std::memcmp(a.vert, b.vert, sizeof a.vert) == 0;
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
    (void)hipSetDevice(c->device);
    const size_t npix = (size_t)rows * (size_t)d->width;
    if (!ensure_state(c, 1, npix, false, 1)) return BC_ERR_NOMEM;
    RenderState& R = *c->rs;
    if (!bc_ensure(c, R.keep, R.keep_cap, npix * (size_t)d->spp * 16, "forward sample buffer")) return BC_ERR_NOMEM;
    if (!bc_ensure(c, c->out_buf, c->out_cap, npix * 28 + 64, "render output")) return BC_ERR_NOMEM;
    int rc = render_plain(c, d, y0, rows, (float*)c->out_buf, nullptr, rays_out, nullptr, (float4*)R.keep);
    if (rc != BC_OK) return rc;
    if (!bc_check(c, hipMemcpy(rgba, c->out_buf, npix * 16, hipMemcpyDeviceToHost), "copy image")) return BC_ERR_DEVICE;
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
    if (!ensure_state(c, paths, npix, false, paths)) return BC_ERR_NOMEM;
    RenderState& R = *c->rs;
    const size_t ng = (size_t)c->nmat * 6;
    if (!bc_ensure(c, c->out_buf, c->out_cap, std::max(npix * 28 + 64, npix * 12 + ng * 4 + 64), "adjoint buffers"))
        return BC_ERR_NOMEM;
    float* d_dl = (float*)c->out_buf;
    float* d_grad = d_dl + npix * 3;
    hipStream_t st = c->stream;
    if (!bc_check(c, hipMemcpyAsync(d_dl, dl, npix * 12, hipMemcpyHostToDevice, st), "upload d loss") ||
        !bc_check(c, hipMemsetAsync(d_grad, 0, ng * 4, st), "clear gradient") ||
        !bc_check(c, hipMemsetAsync(R.B.totals, 0, 16, st), "clear totals"))
        return BC_ERR_DEVICE;
    bcd::SceneBufs S{c->d_nodes, c->d_prims, c->d_ids, c->d_pmat, c->d_mats, c->d_lights, c->d_pdf, c->d_light_of, c->d_err};
    bcd::RenderParams P = make_params(c, d, y0, rows, false);
    bcd::Adj A{};
    A.dl = d_dl;
    A.inv_spp = 1.0f / (float)d->spp;
    A.Lt = R.B.res;
    A.sg = R.sg;
    const int grid = (int)std::min<size_t>((paths + bcd::kBlock - 1) / bcd::kBlock, (size_t)c->num_cus * 8);
    for (int s0 = 0; s0 < d->spp; s0 += spp_pass) {
        P.s0 = s0;
        P.ns = std::min(spp_pass, d->spp - s0);
        hipLaunchKernelGGL(bcd::k_raygen_adj, dim3(grid), dim3(bcd::kBlock), 0, st, P, R.B, (const float4*)R.keep,
                           R.B.res);
        for (int b = 0; b < d->bounces; b++) {
            const int cur = b & 1;
            switch (c->stack) {
            case 8: launch_bounce_adj<8>(S, R.B, P, cur, b, R.guide_tmp, A, d_grad, grid, st); break;
            case 16: launch_bounce_adj<16>(S, R.B, P, cur, b, R.guide_tmp, A, d_grad, grid, st); break;
            case 24: launch_bounce_adj<24>(S, R.B, P, cur, b, R.guide_tmp, A, d_grad, grid, st); break;
            default: launch_bounce_adj<32>(S, R.B, P, cur, b, R.guide_tmp, A, d_grad, grid, st); break;
            }
        }
        if (!bc_check(c, hipGetLastError(), "adjoint launch")) return BC_ERR_DEVICE;
    }
    std::vector<float> g(ng);
    if (!bc_check(c, hipMemcpyAsync(g.data(), d_grad, ng * 4, hipMemcpyDeviceToHost, st), "read gradient") ||
        !bc_check(c, hipStreamSynchronize(st), "backward"))
        return BC_ERR_DEVICE;
    const size_t n3 = (size_t)c->nmat * 3;
    for (size_t i = 0; i < n3; i++) {
        dcolor[i] += g[i];
        demission[i] += g[n3 + i];
    }
    return bc_check_kernel_error(c);
}
