// trace_kernels.hip -- ray-query kernels over the BVH8 (closest hit and any hit).
// The ray batch of bc_trace: 32-byte rays in, 8-byte (t, id) or 4-byte occlusion out.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "trace_device.hpp"

namespace bcd {

// Resident waves per SIMD the register allocation is held to (64 VGPRs for 8 waves, 72 for 7).
// Every loop fits 64 VGPRs; closest hit keeps the ray's index in scratch across its traversal
// loop (one store before, one load after), which costs less than the eighth wave gains.
#define BC_TRACE_WAVES(N) __attribute__((amdgpu_waves_per_eu(N, N)))
// Consecutive workgroups of k_trace that one XCD takes per turn (see xcd_chunk): 4096 rays.
constexpr uint32_t kTraceXcdRun = 64;
#ifndef BC_CLOSEST_WAVES
#define BC_CLOSEST_WAVES 8
#endif

// The LDS traversal stack is dynamic: cap entries per lane, cap * (threads per block) uint2 per
// block.  One ray per thread, one wave per workgroup (kTraceBlock): a finished wave frees its
// slot at once instead of waiting for the other waves of a larger workgroup.
template <bool ANY>
__global__ __launch_bounds__(kTraceBlock) BC_TRACE_WAVES(BC_CLOSEST_WAVES) void k_trace(const bc_node8* __restrict__ nodes,
                                                                           const bc_prim* __restrict__ prims,
                                                                           const uint32_t* __restrict__ ids,
                                                                           const bc_ray* __restrict__ rays, int64_t n,
                                                                           void* __restrict__ out, int cap,
                                                                           int* __restrict__ err,
                                                                           const bc_ray* __restrict__ sorted,
                                                                           const uint32_t* __restrict__ perm,
                                                                           const int* __restrict__ oflag,
                                                                           const int* __restrict__ sel) {
    extern __shared__ uint2 stk[];
    if (sel != nullptr && sel[0] != 0) return;  // the probe found the batch incoherent: k_trace_refill runs it
    const int64_t i = (int64_t)xcd_chunk(blockIdx.x, gridDim.x, kTraceXcdRun) * kTraceBlock + threadIdx.x;
    if (i >= n) return;
    const bool ord = oflag != nullptr && oflag[0] != 0;  // uniform: a scalar load
    const float4* rp = reinterpret_cast<const float4*>((ord ? sorted : rays) + i);
    const float4 a = rp[0], b = rp[1];
    Ray r;
    r.ox = a.x; r.oy = a.y; r.oz = a.z; r.tmin = a.w;
    r.dx = b.x; r.dy = b.y; r.dz = b.z;
    const HitRec h = traverse<ANY>(nodes, prims, ids, 0u, r, b.w, stk + threadIdx.x, kTraceBlock, cap, err);
    const int64_t o = ord ? (int64_t)perm[i] : i;
    if (ANY) {
        reinterpret_cast<int32_t*>(out)[o] = h.rec != 0xffffffffu ? 1 : 0;
    } else {
        bc_hit hr;
        hr.t = h.id >= 0 ? h.t : 0.0f;
        hr.prim = h.id;
        reinterpret_cast<bc_hit*>(out)[o] = hr;
    }
}

__device__ __forceinline__ void load_ray(const bc_ray* __restrict__ rays, int64_t i, Ray& r, float& tmax) {
    const float4* rp = reinterpret_cast<const float4*>(rays + i);
    const float4 a = rp[0], b = rp[1];
    r.ox = a.x; r.oy = a.y; r.oz = a.z; r.tmin = a.w;
    r.dx = b.x; r.dy = b.y; r.dz = b.z;
    tmax = b.w;
}

__device__ __forceinline__ void store_hit(void* __restrict__ out, int64_t i, const HitRec& h, bool any) {
    if (any) {
        reinterpret_cast<int32_t*>(out)[i] = h.rec != 0xffffffffu ? 1 : 0;
    } else {
        bc_hit o;
        o.t = h.id >= 0 ? h.t : 0.0f;
        o.prim = h.id;
        reinterpret_cast<bc_hit*>(out)[i] = o;
    }
}

// Lane refill: each wave owns a run of kRefillRays consecutive rays.  Its lanes start on the
// first 64; whenever at least `refill` lanes have finished, the wave hands them the next rays
// of its run (ballot + prefix count, no atomics), so a wave keeps most lanes busy instead of
// waiting for its slowest ray.  Per-ray results are those of k_trace.
constexpr int kRefillRays = 512;
constexpr int kRefillBlock = 256;  // four waves, each with its own run

template <bool ANY>
__global__ __launch_bounds__(kRefillBlock) BC_TRACE_WAVES(ANY ? 8 : BC_CLOSEST_WAVES) void k_trace_refill(const bc_node8* __restrict__ nodes,
                                                                                  const bc_prim* __restrict__ prims,
                                                                                  const uint32_t* __restrict__ ids,
                                                                                  const bc_ray* __restrict__ rays, int64_t n,
                                                                                  void* __restrict__ out, int refill, int cap,
                                                                                  int* __restrict__ err,
                                                                                  const bc_ray* __restrict__ sorted,
                                                                                  const uint32_t* __restrict__ perm,
                                                                                  const int* __restrict__ oflag,
                                                                                  const int* __restrict__ sel) {
    extern __shared__ uint2 stk[];
    if (sel != nullptr && sel[0] == 0) return;  // the probe found the batch coherent: k_trace runs it
    const bool ord = oflag != nullptr && oflag[0] != 0;  // uniform: a scalar load
    const int lane = (int)(threadIdx.x & 63u);
    // the run's bounds are wave-uniform (scalar registers); lanes hold 32-bit offsets into it
    const uint32_t wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t first = ((int64_t)xcd_chunk(blockIdx.x, gridDim.x, 16u) * (kRefillBlock / 64) + wib) * kRefillRays;
    if (first >= n) return;
    const int cnt = (int)min(n - first, (int64_t)kRefillRays);
    const bc_ray* run = (ord ? sorted : rays) + first;
    int next = 64;  // wave-uniform: the next unassigned ray of the run
    int off = lane;
    bool active = off < cnt;
    TravState st;
    if (active) {
        Ray r;
        float tmax;
        load_ray(run, off, r, tmax);
        trav_init(st, r, tmax, 0u);
    }
    uint2* my = stk + threadIdx.x;
    for (;;) {
        if (active && trav_step<ANY>(st, nodes, prims, ids, my, kRefillBlock, cap, err)) {
            store_hit(out, ord ? (int64_t)perm[first + off] : first + off, trav_result(st, ids, ANY), ANY);
            active = false;
        }
        const uint64_t idle = __ballot(!active);
        const int nidle = __popcll(idle);
        if (nidle == 64 && next >= cnt) break;
        if (nidle >= refill && next < cnt) {
            if (!active) {
                const int k = next + __popcll(idle & ((1ull << lane) - 1ull));
                if (k < cnt) {
                    off = k;
                    Ray r;
                    float tmax;
                    load_ray(run, off, r, tmax);
                    trav_init(st, r, tmax, 0u);
                    active = true;
                }
            }
            next += nidle;
        }
    }
}

template <bool ANY>
static hipError_t launch_refill_t(int cap, const bc_node8* nodes, const bc_prim* prims, const uint32_t* ids, const bc_ray* rays,
                                  int64_t n, void* out, int refill, int* err, hipStream_t st, const bc_ray* sorted,
                                  const uint32_t* perm, const int* oflag, const int* sel) {
    const size_t lds = (size_t)cap * kRefillBlock * sizeof(uint2);
    const int64_t per_block = (int64_t)kRefillRays * (kRefillBlock / 64);
    const int64_t blocks = (n + per_block - 1) / per_block;
    hipLaunchKernelGGL((k_trace_refill<ANY>), dim3((unsigned)blocks), dim3(kRefillBlock), lds, st, nodes, prims, ids, rays,
                       n, out, refill, cap, err, sorted, perm, oflag, sel);
    return hipGetLastError();
}

template <bool ANY>
static hipError_t launch_trace_t(int cap, const bc_node8* nodes, const bc_prim* prims, const uint32_t* ids, const bc_ray* rays,
                                 int64_t n, void* out, int refill, int* err, hipStream_t st, const bc_ray* sorted,
                                 const uint32_t* perm, const int* oflag, const int* sel) {
    if (refill > 0 && sel == nullptr)
        return launch_refill_t<ANY>(cap, nodes, prims, ids, rays, n, out, refill, err, st, sorted, perm, oflag, nullptr);
    const size_t lds = (size_t)cap * kTraceBlock * sizeof(uint2);
    const int64_t blocks = (n + kTraceBlock - 1) / kTraceBlock;
    hipLaunchKernelGGL((k_trace<ANY>), dim3((unsigned)blocks), dim3(kTraceBlock), lds, st, nodes, prims, ids, rays, n, out,
                       cap, err, sorted, perm, oflag, sel);
    if (sel == nullptr) return hipGetLastError();
    // the probe's choice is read by both kernels; the one not chosen returns at once
    return launch_refill_t<ANY>(cap, nodes, prims, ids, rays, n, out, refill, err, st, sorted, perm, oflag, sel);
}

hipError_t launch_trace(int stack, bool any_hit, const bc_node8* nodes, const bc_prim* prims, const uint32_t* ids, const bc_ray* rays,
                        int64_t n, void* out, int refill, int* err, hipStream_t st, const bc_ray* sorted,
                        const uint32_t* perm, const int* oflag, const int* sel) {
    if (n <= 0) return hipSuccess;
    if (oflag && (!sorted || !perm)) return hipErrorInvalidValue;
    if (stack < 1 || stack > kMaxStack) return hipErrorInvalidValue;
    if (sel && refill <= 0) return hipErrorInvalidValue;
    if (refill > 64) refill = 64;
    return any_hit ? launch_trace_t<true>(stack, nodes, prims, ids, rays, n, out, refill, err, st, sorted, perm, oflag, sel)
                   : launch_trace_t<false>(stack, nodes, prims, ids, rays, n, out, refill, err, st, sorted, perm, oflag, sel);
}

// Counting instantiation of k_trace: the same traversal of each ray (same results), plus the
// work it did: counts[i] = (node visits, primitive-record tests, id-table reads, traversal steps).
// The algorithmic bytes of a launch follow from the first three (64 B per node, 36 B per record,
// 4 B per id); steps minus node visits are the steps in which the ray only popped its stack.
template <bool ANY>
__global__ __launch_bounds__(kTraceBlock) void k_trace_count(const bc_node8* __restrict__ nodes,
                                                             const bc_prim* __restrict__ prims,
                                                             const uint32_t* __restrict__ ids,
                                                             const bc_ray* __restrict__ rays, int64_t n,
                                                             void* __restrict__ out, int cap, int* __restrict__ err,
                                                             uint4* __restrict__ counts) {
    extern __shared__ uint2 stk[];
    const int64_t i = (int64_t)blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= n) return;
    Ray r;
    float tmax;
    load_ray(rays, i, r, tmax);
    TravState st;
    trav_init(st, r, tmax, 0u);
    uint32_t steps = 1;
    while (!trav_step<ANY, true>(st, nodes, prims, ids, stk + threadIdx.x, kTraceBlock, cap, err)) steps++;
    const HitRec h = trav_result(st, ids, ANY);
    const uint32_t final_id = (!ANY && h.rec != 0xffffffffu) ? 1u : 0u;
    store_hit(out, i, h, ANY);
    counts[i] = make_uint4(st.n_node, st.n_rec, st.n_id + final_id, steps);
}

hipError_t launch_trace_count(int stack, bool any_hit, const bc_node8* nodes, const bc_prim* prims, const uint32_t* ids,
                              const bc_ray* rays, int64_t n, void* out, int* err, uint4* counts, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (stack < 1 || stack > kMaxStack) return hipErrorInvalidValue;
    const size_t lds = (size_t)stack * kTraceBlock * sizeof(uint2);
    const int64_t blocks = (n + kTraceBlock - 1) / kTraceBlock;
    if (any_hit)
        hipLaunchKernelGGL((k_trace_count<true>), dim3((unsigned)blocks), dim3(kTraceBlock), lds, st, nodes, prims, ids,
                           rays, n, out, stack, err, counts);
    else
        hipLaunchKernelGGL((k_trace_count<false>), dim3((unsigned)blocks), dim3(kTraceBlock), lds, st, nodes, prims, ids,
                           rays, n, out, stack, err, counts);
    return hipGetLastError();
}

// VALU issue probes, no memory access in their loops, so their instruction rate is what the SIMDs
// issue of their instruction mix.  k_valu_probe: the eight-child slab test of a node visit
// (byte-to-float conversions, FMAs, three-operand min/max, compares and mask building) on
// operands held in registers; each thread runs `iters` visits on its own operands and writes one
// word only in a case that does not occur, so nothing is dead.  k_fma_probe: eight independent
// fp32 FMA chains per thread, `iters` rounds of eight FMAs.
__global__ __launch_bounds__(256) void k_fma_probe(uint32_t iters, uint32_t seed, uint32_t* __restrict__ sink) {
    const float b = __uint_as_float(0x3f7ffff0u), c = __uint_as_float(0x3a800000u | (seed & 0xffu));
    float a[8];
#pragma unroll
    for (int k = 0; k < 8; k++) a[k] = (float)(threadIdx.x + (uint32_t)k);
    for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
        for (int k = 0; k < 8; k++) a[k] = __builtin_fmaf(a[k], b, c);
    }
    float t = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) t += a[k];
    if (t == -1.0f) sink[0] = 1u;  // never: the chains stay live
}

__global__ __launch_bounds__(256) void k_valu_probe(uint32_t iters, uint32_t seed, uint32_t* __restrict__ sink) {
    uint32_t w0 = seed ^ (blockIdx.x * 256u + threadIdx.x) * 0x9E3779B9u;
    uint32_t w1 = w0 * 0x85EBCA6Bu + 1u, w2 = w1 * 0xC2B2AE35u + 3u, w3 = w2 * 0x27D4EB2Fu + 5u;
    float ax = __uint_as_float(0x3f800000u | (w0 >> 9)), ay = __uint_as_float(0x3f800000u | (w1 >> 9));
    float az = __uint_as_float(0x3f800000u | (w2 >> 9));
    float ox = -ax, oy = -ay, oz = -az;
    uint32_t acc = 0;
    for (uint32_t it = 0; it < iters; it++) {
        float tn_min = 3.0e38f;
        uint32_t hits = 0;
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const uint32_t sh = (uint32_t)(s & 3) * 8u;
            const uint32_t a = s < 4 ? w0 : w1, b = s < 4 ? w2 : w3;
            const float tnx = __builtin_fmaf(u8f(a, sh), ax, ox), tny = __builtin_fmaf(u8f(b, sh), ay, oy);
            const float tnz = __builtin_fmaf(u8f(a ^ b, sh), az, oz);
            const float tfx = __builtin_fmaf(u8f(b, sh ^ 8u), ax, ox), tfy = __builtin_fmaf(u8f(a, sh ^ 8u), ay, oy);
            const float tfz = __builtin_fmaf(u8f(a + b, sh), az, oz);
            const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
            const float tf = fminf(fminf(tfx, tfy), fminf(tfz, 1.0e30f));
            const bool h = tn <= tf;
            hits |= (h ? 1u : 0u) << s;
            tn_min = fminf(tn_min, h ? tn : 3.0e38f);
        }
        acc += hits + (__float_as_uint(tn_min) >> 16);
        // new operands for the next visit: cheap integer steps, as a traversal's new node words
        w0 = w0 * 1664525u + 1013904223u;
        w1 ^= w0 >> 7;
        w2 += w1 << 3;
        w3 ^= w2 >> 5;
        ox += 0.5f;
    }
    if (acc == 0x12345678u) sink[0] = acc;  // practically never: keeps the loop alive
}

hipError_t launch_valu_probe(int mix, int blocks, uint32_t iters, uint32_t seed, uint32_t* sink, hipStream_t st) {
    if (mix == 1)
        hipLaunchKernelGGL(k_fma_probe, dim3((unsigned)blocks), dim3(256), 0, st, iters, seed, sink);
    else
        hipLaunchKernelGGL(k_valu_probe, dim3((unsigned)blocks), dim3(256), 0, st, iters, seed, sink);
    return hipGetLastError();
}

// Scattered-read probe: each thread reads `iters` 64-byte blocks (four 16-byte loads, the fetch of
// one node) at hashed positions of a buffer of block_mask + 1 blocks, and writes one word only in
// a case that does not occur.  Over a buffer far larger than the L2 (4 MB per XCD) nearly every
// read fills a 128-byte line from the Infinity Cache, the access an incoherent traversal makes, so
// its rate is the ceiling of scattered line fills that traversal is compared with.
__global__ __launch_bounds__(256) void k_line_probe(const uint4* __restrict__ buf, uint32_t block_mask, uint32_t iters,
                                                    uint32_t seed, uint32_t* __restrict__ sink) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    uint32_t acc = 0;
    for (uint32_t it = 0; it < iters; it++) {
        const uint32_t b = bc_mix32((t * 0x9E3779B9u) ^ (it * 0x85EBCA6Bu) ^ seed) & block_mask;
        const uint4* p = buf + (size_t)b * 4u;
        const uint4 a0 = p[0], a1 = p[1], a2 = p[2], a3 = p[3];
        acc += (a0.x ^ a0.y ^ a0.z ^ a0.w) + (a1.x ^ a1.y ^ a1.z ^ a1.w) + (a2.x ^ a2.y ^ a2.z ^ a2.w) +
               (a3.x ^ a3.y ^ a3.z ^ a3.w);
    }
    if (acc == 0x12345678u) sink[0] = acc;  // practically never: keeps every load alive
}

hipError_t launch_line_probe(const uint4* buf, uint32_t block_mask, int blocks, uint32_t iters, uint32_t seed,
                             uint32_t* sink, hipStream_t st) {
    hipLaunchKernelGGL(k_line_probe, dim3((unsigned)blocks), dim3(256), 0, st, buf, block_mask, iters, seed, sink);
    return hipGetLastError();
}

int stack_for_depth(int depth) {
    const int need = depth - 1;
    if (need > kMaxStack) return -1;
    return need < 1 ? 1 : need;
}


}  // namespace bcd
