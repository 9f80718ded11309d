// trace_kernels.hip -- ray-query kernels over the BVH8 (closest hit and any hit).
// The ray batch of bc_trace: 32-byte rays in, 8-byte (t, id) or 4-byte occlusion out.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "trace_device.hpp"

namespace bcd {

// Resident waves per SIMD the register allocation is held to (64 VGPRs for 8 waves, 72 for 7).
// The any-hit refill loop fits 64 without spilling; the other loops fit 72 and would spill at 64.
#define BC_TRACE_WAVES(N) __attribute__((amdgpu_waves_per_eu(N, N)))
// Consecutive workgroups of k_trace that one XCD takes per turn (see xcd_chunk): 4096 rays.
constexpr uint32_t kTraceXcdRun = 64;
#ifndef BC_CLOSEST_WAVES
#define BC_CLOSEST_WAVES 7
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
                                                                           const int* __restrict__ oflag) {
    extern __shared__ uint2 stk[];
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
                                                                                  const int* __restrict__ oflag) {
    extern __shared__ uint2 stk[];
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
static hipError_t launch_trace_t(int cap, const bc_node8* nodes, const bc_prim* prims, const uint32_t* ids, const bc_ray* rays,
                                 int64_t n, void* out, int refill, int* err, hipStream_t st, const bc_ray* sorted,
                                 const uint32_t* perm, const int* oflag) {
    if (refill > 0) {
        const size_t lds = (size_t)cap * kRefillBlock * sizeof(uint2);
        const int64_t per_block = (int64_t)kRefillRays * (kRefillBlock / 64);
        const int64_t blocks = (n + per_block - 1) / per_block;
        hipLaunchKernelGGL((k_trace_refill<ANY>), dim3((unsigned)blocks), dim3(kRefillBlock), lds, st, nodes, prims, ids, rays,
                           n, out, refill, cap, err, sorted, perm, oflag);
        return hipGetLastError();
    }
    const size_t lds = (size_t)cap * kTraceBlock * sizeof(uint2);
    const int64_t blocks = (n + kTraceBlock - 1) / kTraceBlock;
    hipLaunchKernelGGL((k_trace<ANY>), dim3((unsigned)blocks), dim3(kTraceBlock), lds, st, nodes, prims, ids, rays, n, out,
                       cap, err, sorted, perm, oflag);
    return hipGetLastError();
}

hipError_t launch_trace(int stack, bool any_hit, const bc_node8* nodes, const bc_prim* prims, const uint32_t* ids, const bc_ray* rays,
                        int64_t n, void* out, int refill, int* err, hipStream_t st, const bc_ray* sorted,
                        const uint32_t* perm, const int* oflag) {
    if (n <= 0) return hipSuccess;
    if (oflag && (!sorted || !perm)) return hipErrorInvalidValue;
    if (stack < 1 || stack > kMaxStack) return hipErrorInvalidValue;
    if (refill > 64) refill = 64;
    return any_hit ? launch_trace_t<true>(stack, nodes, prims, ids, rays, n, out, refill, err, st, sorted, perm, oflag)
                   : launch_trace_t<false>(stack, nodes, prims, ids, rays, n, out, refill, err, st, sorted, perm, oflag);
}

int stack_for_depth(int depth) {
    const int need = depth - 1;
    if (need > kMaxStack) return -1;
    return need < 1 ? 1 : need;
}


}  // namespace bcd
