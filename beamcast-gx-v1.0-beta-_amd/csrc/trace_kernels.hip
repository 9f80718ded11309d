// trace_kernels.hip -- ray-query kernels over the BVH8 (closest hit and any hit).
// The ray batch of bc_trace: 32-byte rays in, 8-byte (t, id) or 4-byte occlusion out.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "trace_device.hpp"

namespace bcd {

template <bool ANY, int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_trace(const bc_node8* __restrict__ nodes,
                                                       const bc_prim* __restrict__ prims,
                                                       const bc_ray* __restrict__ rays, int64_t n,
                                                       void* __restrict__ out, int* __restrict__ err) {
    __shared__ uint2 stk[STACK * kTraceBlock];
    const int64_t i = (int64_t)blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= n) return;
    const float4* rp = reinterpret_cast<const float4*>(rays + i);
    const float4 a = rp[0], b = rp[1];
    Ray r;
    r.ox = a.x; r.oy = a.y; r.oz = a.z; r.tmin = a.w;
    r.dx = b.x; r.dy = b.y; r.dz = b.z;
    const HitRec h = traverse<ANY, STACK>(nodes, prims, r, b.w, stk + threadIdx.x, kTraceBlock, err);
    if (ANY) {
        reinterpret_cast<int32_t*>(out)[i] = h.id >= 0 ? 1 : 0;
    } else {
        bc_hit o;
        o.t = h.id >= 0 ? h.t : 0.0f;
        o.prim = h.id;
        reinterpret_cast<bc_hit*>(out)[i] = o;
    }
}

template <bool ANY, int STACK>
static hipError_t launch_trace_t(const bc_node8* nodes, const bc_prim* prims, const bc_ray* rays, int64_t n,
                                 void* out, int* err, hipStream_t st) {
    const int64_t blocks = (n + kTraceBlock - 1) / kTraceBlock;
    hipLaunchKernelGGL((k_trace<ANY, STACK>), dim3((unsigned)blocks), dim3(kTraceBlock), 0, st, nodes, prims, rays, n,
                       out, err);
    return hipGetLastError();
}

template <bool ANY>
static hipError_t launch_trace_any(int stack, const bc_node8* nodes, const bc_prim* prims, const bc_ray* rays,
                                   int64_t n, void* out, int* err, hipStream_t st) {
    switch (stack) {
    case 8: return launch_trace_t<ANY, 8>(nodes, prims, rays, n, out, err, st);
    case 16: return launch_trace_t<ANY, 16>(nodes, prims, rays, n, out, err, st);
    case 24: return launch_trace_t<ANY, 24>(nodes, prims, rays, n, out, err, st);
    case 32: return launch_trace_t<ANY, 32>(nodes, prims, rays, n, out, err, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_trace(int stack, bool any_hit, const bc_node8* nodes, const bc_prim* prims, const bc_ray* rays,
                        int64_t n, void* out, int* err, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    return any_hit ? launch_trace_any<true>(stack, nodes, prims, rays, n, out, err, st)
                   : launch_trace_any<false>(stack, nodes, prims, rays, n, out, err, st);
}

int stack_for_depth(int depth) {
    const int need = depth - 1;
    if (need <= 8) return 8;
    if (need <= 16) return 16;
    if (need <= 24) return 24;
    if (need <= 32) return 32;
    return -1;
}


// Persistent variant: a grid sized to the machine; each lane takes the next ray from a
// device counter as soon as its previous ray is finished (one atomic per wave per refill).
template <bool ANY, int STACK, int PPS>
__global__ __launch_bounds__(kTraceBlock) void k_trace_pers(const bc_node8* __restrict__ nodes,
                                                            const bc_prim* __restrict__ prims,
                                                            const bc_ray* __restrict__ rays, int64_t n,
                                                            void* __restrict__ out, int* __restrict__ err,
                                                            unsigned long long* __restrict__ counter) {
    __shared__ uint2 stk[STACK * kTraceBlock];
    uint2* my = stk + threadIdx.x;
    TravState s;
    bool active = false, retired = false;
    int64_t idx = 0;
    const uint32_t lane = __lane_id();
    for (;;) {
        const uint64_t need = __ballot(!active && !retired);
        if (need) {
            const uint32_t leader = (uint32_t)__builtin_ctzll(need);
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(need));
            base = __shfl(base, (int)leader);
            if (!active && !retired) {
                idx = (int64_t)(base + (unsigned long long)__popcll(need & ((1ull << lane) - 1ull)));
                if (idx < n) {
                    const float4* rp = reinterpret_cast<const float4*>(rays + idx);
                    const float4 a = rp[0], b = rp[1];
                    Ray r;
                    r.ox = a.x; r.oy = a.y; r.oz = a.z; r.tmin = a.w;
                    r.dx = b.x; r.dy = b.y; r.dz = b.z;
                    trav_init(s, r, b.w);
                    active = true;
                } else {
                    retired = true;
                }
            }
        }
        if (__ballot(active) == 0) break;
        if (active && trav_step<ANY, STACK, PPS>(s, nodes, prims, my, kTraceBlock, err)) {
            if (ANY) {
                reinterpret_cast<int32_t*>(out)[idx] = s.best.id >= 0 ? 1 : 0;
            } else {
                bc_hit o;
                o.t = s.best.id >= 0 ? s.best.t : 0.0f;
                o.prim = s.best.id;
                reinterpret_cast<bc_hit*>(out)[idx] = o;
            }
            active = false;
        }
    }
}

template <bool ANY, int STACK>
static hipError_t launch_pers(int variant, const bc_node8* nodes, const bc_prim* prims, const bc_ray* rays, int64_t n,
                              void* out, int* err, unsigned long long* counter, int blocks, hipStream_t st) {
    if (variant == 1)
        hipLaunchKernelGGL((k_trace_pers<ANY, STACK, 0>), dim3(blocks), dim3(kTraceBlock), 0, st, nodes, prims, rays, n,
                           out, err, counter);
    else
        hipLaunchKernelGGL((k_trace_pers<ANY, STACK, 1>), dim3(blocks), dim3(kTraceBlock), 0, st, nodes, prims, rays, n,
                           out, err, counter);
    return hipGetLastError();
}

template <bool ANY>
static hipError_t launch_pers_any(int variant, int stack, const bc_node8* nodes, const bc_prim* prims,
                                  const bc_ray* rays, int64_t n, void* out, int* err, unsigned long long* counter,
                                  int blocks, hipStream_t st) {
    switch (stack) {
    case 8: return launch_pers<ANY, 8>(variant, nodes, prims, rays, n, out, err, counter, blocks, st);
    case 16: return launch_pers<ANY, 16>(variant, nodes, prims, rays, n, out, err, counter, blocks, st);
    case 24: return launch_pers<ANY, 24>(variant, nodes, prims, rays, n, out, err, counter, blocks, st);
    case 32: return launch_pers<ANY, 32>(variant, nodes, prims, rays, n, out, err, counter, blocks, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_trace_persistent(int variant, int stack, bool any_hit, const bc_node8* nodes, const bc_prim* prims,
                                   const bc_ray* rays, int64_t n, void* out, int* err, unsigned long long* counter,
                                   int blocks, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(counter, 0, sizeof(unsigned long long), st);
    if (e != hipSuccess) return e;
    return any_hit ? launch_pers_any<true>(variant, stack, nodes, prims, rays, n, out, err, counter, blocks, st)
                   : launch_pers_any<false>(variant, stack, nodes, prims, rays, n, out, err, counter, blocks, st);
}

}  // namespace bcd
