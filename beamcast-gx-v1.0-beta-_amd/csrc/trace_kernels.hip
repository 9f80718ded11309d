// trace_kernels.hip -- ray-query kernels over the BVH8 (closest hit and any hit).
// The ray batch of bc_trace: 32-byte rays in, 8-byte (t, id) or 4-byte occlusion out.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "trace_device.hpp"

namespace bcd {

template <bool ANY, int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_trace(const bc_node8* __restrict__ nodes,
                                                       const bc_prim* __restrict__ prims,
                                                       const uint32_t* __restrict__ ids,
                                                       const bc_ray* __restrict__ rays, int64_t n,
                                                       void* __restrict__ out, int* __restrict__ err) {
    __shared__ uint2 stk[STACK * kTraceBlock];
    const int64_t i = (int64_t)xcd_chunk(blockIdx.x, gridDim.x, 64u) * kTraceBlock + threadIdx.x;
    if (i >= n) return;
    const float4* rp = reinterpret_cast<const float4*>(rays + i);
    const float4 a = rp[0], b = rp[1];
    Ray r;
    r.ox = a.x; r.oy = a.y; r.oz = a.z; r.tmin = a.w;
    r.dx = b.x; r.dy = b.y; r.dz = b.z;
    const HitRec h = traverse<ANY, STACK>(nodes, prims, ids, 0u, r, b.w, stk + threadIdx.x, kTraceBlock, err);
    if (ANY) {
        reinterpret_cast<int32_t*>(out)[i] = h.rec != 0xffffffffu ? 1 : 0;
    } else {
        bc_hit o;
        o.t = h.id >= 0 ? h.t : 0.0f;
        o.prim = h.id;
        reinterpret_cast<bc_hit*>(out)[i] = o;
    }
}

template <bool ANY, int STACK>
static hipError_t launch_trace_t(const bc_node8* nodes, const bc_prim* prims, const uint32_t* ids, const bc_ray* rays, int64_t n,
                                 void* out, int* err, hipStream_t st) {
    const int64_t blocks = (n + kTraceBlock - 1) / kTraceBlock;
    hipLaunchKernelGGL((k_trace<ANY, STACK>), dim3((unsigned)blocks), dim3(kTraceBlock), 0, st, nodes, prims, ids, rays,
                       n, out, err);
    return hipGetLastError();
}

template <bool ANY>
static hipError_t launch_trace_any(int stack, const bc_node8* nodes, const bc_prim* prims, const uint32_t* ids, const bc_ray* rays,
                                   int64_t n, void* out, int* err, hipStream_t st) {
    switch (stack) {
    case 8: return launch_trace_t<ANY, 8>(nodes, prims, ids, rays, n, out, err, st);
    case 16: return launch_trace_t<ANY, 16>(nodes, prims, ids, rays, n, out, err, st);
    case 24: return launch_trace_t<ANY, 24>(nodes, prims, ids, rays, n, out, err, st);
    case 32: return launch_trace_t<ANY, 32>(nodes, prims, ids, rays, n, out, err, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_trace(int stack, bool any_hit, const bc_node8* nodes, const bc_prim* prims, const uint32_t* ids, const bc_ray* rays,
                        int64_t n, void* out, int* err, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    return any_hit ? launch_trace_any<true>(stack, nodes, prims, ids, rays, n, out, err, st)
                   : launch_trace_any<false>(stack, nodes, prims, ids, rays, n, out, err, st);
}

int stack_for_depth(int depth) {
    const int need = depth - 1;
    if (need <= 8) return 8;
    if (need <= 16) return 16;
    if (need <= 24) return 24;
    if (need <= 32) return 32;
    return -1;
}


}  // namespace bcd
