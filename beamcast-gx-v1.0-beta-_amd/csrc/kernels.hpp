// kernels.hpp -- host-callable launchers of the HIP kernels (all asynchronous on `st`).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bc_layout.h"

namespace bcd {

constexpr int kTraceBlock = 64;  // threads per workgroup of the one-ray-per-thread query kernel
// Largest traversal stack (entries per lane): 32 x 256 lanes x 8 bytes = 64 KiB of LDS per block.
constexpr int kMaxStack = 32;

// LDS stack entries per lane that a BVH8 of the given depth needs (depth - 1, at least 1), or
// -1 beyond kMaxStack.
int stack_for_depth(int depth);

// Ray queries: any_hit writes int32 occlusion flags, else bc_hit records.  stack: LDS stack
// entries per lane (stack_for_depth).  *err is set non-zero by the kernel if a traversal stack
// would overflow.
hipError_t launch_trace(int stack, bool any_hit, const bc_node8* nodes, const bc_prim* prims, const uint32_t* ids,
                        const bc_ray* rays, int64_t n, void* out, int refill, int* err, hipStream_t st);

// Refit: records (and materials) of the moved primitives in leaf order from the view's arrays.
hipError_t launch_refit_records(const double* geom, const int32_t* kind, const int32_t* mat, const uint32_t* ids, uint32_t n,
                                bc_prim* prims, int32_t* pmat, int* err, hipStream_t st);
// One level of the device BVH8 refit: recomputes the boxes of the `n` nodes in `list`
// (whose children are already refit) and stores each node's union box in nbox[2 i .. 2 i + 1].
hipError_t launch_refit_level(bc_node8* nodes, const bc_prim* prims, const uint32_t* ids, float4* nbox,
                              const uint32_t* list, uint32_t n, int* err, hipStream_t st);

}  // namespace bcd
