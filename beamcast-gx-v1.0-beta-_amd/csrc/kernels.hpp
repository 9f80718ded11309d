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

// Spatial binning of a ray batch (order.hip): bins = (Morton cell of the origin on a
// 2^bits-per-axis grid over the root box) x (direction octant).  cnt holds G x bins counts,
// total and bstart one entry per bin, sorted and perm one entry per ray; flag[0] is the decision
// (sort or not), flag[1] the probe's count of sampled adjacent pairs in different bins.
struct OrderScratch {
    uint32_t* cnt;
    uint32_t* total;
    uint32_t* bstart;
    bc_ray* sorted;
    uint32_t* perm;
    int* flag;
};
constexpr int kOrderMinBits = 2;
constexpr int kOrderMaxBits = 4;
constexpr int kOrderMaxGroups = 1024;
constexpr int64_t kOrderMaxRays = (int64_t)1 << 31;
size_t order_bins(int bits);
// mode 1: always sort; 2: sort when at least half of the sampled adjacent pairs differ in bin.
// G workgroups count and scatter one contiguous chunk each.
hipError_t launch_order(int bits, int mode, const bc_ray* rays, int64_t n, const bc_node8* nodes, int G,
                        const OrderScratch& s, hipStream_t st);
// One block: flag[0] = 1 when at least half of 4,096 sampled adjacent ray pairs fall into different
// (origin cell on a 2^bits grid over the scene box, direction octant) bins, else 0; flag[1] = the
// pairs that differ.  n >= 2.
hipError_t launch_coherence_probe(int bits, const bc_ray* rays, int64_t n, const bc_node8* nodes, int* flag,
                                  hipStream_t st);

// Ray queries: any_hit writes int32 occlusion flags, else bc_hit records.  stack: LDS stack
// entries per lane (stack_for_depth).  *err is set non-zero by the kernel if a traversal stack
// would overflow.  With oflag non-null and oflag[0] set (by launch_order), the kernels read
// the rays from `sorted` and write the result of sorted ray j to out[perm[j]].
hipError_t launch_trace(int stack, bool any_hit, const bc_node8* nodes, const bc_prim* prims, const uint32_t* ids,
                        const bc_ray* rays, int64_t n, void* out, int refill, int* err, hipStream_t st,
                        const bc_ray* sorted = nullptr, const uint32_t* perm = nullptr, const int* oflag = nullptr,
                        const int* sel = nullptr);
// sel (optional): a device flag written by launch_coherence_probe on the same stream; launch_trace
// then launches both the one-ray-per-thread kernel and the lane-refill kernel (threshold refill),
// and each returns at once unless the flag chose it (nonzero: refill).

// The traversal of launch_trace, one ray per thread, that also writes each ray's node visits,
// primitive-record tests and id-table reads to counts[i] (x, y, z; w = 0).
hipError_t launch_trace_count(int stack, bool any_hit, const bc_node8* nodes, const bc_prim* prims, const uint32_t* ids,
                              const bc_ray* rays, int64_t n, void* out, int* err, uint4* counts, hipStream_t st);
// VALU issue probes: `blocks` workgroups of 256 threads, each thread `iters` register-only
// eight-child slab tests (mix 0) or rounds of eight independent FMAs (mix 1).  sink receives
// nothing in practice (it keeps the loops from being dead).
hipError_t launch_valu_probe(int mix, int blocks, uint32_t iters, uint32_t seed, uint32_t* sink, hipStream_t st);
hipError_t launch_line_probe(const uint4* buf, uint32_t block_mask, int blocks, uint32_t iters, uint32_t seed,
                             uint32_t* sink, hipStream_t st);

// Refit: records (and materials) of the moved primitives in leaf order from the view's arrays.
hipError_t launch_refit_records(const double* geom, const int32_t* kind, const int32_t* mat, const uint32_t* ids, uint32_t n,
                                bc_prim* prims, int32_t* pmat, int* err, hipStream_t st);
// One level of the device BVH8 refit: recomputes the boxes of the `n` nodes in `list`
// (whose children are already refit) and stores each node's union box in nbox[2 i .. 2 i + 1].
hipError_t launch_refit_level(bc_node8* nodes, const bc_prim* prims, const uint32_t* ids, float4* nbox,
                              const uint32_t* list, uint32_t n, int* err, hipStream_t st);

}  // namespace bcd
