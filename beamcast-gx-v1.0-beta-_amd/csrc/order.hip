// order.hip -- spatial binning of ray batches before traversal (bc_trace / bc_trace_device).
//
// Incoherent batches (origins spread over the scene, directions over the sphere) visit nodes
// all over the BVH from neighbouring lanes, so each XCD's 4 MB L2 sees the whole scene.  A
// counting sort by (Morton code of the origin's cell on a 2^bits grid over the root box, direction
// octant) gives each wave rays that start in one cell and head into one octant: they share their
// upper-level nodes, and the XCD that takes a run of consecutive workgroups works on one region of
// the scene.  The traversal then reads rays in sorted order and writes each result to the ray's
// own index, so results are those of the unsorted batch.
//
// Passes (all on the library's stream, no host synchronisation):
//   probe    : decides whether to sort (always, or when at least half of 4,096 adjacent pairs of
//              the batch, at hashed positions, fall into different bins) and writes the flag;
//   count    : G workgroups, one contiguous chunk each, bin histogram in LDS -> cnt[G][NB];
//   colscan  : per bin, exclusive prefix over the workgroups -> cnt, and per-bin totals;
//   binscan  : exclusive prefix over the bins -> bin starts;
//   scatter  : each workgroup places its chunk's rays at bin start + its prefix + LDS rank,
//              writing the sorted ray and its original index.
// Every pass returns at once when the flag says not to sort, and the traversal then reads the
// caller's batch directly.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace bcd {
namespace {

constexpr int kOrderBlock = 1024;
constexpr int kProbePairs = 4096;
constexpr int kProbeBlock = 256;

struct Grid {
    float ox, oy, oz, sx, sy, sz, cmax;
};

// The root node's quantisation grid spans the scene: origin k * 2^e, extent 255 * 2^e per axis.
// Cells per unit length = 2^bits / extent.
__device__ __forceinline__ Grid root_grid(const bc_node8* __restrict__ nodes, int bits) {
    const bc_node8 root = nodes[0];
    Grid g;
    float o[3], s[3];
    const float cells = (float)(1u << bits);
    for (int a = 0; a < 3; a++) {
        const float step = __uint_as_float((uint32_t)(bc_node_e(root, a) + 127) << 23);
        o[a] = (float)bc_node_k(root, a) * step;
        const float c = cells / (255.0f * step);
        s[a] = isfinite(c) ? c : 0.0f;
    }
    g.ox = o[0]; g.oy = o[1]; g.oz = o[2];
    g.sx = s[0]; g.sy = s[1]; g.sz = s[2];
    g.cmax = cells - 1.0f;
    return g;
}

// Bits 0..9 of v moved to bits 0, 3, 6, ..., 27.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v &= 0x3ffu;
    v = (v * 0x00010001u) & 0xff0000ffu;
    v = (v * 0x00000101u) & 0x0f00f00fu;
    v = (v * 0x00000011u) & 0xc30c30c3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// Cell index of one coordinate; NaN and out-of-box origins clamp into the border cells
// (fmaxf returns the non-NaN operand).
__device__ __forceinline__ uint32_t cell(float x, float o, float s, float cmax) {
    return (uint32_t)fminf(fmaxf((x - o) * s, 0.0f), cmax);
}

__device__ __forceinline__ uint32_t bin_of(const float4& a, const float4& b, const Grid& g) {
    const uint32_t m = (spread3(cell(a.x, g.ox, g.sx, g.cmax)) << 2) | (spread3(cell(a.y, g.oy, g.sy, g.cmax)) << 1) |
                       spread3(cell(a.z, g.oz, g.sz, g.cmax));
    const uint32_t oct = (b.x < 0.0f ? 1u : 0u) | (b.y < 0.0f ? 2u : 0u) | (b.z < 0.0f ? 4u : 0u);
    return (m << 3) | oct;
}

__device__ __forceinline__ void load_ray2(const bc_ray* __restrict__ rays, int64_t i, float4& a, float4& b) {
    const float4* p = reinterpret_cast<const float4*>(rays + i);
    a = p[0];
    b = p[1];
}

__device__ __forceinline__ uint32_t bin_at(const bc_ray* __restrict__ rays, int64_t i, const Grid& g) {
    float4 a, b;
    load_ray2(rays, i, a, b);
    return bin_of(a, b, g);
}

// One block.  flag[0] = sort or not; flag[1] = adjacent sampled pairs in different bins.
__global__ __launch_bounds__(kProbeBlock) void k_order_probe(const bc_ray* __restrict__ rays, int64_t n,
                                                             const bc_node8* __restrict__ nodes, int bits, int mode,
                                                             int* __restrict__ flag) {
    __shared__ int diff;
    if (threadIdx.x == 0) diff = 0;
    __syncthreads();
    int d = 0;
    if (mode == 2) {
        const Grid g = root_grid(nodes, bits);
        for (int j = threadIdx.x; j < kProbePairs; j += kProbeBlock) {
            // hashed positions: an even stride would alias with power-of-two image widths and
            // land on the same columns of every row
            const int64_t i = (int64_t)(((uint64_t)(uint32_t)j * 0x9e3779b97f4a7c15ull) % (uint64_t)(n - 1));
            d += bin_at(rays, i, g) != bin_at(rays, i + 1, g) ? 1 : 0;
        }
    }
    atomicAdd(&diff, d);
    __syncthreads();
    if (threadIdx.x == 0) {
        flag[0] = mode == 1 ? 1 : (2 * diff >= kProbePairs ? 1 : 0);
        flag[1] = diff;
    }
}

template <int BITS>
__global__ __launch_bounds__(kOrderBlock) void k_order_count(const bc_ray* __restrict__ rays, int64_t n,
                                                             const bc_node8* __restrict__ nodes, int64_t chunk,
                                                             uint32_t* __restrict__ cnt, const int* __restrict__ flag) {
    constexpr int NB = 1 << (3 * BITS + 3);
    __shared__ uint32_t h[NB];
    if (flag[0] == 0) return;
    for (int k = threadIdx.x; k < NB; k += kOrderBlock) h[k] = 0u;
    __syncthreads();
    const Grid g = root_grid(nodes, BITS);
    const int64_t b0 = (int64_t)blockIdx.x * chunk, b1 = min(n, b0 + chunk);
    for (int64_t i = b0 + threadIdx.x; i < b1; i += kOrderBlock) atomicAdd(&h[bin_at(rays, i, g)], 1u);
    __syncthreads();
    uint32_t* row = cnt + (size_t)blockIdx.x * NB;
    for (int k = threadIdx.x; k < NB; k += kOrderBlock) row[k] = h[k];
}

// cnt[G][NB] -> exclusive prefix over the workgroups within each bin, and the bin totals.
// Block: 64 bins x 16 strips of workgroups.  NB is a multiple of 64.
__global__ __launch_bounds__(1024) void k_order_colscan(uint32_t* __restrict__ cnt, int G, int NB,
                                                        uint32_t* __restrict__ total, const int* __restrict__ flag) {
    __shared__ uint32_t part[16][64];
    if (flag[0] == 0) return;
    const int tx = (int)(threadIdx.x & 63u), ty = (int)(threadIdx.x >> 6);
    const int b = (int)blockIdx.x * 64 + tx;
    const int g0 = ty * G / 16, g1 = (ty + 1) * G / 16;
    uint32_t s = 0;
    for (int gi = g0; gi < g1; gi++) s += cnt[(size_t)gi * NB + b];
    part[ty][tx] = s;
    __syncthreads();
    uint32_t base = 0;
    for (int y = 0; y < ty; y++) base += part[y][tx];
    if (ty == 15) total[b] = base + s;
    for (int gi = g0; gi < g1; gi++) {
        const size_t k = (size_t)gi * NB + b;
        const uint32_t c = cnt[k];
        cnt[k] = base;
        base += c;
    }
}

// Exclusive prefix of the NB bin totals (one block of 1024 threads): the totals are staged in
// LDS with coalesced loads, then each thread scans a run of NB / 1024 of them.
template <int NB>
__global__ __launch_bounds__(1024) void k_order_binscan(const uint32_t* __restrict__ total, uint32_t* __restrict__ bstart,
                                                        const int* __restrict__ flag) {
    constexpr int per = NB >= 1024 ? NB / 1024 : 1;
    __shared__ uint32_t t[NB];
    __shared__ uint32_t s[1024];
    if (flag[0] == 0) return;
    for (int k = threadIdx.x; k < NB; k += 1024) t[k] = total[k];
    __syncthreads();
    const int k0 = (int)threadIdx.x * per, k1 = min(NB, k0 + per);
    uint32_t sum = 0;
    for (int k = k0; k < k1; k++) sum += t[k];
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t v = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = s[threadIdx.x] - sum;
    for (int k = k0; k < k1; k++) {
        const uint32_t c = t[k];
        t[k] = run;
        run += c;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < NB; k += 1024) bstart[k] = t[k];
}

template <int BITS>
__global__ __launch_bounds__(kOrderBlock) void k_order_scatter(const bc_ray* __restrict__ rays, int64_t n,
                                                               const bc_node8* __restrict__ nodes, int64_t chunk,
                                                               const uint32_t* __restrict__ cnt,
                                                               const uint32_t* __restrict__ bstart, bc_ray* __restrict__ sorted,
                                                               uint32_t* __restrict__ perm, const int* __restrict__ flag) {
    constexpr int NB = 1 << (3 * BITS + 3);
    __shared__ uint32_t o[NB];
    if (flag[0] == 0) return;
    const uint32_t* row = cnt + (size_t)blockIdx.x * NB;
    for (int k = threadIdx.x; k < NB; k += kOrderBlock) o[k] = bstart[k] + row[k];
    __syncthreads();
    const Grid g = root_grid(nodes, BITS);
    const int64_t b0 = (int64_t)blockIdx.x * chunk, b1 = min(n, b0 + chunk);
    for (int64_t i = b0 + threadIdx.x; i < b1; i += kOrderBlock) {
        float4 a, b;
        load_ray2(rays, i, a, b);
        const uint32_t pos = atomicAdd(&o[bin_of(a, b, g)], 1u);
        float4* q = reinterpret_cast<float4*>(sorted + pos);
        q[0] = a;
        q[1] = b;
        perm[pos] = (uint32_t)i;
    }
}

template <int BITS>
hipError_t launch_order_t(const bc_ray* rays, int64_t n, const bc_node8* nodes, int G, const OrderScratch& s,
                          hipStream_t st) {
    constexpr int NB = 1 << (3 * BITS + 3);
    const int64_t chunk = (n + G - 1) / G;
    hipLaunchKernelGGL((k_order_count<BITS>), dim3(G), dim3(kOrderBlock), 0, st, rays, n, nodes, chunk, s.cnt, s.flag);
    hipLaunchKernelGGL(k_order_colscan, dim3(NB / 64), dim3(1024), 0, st, s.cnt, G, NB, s.total, s.flag);
    hipLaunchKernelGGL((k_order_binscan<NB>), dim3(1), dim3(1024), 0, st, s.total, s.bstart, s.flag);
    hipLaunchKernelGGL((k_order_scatter<BITS>), dim3(G), dim3(kOrderBlock), 0, st, rays, n, nodes, chunk, s.cnt, s.bstart,
                       s.sorted, s.perm, s.flag);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_coherence_probe(int bits, const bc_ray* rays, int64_t n, const bc_node8* nodes, int* flag,
                                  hipStream_t st) {
    if (bits < kOrderMinBits || bits > kOrderMaxBits || n < 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_order_probe, dim3(1), dim3(kProbeBlock), 0, st, rays, n, nodes, bits, 2, flag);
    return hipGetLastError();
}

size_t order_bins(int bits) { return (size_t)1 << (3 * bits + 3); }

hipError_t launch_order(int bits, int mode, const bc_ray* rays, int64_t n, const bc_node8* nodes, int G,
                        const OrderScratch& s, hipStream_t st) {
    if (bits < kOrderMinBits || bits > kOrderMaxBits || mode < 1 || mode > 2 || n < 2 || n > kOrderMaxRays || G < 1 ||
        G > kOrderMaxGroups)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_order_probe, dim3(1), dim3(kProbeBlock), 0, st, rays, n, nodes, bits, mode, s.flag);
    switch (bits) {
        case 2: return launch_order_t<2>(rays, n, nodes, G, s, st);
        case 3: return launch_order_t<3>(rays, n, nodes, G, s, st);
        default: return launch_order_t<4>(rays, n, nodes, G, s, st);
    }
}

}  // namespace bcd
