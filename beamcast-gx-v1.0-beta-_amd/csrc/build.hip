// build.hip -- device SAH build of the binary BVH, node for node the Go Scene.Build.
//
// Restates buildNodeSAH (renderer.go:526-633) with its helpers nodeBounds, largestAxis,
// centroidAxis and AABB.SurfaceArea (renderer.go:133-139, 475-504) on the GPU, so that the
// resulting Nodes (Box, Left, Right, Start, Count) and Order equal the host build exactly:
//   * node bounds are min/max unions (order-independent; compared through an order-preserving
//     integer key of the double so that -0 sorts below +0, as Go's math.Min/Max treat zeros);
//   * bin index, surface areas and split costs use the same fp64 operations in the same order
//     (compiled without contraction);
//   * the partition reproduces the permutation of Go's two-pointer swap loop
//     (i from the left, swap with j from the right) through a closed form that needs only
//     prefix counts: with R = "bin > split", nR right-side primitives, mid = end - nR,
//     holes H[j] = the R positions left of mid (ascending) and fillers F[j] = the L positions
//     right of mid (descending), the loop sends
//        L left of mid  -> stays            R left of mid (hole j) -> (j ? F[j-1] : end) - 1
//        L right of mid (filler j) -> H[j]  R at mid -> (k ? F[k-1] : end) - 1, k = #holes
//        R right of mid -> one position down;
//   * node indices are the depth-first pre-order of the recursion (left subtree first).
//
// Schedule: nodes with more than kSmall primitives are split level by level over all of the
// level's primitives at once (chunks of kChunk primitives per workgroup: bins in LDS merged with
// global atomics, one thread per node for the 32-bin sweep, then the partition as prefix counts
// and a scatter).  Nodes of at most kSmall primitives are subtree roots: one wave builds each
// subtree with the recursion's own depth-first order, its primitive indices held in LDS.
// Finally subtree sizes give every node its pre-order index and the nodes are emitted.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "api_internal.hpp"

namespace bcd {
namespace {

constexpr int kMaxBins = 32;   // Scene.BVHBins is clamped to [4, 32]
constexpr int kSmall = 256;    // subtrees of at most this many primitives: one wave each
constexpr int kChunk = 4096;   // primitives per workgroup in the level passes
constexpr int kThreads = 256;  // level-pass workgroup
constexpr int kKeys = 12;      // box min xyz, box max xyz, centroid min xyz, centroid max xyz
constexpr int kBinWords = 1 + kKeys;

enum : int32_t { kLevel = 0, kLeaf = 1, kSubtree = 2 };

struct BNode {
    int32_t start, count;
    int32_t left, right;  // temporary ids of the children, -1 for leaves
    int32_t kind;         // kLevel (split in the level passes), kLeaf, kSubtree
    int32_t split, lcount, axis;
    int32_t size, id;     // nodes in the subtree, final pre-order index
    int32_t base, holes;  // subtree roots: offset in the subtree store; partition: holes
    double box[6], cb[6];
};

struct SNode {            // node of a wave-built subtree (links local to the subtree)
    double box[6];
    int32_t left, right, start, count;
};

// Order-preserving int64 key of a double (negative values have their magnitude bits flipped).
__device__ __forceinline__ long long dkey(double d) {
    const long long b = __double_as_longlong(d);
    return b >= 0 ? b : (b ^ 0x7FFFFFFFFFFFFFFFLL);
}
__device__ __forceinline__ double dval(long long k) {
    return __longlong_as_double(k >= 0 ? k : (k ^ 0x7FFFFFFFFFFFFFFFLL));
}
__device__ __forceinline__ long long key_pos_inf() { return 0x7FF0000000000000LL; }
__device__ __forceinline__ long long key_neg_inf() { return (long long)(0xFFF0000000000000ULL ^ 0x7FFFFFFFFFFFFFFFULL); }

// AABB.SurfaceArea (renderer.go:133-139)
__device__ __forceinline__ double surface_area(const double* mn, const double* mx) {
    const double ex = mx[0] - mn[0], ey = mx[1] - mn[1], ez = mx[2] - mn[2];
    if (ex < 0 || ey < 0 || ez < 0) return 0;
    return 2 * (ex * ey + ey * ez + ez * ex);
}

// largestAxis (renderer.go:485-493)
__device__ __forceinline__ int largest_axis(double x, double y, double z) {
    if (y > x && y >= z) return 1;
    if (z > x && z > y) return 2;
    return 0;
}

// The leaf tests buildNodeSAH makes before binning: count <= leaf, then a degenerate
// centroid extent on the largest axis.  Sets the partition axis of the node.
__device__ int classify(BNode& nd, int leaf) {
    nd.axis = 0;
    if (nd.count <= leaf) return kLeaf;
    const double ex = nd.cb[3] - nd.cb[0], ey = nd.cb[4] - nd.cb[1], ez = nd.cb[5] - nd.cb[2];
    nd.axis = largest_axis(ex, ey, ez);
    if (nd.cb[3 + nd.axis] - nd.cb[nd.axis] < 1e-12) return kLeaf;
    return nd.count <= kSmall ? kSubtree : kLevel;
}

// Bin index of a centroid coordinate: int((c - minC) * invExtent * bins), clamped.
__device__ __forceinline__ int bin_of(double c, double minC, double invExtent, int bins) {
    int b = (int)((c - minC) * invExtent * (double)bins);
    return b < 0 ? 0 : (b >= bins ? bins - 1 : b);
}

// The 32-bin sweep of buildNodeSAH: returns the split bin or -1 for a leaf.  bins hold, per
// bin, count and the keys of the box (min xyz, max xyz).
__device__ int sah_split(const int* cnt, const long long (*key)[6], int bins, int count, int leaf,
                         const double* box) {
    int leftCount[kMaxBins - 1], rightCount[kMaxBins - 1];
    double leftArea[kMaxBins - 1], rightArea[kMaxBins - 1];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    int acc = 0;
    for (int i = 0; i + 1 < bins; i++) {
        if (cnt[i] > 0)
            for (int a = 0; a < 3; a++) { mn[a] = fmin(mn[a], dval(key[i][a])); mx[a] = fmax(mx[a], dval(key[i][3 + a])); }
        acc += cnt[i];
        leftCount[i] = acc;
        leftArea[i] = surface_area(mn, mx);
    }
    for (int a = 0; a < 3; a++) { mn[a] = INFINITY; mx[a] = -INFINITY; }
    acc = 0;
    for (int i = bins - 1; i > 0; i--) {
        if (cnt[i] > 0)
            for (int a = 0; a < 3; a++) { mn[a] = fmin(mn[a], dval(key[i][a])); mx[a] = fmax(mx[a], dval(key[i][3 + a])); }
        acc += cnt[i];
        rightCount[i - 1] = acc;
        rightArea[i - 1] = surface_area(mn, mx);
    }
    int best = -1;
    double bestCost = INFINITY;
    for (int i = 0; i + 1 < bins; i++) {
        if (leftCount[i] == 0 || rightCount[i] == 0) continue;
        const double cost = leftArea[i] * (double)leftCount[i] + rightArea[i] * (double)rightCount[i];
        if (cost < bestCost) { bestCost = cost; best = i; }
    }
    const double leafCost = surface_area(box, box + 3) * (double)count;
    if (best < 0 || (count <= leaf * 2 && bestCost >= leafCost * 0.98)) return -1;
    return best;
}

// ---------------------------------------------------------------------------------------------
// Root bounds

__global__ __launch_bounds__(kThreads) void k_root_bounds(const double* __restrict__ pbox, const double* __restrict__ pcen,
                                                          int n, long long* __restrict__ keys) {
    __shared__ long long s[kKeys];
    if (threadIdx.x < kKeys) s[threadIdx.x] = threadIdx.x % 6 < 3 ? key_pos_inf() : key_neg_inf();
    __syncthreads();
    long long k[kKeys];
    for (int q = 0; q < kKeys; q++) k[q] = q % 6 < 3 ? key_pos_inf() : key_neg_inf();
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        for (int a = 0; a < 3; a++) {
            k[a] = min(k[a], dkey(pbox[6 * (size_t)i + a]));
            k[3 + a] = max(k[3 + a], dkey(pbox[6 * (size_t)i + 3 + a]));
            const long long c = dkey(pcen[3 * (size_t)i + a]);
            k[6 + a] = min(k[6 + a], c);
            k[9 + a] = max(k[9 + a], c);
        }
    }
    for (int q = 0; q < kKeys; q++) {
        if (q % 6 < 3) atomicMin(&s[q], k[q]); else atomicMax(&s[q], k[q]);
    }
    __syncthreads();
    if (threadIdx.x < kKeys) {
        const int q = threadIdx.x;
        if (q % 6 < 3) atomicMin(&keys[q], s[q]); else atomicMax(&keys[q], s[q]);
    }
}

__global__ void k_init_bounds(long long* keys) {
    const int q = threadIdx.x;
    if (q < kKeys) keys[q] = q % 6 < 3 ? key_pos_inf() : key_neg_inf();
}

__global__ void k_init_keys(long long* keys, int words) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= words) return;
    const int q = i % kBinWords;  // bin layout: count, then the 12 keys
    keys[i] = q == 0 ? 0 : ((q - 1) % 6 < 3 ? key_pos_inf() : key_neg_inf());
}

// Root node and its classification; the lists get the root when it is not a leaf.
__global__ void k_make_root(BNode* nodes, const long long* keys, int n, int leaf, int* lists) {
    BNode nd;
    memset(&nd, 0, sizeof nd);
    nd.start = 0;
    nd.count = n;
    nd.left = nd.right = -1;
    nd.split = -1;
    for (int q = 0; q < 6; q++) { nd.box[q] = dval(keys[q]); nd.cb[q] = dval(keys[6 + q]); }
    nd.kind = classify(nd, leaf);
    nd.size = 1;
    nd.id = 0;
    nodes[0] = nd;
    // lists[0]: next-level count, lists[1]: subtree-root count, lists[2]: node counter
    lists[0] = 0;
    lists[1] = 0;
    lists[2] = 1;
}

// ---------------------------------------------------------------------------------------------
// Level passes (nodes of more than kSmall primitives)

struct Chunk {
    int32_t li;      // node's index in the level list
    int32_t p0, p1;  // primitive positions [p0, p1)
};

__global__ __launch_bounds__(kThreads) void k_bin(const BNode* __restrict__ nodes, const int* __restrict__ level,
                                                  const Chunk* __restrict__ chunks, const int* __restrict__ ord,
                                                  const double* __restrict__ pbox, const double* __restrict__ pcen,
                                                  uint8_t* __restrict__ binid, long long* __restrict__ gbins, int bins) {
    __shared__ int cnt[kMaxBins];
    __shared__ long long key[kMaxBins][kKeys];
    const Chunk ch = chunks[blockIdx.x];
    const BNode& nd = nodes[level[ch.li]];
    const int axis = nd.axis;
    const double minC = nd.cb[axis], invExtent = 1 / (nd.cb[3 + axis] - nd.cb[axis]);
    for (int i = threadIdx.x; i < kMaxBins * kKeys; i += kThreads) {
        const int q = i % kKeys;
        key[i / kKeys][q] = q % 6 < 3 ? key_pos_inf() : key_neg_inf();
    }
    if (threadIdx.x < kMaxBins) cnt[threadIdx.x] = 0;
    __syncthreads();
    for (int p = ch.p0 + threadIdx.x; p < ch.p1; p += kThreads) {
        const size_t id = (size_t)ord[p];
        const double c[3] = {pcen[3 * id], pcen[3 * id + 1], pcen[3 * id + 2]};
        const int b = bin_of(c[axis], minC, invExtent, bins);
        binid[p] = (uint8_t)b;
        atomicAdd(&cnt[b], 1);
        for (int a = 0; a < 3; a++) {
            atomicMin(&key[b][a], dkey(pbox[6 * id + a]));
            atomicMax(&key[b][3 + a], dkey(pbox[6 * id + 3 + a]));
            const long long ck = dkey(c[a]);
            atomicMin(&key[b][6 + a], ck);
            atomicMax(&key[b][9 + a], ck);
        }
    }
    __syncthreads();
    long long* g = gbins + (size_t)ch.li * kMaxBins * kBinWords;
    for (int i = threadIdx.x; i < bins * kBinWords; i += kThreads) {
        const int b = i / kBinWords, q = i % kBinWords;
        if (cnt[b] == 0) continue;
        if (q == 0) atomicAdd((unsigned long long*)&g[b * kBinWords], (unsigned long long)cnt[b]);
        else if ((q - 1) % 6 < 3) atomicMin(&g[b * kBinWords + q], key[b][q - 1]);
        else atomicMax(&g[b * kBinWords + q], key[b][q - 1]);
    }
}

// One thread per level node: the sweep, then the children (bounds from the bins).
__global__ void k_sah(BNode* __restrict__ nodes, const int* __restrict__ level, int nl, const long long* __restrict__ gbins,
                      int bins, int leaf, int* __restrict__ next, int* __restrict__ subs, int* __restrict__ lists,
                      int node_cap, int* __restrict__ err) {
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nl) return;
    const int v = level[li];
    BNode nd = nodes[v];
    const long long* g = gbins + (size_t)li * kMaxBins * kBinWords;
    int cnt[kMaxBins];
    long long key[kMaxBins][6];
    for (int b = 0; b < bins; b++) {
        cnt[b] = (int)g[b * kBinWords];
        for (int q = 0; q < 6; q++) key[b][q] = g[b * kBinWords + 1 + q];
    }
    const int split = sah_split(cnt, key, bins, nd.count, leaf, nd.box);
    if (split < 0) {
        nd.kind = kLeaf;
        nodes[v] = nd;
        return;
    }
    const int t = atomicAdd(&lists[2], 2);
    if (t + 2 > node_cap) { atomicOr(err, 8); nd.kind = kLeaf; nodes[v] = nd; return; }
    int lcount = 0;
    for (int b = 0; b <= split; b++) lcount += cnt[b];
    for (int side = 0; side < 2; side++) {
        BNode c;
        memset(&c, 0, sizeof c);
        c.start = side ? nd.start + lcount : nd.start;
        c.count = side ? nd.count - lcount : lcount;
        c.left = c.right = -1;
        c.split = -1;
        c.size = 1;
        long long k[kKeys];
        for (int q = 0; q < kKeys; q++) k[q] = q % 6 < 3 ? key_pos_inf() : key_neg_inf();
        const int b0 = side ? split + 1 : 0, b1 = side ? bins : split + 1;
        for (int b = b0; b < b1; b++) {
            if (cnt[b] == 0) continue;
            for (int q = 0; q < kKeys; q++) {
                const long long x = g[b * kBinWords + 1 + q];
                k[q] = q % 6 < 3 ? min(k[q], x) : max(k[q], x);
            }
        }
        for (int q = 0; q < 6; q++) { c.box[q] = dval(k[q]); c.cb[q] = dval(k[6 + q]); }
        c.kind = classify(c, leaf);
        nodes[t + side] = c;
        if (c.kind == kLevel) next[atomicAdd(&lists[0], 1)] = t + side;
        else if (c.kind == kSubtree) subs[atomicAdd(&lists[1], 1)] = t + side;
    }
    nd.split = split;
    nd.lcount = lcount;
    nd.left = t;
    nd.right = t + 1;
    nodes[v] = nd;
}

// Exclusive block scan of one flag per thread (256 threads = 4 waves); returns the rank and
// the block total in *total.
__device__ __forceinline__ int block_rank(bool flag, int* lds4, int* total) {
    const unsigned long long m = __ballot(flag);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) lds4[w] = __popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
    for (int i = 0; i < kThreads / 64; i++) {
        if (i < w) off += lds4[i];
        tot += lds4[i];
    }
    __syncthreads();
    *total = tot;
    return off + r;
}

__global__ __launch_bounds__(kThreads) void k_count_r(const BNode* __restrict__ nodes, const int* __restrict__ level,
                                                      const Chunk* __restrict__ chunks, const uint8_t* __restrict__ binid,
                                                      int* __restrict__ chunk_r) {
    __shared__ int lds4[kThreads / 64];
    const Chunk ch = chunks[blockIdx.x];
    const BNode& nd = nodes[level[ch.li]];
    if (nd.kind != kLevel) return;
    int sum = 0;
    for (int p0 = ch.p0; p0 < ch.p1; p0 += kThreads) {
        const int p = p0 + threadIdx.x;
        int tot;
        block_rank(p < ch.p1 && binid[p] > nd.split, lds4, &tot);
        sum += tot;
    }
    if (threadIdx.x == 0) chunk_r[blockIdx.x] = sum;
}

// One thread per level node: exclusive prefix of its chunks' R counts.
__global__ void k_chunk_scan(const BNode* __restrict__ nodes, const int* __restrict__ level, int nl,
                             const int* __restrict__ chunk_first, int* __restrict__ chunk_r, int* __restrict__ err) {
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nl) return;
    const BNode& nd = nodes[level[li]];
    if (nd.kind != kLevel) return;
    int acc = 0;
    for (int c = chunk_first[li]; c < chunk_first[li + 1]; c++) {
        const int x = chunk_r[c];
        chunk_r[c] = acc;
        acc += x;
    }
    if (acc != nd.count - nd.lcount) atomicOr(err, 16);
}

// Per primitive: rank among the node's R primitives; holes and fillers by index.
__global__ __launch_bounds__(kThreads) void k_holes(BNode* __restrict__ nodes, const int* __restrict__ level,
                                                    const Chunk* __restrict__ chunks, const uint8_t* __restrict__ binid,
                                                    const int* __restrict__ chunk_off, int* __restrict__ rr,
                                                    int* __restrict__ hf, int* __restrict__ err) {
    __shared__ int lds4[kThreads / 64];
    const Chunk ch = chunks[blockIdx.x];
    BNode& nd = nodes[level[ch.li]];
    if (nd.kind != kLevel) return;
    const int start = nd.start, end = nd.start + nd.count, nR = nd.count - nd.lcount, mid = end - nR;
    int carry = chunk_off[blockIdx.x];
    for (int p0 = ch.p0; p0 < ch.p1; p0 += kThreads) {
        const int p = p0 + threadIdx.x;
        const bool isR = p < ch.p1 && binid[p] > nd.split;
        int tot;
        const int r = carry + block_rank(isR, lds4, &tot);
        carry += tot;
        if (p >= ch.p1) continue;
        rr[p] = r;
        if (p < mid && isR) {  // hole r
            if (r < mid - start) hf[start + r] = p; else atomicOr(err, 16);
        }
        if (p >= mid && !isR) {  // filler, counted from the top
            const int j = (end - 1 - p) - (nR - r);
            if (j >= 0 && j < end - mid) hf[mid + j] = p; else atomicOr(err, 16);
        }
        if (p == mid) nd.holes = r;  // R primitives left of mid
    }
}

__global__ __launch_bounds__(kThreads) void k_move(const BNode* __restrict__ nodes, const int* __restrict__ level,
                                                   const Chunk* __restrict__ chunks, const uint8_t* __restrict__ binid,
                                                   const int* __restrict__ rr, const int* __restrict__ hf,
                                                   const int* __restrict__ ord, int* __restrict__ ord2, int* __restrict__ err) {
    const Chunk ch = chunks[blockIdx.x];
    const BNode& nd = nodes[level[ch.li]];
    if (nd.kind != kLevel) return;
    const int start = nd.start, end = nd.start + nd.count, nR = nd.count - nd.lcount, mid = end - nR, k = nd.holes;
    const int seg_end = k > 0 && k <= end - mid ? hf[mid + k - 1] : end;  // lowest filler, or end without holes
    for (int p = ch.p0 + threadIdx.x; p < ch.p1; p += kThreads) {
        const bool isR = binid[p] > nd.split;
        const int r = rr[p];
        int d = -1;
        if (p < mid) {
            if (!isR) d = p;
            else if (r == 0) d = end - 1;
            else if (r - 1 < end - mid) d = hf[mid + r - 1] - 1;
        } else if (!isR) {
            const int j = (end - 1 - p) - (nR - r);
            if (j >= 0 && j < mid - start) d = hf[start + j];
        } else {
            d = p == mid ? seg_end - 1 : p - 1;
        }
        if (d >= start && d < end) ord2[d] = ord[p]; else atomicOr(err, 16);
    }
}

__global__ __launch_bounds__(kThreads) void k_copy_back(const BNode* __restrict__ nodes, const int* __restrict__ level,
                                                        const Chunk* __restrict__ chunks, const int* __restrict__ ord2,
                                                        int* __restrict__ ord) {
    const Chunk ch = chunks[blockIdx.x];
    if (nodes[level[ch.li]].kind != kLevel) return;
    for (int p = ch.p0 + threadIdx.x; p < ch.p1; p += kThreads) ord[p] = ord2[p];
}

// ---------------------------------------------------------------------------------------------
// Subtrees of at most kSmall primitives: one wave (one 64-thread workgroup) each, depth first.

struct StackEntry {
    int16_t s, e;      // local primitive range
    int16_t parent;    // local node index, -1 for the subtree root
    int16_t side;      // 0 left, 1 right
};

__global__ __launch_bounds__(64) void k_subtree(BNode* __restrict__ nodes, const int* __restrict__ subs, int nsub,
                                                int* __restrict__ ord, const double* __restrict__ pbox,
                                                const double* __restrict__ pcen, SNode* __restrict__ store,
                                                int* __restrict__ store_count, int store_cap, int bins, int leaf,
                                                int* __restrict__ err) {
    __shared__ int s_ord[kSmall], s_tmp[kSmall], s_rr[kSmall], s_hf[kSmall];
    __shared__ uint8_t s_bin[kSmall];
    __shared__ int s_cnt[kMaxBins];
    __shared__ long long s_key[kMaxBins][6];
    __shared__ StackEntry s_stack[kSmall];
    __shared__ int s_base, s_split;
    if (blockIdx.x >= (unsigned)nsub) return;
    const int lane = threadIdx.x;
    const int v = subs[blockIdx.x];
    const int gstart = nodes[v].start, n = nodes[v].count;
    if (lane == 0) {
        const int need = 2 * n - 1;
        const int b = atomicAdd(store_count, need);
        s_base = b + need <= store_cap ? b : -1;
        if (s_base < 0) atomicOr(err, 32);
    }
    for (int p = lane; p < n; p += 64) s_ord[p] = ord[gstart + p];
    __syncthreads();
    const int base = s_base;
    if (base < 0) return;
    int top = 0, m = 0;
    if (lane == 0) s_stack[0] = {0, (int16_t)n, -1, 0};
    __syncthreads();
    top = 1;
    while (top > 0) {
        top--;
        const StackEntry se = s_stack[top];
        const int s = se.s, e = se.e, count = e - s, idx = m++;
        // nodeBounds over [s, e)
        long long k[kKeys];
        for (int q = 0; q < kKeys; q++) k[q] = q % 6 < 3 ? key_pos_inf() : key_neg_inf();
        for (int p = s + lane; p < e; p += 64) {
            const size_t id = (size_t)s_ord[p];
            for (int a = 0; a < 3; a++) {
                k[a] = min(k[a], dkey(pbox[6 * id + a]));
                k[3 + a] = max(k[3 + a], dkey(pbox[6 * id + 3 + a]));
                const long long c = dkey(pcen[3 * id + a]);
                k[6 + a] = min(k[6 + a], c);
                k[9 + a] = max(k[9 + a], c);
            }
        }
        for (int off = 32; off > 0; off >>= 1)
            for (int q = 0; q < kKeys; q++) {
                const long long o = __shfl_xor(k[q], off);
                k[q] = q % 6 < 3 ? min(k[q], o) : max(k[q], o);
            }
        BNode nd;
        nd.count = count;
        for (int q = 0; q < 6; q++) { nd.box[q] = dval(k[q]); nd.cb[q] = dval(k[6 + q]); }
        const int kind = classify(nd, leaf);
        int split = -1;
        if (kind != kLeaf) {
            const int axis = nd.axis;
            const double minC = nd.cb[axis], invExtent = 1 / (nd.cb[3 + axis] - nd.cb[axis]);
            if (lane < kMaxBins) {
                s_cnt[lane] = 0;
                for (int q = 0; q < 6; q++) s_key[lane][q] = q < 3 ? key_pos_inf() : key_neg_inf();
            }
            __syncthreads();
            for (int p = s + lane; p < e; p += 64) {
                const size_t id = (size_t)s_ord[p];
                const int b = bin_of(pcen[3 * id + axis], minC, invExtent, bins);
                s_bin[p] = (uint8_t)b;
                atomicAdd(&s_cnt[b], 1);
                for (int a = 0; a < 3; a++) {
                    atomicMin(&s_key[b][a], dkey(pbox[6 * id + a]));
                    atomicMax(&s_key[b][3 + a], dkey(pbox[6 * id + 3 + a]));
                }
            }
            __syncthreads();
            if (lane == 0) s_split = sah_split(s_cnt, s_key, bins, count, leaf, nd.box);
            __syncthreads();
            split = s_split;
        }
        if (lane == 0) {
            SNode out;
            for (int q = 0; q < 6; q++) out.box[q] = nd.box[q];
            out.left = out.right = -1;
            out.start = split < 0 ? gstart + s : 0;
            out.count = split < 0 ? count : 0;
            store[base + idx] = out;
            if (se.parent >= 0) {
                if (se.side) store[base + se.parent].right = idx;
                else store[base + se.parent].left = idx;
            }
        }
        if (split < 0) continue;
        // partition [s, e): R = bin > split, Go's swap-loop permutation in closed form
        int carry = 0;
        for (int p0 = s; p0 < e; p0 += 64) {
            const int p = p0 + lane;
            const bool isR = p < e && s_bin[p] > split;
            const unsigned long long msk = __ballot(isR);
            if (p < e) s_rr[p] = carry + __popcll(msk & ((1ull << lane) - 1ull));
            carry += __popcll(msk);
        }
        const int nR = carry, mid = e - nR;
        __syncthreads();
        if (nR <= 0 || nR >= count) {  // a split has primitives on both sides
            if (lane == 0) atomicOr(err, 64);
            continue;
        }
        for (int p = s + lane; p < e; p += 64) {
            const bool isR = s_bin[p] > split;
            const int r = s_rr[p];
            if (p < mid && isR) {
                if (r < mid - s) s_hf[s + r] = p; else atomicOr(err, 64);
            }
            if (p >= mid && !isR) {
                const int j = (e - 1 - p) - (nR - r);
                if (j >= 0 && j < e - mid) s_hf[mid + j] = p; else atomicOr(err, 64);
            }
        }
        __syncthreads();
        const int kh = s_rr[mid];
        const int seg_end = kh > 0 && kh <= e - mid ? s_hf[mid + kh - 1] : e;
        for (int p = s + lane; p < e; p += 64) {
            const bool isR = s_bin[p] > split;
            const int r = s_rr[p];
            int d = -1;
            if (p < mid) {
                if (!isR) d = p;
                else if (r == 0) d = e - 1;
                else if (r - 1 < e - mid) d = s_hf[mid + r - 1] - 1;
            } else if (!isR) {
                const int j = (e - 1 - p) - (nR - r);
                if (j >= 0 && j < mid - s) d = s_hf[s + j];
            } else {
                d = p == mid ? seg_end - 1 : p - 1;
            }
            if (d >= s && d < e) s_tmp[d] = s_ord[p]; else atomicOr(err, 64);
        }
        __syncthreads();
        for (int p = s + lane; p < e; p += 64) s_ord[p] = s_tmp[p];
        if (lane == 0) {
            s_stack[top] = {(int16_t)mid, (int16_t)e, (int16_t)idx, 1};
            s_stack[top + 1] = {(int16_t)s, (int16_t)mid, (int16_t)idx, 0};
        }
        top += 2;
        __syncthreads();
    }
    for (int p = lane; p < n; p += 64) ord[gstart + p] = s_ord[p];
    if (lane == 0) {
        nodes[v].size = m;
        nodes[v].base = base;
    }
}

// ---------------------------------------------------------------------------------------------
// Pre-order indices and emission

__global__ void k_sizes(BNode* __restrict__ nodes, const int* __restrict__ level, int nl) {
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nl) return;
    BNode& nd = nodes[level[li]];
    nd.size = nd.kind == kLeaf ? 1 : 1 + nodes[nd.left].size + nodes[nd.right].size;
}

__global__ void k_ids(BNode* __restrict__ nodes, const int* __restrict__ level, int nl) {
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nl) return;
    const BNode& nd = nodes[level[li]];
    if (nd.kind == kLeaf) return;
    nodes[nd.left].id = nd.id + 1;
    nodes[nd.right].id = nd.id + 1 + nodes[nd.left].size;
}

__global__ void k_emit_nodes(const BNode* __restrict__ nodes, int nnodes, double* __restrict__ out_box,
                             int32_t* __restrict__ out_link, int out_cap, int* __restrict__ err) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nnodes) return;
    const BNode& nd = nodes[v];
    if (nd.kind == kSubtree) return;
    if (nd.id < 0 || nd.id >= out_cap) { atomicOr(err, 128); return; }
    const size_t o = (size_t)nd.id;
    for (int q = 0; q < 6; q++) out_box[6 * o + q] = nd.box[q];
    int32_t* l = out_link + 4 * o;
    if (nd.kind == kLeaf) { l[0] = -1; l[1] = -1; l[2] = nd.start; l[3] = nd.count; }
    else { l[0] = nodes[nd.left].id; l[1] = nodes[nd.right].id; l[2] = 0; l[3] = 0; }
}

__global__ void k_emit_subtrees(const BNode* __restrict__ nodes, const int* __restrict__ subs, int nsub,
                                const SNode* __restrict__ store, double* __restrict__ out_box, int32_t* __restrict__ out_link,
                                int out_cap, int* __restrict__ err) {
    if (blockIdx.x >= (unsigned)nsub) return;
    const BNode& root = nodes[subs[blockIdx.x]];
    if (root.id < 0 || root.id + root.size > out_cap) {
        if (threadIdx.x == 0) atomicOr(err, 128);
        return;
    }
    for (int i = threadIdx.x; i < root.size; i += blockDim.x) {
        const SNode& sn = store[root.base + i];
        const size_t o = (size_t)(root.id + i);
        for (int q = 0; q < 6; q++) out_box[6 * o + q] = sn.box[q];
        int32_t* l = out_link + 4 * o;
        if (sn.left < 0) { l[0] = -1; l[1] = -1; l[2] = sn.start; l[3] = sn.count; }
        else { l[0] = root.id + sn.left; l[1] = root.id + sn.right; l[2] = 0; l[3] = 0; }
    }
}

__global__ void k_ranges(const BNode* __restrict__ nodes, const int* __restrict__ level, int nl, int2* __restrict__ out) {
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li < nl) out[li] = make_int2(nodes[level[li]].start, nodes[level[li]].count);
}

__global__ void k_iota(int* ord, int n) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) ord[i] = i;
}

inline unsigned grid_of(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

}  // namespace
}  // namespace bcd

// ---------------------------------------------------------------------------------------------
// Host orchestration

namespace {

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    bool alloc(bc_ctx* c, size_t bytes, const char* what) {
        if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) {
            p = nullptr;
            c->err = std::string("out of device memory: ") + what;
            return false;
        }
        return true;
    }
    template <class T> T* as() const { return (T*)p; }
};

}  // namespace

extern "C" int bc_build_bvh(bc_ctx* c, int32_t n, const double* prim_box, const double* prim_centroid, int32_t bins,
                            int32_t leaf, int32_t node_cap, double* node_box, int32_t* node_link, int32_t* order,
                            int32_t* nnode_out) {
    using namespace bcd;
    if (!c || n <= 0 || !prim_box || !prim_centroid || !node_box || !node_link || !order || !nnode_out) return BC_ERR_ARG;
    if (n >= (1 << 30) / 2) { c->err = "bc_build_bvh: too many primitives"; return BC_ERR_ARG; }
    // Scene.Build's clamps (renderer.go:416-427)
    if (bins < 4) bins = 16;
    if (bins > 32) bins = 32;
    if (leaf < 1) leaf = 4;
    if (leaf > 16) leaf = 16;
    for (size_t i = 0; i < 6 * (size_t)n; i++)
        if (!std::isfinite(prim_box[i])) { c->err = "bc_build_bvh: non-finite primitive box"; return BC_ERR_ARG; }
    for (size_t i = 0; i < 3 * (size_t)n; i++)
        if (!std::isfinite(prim_centroid[i])) { c->err = "bc_build_bvh: non-finite centroid"; return BC_ERR_ARG; }
    (void)hipSetDevice(c->device);
    hipStream_t st = c->stream;
    const size_t N = (size_t)n;
    // level-pass node capacity: 2N/kSmall suffices unless splits are very uneven; grown on demand
    int tcap = (int)std::min<size_t>(2 * N, 4 * N / kSmall + 64);
    const int scap = (int)(2 * N);                                        // subtree store
    DevBuf pbox, pcen, ord, ord2, binid, rr, hf, nodes, store, lists, keys, level, next, subs, obox, olink;
    if (!pbox.alloc(c, 48 * N, "primitive boxes") || !pcen.alloc(c, 24 * N, "centroids") || !ord.alloc(c, 4 * N, "order") ||
        !ord2.alloc(c, 4 * N, "order") || !binid.alloc(c, N, "bins") || !rr.alloc(c, 4 * N, "ranks") ||
        !hf.alloc(c, 4 * N, "holes") || !nodes.alloc(c, sizeof(BNode) * (size_t)tcap, "build nodes") ||
        !store.alloc(c, sizeof(SNode) * (size_t)scap, "subtree nodes") || !lists.alloc(c, 16, "counters") ||
        !keys.alloc(c, 8 * kKeys, "root bounds") || !level.alloc(c, 4 * (size_t)tcap, "level list") ||
        !next.alloc(c, 4 * (size_t)tcap, "level list") || !subs.alloc(c, 4 * (size_t)tcap, "subtree list") ||
        !obox.alloc(c, 48 * 2 * N, "node boxes") || !olink.alloc(c, 16 * 2 * N, "node links"))
        return BC_ERR_NOMEM;
    int* lst = lists.as<int>();
    if (!bc_h2d(c, pbox.p, prim_box, 48 * N, st, "upload boxes") ||
        !bc_h2d(c, pcen.p, prim_centroid, 24 * N, st, "upload centroids") ||
        !bc_check(c, hipMemsetAsync(c->d_err, 0, sizeof(int), st), "clear error"))
        return BC_ERR_DEVICE;
    hipLaunchKernelGGL(k_iota, dim3(grid_of(N, kThreads)), dim3(kThreads), 0, st, ord.as<int>(), n);
    hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, st, keys.as<long long>());
    hipLaunchKernelGGL(k_root_bounds, dim3(std::min<unsigned>(grid_of(N, kThreads), 2048u)), dim3(kThreads), 0, st,
                       pbox.as<double>(), pcen.as<double>(), n, keys.as<long long>());
    hipLaunchKernelGGL(k_make_root, dim3(1), dim3(1), 0, st, nodes.as<BNode>(), keys.as<long long>(), n, leaf, lst);
    if (!bc_check(c, hipGetLastError(), "build launch")) return BC_ERR_DEVICE;
    BNode root;
    if (!bc_check(c, hipMemcpyAsync(&root, nodes.p, sizeof root, hipMemcpyDeviceToHost, st), "read root") ||
        !bc_check(c, hipStreamSynchronize(st), "build root"))
        return BC_ERR_DEVICE;
    std::vector<std::vector<int>> levels;  // level-pass node lists, top down
    std::vector<int> cur;
    if (root.kind == kLevel) cur.push_back(0);
    size_t chunk_cap = N / kChunk + (size_t)tcap + 1;
    DevBuf gbins, dch, dfirst, dcr, dranges;
    size_t gbins_cap = 0;
    if (!dch.alloc(c, sizeof(Chunk) * chunk_cap, "chunks") || !dfirst.alloc(c, 4 * ((size_t)tcap + 1), "chunks") ||
        !dcr.alloc(c, 4 * chunk_cap, "chunk counts") || !dranges.alloc(c, 8 * (size_t)tcap, "level ranges"))
        return BC_ERR_NOMEM;
    int counts[3] = {0, 0, 1};  // next-level nodes, subtree roots, level-pass nodes
    // grows every buffer sized by the level-pass node capacity (nodes and subtree roots kept)
    auto grow = [&](int want) -> bool {
        const int cap = (int)std::min<size_t>(2 * N, std::max<size_t>(2 * (size_t)tcap, (size_t)want + 64));
        DevBuf n2, s2;
        if (!n2.alloc(c, sizeof(BNode) * (size_t)cap, "build nodes") || !s2.alloc(c, 4 * (size_t)cap, "subtree list")) return false;
        if (!bc_check(c, hipMemcpyAsync(n2.p, nodes.p, sizeof(BNode) * (size_t)counts[2], hipMemcpyDeviceToDevice, st), "grow") ||
            (counts[1] > 0 &&
             !bc_check(c, hipMemcpyAsync(s2.p, subs.p, 4 * (size_t)counts[1], hipMemcpyDeviceToDevice, st), "grow")) ||
            !bc_check(c, hipStreamSynchronize(st), "grow"))
            return false;
        std::swap(nodes.p, n2.p);
        std::swap(subs.p, s2.p);
        tcap = cap;
        chunk_cap = N / kChunk + (size_t)tcap + 1;
        level.reset(); next.reset(); dch.reset(); dfirst.reset(); dcr.reset(); dranges.reset();
        return level.alloc(c, 4 * (size_t)tcap, "level list") && next.alloc(c, 4 * (size_t)tcap, "level list") &&
               dch.alloc(c, sizeof(Chunk) * chunk_cap, "chunks") && dfirst.alloc(c, 4 * ((size_t)tcap + 1), "chunks") &&
               dcr.alloc(c, 4 * chunk_cap, "chunk counts") && dranges.alloc(c, 8 * (size_t)tcap, "level ranges");
    };
    std::vector<int2> ranges;
    std::vector<Chunk> ch;
    std::vector<int> first;
    while (!cur.empty()) {
        const int nl = (int)cur.size();
        if (counts[2] + 2 * nl > tcap && !grow(counts[2] + 2 * nl)) return BC_ERR_NOMEM;
        levels.push_back(cur);
        // chunk table of this level
        ranges.resize((size_t)nl);
        if (!bc_h2d(c, level.p, cur.data(), 4 * cur.size(), st, "upload level"))
            return BC_ERR_DEVICE;
        hipLaunchKernelGGL(k_ranges, dim3(grid_of((size_t)nl, 64)), dim3(64), 0, st, nodes.as<BNode>(), level.as<int>(), nl,
                           dranges.as<int2>());
        if (!bc_check(c, hipMemcpyAsync(ranges.data(), dranges.p, 8 * (size_t)nl, hipMemcpyDeviceToHost, st), "read level") ||
            !bc_check(c, hipStreamSynchronize(st), "read level"))
            return BC_ERR_DEVICE;
        ch.clear();
        first.assign(1, 0);
        for (int li = 0; li < nl; li++) {
            const int a = ranges[(size_t)li].x, b = a + ranges[(size_t)li].y;
            for (int p = a; p < b; p += kChunk) ch.push_back({li, p, std::min(p + kChunk, b)});
            first.push_back((int)ch.size());
        }
        if (ch.size() > chunk_cap) { c->err = "bc_build_bvh: chunk table overflow"; return BC_ERR_DEVICE; }
        const size_t need = (size_t)nl * kMaxBins * kBinWords * 8;
        if (need > gbins_cap) {
            gbins.reset();
            if (!gbins.alloc(c, need, "bins")) return BC_ERR_NOMEM;
            gbins_cap = need;
        }
        if (!bc_h2d(c, dch.p, ch.data(), sizeof(Chunk) * ch.size(), st, "upload chunks") ||
            !bc_h2d(c, dfirst.p, first.data(), 4 * first.size(), st, "upload chunks") ||
            !bc_check(c, hipMemsetAsync(lst, 0, sizeof(int), st), "clear level count"))
            return BC_ERR_DEVICE;
        const unsigned nch = (unsigned)ch.size();
        const int words = nl * kMaxBins * kBinWords;
        hipLaunchKernelGGL(k_init_keys, dim3(grid_of((size_t)words, kThreads)), dim3(kThreads), 0, st, gbins.as<long long>(), words);
        hipLaunchKernelGGL(k_bin, dim3(nch), dim3(kThreads), 0, st, nodes.as<BNode>(), level.as<int>(), dch.as<Chunk>(),
                           ord.as<int>(), pbox.as<double>(), pcen.as<double>(), binid.as<uint8_t>(), gbins.as<long long>(), bins);
        hipLaunchKernelGGL(k_sah, dim3(grid_of((size_t)nl, 64)), dim3(64), 0, st, nodes.as<BNode>(), level.as<int>(), nl,
                           gbins.as<long long>(), bins, leaf, next.as<int>(), subs.as<int>(), lst, tcap, c->d_err);
        hipLaunchKernelGGL(k_count_r, dim3(nch), dim3(kThreads), 0, st, nodes.as<BNode>(), level.as<int>(), dch.as<Chunk>(),
                           binid.as<uint8_t>(), dcr.as<int>());
        hipLaunchKernelGGL(k_chunk_scan, dim3(grid_of((size_t)nl, 64)), dim3(64), 0, st, nodes.as<BNode>(), level.as<int>(), nl,
                           dfirst.as<int>(), dcr.as<int>(), c->d_err);
        hipLaunchKernelGGL(k_holes, dim3(nch), dim3(kThreads), 0, st, nodes.as<BNode>(), level.as<int>(), dch.as<Chunk>(),
                           binid.as<uint8_t>(), dcr.as<int>(), rr.as<int>(), hf.as<int>(), c->d_err);
        hipLaunchKernelGGL(k_move, dim3(nch), dim3(kThreads), 0, st, nodes.as<BNode>(), level.as<int>(), dch.as<Chunk>(),
                           binid.as<uint8_t>(), rr.as<int>(), hf.as<int>(), ord.as<int>(), ord2.as<int>(), c->d_err);
        hipLaunchKernelGGL(k_copy_back, dim3(nch), dim3(kThreads), 0, st, nodes.as<BNode>(), level.as<int>(), dch.as<Chunk>(),
                           ord2.as<int>(), ord.as<int>());
        if (!bc_check(c, hipGetLastError(), "build level launch")) return BC_ERR_DEVICE;
        if (!bc_check(c, hipMemcpyAsync(counts, lst, sizeof counts, hipMemcpyDeviceToHost, st), "read level counts") ||
            !bc_check(c, hipStreamSynchronize(st), "build level"))
            return BC_ERR_DEVICE;
        const int nnext = counts[0];
        const int kerr = bc_check_kernel_error(c);
        if (kerr != BC_OK) { c->err = "bc_build_bvh: internal capacity or partition check failed"; return kerr; }
        cur.assign((size_t)nnext, 0);
        if (nnext > 0 &&
            (!bc_check(c, hipMemcpyAsync(cur.data(), next.p, 4 * (size_t)nnext, hipMemcpyDeviceToHost, st), "read next level") ||
             !bc_check(c, hipStreamSynchronize(st), "read next level")))
            return BC_ERR_DEVICE;
    }
    // subtree roots: listed by the level passes (the counter accumulates), or the root itself
    int nsub = 0, ntemp = 1;
    {
        if (!bc_check(c, hipMemcpyAsync(counts, lst, sizeof counts, hipMemcpyDeviceToHost, st), "read counts") ||
            !bc_check(c, hipStreamSynchronize(st), "read counts"))
            return BC_ERR_DEVICE;
        ntemp = counts[2];
        nsub = counts[1];
        if (root.kind == kSubtree) {
            const int zero = 0;
            if (!bc_check(c, hipMemcpyAsync(subs.p, &zero, 4, hipMemcpyHostToDevice, st), "subtree list")) return BC_ERR_DEVICE;
            nsub = 1;
        }
    }
    int* store_count = lst + 3;
    if (!bc_check(c, hipMemsetAsync(store_count, 0, sizeof(int), st), "clear store count")) return BC_ERR_DEVICE;
    if (nsub > 0)
        hipLaunchKernelGGL(k_subtree, dim3((unsigned)nsub), dim3(64), 0, st, nodes.as<BNode>(), subs.as<int>(), nsub, ord.as<int>(),
                           pbox.as<double>(), pcen.as<double>(), store.as<SNode>(), store_count, scap, bins, leaf, c->d_err);
    // sizes bottom-up, pre-order indices top-down (one launch per level)
    for (size_t l = levels.size(); l-- > 0;) {
        const int nl = (int)levels[l].size();
        if (!bc_h2d(c, level.p, levels[l].data(), 4 * (size_t)nl, st, "upload level"))
            return BC_ERR_DEVICE;
        hipLaunchKernelGGL(k_sizes, dim3(grid_of((size_t)nl, 64)), dim3(64), 0, st, nodes.as<BNode>(), level.as<int>(), nl);
    }
    for (size_t l = 0; l < levels.size(); l++) {
        const int nl = (int)levels[l].size();
        if (!bc_h2d(c, level.p, levels[l].data(), 4 * (size_t)nl, st, "upload level"))
            return BC_ERR_DEVICE;
        hipLaunchKernelGGL(k_ids, dim3(grid_of((size_t)nl, 64)), dim3(64), 0, st, nodes.as<BNode>(), level.as<int>(), nl);
    }
    const int out_cap = (int)(2 * N);
    hipLaunchKernelGGL(k_emit_nodes, dim3(grid_of((size_t)ntemp, 64)), dim3(64), 0, st, nodes.as<BNode>(), ntemp,
                       obox.as<double>(), olink.as<int32_t>(), out_cap, c->d_err);
    if (nsub > 0)
        hipLaunchKernelGGL(k_emit_subtrees, dim3((unsigned)nsub), dim3(64), 0, st, nodes.as<BNode>(), subs.as<int>(), nsub,
                           store.as<SNode>(), obox.as<double>(), olink.as<int32_t>(), out_cap, c->d_err);
    if (!bc_check(c, hipGetLastError(), "build emit launch")) return BC_ERR_DEVICE;
    if (!bc_check(c, hipMemcpyAsync(&root, nodes.p, sizeof root, hipMemcpyDeviceToHost, st), "read root") ||
        !bc_check(c, hipStreamSynchronize(st), "build"))
        return BC_ERR_DEVICE;
    const int kerr = bc_check_kernel_error(c);
    if (kerr != BC_OK) { c->err = "bc_build_bvh: internal capacity or partition check failed"; return kerr; }
    const int total = root.size;
    *nnode_out = total;
    if (total > node_cap) { c->err = "bc_build_bvh: node_cap too small (" + std::to_string(total) + " nodes)"; return BC_ERR_ARG; }
    if (!bc_d2h(c, node_box, obox.p, 48 * (size_t)total, st, "read node boxes") ||
        !bc_d2h(c, node_link, olink.p, 16 * (size_t)total, st, "read node links") ||
        !bc_d2h(c, order, ord.p, 4 * N, st, "read order"))
        return BC_ERR_DEVICE;
    return BC_OK;
}
