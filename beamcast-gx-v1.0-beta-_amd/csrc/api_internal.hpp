// api_internal.hpp -- the context behind bc_ctx (persistent runtime, scene residency).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/beamcast_hip.h"
#include "bc_layout.h"

struct RenderState;  // integrator buffers (integrator.hip)
struct PostState;    // post-process buffers (post.hip)

struct bc_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err, name;
    // resident scene
    bool has_scene = false;
    uint64_t topo = 0, geom = 0;
    bc_node8* d_nodes = nullptr;
    bc_prim* d_prims = nullptr;
    uint32_t* d_ids = nullptr;  // by record: primitive id | sphere bit
    int32_t* d_pmat = nullptr;  // by record: material
    bc_mat* d_mats = nullptr;
    bc_light* d_lights = nullptr;
    float* d_pdf = nullptr;          // light selection pdf by primitive
    int32_t* d_light_of = nullptr;   // light index by primitive, -1 if none
    size_t cap_ids = 0, cap_pmat = 0;
    size_t cap_nodes = 0, cap_prims = 0, cap_mats = 0, cap_lights = 0, cap_pdf = 0, cap_light_of = 0;
    int64_t nnode8 = 0, scene_bytes = 0;
    int depth8 = 0, nprim = 0, nmat = 0, nlight = 0, stack = 0, nmat_scene = 0;
    int64_t nrec = 0;  // record slots resident: 8 per BVH8 node
    std::vector<bc_light> h_lights;  // light geometry (emission/CDF filled per material update)
    std::vector<double> h_light_area;
    // light table groups normalised separately ([begin, end) into h_lights): one per scene
    std::vector<std::pair<int, int>> light_groups;
    // packed multi-scene batch (bc_batch_upload); empty when a single scene is resident
    struct BatchScene {
        uint32_t root, lbase;  // root node, first light
        int32_t nlight, nprim, nmat, depth;
        int64_t nnode;
    };
    std::vector<BatchScene> batch;
    // device refit of the resident single scene: record ids, node lists by level (deepest
    // first; level_off[k] .. level_off[k+1] is one level), per-node union boxes
    std::vector<uint32_t> h_ids;
    std::vector<uint32_t> level_off;
    uint32_t* d_levels = nullptr;
    float4* d_nbox = nullptr;
    size_t cap_levels = 0, cap_nbox = 0;
    void* d_stage = nullptr;  // refit staging: the view's geometry, kinds and materials
    size_t cap_stage = 0;
    int flatten = 0;       // BVH8 flattening: 0 auto (device for large scenes), 1 host, 2 device
    // ray queries (closest, any hit): refill a wave's lanes once this many are idle (0: one ray
    // per thread).  Any-hit rays end at very different depths, closest-hit rays much less so.
    int trace_refill[2] = {-1, 8};  // closest (-1: automatic), any hit
    void* d_sel = nullptr;          // coherence probe flag of automatic closest-hit refill
    size_t cap_sel = 0;
    // ray binning before queries (order.hip): mode 0 off, 1 always, 2 when a sample of the
    // batch is incoherent; bits = log2 of the origin cells per axis.  Off by default: on the 1M
    // soup it lifts the incoherent L2 hit rate from 0.47 to 0.95, yet with lane refill the
    // traversal plus its passes is slower than the traversal alone (incoherent closest hit 6.36
    // against 7.1 ms).
    int order_mode = 0, order_bits = 4;
    void* d_order = nullptr;  // bin counts, totals, starts, permutation, decision flag
    void* d_sorted = nullptr; // the sorted rays
    size_t cap_order = 0, cap_sorted = 0;
    int* last_order_flag = nullptr;  // decision of the last query, if it ran the binning passes
    // two pinned buffers for large host-to-device copies (upload.cpp), allocated on first use
    static constexpr size_t kPinChunk = size_t(16) << 20;
    void* h_pin[2] = {nullptr, nullptr};
    hipEvent_t pin_ev[2] = {nullptr, nullptr};
    bool pin_busy[2] = {false, false};
    int pin_next = 0;
    // the pinned buffers hold the rows rendered ahead (image in [0], guides in [1]); any other
    // staged copy clears it before it reuses them
    bool pin_ahead = false;
    // a copy of the next chunk's rows rendered ahead, started into pinned buffer pin_pre_buf (image,
    // then guides) when the previous chunk was served; pin_pre_y0 < 0: none.  Cleared as pin_ahead.
    int pin_pre_y0 = -1, pin_pre_rows = 0, pin_pre_buf = 0;
    bool pin_pre_guides = false;
    struct CopyPool* copy_pool = nullptr;  // host threads of the staged copies (upload.cpp)
    int* d_err = nullptr;  // [0] kernel error flag
    // staging for host-pointer ray queries
    void* tmp_in = nullptr;
    void* tmp_out = nullptr;
    size_t cap_in = 0, cap_out = 0;
    // staging for host-pointer renders
    void* out_buf = nullptr;
    size_t out_cap = 0;
    // bumped by every change of the resident scene or its materials (upload, refit, material
    // update, batch upload): rows that bc_render rendered ahead belong to one epoch
    uint64_t scene_epoch = 0;
    // rows rendered ahead by bc_render: rgba, then guides, of rows [ahead_y0, ahead_y0 +
    // ahead_rows) of ahead_desc (integrator.hip); rows from ahead_next on are not yet served
    bool ahead_on = true, ahead_valid = false, ahead_guides = false;
    int ahead_y0 = 0, ahead_rows = 0, ahead_next = 0;
    uint64_t ahead_epoch = 0;
    bc_render_desc ahead_desc{};
    void* ahead_buf = nullptr;
    size_t ahead_cap = 0;
    std::vector<uint64_t> ahead_row_segs;  // path segments of each row of the block
    // integrator
    RenderState* rs = nullptr;
    // post stack
    PostState* ps = nullptr;
    // device address of this device's fixed scratch array (bc_render_noalloc), resolved at
    // bc_create; null with the reason in scratch_err if that failed
    void* scratch = nullptr;
    std::string scratch_err;
};

bool bc_ensure(bc_ctx* c, void*& p, size_t& cap, size_t bytes, const char* what);
bool bc_check(bc_ctx* c, hipError_t e, const char* what);
// Host-to-device copy of a pageable buffer on stream st; large ones are staged through the
// context's pinned buffers by host threads, overlapping the DMA (upload.cpp).  The source may be
// reused as soon as it returns.
bool bc_h2d(bc_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t st, const char* what);
// Device-to-host copy on stream st into pageable memory, staged the same way; returns once dst
// holds the data.
bool bc_d2h(bc_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t st, const char* what);
bool bc_d2h_pair(bc_ctx* c, void* dst0, const void* src0, size_t n0, void* dst1, const void* src1, size_t n1,
                 hipStream_t st, const char* what);
// Copies src0 (n0 bytes) into pinned buffer 0 and src1 (n1 bytes) into pinned buffer 1 and keeps
// them there (sets pin_ahead): 1 when held, 0 when they do not fit or the buffers cannot be
// allocated (nothing copied, no error), -1 on a failed copy.
int bc_d2h_hold(bc_ctx* c, const void* src0, size_t n0, const void* src1, size_t n1, hipStream_t st,
                const char* what);
// Starts an asynchronous copy of src0 (n0 bytes), then src1 (n1 bytes), into one pinned buffer and
// records that buffer's event: the buffer index, -1 when they do not fit or the buffers cannot be
// allocated (nothing started, no error), -2 on a failed call.  Clears pin_ahead and any earlier
// prefetch; the caller records what the buffer holds.
int bc_d2h_prefetch(bc_ctx* c, const void* src0, size_t n0, const void* src1, size_t n1, hipStream_t st,
                    const char* what);
// Host-to-host copy by the context's copy threads (pageable destination, pinned source).
void bc_host_copy(bc_ctx* c, void* dst, const void* src, size_t n);
void bc_pinned_release(bc_ctx* c);
// Material table and per-group light distributions from the material data (api.cpp).
int bc_set_materials(bc_ctx* c, int32_t nmat, const int32_t* mat_kind, const double* mat_data);
int bc_check_kernel_error(bc_ctx* c);
void bc_render_release(bc_ctx* c);
// Resolves the fixed scratch array of bc_render_noalloc on c's device (integrator.hip).
void bc_scratch_init(bc_ctx* c);
void bc_post_release(bc_ctx* c);
