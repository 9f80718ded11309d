// api.cpp -- the C ABI of beamcast_hip.h: context, scene residency, ray queries, render.
// Mirrors the OpenCL host runtime it replaces (gpu_opencl_windows.go:94-723): a persistent
// context with a stream, scene buffers cached by revision, errors returned as status codes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/beamcast_hip.h"
#include "api_internal.hpp"
#include "bvh8_build.hpp"
#include "bvh8_flatten.hpp"
#include "kernels.hpp"

// Scenes of at least this many primitives are flattened on the device by default.
constexpr int kDeviceFlattenMin = 1 << 16;
// Ray batches of at least this many rays may be binned before a query (bc_set_trace_order).
constexpr int64_t kOrderMinRays = 1 << 16;
constexpr int kAutoRefillLanes = 8;  // automatic closest-hit refill threshold (measured best: 8-12)

namespace {
thread_local std::string g_create_error;

// renderOpenCL's deviceName: the marketing name, or when the runtime reports none (some drivers
// leave hipDeviceProp_t::name empty) one made from the architecture and CU count.
std::string device_label(const hipDeviceProp_t& p) {
    const std::string arch = p.gcnArchName;
    if (p.name[0] != '\0') return std::string(p.name) + " (" + arch + ")";
    return "AMD GPU " + arch.substr(0, arch.find(':')) + " (" + std::to_string(p.multiProcessorCount) + " CUs, " +
           arch + ")";
}

bool ok(bc_ctx* c, hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    if (c) c->err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}
void free_dev(void*& p) {
    if (p) hipFree(p);
    p = nullptr;
}
template <class T>
bool upload_vec(bc_ctx* c, T*& dst, size_t& cap_bytes, const std::vector<T>& v, const char* what) {
    size_t bytes = std::max<size_t>(v.size() * sizeof(T), 16);
    if (bytes > cap_bytes) {
        void* p = dst;
        free_dev(p);
        dst = nullptr;
        cap_bytes = 0;
        if (!ok(c, hipMalloc(&p, bytes), what)) return false;
        dst = (T*)p;
        cap_bytes = bytes;
    }
    if (!v.empty() && !bc_h2d(c, dst, v.data(), v.size() * sizeof(T), c->stream, what))
        return false;
    return true;
}
// Emissive primitives of a scene view as NEE lights: geometry and fp64 area (lightWeight,
// renderer.go:1343-1357).  pbase / mbase offset the primitive and material indices (packed
// scenes).
void collect_lights(const bc_scene_view& v, int pbase, int mbase, std::vector<bc_light>& out,
                    std::vector<double>& area_out) {
    for (int i = 0; i < v.nprim; i++) {
        const int m = v.prim_mat[i];
        if (v.mat_kind[m] != 3) continue;
        const double* g = v.prim_geom + 10 * (size_t)i;
        bc_light L;
        std::memset(&L, 0, sizeof L);
        L.prim = pbase + i;
        double area;
        if (v.prim_kind[i] == 0) {
            L.kind = 0;
            area = 4 * M_PI * g[3] * g[3];
            L.g0[0] = (float)g[0]; L.g0[1] = (float)g[1]; L.g0[2] = (float)g[2]; L.g0[3] = (float)g[3];
        } else {
            if (!v.nee_triangles) continue;
            L.kind = 1;
            const double cx = g[4] * g[8] - g[5] * g[7], cy = g[5] * g[6] - g[3] * g[8], cz = g[3] * g[7] - g[4] * g[6];
            area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
            for (int k = 0; k < 3; k++) {
                L.g0[k] = (float)g[k]; L.g1[k] = (float)g[3 + k]; L.g2[k] = (float)g[6 + k];
            }
        }
        L.area = (float)area;
        uint32_t mbits = (uint32_t)(mbase + m);
        std::memcpy(&L.g1[3], &mbits, 4);  // the light's material index
        out.push_back(L);
        area_out.push_back(area);
    }
}

bool check_view(bc_ctx* c, const bc_scene_view& v) {
    if (v.nmat <= 0) { c->err = "HIP backend requires at least one material"; return false; }
    for (int i = 0; i < v.nprim; i++)
        if (v.prim_mat[i] < 0 || v.prim_mat[i] >= v.nmat) { c->err = "primitive material out of range"; return false; }
    return true;
}
}  // namespace

bool bc_ensure(bc_ctx* c, void*& p, size_t& cap, size_t bytes, const char* what) {
    if (bytes <= cap && p) return true;
    free_dev(p);
    cap = 0;
    if (!ok(c, hipMalloc(&p, bytes), what)) return false;
    cap = bytes;
    return true;
}

bool bc_check(bc_ctx* c, hipError_t e, const char* what) { return ok(c, e, what); }

extern "C" {

int bc_device_available(char* name, int cap) {
    int n = 0;
    std::string msg;
    int found = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        msg = "no HIP device";
    } else {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, 0) == hipSuccess && std::strstr(p.gcnArchName, "gfx950")) {
            msg = device_label(p);
            found = 1;
        } else {
            msg = "device 0 is not gfx950";
        }
    }
    if (name && cap > 0) {
        std::snprintf(name, (size_t)cap, "%s", msg.c_str());
    }
    return found;
}

bc_ctx* bc_create(int device_ordinal) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device_ordinal < 0 || device_ordinal >= n) {
        g_create_error = "no HIP device with ordinal " + std::to_string(device_ordinal);
        return nullptr;
    }
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device_ordinal) != hipSuccess || !std::strstr(p.gcnArchName, "gfx950")) {
        g_create_error = "device is not gfx950 (MI355X); this library is built for gfx950 only";
        return nullptr;
    }
    if (hipSetDevice(device_ordinal) != hipSuccess) {
        g_create_error = "hipSetDevice failed";
        return nullptr;
    }
    bc_ctx* c = new bc_ctx();
    c->device = device_ordinal;
    c->name = device_label(p);
    c->num_cus = p.multiProcessorCount;
    if (!ok(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate") ||
        !ok(c, hipEventCreate(&c->ev0), "hipEventCreate") || !ok(c, hipEventCreate(&c->ev1), "hipEventCreate") ||
        !ok(c, hipMalloc((void**)&c->d_err, 64), "hipMalloc(err)") ||
        !ok(c, hipMemset(c->d_err, 0, 64), "hipMemset(err)")) {
        g_create_error = c->err;
        bc_destroy(c);
        return nullptr;
    }
    bc_scratch_init(c);
    return c;
}

void bc_destroy(bc_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    void* ptrs[] = {c->d_nodes, c->d_prims, c->d_ids, c->d_pmat, c->d_mats, c->d_lights, c->d_pdf, c->d_light_of, c->d_err,
                    c->tmp_in, c->tmp_out, c->out_buf, c->ahead_buf, c->d_levels, c->d_nbox, c->d_stage, c->d_order, c->d_sorted,
                    c->d_sel};
    for (void* q : ptrs)
        if (q) hipFree(q);
    bc_render_release(c);
    bc_post_release(c);
    bc_pinned_release(c);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

const char* bc_last_error(const bc_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }
const char* bc_device_name(const bc_ctx* c) { return c ? c->name.c_str() : ""; }

int bc_scene_upload(bc_ctx* c, const bc_scene_view* v, uint64_t topo, uint64_t geom) {
    if (!c || !v) return BC_ERR_ARG;
    if (!check_view(c, *v)) return BC_ERR_SCENE;
    hipSetDevice(c->device);
    std::string err;
    if (!bcx::validate_view(*v, err)) { c->err = err; return BC_ERR_SCENE; }
    const bool on_device = c->flatten == 2 || (c->flatten == 0 && v->nprim >= kDeviceFlattenMin);
    bcx::PreparedView pv;
    const bc_scene_view& w = bcx::prepare_view(*v, pv);  // binary leaves of at most eight primitives
    bcx::Bvh8 b;
    if (!on_device && !bcx::build_bvh8(w, b, err)) { c->err = err; return BC_ERR_SCENE; }
    c->scene_epoch++;
    c->has_scene = false;
    c->batch.clear();
    bcd::FlattenResult fr;
    if (on_device) {
        const int rc = bcd::flatten_device(c, w, fr);
        if (rc != BC_OK) return rc;
    } else {
        fr.nnode8 = (int64_t)b.nodes.size();
        fr.depth = b.depth;
    }
    const int stack = bcd::stack_for_depth(fr.depth);
    if (stack < 0) { c->err = "BVH8 depth " + std::to_string(fr.depth) + " exceeds the traversal stack (33 levels)"; return BC_ERR_SCENE; }
    c->h_lights.clear();
    c->h_light_area.clear();
    collect_lights(*v, 0, 0, c->h_lights, c->h_light_area);
    c->light_groups.assign(1, std::make_pair(0, (int)c->h_lights.size()));
    c->nprim = v->nprim;
    c->nmat_scene = 0;
    {
        const int rc = bc_update_materials(c, v->nmat, v->mat_kind, v->mat_data);
        if (rc != BC_OK) return rc;
    }
    if (on_device) {
        c->level_off = std::move(fr.level_off);
        c->h_ids = std::move(fr.ids);
    } else {
        if (!upload_vec(c, c->d_nodes, c->cap_nodes, b.nodes, "upload nodes") ||
            !upload_vec(c, c->d_prims, c->cap_prims, b.prims, "upload prims") ||
            !upload_vec(c, c->d_ids, c->cap_ids, b.ids, "upload prim ids") ||
            !upload_vec(c, c->d_pmat, c->cap_pmat, b.mats, "upload prim materials"))
            return BC_ERR_DEVICE;
        // node lists by level for device refits (deepest level first)
        std::vector<std::vector<uint32_t>> lv((size_t)b.depth + 1);
        for (size_t i = 0; i < b.node_depth.size(); i++) lv[(size_t)b.node_depth[i]].push_back((uint32_t)i);
        std::vector<uint32_t> flat;
        c->level_off.assign(1, 0u);
        for (int d = b.depth; d >= 1; d--) {
            flat.insert(flat.end(), lv[(size_t)d].begin(), lv[(size_t)d].end());
            c->level_off.push_back((uint32_t)flat.size());
        }
        if (!upload_vec(c, c->d_levels, c->cap_levels, flat, "upload node levels")) return BC_ERR_DEVICE;
        c->h_ids = std::move(b.ids);
    }
    if (!ok(c, hipStreamSynchronize(c->stream), "scene upload")) return BC_ERR_DEVICE;
    c->nnode8 = fr.nnode8;
    c->nrec = (int64_t)BC_REC_PER_NODE * fr.nnode8;
    c->depth8 = fr.depth;
    c->stack = stack;
    c->topo = topo;
    c->geom = geom;
    c->scene_bytes = (int64_t)((size_t)fr.nnode8 * sizeof(bc_node8) + (size_t)c->nrec * (sizeof(bc_prim) + 8));
    c->has_scene = true;
    return BC_OK;
}

int bc_set_trace_refill(bc_ctx* c, int closest_lanes, int any_lanes) {
    if (!c || closest_lanes < -1 || closest_lanes > 64 || any_lanes < 0 || any_lanes > 64) return BC_ERR_ARG;
    c->trace_refill[0] = closest_lanes;
    c->trace_refill[1] = any_lanes;
    return BC_OK;
}

int bc_set_trace_order(bc_ctx* c, int mode, int cell_bits) {
    if (!c || mode < 0 || mode > 2 || cell_bits < bcd::kOrderMinBits || cell_bits > bcd::kOrderMaxBits) return BC_ERR_ARG;
    c->order_mode = mode;
    c->order_bits = cell_bits;
    return BC_OK;
}

int bc_trace_order_last(bc_ctx* c, int* out2) {
    if (!c || !out2) return BC_ERR_ARG;
    out2[0] = 0;
    out2[1] = -1;
    if (!c->last_order_flag) return BC_OK;
    if (!ok(c, hipMemcpyAsync(out2, c->last_order_flag, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream), "read order flag") ||
        !ok(c, hipStreamSynchronize(c->stream), "read order flag"))
        return BC_ERR_DEVICE;
    return BC_OK;
}

int bc_set_render_ahead(bc_ctx* c, int on) {
    if (!c || on < 0 || on > 1) return BC_ERR_ARG;
    c->ahead_on = on != 0;
    c->ahead_valid = false;
    return BC_OK;
}

int bc_set_flatten(bc_ctx* c, int mode) {
    if (!c || mode < 0 || mode > 2) return BC_ERR_ARG;
    c->flatten = mode;
    return BC_OK;
}

// Material table and light distribution (rebuildLightDistribution, renderer.go:363-407) from
// material data; also the cheap path for parameter updates of the differentiable render.
int bc_update_materials(bc_ctx* c, int32_t nmat, const int32_t* mat_kind, const double* mat_data) {
    if (!c || nmat <= 0 || !mat_kind || !mat_data) return BC_ERR_ARG;
    if (!c->batch.empty()) { c->err = "material updates apply to a single resident scene, not a batch"; return BC_ERR_ARG; }
    if (c->nmat_scene && nmat != c->nmat_scene) { c->err = "material count differs from the uploaded scene"; return BC_ERR_ARG; }
    return bc_set_materials(c, nmat, mat_kind, mat_data);
}

}  // extern "C"

int bc_set_materials(bc_ctx* c, int32_t nmat, const int32_t* mat_kind, const double* mat_data) {
    c->scene_epoch++;
    std::vector<bc_mat> mats((size_t)nmat);
    for (int i = 0; i < nmat; i++) {
        const double* d = mat_data + 8 * (size_t)i;
        bc_mat& m = mats[(size_t)i];
        std::memset(&m, 0, sizeof m);
        m.color[0] = (float)d[0]; m.color[1] = (float)d[1]; m.color[2] = (float)d[2]; m.rough = (float)d[3];
        m.ior = (float)d[4];
        m.emission[0] = (float)d[5]; m.emission[1] = (float)d[6]; m.emission[2] = (float)d[7];
        m.kind = mat_kind[i];
    }
    std::vector<bc_light> lights = c->h_lights;
    const size_t nl = lights.size();
    std::vector<float> pdf((size_t)c->nprim, 0.0f);
    std::vector<int32_t> light_of((size_t)c->nprim, -1);
    std::vector<double> cdf(nl);
    for (const auto& grp : c->light_groups) {  // one normalised distribution per scene
        double total = 0;
        for (int i = grp.first; i < grp.second; i++) {
            uint32_t mbits;
            std::memcpy(&mbits, &lights[(size_t)i].g1[3], 4);
            if ((int)mbits >= nmat || mat_kind[mbits] != 3) { c->err = "a light's material is no longer emissive"; return BC_ERR_ARG; }
            const double* md = mat_data + 8 * (size_t)mbits;
            lights[(size_t)i].le[0] = (float)md[5]; lights[(size_t)i].le[1] = (float)md[6]; lights[(size_t)i].le[2] = (float)md[7];
            double wt = (0.2126 * md[5] + 0.7152 * md[6] + 0.0722 * md[7]) * c->h_light_area[(size_t)i];
            total += wt < 1e-9 ? 1e-9 : wt;
            cdf[(size_t)i] = total;
        }
        double prev = 0;
        for (int i = grp.first; i < grp.second; i++) {
            const double cv = cdf[(size_t)i] / total;
            pdf[(size_t)lights[(size_t)i].prim] = (float)(cv - prev);
            prev = cv;
            cdf[(size_t)i] = cv;
        }
        if (grp.second > grp.first) cdf[(size_t)grp.second - 1] = 1;
    }
    for (size_t i = 0; i < nl; i++) {
        lights[i].cdf = (float)cdf[i];
        light_of[(size_t)lights[i].prim] = (int32_t)i;
    }
    (void)hipSetDevice(c->device);
    if (!upload_vec(c, c->d_mats, c->cap_mats, mats, "upload materials") ||
        !upload_vec(c, c->d_lights, c->cap_lights, lights, "upload lights") ||
        !upload_vec(c, c->d_pdf, c->cap_pdf, pdf, "upload light pdf") ||
        !upload_vec(c, c->d_light_of, c->cap_light_of, light_of, "upload light map") ||
        !ok(c, hipStreamSynchronize(c->stream), "material upload"))
        return BC_ERR_DEVICE;
    c->nmat = nmat;
    c->nmat_scene = nmat;
    c->nlight = (int)nl;
    return BC_OK;
}

extern "C" {

int bc_scene_refit(bc_ctx* c, const bc_scene_view* v, uint64_t geom) {
    if (!c || !v) return BC_ERR_ARG;
    if (!c->has_scene || !c->batch.empty()) { c->err = "bc_scene_refit needs a resident single scene"; return BC_ERR_SCENE; }
    c->scene_epoch++;
    if (v->nprim != c->nprim || (int64_t)c->h_ids.size() != c->nrec) {
        c->err = "bc_scene_refit: the primitive count differs from the resident scene (topology changed)";
        return BC_ERR_ARG;
    }
    if (!check_view(c, *v)) return BC_ERR_SCENE;
    for (size_t r = 0; r < c->h_ids.size(); r++) {
        if (c->h_ids[r] == BC_REC_HOLE) continue;
        const int id = (int)(c->h_ids[r] & ~BC_SPHERE_BIT);
        if ((v->prim_kind[id] == 0) != ((c->h_ids[r] & BC_SPHERE_BIT) != 0)) {
            c->err = "bc_scene_refit: a primitive changed kind (topology changed)";
            return BC_ERR_ARG;
        }
    }
    (void)hipSetDevice(c->device);
    // the view's geometry, kinds and materials to the device; the moved records (and their
    // materials) are rewritten there in leaf order
    const size_t n = (size_t)v->nprim;
    void* stg = c->d_stage;
    if (!bc_ensure(c, stg, c->cap_stage, 88 * n, "refit staging")) return BC_ERR_NOMEM;
    c->d_stage = stg;
    double* g = (double*)stg;
    int32_t* kd = (int32_t*)(g + 10 * n);
    int32_t* mt = kd + n;
    if (!bc_h2d(c, g, v->prim_geom, 80 * n, c->stream, "refit upload") ||
        !bc_h2d(c, kd, v->prim_kind, 4 * n, c->stream, "refit upload") ||
        !bc_h2d(c, mt, v->prim_mat, 4 * n, c->stream, "refit upload") ||
        !ok(c, bcd::launch_refit_records(g, kd, mt, c->d_ids, (uint32_t)c->nrec, c->d_prims, c->d_pmat, c->d_err, c->stream),
            "refit records"))
        return BC_ERR_DEVICE;
    void* nb = c->d_nbox;
    if (!bc_ensure(c, nb, c->cap_nbox, (size_t)c->nnode8 * 32, "refit boxes")) return BC_ERR_NOMEM;
    c->d_nbox = (float4*)nb;
    for (size_t k = 0; k + 1 < c->level_off.size(); k++) {
        const uint32_t a = c->level_off[k], n = c->level_off[k + 1] - a;
        if (!ok(c, bcd::launch_refit_level(c->d_nodes, c->d_prims, c->d_ids, c->d_nbox, c->d_levels + a, n, c->d_err,
                                           c->stream), "refit launch"))
            return BC_ERR_DEVICE;
    }
    // light geometry and its distribution (areas change with the primitives)
    c->h_lights.clear();
    c->h_light_area.clear();
    collect_lights(*v, 0, 0, c->h_lights, c->h_light_area);
    c->light_groups.assign(1, std::make_pair(0, (int)c->h_lights.size()));
    const int rc = bc_set_materials(c, v->nmat, v->mat_kind, v->mat_data);
    if (rc != BC_OK) return rc;
    if (!ok(c, hipStreamSynchronize(c->stream), "refit")) return BC_ERR_DEVICE;
    const int krc = bc_check_kernel_error(c);
    if (krc != BC_OK) return krc;
    c->geom = geom;
    return BC_OK;
}

int bc_batch_upload(bc_ctx* c, int n, const bc_scene_view* views) {
    if (!c || n <= 0 || !views) return BC_ERR_ARG;
    for (int s = 0; s < n; s++)
        if (!check_view(c, views[s])) { c->err = "batch scene " + std::to_string(s) + ": " + c->err; return BC_ERR_SCENE; }
    c->scene_epoch++;
    hipSetDevice(c->device);
    // the scenes' BVH8 builds are independent: spread them over host threads
    std::vector<bcx::Bvh8> bv((size_t)n);
    std::vector<std::string> errs((size_t)n);
    std::vector<char> okb((size_t)n, 0);
    std::atomic<int> next{0};
    const unsigned nth = std::max(1u, std::min<unsigned>((unsigned)n, std::min(16u, std::thread::hardware_concurrency())));
    auto work = [&]() {
        for (int s; (s = next.fetch_add(1)) < n;) okb[(size_t)s] = bcx::build_bvh8(views[s], bv[(size_t)s], errs[(size_t)s]);
    };
    std::vector<std::thread> pool;
    try {  // a thread that cannot start leaves its scenes to the others (the queue is shared)
        pool.reserve(nth);
        for (unsigned t = 1; t < nth; t++) pool.emplace_back(work);
    } catch (...) {
    }
    work();
    for (auto& t : pool) t.join();
    int depth = 0;
    for (int s = 0; s < n; s++) {
        if (!okb[(size_t)s]) { c->err = "batch scene " + std::to_string(s) + ": " + errs[(size_t)s]; return BC_ERR_SCENE; }
        depth = std::max(depth, bv[(size_t)s].depth);
    }
    const int stack = bcd::stack_for_depth(depth);
    if (stack < 0) { c->err = "BVH8 depth " + std::to_string(depth) + " exceeds the traversal stack (33 levels)"; return BC_ERR_SCENE; }
    // concatenate: node, record, primitive-id and material indices rebased per scene
    bcx::Bvh8 all;
    std::vector<int32_t> mkind;
    std::vector<double> mdata;
    std::vector<bc_ctx::BatchScene> refs;
    c->has_scene = false;
    c->h_lights.clear();
    c->h_light_area.clear();
    c->light_groups.clear();
    uint32_t pbase = 0;
    for (int s = 0; s < n; s++) {
        const bcx::Bvh8& b = bv[(size_t)s];
        const bc_scene_view& v = views[s];
        const uint32_t nbase = (uint32_t)all.nodes.size();
        const int mbase = (int)mkind.size();
        if ((uint64_t)nbase + b.nodes.size() > (uint64_t)BC_NODE_MAX || (uint64_t)pbase + (uint64_t)v.nprim >= 0x80000000ull) {
            c->err = "batch exceeds 2^24 BVH8 nodes or 2^31 primitives";
            return BC_ERR_SCENE;
        }
        // node links rebased; the record slots follow the nodes (8 per node), so they stay aligned
        for (bc_node8 nd : b.nodes) {
            bc_node_set_link(nd, bc_node_child_base(nd) + nbase, bc_node_imask(nd));
            all.nodes.push_back(nd);
        }
        all.prims.insert(all.prims.end(), b.prims.begin(), b.prims.end());
        for (uint32_t id : b.ids)
            all.ids.push_back(id == BC_REC_HOLE ? id : (((id & ~BC_SPHERE_BIT) + pbase) | (id & BC_SPHERE_BIT)));
        for (int32_t m : b.mats) all.mats.push_back(m < 0 ? m : m + mbase);
        mkind.insert(mkind.end(), v.mat_kind, v.mat_kind + v.nmat);
        mdata.insert(mdata.end(), v.mat_data, v.mat_data + 8 * (size_t)v.nmat);
        const int l0 = (int)c->h_lights.size();
        collect_lights(v, (int)pbase, mbase, c->h_lights, c->h_light_area);
        c->light_groups.emplace_back(l0, (int)c->h_lights.size());
        refs.push_back({nbase, (uint32_t)l0, (int32_t)c->h_lights.size() - l0, v.nprim, v.nmat, b.depth,
                        (int64_t)b.nodes.size()});
        pbase += (uint32_t)v.nprim;
    }
    c->nprim = (int)pbase;
    c->nmat_scene = 0;
    c->batch.clear();
    {
        const int rc = bc_set_materials(c, (int32_t)mkind.size(), mkind.data(), mdata.data());
        if (rc != BC_OK) return rc;
    }
    if (!upload_vec(c, c->d_nodes, c->cap_nodes, all.nodes, "upload nodes") ||
        !upload_vec(c, c->d_prims, c->cap_prims, all.prims, "upload prims") ||
        !upload_vec(c, c->d_ids, c->cap_ids, all.ids, "upload prim ids") ||
        !upload_vec(c, c->d_pmat, c->cap_pmat, all.mats, "upload prim materials") ||
        !ok(c, hipStreamSynchronize(c->stream), "batch upload"))
        return BC_ERR_DEVICE;
    c->nnode8 = (int64_t)all.nodes.size();
    c->nrec = (int64_t)all.prims.size();
    c->depth8 = depth;
    c->stack = stack;
    c->topo = c->geom = 0;
    c->scene_bytes = (int64_t)(all.nodes.size() * sizeof(bc_node8) + all.prims.size() * (sizeof(bc_prim) + 8));
    c->batch = refs;
    return BC_OK;
}

int bc_scene_stats(const bc_ctx* c, int64_t* o) {
    if (!c || !o) return BC_ERR_ARG;
    o[0] = c->nnode8; o[1] = c->depth8; o[2] = c->nprim; o[3] = c->scene_bytes;
    return BC_OK;
}

int bc_scene_arrays(bc_ctx* c, void* nodes, void* prims, uint32_t* ids) {
    if (!c) return BC_ERR_ARG;
    if (!c->has_scene && c->batch.empty()) { c->err = "no scene uploaded"; return BC_ERR_SCENE; }
    (void)hipSetDevice(c->device);
    if (nodes && !ok(c, hipMemcpy(nodes, c->d_nodes, (size_t)c->nnode8 * sizeof(bc_node8), hipMemcpyDeviceToHost), "read nodes"))
        return BC_ERR_DEVICE;
    if (prims && !ok(c, hipMemcpy(prims, c->d_prims, (size_t)c->nrec * sizeof(bc_prim), hipMemcpyDeviceToHost), "read records"))
        return BC_ERR_DEVICE;
    if (ids && !ok(c, hipMemcpy(ids, c->d_ids, (size_t)c->nrec * sizeof(uint32_t), hipMemcpyDeviceToHost), "read ids"))
        return BC_ERR_DEVICE;
    return BC_OK;
}

int bc_trace_device(bc_ctx* c, const void* d_rays, int64_t n, int any_hit, void* d_out) {
    if (!c || n < 0 || (n > 0 && (!d_rays || !d_out))) return BC_ERR_ARG;
    if (!c->has_scene) { c->err = "no scene uploaded"; return BC_ERR_SCENE; }
    const bc_ray* rays = (const bc_ray*)d_rays;
    const bc_ray* sorted = nullptr;
    const uint32_t* perm = nullptr;
    const int* oflag = nullptr;
    c->last_order_flag = nullptr;
    if (c->order_mode != 0 && n >= kOrderMinRays && n <= bcd::kOrderMaxRays) {
        // one contiguous chunk of at least 16,384 rays per workgroup, at most one workgroup per CU
        const int G = (int)std::min<int64_t>(std::min(c->num_cus, bcd::kOrderMaxGroups), (n + 16383) / 16384);
        const size_t nb = bcd::order_bins(c->order_bits);
        const size_t words = (size_t)G * nb + 2 * nb + (size_t)n + 2;
        if (!bc_ensure(c, c->d_order, c->cap_order, words * sizeof(uint32_t), "hipMalloc(ray order)") ||
            !bc_ensure(c, c->d_sorted, c->cap_sorted, (size_t)n * sizeof(bc_ray), "hipMalloc(sorted rays)"))
            return BC_ERR_NOMEM;
        uint32_t* w = (uint32_t*)c->d_order;
        bcd::OrderScratch s;
        s.cnt = w;
        s.total = s.cnt + (size_t)G * nb;
        s.bstart = s.total + nb;
        s.perm = s.bstart + nb;
        s.flag = (int*)(s.perm + n);
        s.sorted = (bc_ray*)c->d_sorted;
        if (!ok(c, bcd::launch_order(c->order_bits, c->order_mode, rays, n, c->d_nodes, G, s, c->stream), "ray order launch"))
            return BC_ERR_DEVICE;
        sorted = s.sorted;
        perm = s.perm;
        oflag = s.flag;
        c->last_order_flag = s.flag;
    }
    int refill = c->trace_refill[any_hit != 0 ? 1 : 0];
    const int* sel = nullptr;
    if (refill < 0) {
        // automatic (closest hit): lane refill for incoherent batches, one ray per thread for
        // coherent ones (where refilled lanes lose the wave's shared nodes), chosen on the device
        refill = kAutoRefillLanes;
        if (oflag == nullptr && n >= kOrderMinRays) {
            if (!bc_ensure(c, c->d_sel, c->cap_sel, 2 * sizeof(int), "hipMalloc(coherence flag)")) return BC_ERR_NOMEM;
            if (!ok(c, bcd::launch_coherence_probe(c->order_bits, rays, n, c->d_nodes, (int*)c->d_sel, c->stream),
                    "coherence probe"))
                return BC_ERR_DEVICE;
            sel = (const int*)c->d_sel;
        } else {
            refill = 0;
        }
    }
    hipError_t e = bcd::launch_trace(c->stack, any_hit != 0, c->d_nodes, c->d_prims, c->d_ids, rays, n, d_out, refill,
                                     c->d_err, c->stream, sorted, perm, oflag, sel);
    return ok(c, e, "trace launch") ? BC_OK : BC_ERR_DEVICE;
}

int bc_trace_counts(bc_ctx* c, const void* d_rays, int64_t n, int any_hit, void* d_out, uint32_t* d_counts) {
    if (!c || n < 0 || (n > 0 && (!d_rays || !d_out || !d_counts))) return BC_ERR_ARG;
    if (!c->has_scene) { c->err = "no scene uploaded"; return BC_ERR_SCENE; }
    hipSetDevice(c->device);
    hipError_t e = bcd::launch_trace_count(c->stack, any_hit != 0, c->d_nodes, c->d_prims, c->d_ids, (const bc_ray*)d_rays,
                                           n, d_out, c->d_err, (uint4*)d_counts, c->stream);
    return ok(c, e, "counting trace launch") ? BC_OK : BC_ERR_DEVICE;
}

int bc_valu_probe(bc_ctx* c, int mix, int blocks, int iters, float* ms) {
    if (!c || mix < 0 || mix > 1 || blocks <= 0 || iters <= 0 || !ms) return BC_ERR_ARG;
    hipSetDevice(c->device);
    if (!ok(c, hipEventRecord(c->ev0, c->stream), "probe") ||
        !ok(c, bcd::launch_valu_probe(mix, blocks, (uint32_t)iters, 0x5EEDu, (uint32_t*)c->d_err + 8, c->stream),
            "probe launch") ||
        !ok(c, hipEventRecord(c->ev1, c->stream), "probe") || !ok(c, hipEventSynchronize(c->ev1), "probe") ||
        !ok(c, hipEventElapsedTime(ms, c->ev0, c->ev1), "probe"))
        return BC_ERR_DEVICE;
    return BC_OK;
}

int bc_line_probe(bc_ctx* c, const void* d_buf, uint64_t bytes, int blocks, int iters, float* ms) {
    // a power-of-two number of 64-byte blocks, indexable by 32 bits
    if (!c || !d_buf || bytes < 64 || (bytes & (bytes - 1)) != 0 || bytes / 64 - 1 > 0xFFFFFFFFull ||
        (reinterpret_cast<uintptr_t>(d_buf) & 15) != 0 || blocks <= 0 || iters <= 0 || !ms)
        return BC_ERR_ARG;
    hipSetDevice(c->device);
    if (!ok(c, hipEventRecord(c->ev0, c->stream), "probe") ||
        !ok(c, bcd::launch_line_probe(static_cast<const uint4*>(d_buf), (uint32_t)(bytes / 64 - 1), blocks,
                                      (uint32_t)iters, 0x5EEDu, (uint32_t*)c->d_err + 8, c->stream),
            "probe launch") ||
        !ok(c, hipEventRecord(c->ev1, c->stream), "probe") || !ok(c, hipEventSynchronize(c->ev1), "probe") ||
        !ok(c, hipEventElapsedTime(ms, c->ev0, c->ev1), "probe"))
        return BC_ERR_DEVICE;
    return BC_OK;
}

int bc_trace(bc_ctx* c, const float* rays, int64_t n, int any_hit, void* out) {
    if (!c || n < 0 || (n > 0 && (!rays || !out))) return BC_ERR_ARG;
    if (n == 0) return BC_OK;
    hipSetDevice(c->device);
    const size_t in_b = (size_t)n * sizeof(bc_ray), out_b = (size_t)n * (any_hit ? 4 : sizeof(bc_hit));
    if (!bc_ensure(c, c->tmp_in, c->cap_in, in_b, "hipMalloc(rays)") ||
        !bc_ensure(c, c->tmp_out, c->cap_out, out_b, "hipMalloc(hits)"))
        return BC_ERR_NOMEM;
    if (!bc_h2d(c, c->tmp_in, rays, in_b, c->stream, "copy rays")) return BC_ERR_DEVICE;
    int rc = bc_trace_device(c, c->tmp_in, n, any_hit, c->tmp_out);
    if (rc != BC_OK) return rc;
    if (!bc_d2h(c, out, c->tmp_out, out_b, c->stream, "copy hits")) return BC_ERR_DEVICE;
    return bc_check_kernel_error(c);
}

void* bc_alloc(bc_ctx* c, uint64_t bytes) {
    if (!c) return nullptr;
    hipSetDevice(c->device);
    void* p = nullptr;
    if (!ok(c, hipMalloc(&p, std::max<uint64_t>(bytes, 16)), "bc_alloc")) return nullptr;
    return p;
}
int bc_free(bc_ctx* c, void* p) {
    if (!c) return BC_ERR_ARG;
    return ok(c, hipFree(p), "bc_free") ? BC_OK : BC_ERR_DEVICE;
}
int bc_upload(bc_ctx* c, void* d, const void* s, uint64_t bytes) {
    if (!c) return BC_ERR_ARG;
    if (!bc_h2d(c, d, s, bytes, c->stream, "bc_upload") ||
        !ok(c, hipStreamSynchronize(c->stream), "bc_upload"))
        return BC_ERR_DEVICE;
    return BC_OK;
}
int bc_download(bc_ctx* c, void* d, const void* s, uint64_t bytes) {
    if (!c) return BC_ERR_ARG;
    if (!bc_d2h(c, d, s, bytes, c->stream, "bc_download")) return BC_ERR_DEVICE;
    return bc_check_kernel_error(c);
}
int bc_copy_device(bc_ctx* c, void* d_dst, const void* d_src, uint64_t bytes) {
    if (!c) return BC_ERR_ARG;
    return ok(c, hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, c->stream), "bc_copy_device") ? BC_OK
                                                                                                          : BC_ERR_DEVICE;
}
int bc_synchronize(bc_ctx* c) {
    if (!c) return BC_ERR_ARG;
    if (!ok(c, hipStreamSynchronize(c->stream), "synchronize")) return BC_ERR_DEVICE;
    return bc_check_kernel_error(c);
}
int bc_timer_start(bc_ctx* c) {
    if (!c) return BC_ERR_ARG;
    return ok(c, hipEventRecord(c->ev0, c->stream), "timer start") ? BC_OK : BC_ERR_DEVICE;
}
int bc_timer_stop(bc_ctx* c, float* ms) {
    if (!c || !ms) return BC_ERR_ARG;
    if (!ok(c, hipEventRecord(c->ev1, c->stream), "timer stop") || !ok(c, hipEventSynchronize(c->ev1), "timer sync") ||
        !ok(c, hipEventElapsedTime(ms, c->ev0, c->ev1), "timer elapsed"))
        return BC_ERR_DEVICE;
    return BC_OK;
}

}  // extern "C"

int bc_check_kernel_error(bc_ctx* c) {
    int h = 0;
    if (!ok(c, hipMemcpy(&h, c->d_err, sizeof(int), hipMemcpyDeviceToHost), "read kernel error flag")) return BC_ERR_DEVICE;
    if (h) {
        c->err = "kernel reported an error (traversal stack overflow or queue overflow)";
        hipMemset(c->d_err, 0, sizeof(int));
        return BC_ERR_DEVICE;
    }
    return BC_OK;
}
