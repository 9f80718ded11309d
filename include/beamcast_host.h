/*
 * beamcast_host.h -- C entry points of the host-side scene mirror and workload generators.
 *
 * The Go renderer owns Scene / Camera / RenderSettings in the shipped product and hands
 * the C ABI of beamcast_hip.h a bc_scene_view.  This header exposes the C++ mirror of
 * those Go types (same names and semantics) so that callers without Go -- the tests,
 * the benchmark -- can build scenes exactly as Scene.Build would, and generate the
 * seeded synthetic workloads the benchmark and tests run on.
 */
#ifndef BEAMCAST_HOST_H
#define BEAMCAST_HOST_H
#include <stdint.h>

#include "beamcast_hip.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct bch_scene bch_scene;

/* Scene (renderer.go:225-241) */
bch_scene* bch_scene_new(void);
void bch_scene_free(bch_scene* s);
/* Scene.AddMaterial (renderer.go:243): data8 = Color.xyz, Roughness, IOR, Emission.xyz */
int bch_add_material(bch_scene* s, int kind, const double* data8);
/* Scene.AddSphere / AddTriangle (renderer.go:246-270): degenerate input is dropped */
void bch_add_sphere(bch_scene* s, const double* center, double radius, int mat);
void bch_add_triangle(bch_scene* s, const double* a, const double* b, const double* c, int mat);
/* Scene.Build with BVHMode SAH (renderer.go:409-449); bins/leaf clamped as there */
void bch_build(bch_scene* s, int bins, int leaf);
/* Scene.Build with BVHMode "Median" (renderer.go:409-449, buildNodeMedian :506-524); leaf clamped
 * to [1, 16] as there.  The order among equal centroids is not pinned (Go's sort.Slice is
 * unstable). */
void bch_build_median(bch_scene* s, int leaf);
/* Scene.Build with the tree made on the GPU (bc_build_bvh): the same Nodes and Order as
 * bch_build; returns the bc_build_bvh status (message via bc_last_error(ctx)) */
int bch_build_device(bch_scene* s, bc_ctx* ctx, int bins, int leaf);
/* The host BVH8 flattening of bc_scene_upload, run without a device: out4 = wide nodes, depth,
 * leaf slots, and 1 if every primitive lies in exactly one leaf slot with its record at
 * 8 * node + slot (holes elsewhere), every node is reached once, every grid origin is k 2^e with
 * |k| < 2^20, and every record's exact box lies inside the quantised boxes of its leaf slot and of
 * all internal slots above it (else 0).  Returns 0, or -1 if the view cannot be encoded.  No
 * counterpart in the reference (its GPU path uploads the binary tree). */
int bch_flatten_check(bch_scene* s, int64_t* out4);
/* Primitive.Box (min xyz, max xyz) and Primitive.Centroid of every primitive: nprim x 6, x 3 */
void bch_export_bounds(const bch_scene* s, double* box6, double* centroid3);
/* Scene.Refit (renderer.go:451-473) */
void bch_refit(bch_scene* s);
/* counts: nprim, nnode, nmat; revisions: topology, geometry */
void bch_counts(const bch_scene* s, int32_t* out3);
void bch_revisions(const bch_scene* s, uint64_t* out2);
/* view for bc_scene_upload; valid until the scene is changed or freed */
const bc_scene_view* bch_view(bch_scene* s, int nee_triangles);
/* primitives as the Go fields A,B,C (triangles) or Center,Radius (spheres): nprim x 9 */
void bch_export_prims(const bch_scene* s, int32_t* kind, int32_t* mat, double* geom9);
void bch_export_materials(const bch_scene* s, int32_t* kind, double* data8);
/* translate primitives [start, start+count) by d (translatePrimitive, renderer.go:282-303) */
void bch_translate(bch_scene* s, int start, int count, const double* d);

/* LoadOBJ (renderer.go:2996-3069), parsed in parallel on `threads` host threads (0 = all):
 * the same primitives in the same order, one diffuse material (.72, .76, .82), and camera_out
 * (may be NULL) = pos.xyz, target.xyz, up.xyz, fov as LoadOBJ frames it.  The scene is not built
 * (bch_build or bch_build_device).  NULL on error, with LoadOBJ's message in err. */
bch_scene* bch_load_obj(const char* path, double* camera_out, char* err, int err_cap, int threads);

/* OBJ straight to a resident device scene: bch_load_obj (parallel parse), the SAH tree built on
 * the device (bch_build_device with bins/leaf), then bc_scene_upload on ctx.  Returns the scene
 * (the Go Scene mirror, for the CPU paths) or NULL with the message in err. */
bch_scene* bch_ingest_obj(bc_ctx* ctx, const char* path, int bins, int leaf, int threads, double* camera_out,
                          char* err, int err_cap);
/* Writes the triangles of s as OBJ text (three "v" lines with %.17g coordinates and one "f" line
 * per triangle, every other face with relative indices): an ingestion input that loads back to
 * the same primitives.  0 ok, -1 bad arguments, -2 I/O error. */
int bch_write_obj(const bch_scene* s, const char* path);

/* Beamcast GX triangle page store (BGXPAGE1, PagedTriangleStore, gx_stream.cpp:9-55): records
 * laid out as hx::PackedTriangle (four 16-byte float vectors and the material, 80 bytes). */
typedef struct bch_packed_triangle {
    float v0[4], e1[4], e2[4], normal[4];
    uint32_t material, pad[3];
} bch_packed_triangle;
typedef struct bch_pages bch_pages;
/* PagedTriangleStore::write: 0 ok, -1 bad arguments, -2 I/O error */
int bch_pages_write(const char* path, const bch_packed_triangle* tris, int64_t n, uint32_t triangles_per_page);
/* PagedTriangleStore constructor: NULL with the message in err on a missing or invalid file */
bch_pages* bch_pages_open(const char* path, char* err, int err_cap);
void bch_pages_close(bch_pages* p);
int64_t bch_pages_count(const bch_pages* p);
int bch_pages_info(const bch_pages* p, int64_t page, uint64_t* file_offset, uint32_t* triangle_count);
/* load_page into out (triangle_count records): 0 ok, -1 bad page, -2/-3 I/O or truncated */
int bch_pages_load(const bch_pages* p, int64_t page, bch_packed_triangle* out);
/* append_page_to_scene for pages [first, first + npages): AddTriangle(v0, v0 + e1, v0 + e2,
 * material) with the sums in float as there */
int bch_pages_append(bch_scene* s, const bch_pages* p, int64_t first, int64_t npages);

/* Quality(name) (renderer.go:1064-1117), or UltraRealismSettings(backend) (:1119-1135) when backend
 * is not NULL (then name is ignored): out20 = Width, Height, SPP, Bounces, Denoise, MIS,
 * PowerLightSampling, AdaptiveSampling, ProgressiveUpdates, Sampler (0 Random, 1 Halton, 2 R2),
 * MSAA, TSAA, TAA, TXAA, FXAA, HBAO, HBAOPlus, RRDepth, Backend (0 Auto, 1 CPU, 2 OpenCL GPU),
 * BVHBins; *firefly = FireflyClamp; label (cap bytes, may be NULL) = the Quality label.  Returns 0,
 * or -1 for an unknown backend name ("Auto", "CPU", "OpenCL GPU"). */
int bch_quality(const char* name, const char* backend, int32_t* out20, double* firefly, char* label, int cap);
/* effectiveSPP (renderer.go:2395-2417) of Quality(name) with its SPP, MSAA and TSAA replaced */
int bch_effective_spp(const char* name, int spp, int msaa, int tsaa);

/* The RenderSettings fields the GPU render path reads (renderer.go:1027-1062). */
typedef struct bch_settings {
    int32_t width, height, spp, bounces, rr_depth;
    int32_t sampler;                 /* 0 Random, 1 Halton, 2 R2 */
    int32_t mis, power_lights;
    int32_t msaa, tsaa, taa, txaa, hbao, hbao_plus, debug_view, denoise;  /* effectiveSPP, guides */
    int32_t progressive, publish_hz; /* ProgressiveUpdates, PublishHz: the chunk height */
    int32_t upscale;                 /* internalSize: 0 Off, 1 Ultra Quality, 2 Balanced, 3 Performance */
    double firefly_clamp;
} bch_settings;
/* Quality(name) (renderer.go:1064-1117) as bch_settings; 0, or BC_ERR_ARG without an output */
int bch_settings_quality(const char* name, bch_settings* out);
/* The render description renderOpenCL's host loop renders a frame with (internalSize, effectiveSPP,
 * sampler and light settings, Camera.basis, the kernel's seed 0x9e3779b9); camera10 = pos.xyz,
 * target.xyz, up.xyz, fov. */
int bch_render_desc_for(const double* camera10, const bch_settings* settings, bc_render_desc* out);
/* Guide (renderer.go:726-731) as renderOpenCL decodes it from the compact guides
 * (gpu_opencl_windows.go:686-716). */
typedef struct bch_guide {
    double p[3], n[3], albedo[3], depth;
    int32_t valid, pad;
} bch_guide;
/* onChunk (renderer.go:2657-2677): rows [y0, y0+rows) of the linear image, rows*W*3 doubles */
typedef void (*bch_chunk_fn)(void* user, int32_t y0, int32_t rows, const double* rgb, double progress);
/* renderOpenCL (gpu_opencl_windows.go:565-723) over the HIP backend: one process-wide runtime
 * (context, lock and scene cache, as persistentCL), ensureSceneLocked's choice keyed on the scene
 * and its revisions (*scene_action: 0 resident scene reused, 1 refit on the device because only
 * GPUGeometryRevision changed, 2 full upload), the frame rendered in row chunks of 64 rows (32 at
 * PublishHz >= 60, 128 at <= 15, all rows without progressive updates) with on_chunk called after
 * each, the cancel flag polled before every chunk (BC_ERR_CANCELLED, "render cancelled"), and
 * the outputs decoded as there: linear (W*H*3 doubles), guides (W*H, may be NULL; filled when the
 * settings need guides: denoise, HBAO, TAA/TSAA/TXAA or a debug view), the path-segment count and
 * the device name.  Errors return a status with the message in err. */
int bch_render_hip(bch_scene* s, const double* camera10, const bch_settings* settings, const volatile int32_t* cancel,
                   bch_chunk_fn on_chunk, void* user, double* linear, bch_guide* guides, int64_t* rays_out,
                   int32_t* scene_action, char* device_name, int name_cap, char* err, int err_cap);
/* Destroys the process-wide runtime of bch_render_hip (the next call creates it again). */
void bch_render_hip_release(void);

/* Camera.basis (renderer.go:947-960): out12 = origin, lower, horiz, vert */
void bch_camera_basis(const double* pos, const double* target, const double* up, double fov,
                      double aspect, double* out12);

/* Seeded workloads.  camera_out (may be NULL) receives pos.xyz, target.xyz, up.xyz, fov.
 *   1  closed box: 5 walls, 2 rotated boxes, ceiling quad light (36 triangles)
 *   2  triangle soup: `param` triangles, centroids U[0,1]^3, edges U(.002,.02)
 *   3  Material Showcase (renderer.go:2944-2959)
 *   4  Cornell-like Room (renderer.go:2960-2979)
 *   5  Stress Grid (renderer.go:2980-2995)
 *   6  heightfield terrain: param x param grid cells, fBm heights, 64 materials
 *   7  closed box retessellated to ~param triangles with 64 materials
 *   8  scene `param` of the multi-scene batch: a log-uniform [1e4, 2e6] primitive budget filled
 *      by instanced boxes (12 triangles each), analytic spheres and 2-16 emissive quads (between
 *      the budget and 11 primitives more), diffuse/metal/glass palette
 * The scene is built (SAH, bins, leaf) before it is returned. */
bch_scene* bch_workload_scene(int kind, uint64_t seed, int64_t param, int bins, int leaf, double* camera_out);
/* Rays for the traversal workload (n x 8 floats: o.xyz, tmin, d.xyz, tmax):
 *   set 0 coherent pinhole primaries through pixel centres of a width x width image
 *   set 1 incoherent: origins U[0,1]^3, directions uniform on the sphere
 * any_hit selects the any-hit tmax distribution.  first/count select a sub-range of the
 * full set, so a sample can be regenerated without the rest. */
int bch_workload_rays(int set, uint64_t seed, int64_t width, int any_hit, int64_t first,
                      int64_t count, float* out);

#ifdef __cplusplus
}
#endif
#endif
