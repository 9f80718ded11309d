/*
 * oracle.h -- CPU restatement of the Beamcast Go renderer's path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, its C ABI,
 * the host-side scene mirror) links, loads or calls this code.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, and only as
 * the checker (or the timed CPU comparator), never as the thing measured or shipped.
 *
 * What it restates (reference = windows_portable/ of Beamcast GX 5.9):
 *   - Scene.AddTriangle / AddSphere derived fields        renderer.go:238-270
 *   - Scene.Build + buildNodeSAH (binned SAH, one axis)   renderer.go:409-449, 526-633
 *   - rebuildLightDistribution (power-weighted CDF)       renderer.go:363-407
 *   - threadedBVHEscapes                                  renderer.go:3251-3276
 *   - AABB.HitAccel slab test                             renderer.go:147-210
 *   - triHit / triAnyHit / sphereHitAccel                 renderer.go:650-770
 *   - Scene.Hit (16-entry stack + recursion) / Occluded   renderer.go:777-939
 *   - Camera.basis                                        renderer.go:947-960
 *   - trace / directLight / chooseLight / powerHeuristic  renderer.go:1262-1563
 *   - radicalInverse / buildSampleSequence / pixelSampleRotation  renderer.go:1844-1915
 *   - adaptivePixel camera jitter (adaptive sampling off) renderer.go:1928-1970
 *   - compact guide packing                               gpu_opencl_windows.go:787-797
 *
 * Two precisions of the integrator and traversal are provided:
 *   fp64 -- the Go CPU semantics (Scene.Hit stack order, later-tested equal-t wins);
 *   fp32 -- the same algorithm evaluated in IEEE single precision with the exact
 *           operation order used by the HIP kernels (no contraction, correctly
 *           rounded divide/sqrt), with a deterministic equal-t tie rule
 *           (smaller primitive index wins) so that the answer does not depend on
 *           traversal order.
 * Random numbers come from a counter-based hash keyed on (seed, pixel, sample,
 * dimension) instead of the time-seeded math/rand streams of renderer.go:2741;
 * see orc_rng_u32.
 */
#ifndef BEAMCAST_ORACLE_H
#define BEAMCAST_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

/* Material kinds follow renderer.go:86-91 (MatDiffuse, MatMetal, MatGlass, MatEmissive). */
enum { ORC_DIFFUSE = 0, ORC_METAL = 1, ORC_GLASS = 2, ORC_EMISSIVE = 3 };
/* Primitive kinds follow renderer.go:101-104 (PrimSphere, PrimTriangle). */
enum { ORC_SPHERE = 0, ORC_TRIANGLE = 1 };

/* mat_data: nmat x 8 doubles = color.xyz, roughness, ior, emission.xyz
 * prim_data: nprim x 9 doubles; triangle = a.xyz b.xyz c.xyz ; sphere = center.xyz radius 0...
 * Degenerate triangles (|e1 x e2|^2 < 1e-18) and spheres with r <= 0 are dropped exactly
 * as AddTriangle / AddSphere drop them (renderer.go:246-260), which shifts later indices.
 * nee_triangles != 0 additionally registers emissive triangles as NEE lights (an extension;
 * the Go CPU path samples emissive spheres only, renderer.go:253-255). */
orc_scene* orc_scene_new(int nmat, const int32_t* mat_kind, const double* mat_data,
                         int nprim, const int32_t* prim_kind, const int32_t* prim_mat,
                         const double* prim_data, int bins, int leaf, int nee_triangles);
void orc_scene_free(orc_scene* s);
int  orc_scene_nprim(const orc_scene* s);
int  orc_scene_nnode(const orc_scene* s);
int  orc_scene_fallback_sorts(const orc_scene* s);
/* node_box nnode x 6 (min.xyz max.xyz), node_link nnode x 4 (left,right,start,count), order nprim */
void orc_scene_nodes(const orc_scene* s, double* node_box, int32_t* node_link, int32_t* order);
void orc_scene_escapes(const orc_scene* s, int32_t* escape);
/* Light table: count, then light prim indices, normalised CDF, PDF by primitive. */
int  orc_scene_nlights(const orc_scene* s);
void orc_scene_lights(const orc_scene* s, int32_t* light_prims, double* cdf, double* pdf_by_prim);

/* AABB.HitAccel on one box; ray = o.xyz d.xyz. Returns hit, writes near. */
int orc_aabb_hit(const double* box6, const double* ray6, double tmin, double tmax, double* near_out);

/* Ray batches: rays are n x 8: o.xyz tmin d.xyz tmax.
 * Closest hit outputs: prim (Scene.Primitives index, -1 on miss) and t. */
void orc_hit64(const orc_scene* s, int64_t n, const double* rays, int32_t* prim, double* t);
void orc_hit64_threaded(const orc_scene* s, int64_t n, const double* rays, int32_t* prim, double* t);
void orc_occluded64(const orc_scene* s, int64_t n, const double* rays, int32_t* occ);
void orc_brute64(const orc_scene* s, int64_t n, const double* rays, int32_t* prim, double* t);
/* fp32 variants over the fp32 records (A,E1,E2 rounded from the fp64 scene).  counts (may be
 * NULL) receives per-ray node fetches and primitive tests of the threaded traversal on the
 * reference 32-byte node layout. nthreads <= 0 uses all online CPUs. */
void orc_hit32(const orc_scene* s, int64_t n, const float* rays, int32_t* prim, float* t,
               uint32_t* node_visits, uint32_t* prim_tests, int nthreads);
void orc_any32(const orc_scene* s, int64_t n, const float* rays, int32_t* occ,
               uint32_t* node_visits, uint32_t* prim_tests, int nthreads);
void orc_brute32(const orc_scene* s, int64_t n, const float* rays, int32_t* prim, float* t);
/* Multi-threaded fp64 Scene.Hit / Occluded, for timing the CPU comparator. Returns seconds. */
double orc_time_hit64(const orc_scene* s, int64_t n, const float* rays, int any_hit,
                      int nthreads, int32_t* prim_out);

/* Sampling primitives */
double   orc_radical_inverse(uint64_t n, uint64_t base);
void     orc_pixel_rotation(int x, int y, double* rx, double* ry);
/* camera jitter of sample smp at pixel (x, y) of a W-wide image: sampler 1 Halton, 2 R2 (both
 * with the pixel's Cranley-Patterson rotation), else counter-based uniforms */
void     orc_camera_jitter(int sampler, uint32_t seed, int W, int x, int y, int smp, double* jx, double* jy);
double   orc_power_heuristic(double a, double b);
uint32_t orc_path_key(uint32_t seed, uint32_t pixel, uint32_t sample);
uint32_t orc_rng_u32(uint32_t key, uint32_t dim);
float    orc_rng_float(uint32_t key, uint32_t dim);
/* cosine hemisphere sample for normal n using the integrator's bounded disk rejection */
void     orc_cos_hemisphere64(const double* n, uint32_t key, uint32_t dim0, double* out);

/* Camera.basis (renderer.go:947-960): writes origin, lower, horiz, vert (12 doubles). */
void orc_camera_basis(const double* pos, const double* target, const double* up, double fov,
                      double aspect, double* out12);

/* Render description shared with the C ABI (same field meaning as bc_render_desc). */
typedef struct orc_render_desc {
    int32_t width, height, spp, bounces, rr_depth;
    int32_t sampler;        /* 0 random, 1 halton, 2 r2 */
    int32_t mis, power_lights;
    int32_t sky_mode;       /* 0 renderer.go background gradient, 1 constant sky_color */
    float   firefly_clamp;
    uint32_t seed;
    float   sky_color[3];
    double  origin[3], lower[3], horiz[3], vert[3];
    int32_t restir_candidates;  /* >0: resampled direct lighting over {sun, sky, emitters} */
    float   sun_dir[3];         /* unit direction towards the sun */
    float   sun_irradiance[3];
    int32_t restir_spatial;     /* >0: spatial reuse at primary vertices over that many neighbours */
    int32_t restir_radius;      /* pixel radius of the neighbours (0 = 16) */
} orc_render_desc;

/* Renders rows [y0, y0+rows).  rgba: rows*W*4 (rgb mean radiance, w = path segments),
 * guides: rows*W*3 u32 (may be NULL), per-pixel sums as doubles for fp64 mode.
 * Returns the number of path segments (reference-compatible ray count).
 * shadow_rays (may be NULL) receives the number of NEE shadow rays traced. */
int64_t orc_render64(const orc_scene* s, const orc_render_desc* d, int y0, int rows,
                     double* rgba, uint32_t* guides, int64_t* shadow_rays, int nthreads);
int64_t orc_render32(const orc_scene* s, const orc_render_desc* d, int y0, int rows,
                     float* rgba, uint32_t* guides, int64_t* shadow_rays, int nthreads);

/* Post stack (oracle_post.c): denoise -> FXAA -> TXAA -> toRGBA on a rendered frame.
 * rgba: W*H*4 floats (bc_render layout); guides: W*H*3 compact guide words (may be NULL when
 * denoise_passes == 0).  linear_out: W*H*3 doubles (may be NULL); pixels_out: W*H RGBA8 words
 * (may be NULL).  Returns 0, or -1 when out of memory. */
int orc_post(int w, int h, const float* rgba, const uint32_t* guides, int denoise_passes, int fxaa_level,
             int txaa_level, double exposure, double gamma, int hdr, double* linear_out, uint32_t* pixels_out);
uint32_t orc_to_rgba(const double* c, double exposure, double gamma, int hdr);
void orc_decode_oct_normal(uint32_t qx, uint32_t qy, double* n);
/* single primitives (known-answer tests): triHit / sphereHitAccel in fp64, the GPU's fp32 tests */
int orc_tri_hit64(const double* g9, const double* ray6, double tmin, double tmax, double* t_out);
int orc_sphere_hit64(const double* c4, const double* ray6, double tmin, double tmax, double* t_out);
int orc_tri_test32(const float* g9, const float* o, const float* d, float tmin, float tmax, float* t_out);
int orc_sphere_test32(const double* c4, const float* o, const float* d, float tmin, float tmax, float* t_out);
/* packGuideNormalMat / packJitter (gpu_opencl_windows.go:787-797) */
uint32_t orc_pack_guide32(const float* n3, int mat);
uint32_t orc_pack_jitter(float ju, float jv);

#ifdef __cplusplus
}
#endif
#endif
