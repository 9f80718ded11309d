/*
 * beamcast_hip.h -- C ABI of the MI355X path-tracing backend for the Beamcast Go renderer.
 *
 * This is the drop-in boundary that replaces the OpenCL GPU Dataflow path
 * (windows_portable/gpu_opencl_windows.go, stub gpu_opencl_stub.go).  A cgo shim calls
 * these entry points where the Go code today calls renderOpenCL / openCLGPUAvailable.
 * Plain pointers and sizes only; every call returns 0 on success or a negative status,
 * with a message retrievable from bc_last_error().  A context is not thread-safe: the
 * caller serialises calls on it, as persistentCL.mu does today (gpu_opencl_windows.go:570-572).
 */
#ifndef BEAMCAST_HIP_H
#define BEAMCAST_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct bc_ctx bc_ctx;

enum {
    BC_OK = 0,
    BC_ERR_ARG = -1,        /* invalid argument / size */
    BC_ERR_DEVICE = -2,     /* HIP runtime failure */
    BC_ERR_SCENE = -3,      /* scene missing or not encodable */
    BC_ERR_CANCELLED = -4,  /* cancel flag observed between chunks */
    BC_ERR_NOMEM = -5
};

/* Replaces openCLGPUAvailable() (gpu_opencl_windows.go:172; stub gpu_opencl_stub.go:10-12).
 * Returns 1 and writes the device name when a gfx950 device is usable, else 0 and a reason. */
int bc_device_available(char* name, int cap);

/* Replaces the persistentCL singleton set-up (gpu_opencl_windows.go:94-131, initLocked :224-316).
 * Returns NULL on failure; bc_last_error(NULL) then gives the reason. */
bc_ctx* bc_create(int device_ordinal);
void bc_destroy(bc_ctx* ctx);
const char* bc_last_error(const bc_ctx* ctx);
/* Device name, as renderOpenCL's deviceName result (gpu_opencl_windows.go:565-566). */
const char* bc_device_name(const bc_ctx* ctx);

/* Caller-owned view of a built Go Scene (renderer.go:225-241): the SAH BVH of
 * Scene.Build (renderer.go:409-449) and the primitives/materials it indexes. */
typedef struct bc_scene_view {
    int32_t nnode;              /* len(Scene.Nodes) */
    const double* node_box;     /* nnode x 6: Box.Min.xyz, Box.Max.xyz */
    const int32_t* node_link;   /* nnode x 4: Left, Right, Start, Count */
    int32_t nprim;              /* len(Scene.Primitives) == len(Scene.Order) */
    const int32_t* order;       /* Scene.Order */
    const int32_t* prim_kind;   /* 0 PrimSphere, 1 PrimTriangle */
    const int32_t* prim_mat;    /* Primitive.Material */
    const double* prim_geom;    /* nprim x 10: triangle A.xyz E1.xyz E2.xyz 0; sphere Center.xyz Radius 0.. */
    int32_t nmat;               /* len(Scene.Materials) */
    const int32_t* mat_kind;    /* 0 diffuse, 1 metal, 2 glass, 3 emissive */
    const double* mat_data;     /* nmat x 8: Color.xyz, Roughness, IOR, Emission.xyz */
    int32_t nee_triangles;      /* 1: emissive triangles are NEE lights too (Go samples spheres only) */
} bc_scene_view;

/* Replaces ensureSceneLocked / uploadFullSceneLocked / updateGeometryLocked
 * (gpu_opencl_windows.go:379-495) including validateGPUSceneEncoding (:470-482).
 * Every call flattens and uploads the given scene; the caller decides when to call, keyed by
 * (scene, topo_rev, geom_rev) as ensureSceneLocked does.  The revisions are recorded with the
 * scene (bc_backward checks that the scene did not change since bc_forward).  A scene may take
 * at most 2^24 BVH8 nodes (about 90 million primitives); a larger one is BC_ERR_SCENE. */
int bc_scene_upload(bc_ctx* ctx, const bc_scene_view* scene, uint64_t topo_rev, uint64_t geom_rev);
/* Replaces the material part of uploadFullSceneLocked (gpu_opencl_windows.go:384-396) and the
 * light distribution of rebuildLightDistribution (renderer.go:363-407) without touching the
 * BVH: new Color/Roughness/IOR/Emission for the uploaded scene (same material count and kinds
 * of the lights).  The differentiable render updates its parameters through this call. */
int bc_update_materials(bc_ctx* ctx, int32_t nmat, const int32_t* mat_kind, const double* mat_data);
/* Device SAH build of the binary BVH (opt-in replacement of the host build): the Nodes and
 * Order of Scene.Build in SAH mode (renderer.go:409-449, buildNodeSAH :526-633), node for node
 * and index for index equal to the host build (same single-axis binned sweep, 0.98 leaf rule,
 * two-pointer partition order and depth-first node numbering).  Inputs: nprim primitive boxes
 * (Primitive.Box: min xyz, max xyz) and centroids (Primitive.Centroid), BVHBins and
 * BVHLeafSize (clamped as Build does: bins to [4, 32] with < 4 meaning 16, leaf to [1, 16]).
 * Outputs: node_box (node_cap x 6: Box min, max), node_link (node_cap x 4: Left, Right, Start,
 * Count), order (nprim), *nnode; node_cap = 2 * nprim - 1 always suffices.  Non-finite input is
 * rejected (BC_ERR_ARG).  The median mode and the light distribution stay with the caller. */
int bc_build_bvh(bc_ctx* ctx, int32_t nprim, const double* prim_box, const double* prim_centroid, int32_t bins,
                 int32_t leaf_size, int32_t node_cap, double* node_box, int32_t* node_link, int32_t* order,
                 int32_t* nnode);
/* Geometry-only update of the resident single scene: primitives moved, topology unchanged (same
 * primitive count, kinds and Go BVH leaf membership).  Replaces updateGeometryLocked
 * (gpu_opencl_windows.go:447-468), which re-packs and re-uploads every node after Scene.Refit
 * (renderer.go:451-473): here only the primitive records are uploaded and the BVH8 node boxes are
 * refit on the device bottom-up, one launch per level, with the same conservative quantisation as
 * bc_scene_upload.  Lights and their distribution are rebuilt from the new geometry.  Traversal
 * stays exact (the hit set of a ray does not depend on the box tightness); geom_rev is recorded. */
int bc_scene_refit(bc_ctx* ctx, const bc_scene_view* scene, uint64_t geom_rev);
/* Ray-query schedule of bc_trace / bc_trace_device (no counterpart in the reference), for
 * closest-hit and any-hit queries: with lanes > 0 each wave owns a run of rays and gives new rays
 * to its finished lanes once `lanes` of its 64 are idle; 0 runs one ray per thread to completion.
 * closest_lanes = -1 (the default) chooses per batch of 65,536 rays or more, on the device: a
 * one-block probe of 4,096 adjacent ray pairs (origin cell and direction octant, as ray binning
 * bins them) picks refill at 8 lanes when at least half of the pairs differ (incoherent rays) and
 * one ray per thread otherwise (coherent rays, whose waves share upper nodes); smaller batches run
 * one ray per thread.  Defaults -1 (closest) and 8 (any hit).  The renders use the same settings:
 * shadow rays the any-hit one; the second segment's extension rays of the plain schedule the
 * closest-hit one (-1: 8 lanes).  Results do not depend on it. */
int bc_set_trace_refill(bc_ctx* ctx, int closest_lanes, int any_lanes);
/* Ray binning of bc_trace / bc_trace_device (no counterpart in the reference): batches of 65,536
 * rays or more can be counting-sorted on the device by (Morton cell of the origin on a grid of
 * 2^cell_bits cells per axis over the scene box, direction octant) before traversal, so that a
 * wave's rays start in one cell and share their upper nodes; results are written back to each
 * ray's own index.  mode 0 off (default), 1 always, 2 when at least half of 4,096 sampled
 * adjacent ray pairs of the batch fall into different bins.  cell_bits 2..4 (default 4).  Results
 * do not depend on it; it keeps 36 bytes per ray of scratch. */
int bc_set_trace_order(bc_ctx* ctx, int mode, int cell_bits);
/* Decision of the last bc_trace / bc_trace_device call (synchronises the stream): out2[0] = 1 if
 * its rays were binned, out2[1] = sampled adjacent pairs (of 4,096) in different bins, or -1 when
 * the call ran no binning pass (mode 0 or fewer than 65,536 rays).  No counterpart in the
 * reference. */
int bc_trace_order_last(bc_ctx* ctx, int* out2);
/* Where bc_scene_upload flattens the binary tree into BVH8 nodes (no counterpart in the
 * reference): 0 automatic (the device for scenes of 65,536 primitives or more), 1 the host, 2 the
 * device.  Both produce the same bytes; the device path uploads the view's arrays instead. */
int bc_set_flatten(bc_ctx* ctx, int mode);
/* Statistics of the uploaded scene: nodes8, depth8, prims, bytes resident. */
int bc_scene_stats(const bc_ctx* ctx, int64_t* out4);
/* Copies the resident BVH8 to host memory for inspection (no counterpart in the reference, which
 * never reads its packed buffers back): nodes (nodes8 x 64 bytes, the bc_node8 layout of
 * csrc/bc_layout.h), records (8 x nodes8 record slots of 36 bytes, nine floats: the record of
 * leaf slot s of node i at 8 i + s, zero elsewhere) and record ids (8 x nodes8 uint32, primitive
 * index | 0x80000000 for spheres, 0xFFFFFFFF in slots without a primitive), nodes8 from
 * bc_scene_stats; NULL skips an array. */
int bc_scene_arrays(bc_ctx* ctx, void* nodes, void* records, uint32_t* ids);

/* Render description: the RenderSettings fields the path reads (renderer.go:1027-1062), the
 * camera basis renderOpenCL passes as kernel args 11-14 (gpu_opencl_windows.go:580-612) in
 * fp64, and this backend's own knobs. */
typedef struct bc_render_desc {
    int32_t width, height;      /* internalSize (renderer.go:2066-2079) */
    int32_t spp;                /* effectiveSPP (renderer.go:2395-2417) */
    int32_t bounces;            /* path segments per sample */
    int32_t rr_depth;           /* RRDepth; < 1 means 4 (renderer.go:1550-1553) */
    int32_t sampler;            /* 0 Random, 1 Halton (+ Cranley-Patterson rotation), 2 R2 */
    int32_t mis;                /* power-heuristic MIS */
    int32_t power_lights;       /* power-weighted light selection */
    int32_t sky_mode;           /* 0 background() gradient of renderer.go:1326, 1 constant sky_color */
    float firefly_clamp;        /* FireflyClamp; <= 0 off */
    uint32_t seed;              /* base seed of the counter-based generator */
    float sky_color[3];
    double origin[3], lower[3], horiz[3], vert[3];   /* Camera.basis (renderer.go:947-960) */
    int32_t mode;               /* 0 plain wavefront, 1 persistent + graph */
    int32_t samples_per_pass;   /* 0 = automatic */
    int32_t reserved[6];        /* mode 1: [0] pool slots, [1] iterations per graph (0 for either =
                                   autotune, which may pick the plain schedule), [2] slots per
                                   segment (0 = 256; multiple of 256);
                                   any mode: [3] path segments per XCD turn (0 = 1);
                                   mode 1: [4] 1 = shade each segment's paths sorted by material,
                                   [5] at most this many samples per sample window (0 = as many
                                   as the per-sample result buffer holds) */
    /* Direct lighting by resampled importance sampling (ReSTIR DI initial candidates; an
     * extension, the Go integrator has one-sample NEE): restir_candidates > 0 draws that many
     * candidates per diffuse vertex from {sun, sky, emitters} and keeps one in a reservoir. */
    int32_t restir_candidates;
    float sun_dir[3];           /* direction towards the sun */
    float sun_irradiance[3];    /* rgb; 0 = no sun (needs restir_candidates > 0) */
    /* ReSTIR DI spatial reuse at primary vertices (mode 0, single scene, needs restir_candidates
     * > 0): each pixel combines its reservoir with up to restir_spatial (<= 16) neighbours'
     * reservoirs of the same sample index, picked within restir_radius (1..64, 0 = 16) pixels
     * among the rendered rows, before its shadow ray.  0 = off. */
    int32_t restir_spatial;
    int32_t restir_radius;
} bc_render_desc;

/* Replaces renderOpenCL's chunk loop (gpu_opencl_windows.go:565-723): renders rows
 * [y0, y0+rows) into caller-owned host buffers in the reference output layout:
 *   rgba   rows*W*4 floats: mean radiance rgb, w = path segments (gpu_opencl_windows.go:820)
 *   guides rows*W*3 u32 (may be NULL): depth bits, oct-normal|material|valid, jitter
 *          (gpu_opencl_windows.go:787-797, decoded at :688-717)
 * rays_out receives the path-segment count (Meta.Rays).  cancel (may be NULL) is read once, when
 * the call starts: a set flag returns BC_ERR_CANCELLED before anything is rendered.  A frame is
 * cancelled between chunks, as renderOpenCL polls its flag before each chunk
 * (gpu_opencl_windows.go:644-646); bch_render_hip (beamcast_host.h) does that.
 * Render-ahead (on unless bc_set_render_ahead turns it off; not with spatial reuse, whose pixels
 * read their neighbours among the rendered rows): a call whose rows are not already rendered
 * renders them together with the rest of the frame when that fits one window of per-sample
 * results (mode 1: W x rows x spp <= 2^27 samples) or 2^23 pixels (mode 0), else with the
 * following rows up to the multiple of `rows` that fits, into a buffer of the context; later
 * calls with the same description for rows inside that block, with no scene or material change
 * in between, copy them out, each row once and in increasing order (a call for rows before the
 * end of the rows last served, as a new frame's first chunk, renders anew).  Every pixel depends
 * only on the description, the scene and its index, so the output is what rendering each chunk
 * alone gives; the GPU runs one render per block instead of one per chunk, with no host round
 * trip between.  rays_out is then the sum of the rows' per-pixel segment counts (rgba w), as
 * renderOpenCL sums them. */
int bc_render(bc_ctx* ctx, const bc_render_desc* desc, int y0, int rows, float* rgba,
              uint32_t* guides, uint64_t* rays_out, const volatile int32_t* cancel);
/* No counterpart in the reference: render-ahead of bc_render, 1 on (the default) or 0 off;
 * either call drops rows rendered ahead. */
int bc_set_render_ahead(bc_ctx* ctx, int on);
/* Same, with rgba/guides in device memory (no host copy). shadow_rays_out may be NULL.
 * mode 1 (persistent): a pool of path slots refilled in place as paths end, the bounce loop
 * replayed as a captured hipGraph; reserved[0] = pool slots, reserved[1] = iterations per
 * graph, both 0 = chosen by the autotuner on first use for this image shape and scene. */
int bc_render_device(bc_ctx* ctx, const bc_render_desc* desc, int y0, int rows, float* d_rgba,
                     uint32_t* d_guides, uint64_t* rays_out, uint64_t* shadow_rays_out);

/* No counterpart in the reference: bc_render_device for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Renders rows [y0, y0+rows) with the plain schedule (desc->mode must be 0, restir_spatial 0)
 * into the caller's d_rgba / d_guides (device memory, bc_render_device's layout) with results
 * bit-identical to bc_render_device.  It calls no hipMalloc, hipFree or other allocating routine:
 * its path state, pixel sums, queue lengths and tallies live in one fixed device array of under
 * 49 MiB that each GPU holds from the library's first use until the process ends.  Limits:
 * rows * width <= 262144 pixels per call (a row chunk of a larger frame, as renderOpenCL renders
 * them); samples are rendered in passes of at most 262144 paths.  Larger calls return
 * BC_ERR_ARG.  Every call on one GPU shares that GPU's array and leaves its contents behind (the
 * next call clears what must start at zero): calls that overlap on one GPU, from several
 * contexts or threads, race and must be serialised by the caller. */
int bc_render_noalloc(bc_ctx* ctx, const bc_render_desc* desc, int y0, int rows, float* d_rgba,
                      uint32_t* d_guides, uint64_t* rays_out, uint64_t* shadow_out);

/* Differentiable render (path-replay backpropagation; no counterpart in the reference).
 * bc_forward renders like bc_render (plain schedule, host rgba) and keeps each sample's path
 * radiance; bc_backward replays every sample from its key and accumulates (+=) the gradient of
 * the loss wrt every material's Color (nmat x 3) and Emission (nmat x 3), given
 * dloss_dimage (rows*W*3 floats, d loss / d mean pixel radiance).  Russian roulette and the
 * power-weighted light selection are held fixed in the derivative; firefly_clamp must be <= 0. */
int bc_forward(bc_ctx* ctx, const bc_render_desc* desc, int y0, int rows, float* rgba, uint64_t* rays_out);
int bc_backward(bc_ctx* ctx, const bc_render_desc* desc, int y0, int rows, const float* dloss_dimage,
                float* dcolor, float* demission);

/* Multi-scene batch (no counterpart in the reference, whose backend holds one scene): packs n
 * scenes into one device arena (BVH8 nodes, records, materials and light tables concatenated and
 * rebased; each scene keeps its own light distribution).  Replaces the resident scene or batch;
 * bc_render / bc_trace need a single scene (bc_scene_upload) again afterwards. */
int bc_batch_upload(bc_ctx* ctx, int n, const bc_scene_view* scenes);
/* Renders every scene of the resident batch: descs[i] (full frames) into d_rgba[i] (device,
 * W_i x H_i x 4 floats, the layout of bc_render).  descs[0].mode selects the schedule: 0 renders
 * the scenes one after another with the plain wavefront schedule, 1 runs one persistent schedule
 * (pool, graph and segment from descs[0].reserved, or autotuned) over the samples of all scenes.
 * Images equal those of rendering each scene alone.  rays_out / shadow_out: totals (may be NULL). */
int bc_batch_render_device(bc_ctx* ctx, int n, const bc_render_desc* descs, float* const* d_rgba,
                           uint64_t* rays_out, uint64_t* shadow_out);

/* Post stack on the device (compact guides + denoise/AA/tone map): restates, in fp32, the part of
 * RenderState.Start after the render (renderer.go:2842-2895) with HBAO, temporal reuse, bloom,
 * debug views and upscaling off: atrousDenoise with the compact guides (renderer.go:2228-2287,
 * guide decode gpu_opencl_windows.go:552-562, 686-716), applyFXAA (:2557-2586),
 * applyTXAAResolve (:2588-2595, edgeAwareSharpen :2146-2183), toRGBA (:2055-2064). */
typedef struct bc_post_desc {
    int32_t denoise_passes;  /* atrousDenoise iterations (Start: 2 + min(3, max(TAA, TXAA))); 0 = off */
    int32_t fxaa;            /* applyFXAA level, 0 = off */
    int32_t txaa;            /* applyTXAAResolve level, 0 = off */
    int32_t hdr;             /* toneMapHDR mode (0 clamp, 1 Reinhard, 3 ACES fit, other filmic) */
    float exposure, gamma;   /* toRGBA */
    int32_t reserved[4];
} bc_post_desc;
/* d_rgba: W*H*4 floats as bc_render_device writes them; d_guides: its compact guides (needed
 * when denoise_passes > 0).  Writes RGBA8 words (0xAABBGGRR, toRGBA's packing) to d_pixels and/or
 * the processed linear frame (same float4 layout) to d_linear; either may be NULL, not both. */
int bc_postprocess_device(bc_ctx* ctx, const bc_post_desc* post, int width, int height, const float* d_rgba,
                          const uint32_t* d_guides, uint32_t* d_pixels, float* d_linear);
/* Whole frame: bc_render_device + bc_postprocess_device, then only the RGBA8 pixels (4 B/px;
 * 28 B/px without the device post stack) and, if linear is not NULL, the processed linear frame
 * come back to the host. */
int bc_render_post(bc_ctx* ctx, const bc_render_desc* desc, const bc_post_desc* post, uint32_t* pixels,
                   float* linear, uint64_t* rays_out);

/* Schedule last used by a mode-1 render, 8 ints: pool slots, iterations per graph, traversal
 * stack depth, sample windows, slots per segment, material-sorted shading (0/1), 1 if the
 * autotuner chose the plain schedule (the other fields then give the best persistent shape it
 * found), 0.  No counterpart in the reference (its kernel has one fixed shape). */
int bc_render_tuning(const bc_ctx* ctx, int32_t* out8);

/* Ray queries against the uploaded scene (the traversal of gpu_opencl_windows.go:758-777
 * and of Scene.Hit / Scene.Occluded, renderer.go:777-895).  rays: n x 8 floats
 * (o.xyz, tmin, d.xyz, tmax).  Closest hit writes n x (float t, int32 prim) with prim = -1
 * on a miss; any hit writes n int32 (1 occluded, 0 clear). Host pointers. */
int bc_trace(bc_ctx* ctx, const float* rays, int64_t n, int any_hit, void* out);
/* Same with device pointers, enqueued on the context stream (asynchronous). */
int bc_trace_device(bc_ctx* ctx, const void* d_rays, int64_t n, int any_hit, void* d_out);

/* Diagnostics with no counterpart in the reference (the roofline accounting of the benchmark).
 * bc_trace_counts: bc_trace_device's query (same results in d_out), one ray per thread, that also
 * writes per ray d_counts[4i .. 4i+3] = node visits (64-byte bc_node8 each), primitive-record tests
 * (36-byte records), id-table reads (4 bytes), traversal steps (a step visits at most one node;
 * the others only pop the stack).  Asynchronous on the context stream.
 * bc_valu_probe: times (HIP events, *ms) `blocks` workgroups of 256 threads that each run `iters`
 * register-only eight-child slab tests, the VALU instruction mix of a node visit (mix 0), or
 * `iters` rounds of eight independent fp32 FMAs (mix 1), to calibrate the VALU issue rates the
 * traversal is compared with.
 * bc_line_probe: times `blocks` workgroups of 256 threads that each read `iters` 64-byte blocks
 * (four 16-byte loads, one node fetch) at hashed positions of the device buffer d_buf of `bytes`
 * bytes (a power of two, at least 64; 16-byte aligned; its contents do not matter): over a buffer
 * far larger than the L2, the rate of scattered line fills from the Infinity Cache that an
 * incoherent traversal is compared with.  BC_ERR_ARG for any other buffer. */
int bc_trace_counts(bc_ctx* ctx, const void* d_rays, int64_t n, int any_hit, void* d_out, uint32_t* d_counts);
int bc_valu_probe(bc_ctx* ctx, int mix, int blocks, int iters, float* ms);
int bc_line_probe(bc_ctx* ctx, const void* d_buf, uint64_t bytes, int blocks, int iters, float* ms);

/* Device memory and stream helpers for callers that keep inputs resident in HBM. */
void* bc_alloc(bc_ctx* ctx, uint64_t bytes);
int bc_free(bc_ctx* ctx, void* d_ptr);
int bc_upload(bc_ctx* ctx, void* d_dst, const void* src, uint64_t bytes);
int bc_download(bc_ctx* ctx, void* dst, const void* d_src, uint64_t bytes);
int bc_copy_device(bc_ctx* ctx, void* d_dst, const void* d_src, uint64_t bytes);  /* stream-ordered */
int bc_synchronize(bc_ctx* ctx);
/* Stream-ordered timing: record a start marker, then read the elapsed ms since it. */
int bc_timer_start(bc_ctx* ctx);
int bc_timer_stop(bc_ctx* ctx, float* ms);

#ifdef __cplusplus
}
#endif
#endif
