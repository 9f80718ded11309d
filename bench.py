#!/usr/bin/env python3
"""Traversal benchmark of the MI355X backend (one GPU).

Workload (the traversal microbenchmark): a seeded 1,000,000-triangle soup (centroids
uniform in the unit cube, edge lengths U(0.002, 0.02)), SAH BVH with 32 bins and leaf size 4
built on the host exactly as Scene.Build does, flattened to the quantised BVH8, and two ray
sets of 4096 x 4096 = 16,777,216 rays each: coherent pinhole primaries through pixel centres
and incoherent rays (origins uniform in the cube, directions uniform on the sphere).  One
step = the four launches closest-hit and any-hit over both sets = 67,108,864 ray queries,
inputs resident in HBM.  value = ray queries per second of the whole step, in Mrays/s.

Also reported: the roofline of the dominant kernel -- algorithmic bytes of the BVH8 traversal
that runs (64 B per node visit, 36 B per primitive-record test, 4 B per id-table read, the ray and
its result), counted per ray by a counting instantiation of the same traversal over the whole
launch, over the kernel's HIP-event duration, against the 8 TB/s HBM peak -- with the VALU issue
rate of the same kernel against a calibrated peak as the second roofline; and the CPU baseline:
the oracle's fp64 restatement of Scene.Hit / Scene.Occluded (renderer.go:777-895) on every host
core this job may use (affinity and cgroup quota, both recorded with the CPU model) over a
bounded sample of the same rays.
"""
import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "beamcast-gx-v1.0-beta-_amd"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

METRIC = "path-samples/s & Mrays/s on 1 MI355X vs host CPU; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gen_rays(bc, set_id, width, any_hit, threads):
    n = width * width
    out = np.empty((n, 8), np.float32)
    chunks = np.linspace(0, n, threads * 4 + 1).astype(np.int64)

    def work(i):
        a, b = int(chunks[i]), int(chunks[i + 1])
        rc = bc.lib().bch_workload_rays(set_id, 0x5EED, width, 1 if any_hit else 0, a, b - a,
                                        bc.c_void_p(out[a:b].ctypes.data))
        if rc != 0:
            raise RuntimeError("ray generation failed")

    with cf.ThreadPoolExecutor(threads) as ex:
        list(ex.map(work, range(len(chunks) - 1)))
    return out


def fixed_schedule(args, d):
    """--schedule pool,iters,segment,sort: the persistent schedule given explicitly (no autotune)."""
    if args.schedule:
        pool, iters, seg, sort = (int(x) for x in args.schedule.split(","))
        d.reserved[0], d.reserved[1], d.reserved[2], d.reserved[4] = pool, iters, seg, sort
    return d


def go_facing(dev, desc, reps=1):
    """The renderer's own path to the GPU: bc_render into host buffers in 64-row chunks with the
    compact guides, as renderOpenCL chunks a frame (gpu_opencl_windows.go:631-641) and the cgo
    shim calls it (one image and one guide chunk buffer per frame, reused by every chunk); seconds
    per frame after one untimed frame."""
    def frame():
        # one chunk buffer of each kind per frame, reused by every chunk, as the shim mallocs them
        raw = np.empty((64, desc.width, 4), np.float32)
        raw_guide = np.empty((64, desc.width, 3), np.uint32)
        for y0 in range(0, desc.height, 64):
            rows = min(64, desc.height - y0)
            dev.render(desc, y0=y0, rows=rows, guides=True, out=raw[:rows], guides_out=raw_guide[:rows])

    frame()
    t0 = time.perf_counter()
    for _ in range(reps):
        frame()
    return (time.perf_counter() - t0) / reps


def bench_render(bc, dev, args, threads):
    """Closed box 256x256, 16 spp, 4 segments: path-samples/s of the wavefront integrator
    (rgb + guides left in HBM) next to the oracle's fp64 integrator on the host cores."""
    box = bc.HostScene.workload(bc.CLOSED_BOX)
    dev.upload(box, nee_triangles=True)
    w = h = 256
    spp, bounces = 16, 4
    d = bc.render_desc(box, w, h, spp, bounces, rr_depth=3, sampler=1, mis=True, power=True)
    d_rgba = dev.alloc(w * h * 16)
    d_g = dev.alloc(w * h * 12)
    for _ in range(2):
        dev.render_device(d, 0, h, d_rgba, d_g)
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        segs, shadow = dev.render_device(d, 0, h, d_rgba, d_g)
    sec = (time.perf_counter() - t0) / reps
    dev.free(d_rgba)
    dev.free(d_g)
    samples = w * h * spp
    out = {"workload": "closed box (36 tris, quad light), 256x256, 16 spp Halton, 4 segments, MIS + power NEE "
                       "over emissive triangles, RR from bounce 3, rgb + guides",
           "ms_per_frame": round(sec * 1e3, 3), "path_samples_per_s": round(samples / sec, 1),
           "mrays_s_segments": round(segs / sec / 1e6, 2), "mrays_s_total": round((segs + shadow) / sec / 1e6, 2),
           "segments_per_sample": round(segs / samples, 4)}
    # batched-persistent schedule (in-place regeneration, hipGraph replay, autotuned shape)
    dp = fixed_schedule(args, bc.render_desc(box, w, h, spp, bounces, rr_depth=3, sampler=1, mis=True, power=True,
                                             mode=1))
    d_rgba = dev.alloc(w * h * 16)
    d_g = dev.alloc(w * h * 12)
    dev.render_device(dp, 0, h, d_rgba, d_g)  # autotunes and captures once
    t0 = time.perf_counter()
    for _ in range(reps):
        dev.render_device(dp, 0, h, d_rgba, d_g)
    psec = (time.perf_counter() - t0) / reps
    dev.free(d_rgba)
    dev.free(d_g)
    out["persistent"] = {"ms_per_frame": round(psec * 1e3, 3), "path_samples_per_s": round(samples / psec, 1),
                         "over_plain": round(sec / psec, 3), "tuning": dev.tuning()}
    gsec = go_facing(dev, d, reps=5)
    out["go_facing"] = {"ms_per_frame": round(gsec * 1e3, 3), "path_samples_per_s": round(samples / gsec, 1),
                        "path": "bc_render, 64-row chunks, rgba + compact guides copied to host per chunk",
                        "over_device_resident_time": round(gsec / sec, 3)}
    if not args.no_cpu:
        import oracle as orc

        o = orc.Scene.from_host(box, nee_triangles=True)
        od = orc.desc_from(d)
        rows = 16
        t0 = time.perf_counter()
        cpu_img = o.render(od, precision=64, y0=120, rows=rows, threads=threads)[0]
        csec = time.perf_counter() - t0
        cpu_rate = w * rows * spp / csec
        out["cpu_baseline"] = {"path_samples_per_s": round(cpu_rate, 1), "cores": threads, "kind": "port",
                               "sample": f"rows 120-{120 + rows} of the same frame, fp64 trace restatement",
                               "host": cpu_info()}
        out["gpu_over_cpu"] = round(samples / sec / cpu_rate, 1)
        # matched image error: both 16-spp bands against an independent-seed fp64 render at 4096 spp
        dref = bc.render_desc(box, w, h, 4096, bounces, rr_depth=3, sampler=1, mis=True, power=True, seed=0x51ED)
        ref = o.render(orc.desc_from(dref), precision=64, y0=120, rows=rows, threads=threads)[0][..., :3]
        gpu_img = dev.render(d, y0=120, rows=rows)[0]
        scale = float(ref.mean())
        e_gpu = float(np.sqrt(np.mean((gpu_img[..., :3] - ref) ** 2))) / scale
        e_cpu = float(np.sqrt(np.mean((cpu_img[..., :3] - ref) ** 2))) / scale
        out["matched_rmse"] = {"reference": f"fp64 oracle, 4096 spp, independent seed, rows 120-{120 + rows}",
                               "rel_rmse_gpu": round(e_gpu, 5), "rel_rmse_cpu": round(e_cpu, 5),
                               "cpu_over_gpu": round(e_cpu / e_gpu, 4)}
    return out


def bench_terrain(bc, dev, args, threads):
    """Procedural terrain: a (cells x cells) fBm heightfield (2 x 2047^2 = 8,380,418 triangles at the
    default), 64 diffuse materials, sun from (0.3, 0.8, 0.2) with irradiance 3 plus a constant sky 0.3;
    1920x1080, 64 spp, 8 segments, resampled direct lighting with 32 candidates.  Plain wavefront
    schedule vs the persistent schedule (in-place regeneration, hipGraph replay, autotuned pool,
    segment and graph length), and the oracle's fp64 integrator on a bounded band of rows."""
    t0 = time.time()
    s = bc.HostScene.workload(bc.TERRAIN, param=args.terrain_cells, bins=32, leaf=4)
    build = time.time() - t0
    t0 = time.time()
    dev.upload(s)
    upload = time.time() - t0
    log(f"[bench] terrain: host SAH build {build:.1f}s, BVH8 flatten + upload {upload:.1f}s, {dev.stats()}")
    w, h, spp, bounces = 1920, 1080, args.terrain_spp, 8
    kw = dict(rr_depth=3, sampler=1, sky_mode=1, sky_color=(0.3, 0.3, 0.3), restir=32, sun_dir=(0.3, 0.8, 0.2),
              sun=(3.0, 3.0, 3.0))
    d = bc.render_desc(s, w, h, spp, bounces, **kw)
    d_rgba = dev.alloc(w * h * 16)
    d_g = dev.alloc(w * h * 12)
    dp = fixed_schedule(args, bc.render_desc(s, w, h, spp, bounces, mode=1, **kw))
    if args.profile_light:  # one plain and one persistent frame, for counter runs
        segs, shadow = dev.render_device(d, 0, h, d_rgba)
        psegs, _ = dev.render_device(dp, 0, h, d_rgba)
        dev.free(d_rgba)
        dev.free(d_g)
        return {"profile_light": True, "segments": segs, "shadow_rays": shadow, "persistent_segments": psegs,
                "tuning": dev.tuning()}
    dev.render_device(d, 0, h, d_rgba, d_g)
    t0 = time.perf_counter()
    segs, shadow = dev.render_device(d, 0, h, d_rgba)
    sec = time.perf_counter() - t0
    # device post stack on the rendered frame: denoise (3 passes), FXAA 2, TXAA 1, toRGBA
    post = bc.post_desc(denoise=3, fxaa=2, txaa=1, hdr=2)
    d_px = dev.alloc(w * h * 4)
    dev.postprocess_device(post, w, h, d_rgba, d_g, d_px)
    dev.timer_start()
    for _ in range(5):
        dev.postprocess_device(post, w, h, d_rgba, d_g, d_px)
    post_ms = dev.timer_stop() / 5
    dev.free(d_px)
    dev.free(d_g)
    # ReSTIR DI spatial reuse at the primary vertices (plain schedule): 4 neighbours within 16 px
    dsp = bc.render_desc(s, w, h, spp, bounces, spatial=4, radius=16, **kw)
    t0 = time.perf_counter()
    ssegs, _ = dev.render_device(dsp, 0, h, d_rgba)
    ssec = time.perf_counter() - t0
    t0 = time.perf_counter()
    dev.render_device(dp, 0, h, d_rgba)  # autotunes and captures once
    tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    psegs, _ = dev.render_device(dp, 0, h, d_rgba)
    psec = time.perf_counter() - t0
    ptune = dev.tuning()  # the timed frame's shape (the chunked renders below tune their own)
    dev.free(d_rgba)
    samples = w * h * spp
    assert psegs == segs
    g_plain = go_facing(dev, d)
    g_pers = go_facing(dev, dp)
    log(f"[bench] terrain Go-facing chunks: plain {g_plain * 1e3:.1f} ms, persistent {g_pers * 1e3:.1f} ms")
    log(f"[bench] terrain: plain {sec * 1e3:.1f} ms, persistent {psec * 1e3:.1f} ms (autotune + first {tune:.1f}s)")
    out = {"workload": f"terrain {args.terrain_cells}^2 cells ({s.counts()[0]} tris, 64 materials), sun + sky 0.3, "
                       f"{w}x{h}, {spp} spp Halton, {bounces} segments, RR from bounce 3, resampled direct "
                       "lighting 32 candidates",
           "ms_per_frame": round(sec * 1e3, 2), "path_samples_per_s": round(samples / sec, 1),
           "mrays_s_segments": round(segs / sec / 1e6, 2), "mrays_s_total": round((segs + shadow) / sec / 1e6, 2),
           "segments_per_sample": round(segs / samples, 4),
           "persistent": {"ms_per_frame": round(psec * 1e3, 2), "path_samples_per_s": round(samples / psec, 1),
                          "over_plain": round(sec / psec, 3), "tuning": ptune,
                          "autotune_and_first_frame_s": round(tune, 2)},
           "post": {"ms_per_frame": round(post_ms, 3), "stages": "a-trous denoise 3 passes, FXAA 2, TXAA 1, toRGBA",
                    "readback_bytes_per_pixel": {"with_device_post": 4, "without": 28}},
           "spatial_reuse": {"ms_per_frame": round(ssec * 1e3, 2), "neighbours": 4, "radius_px": 16,
                             "over_plain_time": round(ssec / sec, 3)},
           "go_facing": {"path": "bc_render, 64-row chunks, rgba + compact guides copied to host per chunk",
                         "plain_ms_per_frame": round(g_plain * 1e3, 2),
                         "persistent_ms_per_frame": round(g_pers * 1e3, 2),
                         "plain_path_samples_per_s": round(samples / g_plain, 1),
                         "persistent_path_samples_per_s": round(samples / g_pers, 1)}}
    if not args.no_cpu:
        import oracle as orc

        t0 = time.time()
        o = orc.Scene.from_host(s, bins=32, leaf=4)
        log(f"[bench] terrain oracle build {time.time() - t0:.1f}s")
        od = orc.desc_from(d)
        y0, rows = h // 2, 1
        t0 = time.perf_counter()
        o.render(od, precision=64, y0=y0, rows=rows, threads=threads)
        one = time.perf_counter() - t0
        rows = int(max(1, min(64, args.cpu_seconds / 2 / max(one, 1e-3))))
        t0 = time.perf_counter()
        o.render(od, precision=64, y0=y0, rows=rows, threads=threads)
        csec = time.perf_counter() - t0
        cpu_rate = w * rows * spp / csec
        out["cpu_baseline"] = {"path_samples_per_s": round(cpu_rate, 1), "cores": threads, "kind": "port",
                               "sample": f"rows {y0}-{y0 + rows} of the same frame, fp64 trace restatement, "
                                         f"{csec:.1f} s", "host": cpu_info()}
        out["gpu_over_cpu"] = round(samples / sec / cpu_rate, 1)
        del o
    return out


BATCH_RES = ((640, 360), (1280, 720), (2560, 1440))


def batch_resolution(i, seed=0xC5):
    """Resolution of batch scene i: one of 640x360, 1280x720, 2560x1440 by a seeded hash."""
    h = (seed * 0x9E3779B97F4A7C15 + (i + 1) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    h ^= h >> 31
    return BATCH_RES[(h * 0x94D049BB133111EB >> 61) % 3]


def bench_batch(bc, dev, args, threads):
    """Multi-scene batch: 32 seeded scenes (log-uniform 1e4..2e6 primitives of instanced boxes and
    spheres, 2-16 quad lights, resolutions from {640x360, 1280x720, 2560x1440}), Ultra transport
    (256 spp by default, 16 segments, RR from bounce 6, no clamp).  Plain: the scenes one after
    another; persistent: one schedule over all scenes' samples.  The persistent shape is autotuned
    on an 8-spp render of the same frames: the autotuner's trials hold one sample per pixel at this
    frame size whatever the spp, so they are the trials a 256-spp render would run."""
    n = args.batch_scenes
    t0 = time.time()
    with cf.ThreadPoolExecutor(threads) as ex:
        scenes = list(ex.map(lambda i: bc.HostScene.workload(bc.BATCH_SCENE, seed=0xC5, param=i), range(n)))
    build = time.time() - t0
    t0 = time.time()
    dev.upload_batch(scenes, nee_triangles=True)
    upload = time.time() - t0
    tris = sum(sc.counts()[0] for sc in scenes)
    log(f"[bench] batch: {n} scenes, {tris} primitives, host build {build:.1f}s, pack + upload {upload:.1f}s")
    descs = []
    for i, sc in enumerate(scenes):
        w, h = batch_resolution(i)
        descs.append(bc.render_desc(sc, w, h, args.batch_spp, 16, rr_depth=6, sampler=1, mis=True, power=True))
    pix = sum(d.width * d.height for d in descs)
    bufs = [dev.alloc(d.width * d.height * 16) for d in descs]
    out = {"workload": f"{n} batch scenes ({tris} primitives, {pix} pixels), {args.batch_spp} spp Halton, 16 segments, "
                       "RR from bounce 6, no clamp, MIS + power NEE over emissive quads",
           "is_c5": n == 32 and args.batch_spp == 256}
    spp = args.batch_spp
    for mode in (0, 1):
        tuning = None
        for d in descs:
            d.mode = mode
            d.spp = 1 if mode == 0 else min(8, spp)
        # untimed: plain, the passes' buffers at one sample; persistent, the autotuner at 8 spp
        t0 = time.perf_counter()
        if mode == 0 or (spp != 1 and not args.schedule):
            dev.render_batch_device(descs, bufs)
        first = time.perf_counter() - t0
        if mode == 1 and args.schedule:
            for d in descs:
                fixed_schedule(args, d)
            tuning = dict(zip(("pool", "iters_per_graph", "segment", "material_sort"),
                              (int(x) for x in args.schedule.split(","))))
        elif mode == 1:
            tuning = dev.tuning()
            for d in descs:
                d.mode = 0 if tuning["plain"] else 1
                d.reserved[0], d.reserved[1] = tuning["pool"], tuning["iters_per_graph"]
                d.reserved[2], d.reserved[4] = tuning["segment"], tuning["material_sort"]
        for d in descs:
            d.spp = spp
        t0 = time.perf_counter()
        segs, shadow = dev.render_batch_device(descs, bufs)
        sec = time.perf_counter() - t0
        key = "plain" if mode == 0 else "persistent"
        out[key] = {"seconds": round(sec, 4), "path_samples_per_s": round(pix * spp / sec, 1),
                    "mrays_s_total": round((segs + shadow) / sec / 1e6, 2),
                    "segments_per_sample": round(segs / (pix * spp), 4)}
        if mode == 1:
            out[key]["tuning"] = dict(tuning or {}, windows=dev.tuning()["windows"], given=bool(args.schedule))
            out[key]["autotune_at_8spp_s"] = round(first, 2)
        log(f"[bench] batch {key}: {sec:.3f}s")
    out["persistent_over_plain"] = round(out["plain"]["seconds"] / out["persistent"]["seconds"], 3)
    for b in bufs:
        dev.free(b)
    return out


def bench_build(bc, dev, args):
    """Scene.Build (SAH, 32 bins, leaf 4) on the host against the same tree built on the device
    (bc_build_bvh): the 1M-triangle soup and the 8.4M-triangle terrain.  The device time is the
    whole call: upload of the primitive bounds, the build, download of nodes and order."""
    out = {}
    for name, kind, param in (("soup", bc.SOUP, args.tris), ("terrain", bc.TERRAIN, args.terrain_cells)):
        s = bc.HostScene.workload(kind, param=param, bins=32, leaf=4)
        t0 = time.perf_counter()
        s.build(32, 4)
        host = time.perf_counter() - t0
        box, cen = s.bounds()
        dev.build_bvh(box, cen, 32, 4)  # first call: allocations
        t0 = time.perf_counter()
        nbox, _, _ = dev.build_bvh(box, cen, 32, 4)
        gpu = time.perf_counter() - t0
        n = s.counts()
        if len(nbox) != n[1]:
            raise RuntimeError(f"device build of {name}: {len(nbox)} nodes, host {n[1]}")
        out[name] = {"primitives": n[0], "nodes": n[1], "host_build_s": round(host, 3), "device_build_s": round(gpu, 4),
                     "host_over_device": round(host / gpu, 1)}
        log(f"[bench] build {name}: host {host:.2f}s, device {gpu * 1e3:.1f} ms ({n[0]} primitives)")
    return out


def bench_ingest(bc, dev, args):
    """OBJ file -> resident BVH8 on the GPU for the 1M-triangle soup (204 MB of OBJ text): the
    reference's order of work (serial LoadOBJ parse, host Scene.Build, flatten + upload) against
    bch_ingest_obj (parallel parse, device SAH build, flatten + upload)."""
    import tempfile

    src = bc.HostScene.workload(bc.SOUP, param=args.tris, bins=32, leaf=4)
    threads = cpu_info()["threads"]
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "soup.obj")
        src.write_obj(path)
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        s, _ = bc.HostScene.load_obj(path, threads=1)
        t1 = time.perf_counter()
        s.build(32, 4)
        t2 = time.perf_counter()
        dev.upload(s)
        t3 = time.perf_counter()
        del s
        t4 = time.perf_counter()
        s2, _ = bc.HostScene.load_obj(path, threads=threads)
        t5 = time.perf_counter()
        s2.build_device(dev, 32, 4)
        t6 = time.perf_counter()
        dev.upload(s2)
        t7 = time.perf_counter()
    serial = t3 - t0
    fast = t7 - t4
    log(f"[bench] ingest: serial parse {t1 - t0:.2f}s + host build {t2 - t1:.2f}s + upload {t3 - t2:.2f}s; "
        f"parallel parse {t5 - t4:.2f}s + device build {t6 - t5:.2f}s + upload {t7 - t6:.2f}s")
    return {"workload": f"{args.tris}-triangle soup as OBJ text ({size / 1e6:.0f} MB) -> resident BVH8",
            "serial_host": {"parse_s": round(t1 - t0, 3), "build_s": round(t2 - t1, 3), "upload_s": round(t3 - t2, 3),
                            "total_s": round(serial, 3)},
            "ingest": {"threads": threads, "parse_s": round(t5 - t4, 3), "build_s": round(t6 - t5, 3),
                       "upload_s": round(t7 - t6, 3), "total_s": round(fast, 3)},
            "speedup": round(serial / fast, 2)}


def bench_refit(bc, dev, args):
    """Geometry-only update of the traversal soup: half of the triangles move; a full flatten +
    upload (bc_scene_upload) against the device refit (bc_scene_refit), both timed around the C
    call with the host view already packed."""
    lib = bc.lib()
    s = bc.HostScene.workload(bc.SOUP, param=args.tris)
    view = s.view()
    topo, geom = s.revisions()
    t0 = time.perf_counter()
    dev.upload_view(view, topo, geom)
    full = time.perf_counter() - t0
    s.translate(0, args.tris // 2, (0.25, -0.1, 0.05))
    s.refit()
    view = s.view()
    geom = s.revisions()[1]
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        dev._check(lib.bc_scene_refit(dev.ctx, view, geom))
        times.append(time.perf_counter() - t0)
    return {"workload": f"{args.tris}-triangle soup, {args.tris // 2} triangles translated",
            "full_upload_ms": round(full * 1e3, 2), "refit_ms": round(min(times) * 1e3, 2),
            "full_over_refit": round(full / min(times), 2)}


def bench_adjoint(bc, dev):
    """Differentiable-render config: closed box retessellated to ~200k triangles with 64 materials,
    512x512, 32 spp, 6 segments; time of one forward and one path-replay backward pass."""
    s = bc.HostScene.workload(bc.BOX_RETESS, param=200_000)
    kind, mat, geom, mk, md = s.export()
    dev.upload(s, nee_triangles=True)
    d = bc.render_desc(s, 512, 512, 32, 6, rr_depth=6, sampler=1, mis=True, power=True, seed=7)
    target, _ = dev.forward(d)
    md2 = md.copy()
    md2[mk == bc.DIFFUSE, 0:3] = 0.5
    md2[mk == bc.EMISSIVE, 5:8] *= 0.5
    dev.update_materials(mk, md2)
    img, _ = dev.forward(d)
    dl = 2 * (img[..., :3] - target[..., :3])
    dev.backward(d, dl, len(mk))
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        img, _ = dev.forward(d)
    fwd = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        dev.backward(d, dl, len(mk))
    bwd = (time.perf_counter() - t0) / reps
    return {"workload": f"closed box retessellated ({s.counts()[0]} tris, {len(mk)} materials), 512x512, 32 spp, "
                        "6 segments, L2 image loss",
            "forward_ms": round(fwd * 1e3, 2), "backward_ms": round(bwd * 1e3, 2),
            "adjoint_over_forward": round(bwd / fwd, 3),
            "forward_path_samples_per_s": round(512 * 512 * 32 / fwd, 1)}


PROFILE_SUMMARY = os.path.join(ROOT, "profiles", "k_trace_soup1M_16Mrays.json")
# MI355X's published fp32 vector rate, 157.3 TFLOP/s, in wave64 FMA instructions per second
# (64 lanes x 2 flops each): one per SIMD every 2 clocks, 1024 SIMDs at 2.4 GHz
VALU_PEAK = 157.3e12 / 128
LINE_PROBE_BYTES = 64 << 20


def roofline_entry(launch, achieved, alg_bytes, alg_ref_bytes, kernel_ms, probe_cfg, probe_ms, line_cfg,  # noqa: C901
                   line_ms):
    """Roofline of the dominant kernel (the slowest launch of the step).

    achieved: the BVH8 traversal's algorithmic bytes per launch over its HIP-event duration, against
    the HBM peak.  traffic: HBM-side bytes per launch of that kernel from the committed rocprofv3
    counter summary (FETCH_SIZE and WRITE_SIZE passes, each scaled by what it reads for a 1 GiB
    device copy).  valu: the same kernel's VALU wave-instructions per second (SQ_INSTS_VALU of the
    committed summary over this run's duration) against the published fp32 vector issue rate, next
    to the rates this run measures for the FMA probe and the slab-test probe (the node visit's
    instruction mix alone, on registers), and the SIMD clocks per VALU instruction the summary's
    counter run gives for each (GRBM_GUI_ACTIVE per XCD over VALU instructions per SIMD).  limiter:
    the same launch's texture-addresser and texture-data busy fractions and VALU instructions per
    SIMD clock from the summary's memory-pipe counter run, next to the probes' VALU rates.
    line_fill: the launch's L2 fill bytes (FETCH_SIZE of the summary) per second of this run against
    those of the scattered-read probe (64-byte reads at hashed positions of a 64 MiB buffer; its
    FETCH_SIZE from the same summary, its duration from this run)."""
    out = {"bound": "hbm", "kernel": f"BVH8 traversal ({launch})", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
           "alg_bytes_per_ray": round(alg_bytes, 1), "ref_layout_alg_bytes_per_ray": round(alg_ref_bytes, 1),
           "valu": None}
    try:
        with open(PROFILE_SUMMARY) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return out
    k = prof.get("launches", {}).get(launch)
    if k:
        out["traffic"] = k["traffic_bytes"]
        out["traffic_source"] = os.path.relpath(PROFILE_SUMMARY, ROOT)
        out["traffic_gbs"] = round(k["traffic_bytes"] / (kernel_ms * 1e-3) / 1e9, 1)
        out["traffic_frac"] = round(out["traffic_gbs"] / HBM_PEAK_GBS, 4)
        out["l2_hit"] = k.get("l2_hit")
        out["active_lanes_per_valu_instr"] = k.get("active_lanes")
    lp = prof.get("line_probe")
    if k and lp and (lp.get("blocks"), lp.get("iters"), lp.get("bytes")) == (*line_cfg, LINE_PROBE_BYTES):
        fill = k["fetch_bytes"] / (kernel_ms * 1e-3) / 1e9
        ceil = lp["fetch_bytes"] / (line_ms * 1e-3) / 1e9
        reads = line_cfg[0] * 256 * line_cfg[1]
        out["line_fill"] = {"achieved": round(fill, 1), "ceiling": round(ceil, 1), "unit": "GB/s",
                            "frac": round(fill / ceil, 4),
                            "basis": "L2 fill bytes (FETCH_SIZE) per second: the launch against the scattered-read "
                                     "probe (64-byte reads at hashed positions of a 64 MiB buffer)",
                            "probe_ms": round(line_ms, 3),
                            "probe_delivered_gbs": round(reads * 64 / (line_ms * 1e-3) / 1e9, 1),
                            "probe_fill_bytes_per_read": round(lp["fetch_bytes"] / reads, 1),
                            "source": os.path.relpath(PROFILE_SUMMARY, ROOT)}
    pf, pm = prof.get("fma_probe"), prof.get("valu_probe")
    ok = all(p and (p.get("blocks"), p.get("iters")) == tuple(probe_cfg) for p in (pf, pm))
    if k and ok and k.get("valu_insts_per_wave"):
        rate = k["valu_insts_per_wave"] * k["waves"] / (kernel_ms * 1e-3)
        fma = pf["valu_insts_per_wave"] * pf["waves"] / (probe_ms[1] * 1e-3)
        mix = pm["valu_insts_per_wave"] * pm["waves"] / (probe_ms[0] * 1e-3)
        cyc = {name: (x.get("valu_class_per_wave") or {}).get("simd_cycles_per_instr")
               for name, x in (("launch", k), ("fma_probe", pf), ("slab_mix_probe", pm))}
        out["valu"] = {"bound": "valu", "achieved": round(rate / 1e9, 2), "peak": round(VALU_PEAK / 1e9, 2),
                       "unit": "G wave-instr/s", "frac": round(rate / VALU_PEAK, 4),
                       "peak_basis": "fp32 vector 157.3 TFLOP/s: one wave64 FMA per SIMD per 2 clocks, 1024 SIMDs, 2.4 GHz",
                       "fma_probe_rate": round(fma / 1e9, 2), "over_fma_probe": round(rate / fma, 4),
                       "slab_mix_rate": round(mix / 1e9, 2), "over_slab_mix": round(rate / mix, 4),
                       "simd_cycles_per_instr": cyc,
                       "f32_arith_share": (k.get("valu_class_per_wave") or {}).get("f32_arith_share"),
                       "probe_ms": {"fma": round(probe_ms[1], 3), "slab_mix": round(probe_ms[0], 3)},
                       "source": os.path.relpath(PROFILE_SUMMARY, ROOT)}
    mp = (k or {}).get("memory_pipe")
    if mp:
        # the two limits the counter runs show saturated together: the texture-data return path
        # (busy fraction of clocks, per CU) and VALU issue against the slab-mix probe's own rate
        mix_clk = pm.get("valu_per_simd_clock") if pm else None
        out["limiter"] = {"td_busy": mp["td_busy"], "ta_busy": mp["ta_busy"],
                          "td_clocks_per_ray": mp["td_clocks_per_ray"], "tcp_accesses_per_ray": mp["tcp_accesses_per_ray"],
                          "valu_per_simd_clock": mp["valu_per_simd_clock"], "slab_mix_probe_valu_per_simd_clock": mix_clk,
                          "fma_probe_valu_per_simd_clock": pf.get("valu_per_simd_clock") if pf else None,
                          "source": os.path.relpath(PROFILE_SUMMARY, ROOT)}
    return out


def cpu_info():
    """Host cores this process may run on: the affinity set, limited by the cgroup CPU quota when
    one is set (a quota of q CPUs lets the job use q cores' worth however many it sees), plus the
    CPU model, for the CPU-baseline records."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = aff if quota is None else max(1, min(aff, int(quota + 0.5)))
    return {"threads": threads, "nproc": os.cpu_count(), "affinity": aff, "cgroup_cpu_quota": quota, "model": model}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=4096)
    ap.add_argument("--tris", type=int, default=1_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline duration")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-render", action="store_true", help="skip the render, terrain and adjoint measurements")
    ap.add_argument("--no-terrain", action="store_true", help="skip the terrain render measurement")
    ap.add_argument("--terrain-cells", type=int, default=2047)
    ap.add_argument("--terrain-spp", type=int, default=64)
    ap.add_argument("--no-batch", action="store_true", help="skip the multi-scene batch measurement")
    ap.add_argument("--only", choices=("render", "terrain", "batch", "adjoint", "refit", "build", "ingest"),
                    help="run this render leg alone (for profiling) and print its result")
    ap.add_argument("--batch-scenes", type=int, default=32)
    ap.add_argument("--batch-spp", type=int, default=256, help="samples per pixel of the batch leg (Ultra: 256)")
    ap.add_argument("--profile-light", action="store_true",
                    help="terrain leg: one plain and one persistent frame only (for counter runs)")
    ap.add_argument("--schedule", default=None,
                    help="pool,iters,segment,sort: run the persistent legs with this schedule instead of the "
                         "autotuner's (for profiling runs)")
    ap.add_argument("--any-refill", type=int, default=8,
                    help="any-hit queries and integrator shadow rays: refill a wave's lanes once this many are idle "
                         "(0: one ray per thread)")
    ap.add_argument("--closest-refill", type=int, default=-1,
                    help="closest-hit queries: refill a wave's lanes once this many are idle (0: one ray per thread, "
                         "-1: chosen per batch by the library's coherence probe)")
    ap.add_argument("--trace-order", type=int, default=0, choices=(0, 1, 2),
                    help="ray binning before queries: 0 off, 1 always, 2 when the batch is incoherent")
    ap.add_argument("--order-bits", type=int, default=4, choices=(2, 3, 4), help="log2 of origin cells per axis")
    ap.add_argument("--copy-check", action="store_true",
                    help="also run one 1 GiB device-to-device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this backend runs on one GPU; the multi-GPU split (per-tile or per-sample "
                                   "ranges with one framebuffer all-reduce) is not built, so --gpus must be 1"}))
        sys.exit(2)

    import beamcast as bc

    cinfo = cpu_info()
    threads = cinfo["threads"]
    log(f"[bench] host: {cinfo}")
    if args.only:
        dev = bc.Device(0)
        dev.set_trace_refill(args.closest_refill, args.any_refill)
        legs = {"render": lambda: bench_render(bc, dev, args, threads), "terrain": lambda: bench_terrain(bc, dev, args, threads),
                "batch": lambda: bench_batch(bc, dev, args, threads), "adjoint": lambda: bench_adjoint(bc, dev),
                "refit": lambda: bench_refit(bc, dev, args), "build": lambda: bench_build(bc, dev, args),
                "ingest": lambda: bench_ingest(bc, dev, args)}
        out = legs[args.only]()
        dev.close()
        print(json.dumps({"leg": args.only, args.only: out}))
        return
    t0 = time.time()
    scene = bc.HostScene.workload(bc.SOUP, seed=0x5EED, param=args.tris, bins=32, leaf=4)
    log(f"[bench] host SAH build of {args.tris} triangles: {time.time() - t0:.2f}s")
    dev = bc.Device(0)
    dev.set_trace_order(args.trace_order, args.order_bits)
    dev.set_trace_refill(args.closest_refill, args.any_refill)
    t0 = time.time()
    dev.upload(scene)
    st = dev.stats()
    log(f"[bench] BVH8 flatten + upload: {time.time() - t0:.2f}s, {st}")

    t0 = time.time()
    sets = {}
    for set_id, sname in ((0, "coherent"), (1, "incoherent")):
        for any_hit in (False, True):
            sets[(sname, any_hit)] = gen_rays(bc, set_id, args.width, any_hit, threads)
    n = args.width * args.width
    log(f"[bench] generated 4 x {n} rays: {time.time() - t0:.2f}s")

    if args.copy_check:
        gib = 1 << 30
        a, b = dev.alloc(gib), dev.alloc(gib)
        dev.copy_device(b, a, gib)
        dev.sync()
        dev.free(a)
        dev.free(b)
        log("[bench] copied 1 GiB device to device (counter calibration)")

    d_rays, d_out = {}, {}
    for k, r in sets.items():
        d_rays[k] = dev.to_device(r)
        d_out[k] = dev.alloc(n * (4 if k[1] else 8))
    launches = list(sets.keys())

    def step():
        for k in launches:
            dev.trace_device(d_rays[k], n, k[1], d_out[k])

    for _ in range(args.warmup):
        step()
    dev.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    dev.sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = 4 * n * args.steps / wall / 1e6
    log(f"[bench] {args.steps} steps: {ms_per_step:.3f} ms/step, {value:.1f} Mrays/s")

    # per-kernel durations with stream-ordered HIP events, back-to-back launches
    kernel_ms, binned = {}, {}
    reps = max(3, args.steps // 2)
    for k in launches:
        dev.timer_start()
        for _ in range(reps):
            dev.trace_device(d_rays[k], n, k[1], d_out[k])
        kernel_ms[k] = dev.timer_stop() / reps
        binned[k] = dev.trace_order_last()[0]
    # results of the last launch, for a sanity summary
    hits = {}
    for k in launches:
        if k[1]:
            hits[k] = float(dev.download(d_out[k], np.empty(n, np.int32)).mean())
        else:
            h = dev.download(d_out[k], np.empty(n, dtype=[("t", np.float32), ("prim", np.int32)]))
            hits[k] = float((h["prim"] >= 0).mean())
    dominant = max(kernel_ms, key=kernel_ms.get)

    # algorithmic bytes of the BVH8 traversal that runs, counted for every ray of each launch by
    # the counting instantiation of the same traversal (whose results must equal the timed
    # launch's): 32 B ray + 8 B (closest) or 4 B (any hit) result + 64 B per node visit + 36 B per
    # primitive-record test + 4 B per id-table read
    d_cnt, d_chk = dev.alloc(n * 16), dev.alloc(n * 8)
    alg, work = {}, {}
    for k in launches:
        out_b = 4 if k[1] else 8
        dev.trace_counts(d_rays[k], n, k[1], d_chk, d_cnt)
        cnt = dev.download(d_cnt, np.empty((n, 4), np.uint32))
        c = cnt.sum(axis=0, dtype=np.int64) / n
        # steps of the slowest ray of each 64-ray wave (k_trace runs one ray per lane to completion)
        wave_steps = float(cnt[: n // 64 * 64, 3].reshape(-1, 64).max(axis=1).mean())
        del cnt
        a = dev.download(d_chk, np.empty(n * out_b, np.uint8))
        b = dev.download(d_out[k], np.empty(n * out_b, np.uint8))
        if not np.array_equal(a, b):
            raise RuntimeError("the counting traversal's results differ from the timed launch's")
        alg[k] = 32 + out_b + 64 * c[0] + 36 * c[1] + 4 * c[2]
        work[k] = {"node_visits_per_ray": round(float(c[0]), 3), "record_tests_per_ray": round(float(c[1]), 3),
                   "id_reads_per_ray": round(float(c[2]), 3), "steps_per_ray": round(float(c[3]), 3),
                   "steps_per_wave": round(wave_steps, 2)}
    dev.free(d_cnt)
    dev.free(d_chk)
    achieved = alg[dominant] * n / (kernel_ms[dominant] * 1e-3) / 1e9

    # VALU issue probes, 8 waves per SIMD, no memory access: eight independent FMA chains per
    # thread (the issue ceiling the traversal's VALU rate is compared with) and the node visit's
    # slab-test instruction mix on registers (what that mix alone issues)
    probe_cfg = (2048, 4000)
    probe_ms = {}
    for mix in (1, 0):
        dev.valu_probe(*probe_cfg, mix=mix)
        probe_ms[mix] = min(dev.valu_probe(*probe_cfg, mix=mix) for _ in range(3))
    # scattered 64-byte reads (one node fetch each) over 64 MiB, the size of the soup's scene: the
    # rate of line fills from the Infinity Cache that the incoherent launch is compared with
    line_cfg = (8192, 64)
    d_line = dev.alloc(LINE_PROBE_BYTES)
    dev.line_probe(d_line, LINE_PROBE_BYTES, *line_cfg)
    line_ms = min(dev.line_probe(d_line, LINE_PROBE_BYTES, *line_cfg) for _ in range(3))
    dev.free(d_line)

    # the reference layout's algorithmic bytes on the same rays (32 + out + 32 per threaded node
    # fetch + 44 per primitive test, counted by the oracle's threaded traversal over a sample), as
    # what the reference's OpenCL kernel would have moved
    import oracle as orc

    t0 = time.time()
    oscene = orc.Scene.from_host(scene, bins=32, leaf=4)
    log(f"[bench] oracle scene build: {time.time() - t0:.2f}s")
    sample_n = 65536
    alg_ref = {}
    for k in launches:
        idx = np.linspace(0, n - 1, sample_n).astype(np.int64)
        smp = sets[k][idx]
        if k[1]:
            _, nv, pt = oscene.any32(smp, threads=threads, counts=True)
        else:
            _, _, nv, pt = oscene.hit32(smp, threads=threads, counts=True)
        out_b = 4 if k[1] else 8
        alg_ref[k] = 32 + out_b + 32 * float(nv.mean()) + 44 * float(pt.mean())

    cpu = None
    if not args.no_cpu:
        # bounded sample of the same workload, equal parts of the four launches
        probe = np.concatenate([sets[k][:: n // 16384][:16384] for k in launches])
        secs = oscene.time_hit64(probe[: len(probe) // 2]) + oscene.time_hit64(probe[len(probe) // 2:], any_hit=True)
        rate = len(probe) / max(secs, 1e-9)
        m = int(min(n, max(16384, rate * args.cpu_seconds / 4)))
        step_idx = max(1, n // m)
        secs = 0.0
        total = 0
        for k in launches:
            smp = sets[k][::step_idx][:m]
            secs += oscene.time_hit64(smp, any_hit=k[1], threads=threads)
            total += len(smp)
        cpu_rate = total / secs / 1e6
        cpu = {"value": round(cpu_rate, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"{total} rays ({m} from each of the 4 launches, strided), fp64 Scene.Hit/Occluded "
                         f"restatement, {secs:.1f} s",
               "host": cinfo, "gpu_over_cpu": round(value / cpu_rate, 1)}
        log(f"[bench] CPU baseline: {cpu_rate:.2f} Mrays/s on {threads} threads ({secs:.1f}s)")

    render = None if args.no_render else bench_render(bc, dev, args, threads)
    adjoint = None if args.no_render else bench_adjoint(bc, dev)
    terrain = None if (args.no_render or args.no_terrain) else bench_terrain(bc, dev, args, threads)
    batch = None if (args.no_render or args.no_batch) else bench_batch(bc, dev, args, threads)
    refit = None if args.no_render else bench_refit(bc, dev, args)
    build = None if (args.no_render or args.no_terrain) else bench_build(bc, dev, args)
    ingest = None if args.no_render else bench_ingest(bc, dev, args)

    names = {("coherent", False): "coherent_closest", ("coherent", True): "coherent_any",
             ("incoherent", False): "incoherent_closest", ("incoherent", True): "incoherent_any"}
    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"traversal microbench: {args.tris}-triangle seeded soup (edges U(0.002,0.02)), SAH 32 bins "
                        f"leaf 4 -> quantised BVH8; 2 x {n} rays (coherent {args.width}^2 pinhole primaries, "
                        f"incoherent uniform origins/directions); closest hit + any hit per step",
            "rays_per_step": 4 * n,
            "bvh8_nodes": int(st["nodes8"]),
            "bvh8_depth": int(st["depth8"]),
            "scene_bytes": int(st["bytes"]),
        },
        "roofline": roofline_entry(names[dominant], achieved, alg[dominant], alg_ref[dominant], kernel_ms[dominant],
                                   probe_cfg, probe_ms, line_cfg, line_ms),
        "cpu_baseline": cpu,
        "render": render,
        "adjoint": adjoint,
        "terrain": terrain,
        "batch": batch,
        "refit": refit,
        "build": build,
        "ingest": ingest,
        "kernels": {names[k]: {"ms": round(kernel_ms[k], 4), "mrays_s": round(n / kernel_ms[k] / 1e3, 1),
                               "alg_bytes_per_ray": round(alg[k], 1),
                               "alg_gbs": round(alg[k] * n / (kernel_ms[k] * 1e-3) / 1e9, 1),
                               "ref_layout_alg_bytes_per_ray": round(alg_ref[k], 1), **work[k],
                               "hit_fraction": round(hits[k], 4), "binned": binned[k]}
                    for k in launches},
    }
    for k in launches:
        dev.free(d_rays[k])
        dev.free(d_out[k])
    dev.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
