"""GPU parity of the wavefront integrator against the oracle (through the C ABI).

Bar (per pixel, identical sample indices): radiance within 1e-4 relative of the oracle's fp32
restatement of trace (renderer.go:1484-1563), |gpu - ref| <= 1e-4 * |ref| + 1e-6; path-segment
counts and compact guides identical.  Against the fp64 Go semantics: same-seed images agree to
a small relative RMSE, and converged means agree.
"""
import numpy as np
import pytest

import beamcast as bc
import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = bc.Device(0)
    yield d
    d.close()


def parity(dev, scene, desc, nee=False, guides=True, bins=32):
    dev.upload(scene, nee_triangles=nee)
    img, g, rays = dev.render(desc, guides=guides)
    o = orc.Scene.from_host(scene, bins=bins, nee_triangles=nee)
    ref, rg, rrays, _ = o.render(orc.desc_from(desc), precision=32, guides=guides)
    assert rays == rrays, f"path segments {rays} vs {rrays}"
    assert np.array_equal(img[..., 3], ref[..., 3])
    err = np.abs(img[..., :3] - ref[..., :3])
    bad = err > 1e-4 * np.abs(ref[..., :3]) + 1e-6
    assert not bad.any(), f"{bad.sum()} channel values outside 1e-4 rel; max err {err.max()}"
    if guides:
        assert np.array_equal(g, rg)
    return img, ref, o


def test_closed_box_parity(dev):
    print("case: closed box 64x64 8spp", flush=True)
    s = bc.HostScene.workload(bc.CLOSED_BOX)
    d = bc.render_desc(s, 64, 64, 8, 4, rr_depth=3, sampler=1, mis=True, power=True)
    img, _, _ = parity(dev, s, d, nee=True)
    assert img[..., :3].mean() > 0.05


@pytest.mark.parametrize("kind,b", [(bc.SHOWCASE, 6), (bc.CORNELL, 5), (bc.STRESS, 6)])
def test_builtin_scene_parity(dev, kind, b):
    print(f"case: scene {kind}", flush=True)
    s = bc.HostScene.workload(kind, bins=16, leaf=4)
    d = bc.render_desc(s, 48, 32, 6, b, rr_depth=4, sampler=1, mis=True, power=True, firefly_clamp=24)
    parity(dev, s, d, bins=16)


@pytest.mark.parametrize("sampler,mis,power", [(0, False, False), (2, True, False), (1, False, True)])
def test_sampler_and_light_modes(dev, sampler, mis, power):
    print(f"case: sampler {sampler} mis {mis} power {power}", flush=True)
    s = bc.HostScene.workload(bc.CORNELL, bins=16, leaf=4)
    d = bc.render_desc(s, 40, 24, 4, 5, rr_depth=2, sampler=sampler, mis=mis, power=power, seed=1234)
    parity(dev, s, d, bins=16, guides=False)


def test_constant_sky_and_multipass(dev):
    """samples_per_pass forces several passes; accumulation order stays the sample order."""
    print("case: multipass", flush=True)
    s = bc.HostScene.workload(bc.BOX_RETESS, param=3000)
    d = bc.render_desc(s, 32, 32, 9, 4, rr_depth=2, sky_mode=1, sky_color=(0.3, 0.3, 0.3), samples_per_pass=2)
    parity(dev, s, d, nee=True)


def test_row_chunks_equal_full_frame(dev):
    """renderOpenCL's row chunking (gpu_opencl_windows.go:631-721): chunks reproduce the frame."""
    s = bc.HostScene.workload(bc.CLOSED_BOX)
    dev.upload(s, nee_triangles=True)
    d = bc.render_desc(s, 40, 30, 4, 4, rr_depth=3)
    full, gf, rays = dev.render(d, guides=True)
    parts, gp, total = [], [], 0
    for y0 in range(0, 30, 8):
        img, g, r = dev.render(d, y0=y0, rows=min(8, 30 - y0), guides=True)
        parts.append(img)
        gp.append(g)
        total += r
    assert np.array_equal(np.concatenate(parts), full) and np.array_equal(np.concatenate(gp), gf)
    assert total == rays


def test_cancel_and_bad_arguments(dev):
    s = bc.HostScene.workload(bc.CLOSED_BOX)
    dev.upload(s)
    d = bc.render_desc(s, 16, 16, 1, 2)
    with pytest.raises(bc.BeamcastError, match="cancelled"):
        dev.render(d, cancel=True)
    with pytest.raises(bc.BeamcastError):
        dev.render(d, y0=10, rows=10)
    d.spp = 0
    with pytest.raises(bc.BeamcastError):
        dev.render(d)


def test_published_path_length_on_gpu(dev):
    """GIGA GPU probe workload (benchmark.go:218-235): Showcase 640x360, 4 spp, 6 bounces, MIS and
    power sampling off, RRDepth 4 -> 1,767,805 segments over 921,600 samples (GIGA_SAMPLE_THIS_ENV.txt:60)."""
    s = bc.HostScene.workload(bc.SHOWCASE, bins=16, leaf=4)
    dev.upload(s)
    d = bc.render_desc(s, 640, 360, 4, 6, rr_depth=4, sampler=0, mis=False, power=False, seed=77)
    _, _, rays = dev.render(d)
    assert abs(rays / 921600 / (1767805 / 921600) - 1) < 0.003


def test_same_seed_fp64_and_converged_mean(dev):
    """Against the fp64 Go semantics: same seeds give nearly the same image; the mean of a
    converged render matches an independent-seed fp64 render."""
    s = bc.HostScene.workload(bc.CLOSED_BOX)
    dev.upload(s, nee_triangles=True)
    o = orc.Scene.from_host(s, nee_triangles=True)
    d = bc.render_desc(s, 48, 48, 64, 4, rr_depth=3)
    img, _, _ = dev.render(d)
    ref64, _, _, _ = o.render(orc.desc_from(d), precision=64)
    rel_rmse = np.sqrt(np.mean((img[..., :3] - ref64[..., :3]) ** 2)) / ref64[..., :3].mean()
    assert rel_rmse < 0.02
    d2 = bc.render_desc(s, 48, 48, 512, 4, rr_depth=3, seed=99)
    img2, _, _ = dev.render(d2)
    assert abs(img2[..., :3].mean() / ref64[..., :3].mean() - 1) < 0.01


def test_converged_4096spp_rmse(dev):
    """Converged images at 4096 spp against fp64 CPU renders with independent seeds: the GPU
    image is as close to an fp64 render as a second fp64 render is (no bias beyond noise)."""
    print("case: closed box 64x64 4096 spp vs fp64", flush=True)
    s = bc.HostScene.workload(bc.CLOSED_BOX)
    dev.upload(s, nee_triangles=True)
    o = orc.Scene.from_host(s, nee_triangles=True)
    kw = dict(rr_depth=3, sampler=1, mis=True, power=True)
    gpu, _, _ = dev.render(bc.render_desc(s, 64, 64, 4096, 4, seed=11, **kw))
    ref_a = o.render(orc.desc_from(bc.render_desc(s, 64, 64, 4096, 4, seed=22, **kw)), precision=64)[0]
    ref_b = o.render(orc.desc_from(bc.render_desc(s, 64, 64, 4096, 4, seed=33, **kw)), precision=64)[0]
    scale = ref_a[..., :3].mean()
    e_gpu = np.sqrt(np.mean((gpu[..., :3] - ref_a[..., :3]) ** 2)) / scale
    e_cpu = np.sqrt(np.mean((ref_b[..., :3] - ref_a[..., :3]) ** 2)) / scale
    assert e_gpu < 0.05
    # two independent 64x64 fp64 renders differ by a relative RMSE with a spread of about 3%
    assert 0.85 < e_gpu / e_cpu < 1.18, f"gpu rmse {e_gpu:.5f} vs fp64 pair {e_cpu:.5f}"
    assert abs(gpu[..., :3].mean() / scale - 1) < 0.003


_SHOWCASE_REF = {}


def showcase_oracle(s, d):
    """The oracle's fp32 render of the persistent-schedule case (computed once per module)."""
    if "ref" not in _SHOWCASE_REF:
        ref, rg, rrays, _ = orc.Scene.from_host(s, bins=16, leaf=4).render(orc.desc_from(d), precision=32,
                                                                            guides=True)
        _SHOWCASE_REF["ref"] = (ref, rg, rrays)
    return _SHOWCASE_REF["ref"]


def assert_oracle(img, rays, ref, rrays, g=None, rg=None):
    assert rays == rrays, f"path segments {rays} vs {rrays}"
    assert np.array_equal(img[..., 3], ref[..., 3])
    err = np.abs(img[..., :3] - ref[..., :3])
    bad = err > 1e-4 * np.abs(ref[..., :3]) + 1e-6
    assert not bad.any(), f"{bad.sum()} channel values outside 1e-4 rel; max err {err.max()}"
    if g is not None:
        assert np.array_equal(g, rg)


@pytest.mark.parametrize("pool,iters,segment,msort", [(4096, 4, 0, 0), (50000, 8, 1024, 0), (3000, 6, 512, 1),
                                                      (20000, 4, 256, 1), (0, 0, 0, 0)])
def test_persistent_schedule_equals_plain(dev, pool, iters, segment, msort):
    """In-place regeneration, hipGraph replay and the autotuner (pool=0) change the schedule but
    not the image: per-sample results are reduced in sample order, so rgba is bit-identical to
    the plain schedule's, and within 1e-4 of the oracle's fp32 render of the same samples."""
    print(f"case: persistent pool={pool} iters={iters} segment={segment} material sort={msort}", flush=True)
    s = bc.HostScene.workload(bc.SHOWCASE, bins=16, leaf=4)
    dev.upload(s)
    d = bc.render_desc(s, 64, 40, 6, 6, rr_depth=2, firefly_clamp=20)
    plain, gp, rp = dev.render(d, guides=True)
    dp = bc.render_desc(s, 64, 40, 6, 6, rr_depth=2, firefly_clamp=20, mode=1, pool=pool, iters=iters,
                        segment=segment, material_sort=msort)
    pers, gq, rq = dev.render(dp, guides=True)
    assert rp == rq
    assert np.array_equal(plain, pers) and np.array_equal(gp, gq)
    ref, rg, rrays = showcase_oracle(s, d)
    assert_oracle(pers, rq, ref, rrays, gq, rg)
    t = dev.tuning()
    assert t["pool"] > 0 and t["iters_per_graph"] >= 2 and t["pool"] % t["segment"] == 0
    if pool:
        assert t["segment"] == (segment or 256) and t["plain"] == 0
    # a second call reuses the captured graph (and the tuned shape)
    again, _, _ = dev.render(dp, guides=True)
    assert np.array_equal(again, pers)


@pytest.mark.parametrize("window,pool,iters", [(2, 4096, 4), (1, 3000, 6), (4, 0, 0)])
def test_persistent_sample_windows(dev, window, pool, iters):
    """Samples split into windows (reserved[5] caps the samples per window, as the per-sample
    result buffer does for large frames): the per-window parameter rewrite, the accumulator reset
    of the first window and the per-window reductions give the plain image bit for bit and the
    oracle's within 1e-4."""
    print(f"case: persistent windows of {window} samples, pool {pool}", flush=True)
    s = bc.HostScene.workload(bc.SHOWCASE, bins=16, leaf=4)
    dev.upload(s)
    d = bc.render_desc(s, 64, 40, 6, 6, rr_depth=2, firefly_clamp=20)
    plain, gp, rp = dev.render(d, guides=True)
    dp = bc.render_desc(s, 64, 40, 6, 6, rr_depth=2, firefly_clamp=20, mode=1, pool=pool, iters=iters,
                        window=window)
    pers, gq, rq = dev.render(dp, guides=True)
    t = dev.tuning()
    if not t["plain"]:
        assert t["windows"] == -(-6 // window), t
    assert rp == rq and np.array_equal(plain, pers) and np.array_equal(gp, gq)
    ref, rg, rrays = showcase_oracle(s, d)
    assert_oracle(pers, rq, ref, rrays, gq, rg)
    # a row chunk with windows: rows 7..29 against the oracle's render of the same rows
    pc, _, rc = dev.render(dp, y0=7, rows=22)
    refc, _, rrc, _ = orc.Scene.from_host(s, bins=16, leaf=4).render(orc.desc_from(d), precision=32, y0=7, rows=22)
    assert_oracle(pc, rc, refc, rrc)
    with pytest.raises(bc.BeamcastError, match="reserved"):
        bad = bc.render_desc(s, 16, 16, 2, 2, mode=1, pool=4096, iters=4, window=-1)
        dev.render(bad)


@pytest.mark.parametrize("w,h,spp,spp_whole", [(512, 300, 32, 27), (4096, 2, 1030, 512)])
def test_plain_row_bands(dev, w, h, spp, spp_whole):
    """A plain render whose samples exceed one pass (2^22 paths) runs band by band: bands of whole
    rows holding all their samples (512x300 at 32 spp: bands of 256 and 44 rows, one pass each),
    or single rows in several passes when a row's samples exceed a pass (4096 wide at 1030 spp:
    passes of 1024 and 6 samples).  The image, guides and segment count equal those of
    whole-frame passes (samples_per_pass) bit for bit, and rows across the first band boundary
    match the oracle."""
    print(f"case: plain row bands {w}x{h} {spp} spp", flush=True)
    s = bc.HostScene.workload(bc.SHOWCASE, bins=16, leaf=4)
    dev.upload(s)
    bounces = 3 if spp < 100 else 1
    kw = dict(rr_depth=2, firefly_clamp=20)
    img, g, rays = dev.render(bc.render_desc(s, w, h, spp, bounces, **kw), guides=True)
    whole = bc.render_desc(s, w, h, spp, bounces, samples_per_pass=spp_whole, **kw)
    img2, g2, rays2 = dev.render(whole, guides=True)
    assert np.array_equal(img, img2) and np.array_equal(g, g2) and rays == rays2
    if h > 256:
        o = orc.Scene.from_host(s, bins=16)
        ref, rg, _, _ = o.render(orc.desc_from(whole), precision=32, y0=254, rows=4, guides=True)
        got = img[254:258]
        assert np.array_equal(got[..., 3], ref[..., 3])
        err = np.abs(got[..., :3] - ref[..., :3])
        assert not (err > 1e-4 * np.abs(ref[..., :3]) + 1e-6).any(), err.max()
        assert np.array_equal(g[254:258], rg)


@pytest.mark.parametrize("mode", [1, 0])
def test_render_ahead_chunks(dev, mode):
    """bc_render renders the rows after a chunk with it and serves the following chunks from
    them (persistent and plain schedules): each chunk equals the chunk rendered alone and its
    slice of the whole frame, the per-chunk segment counts add up to the frame's, and a material
    update, a refit, another description or a later request for guides is never answered with
    stale rows."""
    print(f"case: render-ahead of row chunks, schedule {mode}, showcase 64x40", flush=True)
    s = bc.HostScene.workload(bc.SHOWCASE, bins=16, leaf=4)
    dev.upload(s)
    sched = dict(mode=1, pool=4096, iters=4) if mode == 1 else dict(mode=0)
    dp = bc.render_desc(s, 64, 40, 6, 6, rr_depth=2, firefly_clamp=20, **sched)
    full, gfull, rfull = dev.render(dp, guides=True)

    def chunks(rows=8, guides=True):
        return [dev.render(dp, y0=y, rows=min(rows, 40 - y), guides=guides) for y in range(0, 40, rows)]

    ahead = chunks()
    assert np.array_equal(np.concatenate([c[0] for c in ahead]), full)
    assert np.array_equal(np.concatenate([c[1] for c in ahead]), gfull)
    assert sum(c[2] for c in ahead) == rfull
    # one pair of chunk buffers reused by every chunk, as the cgo shim renders a frame
    raw, raw_guide = np.empty((8, 64, 4), np.float32), np.empty((8, 64, 3), np.uint32)
    for y in range(0, 40, 8):
        a, g, _ = dev.render(dp, y0=y, rows=8, guides=True, out=raw, guides_out=raw_guide)
        assert a is raw and g is raw_guide
        assert np.array_equal(raw, full[y:y + 8]) and np.array_equal(raw_guide, gfull[y:y + 8])
    with pytest.raises(ValueError):
        dev.render(dp, y0=0, rows=8, out=np.empty((8, 64, 4), np.float64))
    dev.set_render_ahead(False)
    alone = chunks()
    dev.set_render_ahead(True)
    for a, b in zip(ahead, alone):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
    # the block is read back once into the pinned staging buffers; a staged upload and download
    # between two chunks reuse both buffers, and the next chunk is then read from the device
    dev.render(dp, y0=0, rows=8, guides=True)
    big = np.random.default_rng(3).integers(0, 256, size=(17 << 20) + 5, dtype=np.uint8)
    p = dev.to_device(big)
    back = np.zeros_like(big)
    dev.download(p, back)
    dev.free(p)
    assert np.array_equal(back, big)
    got = dev.render(dp, y0=8, rows=8, guides=True)
    assert np.array_equal(got[0], ahead[1][0]) and np.array_equal(got[1], ahead[1][1]) and got[2] == ahead[1][2]
    got = dev.render(dp, y0=16, rows=24, guides=True)
    assert np.array_equal(got[0], full[16:]) and np.array_equal(got[1], gfull[16:])

    def fresh(desc, y0, rows, guides=False):
        dev.set_render_ahead(False)
        out = dev.render(desc, y0=y0, rows=rows, guides=guides)
        dev.set_render_ahead(True)
        return out

    # material update between two chunks of one block
    kind, mat, geom, mk, md = s.export()
    dev.render(dp, y0=0, rows=8)
    md2 = md.copy()
    md2[:, 0:3] *= 0.5
    dev.update_materials(mk, md2)
    got = dev.render(dp, y0=8, rows=8)
    want = fresh(dp, 8, 8)
    assert np.array_equal(got[0], want[0]) and got[2] == want[2]
    assert not np.array_equal(got[0], ahead[1][0])
    dev.update_materials(mk, md)
    # refit between two chunks
    dev.render(dp, y0=0, rows=8)
    s.translate(0, s.counts()[0] // 2, (0.05, 0.0, 0.0))
    dev.refit(s)
    got = dev.render(dp, y0=16, rows=8)
    want = fresh(dp, 16, 8)
    assert np.array_equal(got[0], want[0]) and got[2] == want[2]
    # another description (seed) with the block of the first still held
    dev.render(dp, y0=0, rows=8)
    d2 = bc.render_desc(s, 64, 40, 6, 6, rr_depth=2, firefly_clamp=20, **sched)
    d2.seed = dp.seed + 1
    got = dev.render(d2, y0=8, rows=8)
    assert np.array_equal(got[0], fresh(d2, 8, 8)[0])
    # guides asked for after a block rendered without them
    dev.render(dp, y0=0, rows=8)
    got = dev.render(dp, y0=8, rows=8, guides=True)
    want = fresh(dp, 8, 8, guides=True)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    if mode == 0:
        # spatial reuse reads neighbouring reservoirs among the rendered rows: a chunk is never
        # rendered with the rows after it
        ds = bc.render_desc(s, 64, 40, 2, 3, rr_depth=2, restir=8, spatial=4, radius=6)
        dev.render(ds, y0=0, rows=8)
        got = dev.render(ds, y0=8, rows=8)
        assert np.array_equal(got[0], fresh(ds, 8, 8)[0])


SUN = dict(sun_dir=(0.3, 0.8, 0.2), sun=(3.0, 3.0, 3.0))


@pytest.mark.parametrize("mode", [0, 1])
def test_shadow_refill_equals_one_ray_per_thread(dev, mode):
    """Shadow rays with lane refill (a wave hands its idle lanes the next shadow rays of its runs,
    across segments) give the image of one shadow ray per thread, bit for bit, in both schedules."""
    print(f"case: shadow refill, schedule {mode}", flush=True)
    s = bc.HostScene.workload(bc.TERRAIN, param=40, bins=16, leaf=4)
    dev.upload(s)
    kw = dict(rr_depth=3, sampler=1, sky_mode=1, sky_color=(0.3, 0.3, 0.3), restir=8, mode=mode, **SUN)
    if mode:
        kw.update(pool=6000, iters=4)
    imgs = []
    for lanes in (0, 8, 64, 1):
        dev.set_trace_refill(0, lanes)
        img, _, rays = dev.render(bc.render_desc(s, 72, 40, 5, 6, **kw))
        imgs.append((img, rays))
    dev.set_trace_refill(-1, 8)
    for img, rays in imgs[1:]:
        assert rays == imgs[0][1] and np.array_equal(img, imgs[0][0])


@pytest.mark.parametrize("kind,param,sun,nee", [(bc.CLOSED_BOX, 0, False, True), (bc.TERRAIN, 48, True, False),
                                                (bc.SHOWCASE, 0, True, False)])
def test_restir_parity(dev, kind, param, sun, nee):
    """Resampled direct lighting (32 candidates over sun, emitters and sky) matches the oracle per
    pixel, including metal and glass vertices (Showcase) and the skipped BSDF-sampled light."""
    print(f"case: restir scene {kind} param {param} sun {sun}", flush=True)
    s = bc.HostScene.workload(kind, param=param, bins=16, leaf=4)
    extra = SUN if sun else {}
    d = bc.render_desc(s, 48, 32, 4, 6, rr_depth=3, sampler=1, sky_mode=1, sky_color=(0.3, 0.3, 0.3), restir=32,
                       **extra)
    img, _, _ = parity(dev, s, d, nee=nee, bins=16)
    assert img[..., :3].mean() > 0.02


def test_restir_persistent_equals_plain(dev):
    print("case: restir persistent terrain 32 cells", flush=True)
    s = bc.HostScene.workload(bc.TERRAIN, param=32, bins=16, leaf=4)
    dev.upload(s)
    kw = dict(rr_depth=3, sampler=1, sky_mode=1, sky_color=(0.3, 0.3, 0.3), restir=16, **SUN)
    plain, _, rp = dev.render(bc.render_desc(s, 64, 48, 5, 8, **kw))
    pers, _, rq = dev.render(bc.render_desc(s, 64, 48, 5, 8, mode=1, pool=5000, iters=4, **kw))
    assert rp == rq and np.array_equal(plain, pers)


def test_restir_argument_errors(dev):
    print("case: restir argument errors", flush=True)
    s = bc.HostScene.workload(bc.CLOSED_BOX)
    dev.upload(s, nee_triangles=True)
    with pytest.raises(bc.BeamcastError):
        dev.render(bc.render_desc(s, 16, 16, 1, 2, **SUN))          # a sun needs restir > 0
    with pytest.raises(bc.BeamcastError):
        dev.render(bc.render_desc(s, 16, 16, 1, 2, restir=65))       # at most 64 candidates
    with pytest.raises(bc.BeamcastError):
        dev.forward(bc.render_desc(s, 16, 16, 1, 2, restir=8))       # adjoint: one-sample NEE only


@pytest.mark.parametrize("kind,param,sun,nee,k,radius", [(bc.CLOSED_BOX, 0, False, True, 8, 8),
                                                          (bc.TERRAIN, 32, True, False, 4, 16),
                                                          (bc.SHOWCASE, 0, True, False, 16, 3)])
def test_spatial_reuse_parity(dev, kind, param, sun, nee, k, radius):
    """ReSTIR DI spatial reuse at primary vertices matches the oracle per pixel: the same
    neighbours, similarity tests, measure conversions and streaming choice in fp32."""
    print(f"case: spatial reuse scene {kind} k {k} radius {radius}", flush=True)
    s = bc.HostScene.workload(kind, param=param, bins=16, leaf=4)
    extra = SUN if sun else {}
    d = bc.render_desc(s, 48, 32, 4, 5, rr_depth=3, sampler=1, sky_mode=1, sky_color=(0.3, 0.3, 0.3), restir=8,
                       spatial=k, radius=radius, **extra)
    img, _, _ = parity(dev, s, d, nee=nee, bins=16)
    assert img[..., :3].mean() > 0.02


def test_spatial_reuse_row_chunks(dev):
    """Neighbours come from the rendered rows: a chunk render equals the oracle's render of the
    same rows."""
    print("case: spatial reuse in a row chunk", flush=True)
    s = bc.HostScene.workload(bc.CLOSED_BOX)
    dev.upload(s, nee_triangles=True)
    d = bc.render_desc(s, 40, 40, 3, 3, rr_depth=3, sampler=1, restir=8, spatial=8, radius=6)
    img, _, rays = dev.render(d, y0=13, rows=9)
    ref, _, rrays, _ = orc.Scene.from_host(s, nee_triangles=True).render(orc.desc_from(d), precision=32, y0=13, rows=9)
    assert rays == rrays
    err = np.abs(img[..., :3] - ref[..., :3])
    assert not (err > 1e-4 * np.abs(ref[..., :3]) + 1e-6).any()


def test_spatial_reuse_errors(dev):
    print("case: spatial reuse argument errors", flush=True)
    s = bc.HostScene.workload(bc.CLOSED_BOX)
    dev.upload(s, nee_triangles=True)
    with pytest.raises(bc.BeamcastError, match="spatial"):
        dev.render(bc.render_desc(s, 16, 16, 1, 2, spatial=4))                   # needs restir > 0
    with pytest.raises(bc.BeamcastError, match="spatial"):
        dev.render(bc.render_desc(s, 16, 16, 1, 2, restir=8, spatial=4, mode=1))  # plain schedule only
    with pytest.raises(bc.BeamcastError):
        dev.render(bc.render_desc(s, 16, 16, 1, 2, restir=8, spatial=17))


@pytest.mark.parametrize("mode", [0, 1])
def test_render_ahead_block_beyond_pinned_buffers(dev, mode):
    """A block whose image exceeds one 16 MB pinned staging buffer (1280x1000: 20.5 MB of rgba)
    is not held on the host: each 64-row chunk is read from the device block, the next one's copy
    started into a pinned buffer as the last one is served.  Chunks equal the whole-frame render
    bit for bit, with guides, on both schedules, also when a staged copy reuses that buffer
    between two chunks and when the chunk heights change."""
    print(f"case: render-ahead block beyond the pinned buffers, schedule {mode}, 1280x1000", flush=True)
    s = bc.HostScene.workload(bc.SHOWCASE, bins=16, leaf=4)
    dev.upload(s)
    sched = dict(mode=1, pool=1 << 18, iters=4) if mode == 1 else dict(mode=0)
    d = bc.render_desc(s, 1280, 1000, 1, 2, rr_depth=2, firefly_clamp=20, **sched)
    dev.set_render_ahead(False)
    full, gfull, rfull = dev.render(d, guides=True)
    dev.set_render_ahead(True)
    raw, raw_guide = np.empty((64, 1280, 4), np.float32), np.empty((64, 1280, 3), np.uint32)
    total = 0
    for y in range(0, 1000, 64):
        rows = min(64, 1000 - y)
        a, g, r = dev.render(d, y0=y, rows=rows, guides=True, out=raw[:rows], guides_out=raw_guide[:rows])
        assert np.array_equal(a, full[y:y + rows]) and np.array_equal(g, gfull[y:y + rows]), f"rows {y}.."
        total += r
        if y == 128:
            # a staged upload and download between two chunks reuse the pinned buffer that holds
            # the next chunk's copy
            big = np.random.default_rng(5).integers(0, 256, size=(33 << 20) + 7, dtype=np.uint8)
            p = dev.to_device(big)
            back = np.zeros_like(big)
            dev.download(p, back)
            dev.free(p)
            assert np.array_equal(back, big)
    assert total == rfull
    # chunks of another height than the one copied ahead, and without guides
    for y, rows in ((0, 64), (64, 32), (96, 64)):
        a, g, _ = dev.render(d, y0=y, rows=rows)
        assert g is None and np.array_equal(a, full[y:y + rows])
