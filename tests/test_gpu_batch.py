"""Packed multi-scene batches (bc_batch_upload / bc_batch_render_device) through the C ABI.

A batch render must equal rendering each scene alone (same descriptions), bit for bit, under
both schedules: scene after scene (plain) and one persistent schedule over all scenes' samples,
also split into sample windows; and every scene of the batch must match the oracle's fp32 render
of that scene within 1e-4 (per pixel, same sample indices).  The batch-scene workload (boxes,
spheres, metal, glass, quad lights) gets its own oracle parity case too.
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


def scenes_and_descs():
    box = bc.HostScene.workload(bc.CLOSED_BOX)
    show = bc.HostScene.workload(bc.SHOWCASE, bins=16, leaf=4)
    b1 = bc.HostScene.workload(bc.BATCH_SCENE, seed=0xC5, param=1)
    terr = bc.HostScene.workload(bc.TERRAIN, param=24, bins=16, leaf=4)
    descs = [
        bc.render_desc(box, 40, 32, 6, 5, rr_depth=3, sampler=1),
        bc.render_desc(show, 48, 24, 3, 6, rr_depth=2, sampler=2, firefly_clamp=20, seed=77),
        bc.render_desc(b1, 64, 36, 4, 8, rr_depth=4, sampler=1),
        bc.render_desc(terr, 32, 32, 5, 4, rr_depth=3, sky_mode=1, sky_color=(0.3, 0.3, 0.3), restir=16,
                       sun_dir=(0.3, 0.8, 0.2), sun=(3.0, 3.0, 3.0)),
    ]
    return [box, show, b1, terr], descs


_ORACLE = {}


def oracle_images(scenes, descs):
    """The oracle's fp32 render of every batch scene on its own (computed once per module)."""
    if "imgs" not in _ORACLE:
        out = []
        for s, d in zip(scenes, descs):
            # hits (equal distances resolved to the smaller index) do not depend on the tree
            o = orc.Scene.from_host(s, nee_triangles=True)
            ref, _, rrays, _ = o.render(orc.desc_from(d), precision=32)
            out.append((ref, rrays))
        _ORACLE["imgs"] = out
    return _ORACLE["imgs"]


@pytest.mark.parametrize("mode,pool,iters,segment,msort,window", [(0, 0, 0, 0, 0, 0), (1, 3000, 4, 0, 0, 0),
                                                                  (1, 20000, 8, 1024, 1, 0), (1, 0, 0, 0, 0, 0),
                                                                  (1, 3000, 4, 0, 0, 2), (1, 0, 0, 0, 0, 1)])
def test_batch_equals_single_scene_renders(dev, mode, pool, iters, segment, msort, window):
    """Batch = each scene alone, bit for bit, and every scene of the batch within 1e-4 of the
    oracle's render of that scene; window > 0 splits the batch's samples into windows of at most
    that many samples (reserved[5]), one sample-window loop over all scenes."""
    print(f"case: batch mode {mode} pool {pool} iters {iters} segment {segment} material sort {msort} "
          f"window {window}", flush=True)
    scenes, descs = scenes_and_descs()
    singles, srays = [], 0
    for s, d in zip(scenes, descs):
        dev.upload(s, nee_triangles=True)
        img, _, r = dev.render(d)
        singles.append(img)
        srays += r
    dev.upload_batch(scenes, nee_triangles=True)
    for d in descs:
        d.mode = mode
        d.reserved[0], d.reserved[1], d.reserved[2], d.reserved[4] = pool, iters, segment, msort
        d.reserved[5] = window
    imgs, rays = dev.render_batch(descs)
    assert rays == srays
    for k, (a, b) in enumerate(zip(imgs, singles)):
        assert np.array_equal(a, b), f"scene {k} differs"
    if mode == 1 and window and not dev.tuning()["plain"]:
        assert dev.tuning()["windows"] == -(-max(d.spp for d in descs) // window)
    for k, (img, (ref, _)) in enumerate(zip(imgs, oracle_images(scenes, descs))):
        err = np.abs(img[..., :3] - ref[..., :3])
        bad = err > 1e-4 * np.abs(ref[..., :3]) + 1e-6
        assert not bad.any(), f"scene {k}: {bad.sum()} values outside 1e-4; max err {err.max()}"
        assert np.array_equal(img[..., 3], ref[..., 3]), f"scene {k}: path segments differ"
    # single-scene entry points refuse a resident batch
    with pytest.raises(bc.BeamcastError):
        dev.render(descs[0])


def test_batch_scene_parity(dev):
    """The batch-scene workload against the oracle (metal, glass, spheres, emissive quads)."""
    print("case: batch scene workload parity", flush=True)
    s = bc.HostScene.workload(bc.BATCH_SCENE, seed=0xC5, param=3, bins=16, leaf=4)
    dev.upload(s, nee_triangles=True)
    d = bc.render_desc(s, 48, 27, 4, 8, rr_depth=4, sampler=1, seed=11)
    img, _, rays = dev.render(d)
    o = orc.Scene.from_host(s, bins=16, leaf=4, nee_triangles=True)
    ref, _, rrays, _ = o.render(orc.desc_from(d), precision=32)
    assert rays == rrays
    err = np.abs(img[..., :3] - ref[..., :3])
    assert not (err > 1e-4 * np.abs(ref[..., :3]) + 1e-6).any(), err.max()
    assert img[..., :3].mean() > 0.01


def test_batch_argument_errors(dev):
    print("case: batch argument errors", flush=True)
    scenes, descs = scenes_and_descs()
    dev.upload_batch(scenes[:2], nee_triangles=True)
    with pytest.raises(bc.BeamcastError):
        dev.render_batch(descs[:1])  # one description per scene
