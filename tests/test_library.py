"""CPU checks of the product library: it loads, exports every declared symbol, and its host
scene mirror builds the same SAH BVH as the oracle's restatement of Scene.Build."""
import os
import re

import numpy as np
import pytest

import beamcast as bc
import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bch?_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header", ["beamcast_hip.h", "beamcast_host.h"])
def test_library_exports_every_declared_symbol(header):
    lib = bc.lib()
    names = declared(header)
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), f"{n} declared in {header} but not exported"


def test_no_device_reports_cleanly():
    ok, msg = bc.device_available()
    if not ok:
        assert msg
        with pytest.raises(bc.BeamcastError):
            bc.Device(0)


@pytest.mark.parametrize("kind,param,bins", [(bc.CLOSED_BOX, 0, 32), (bc.SHOWCASE, 0, 16), (bc.CORNELL, 0, 16),
                                             (bc.STRESS, 0, 16), (bc.SOUP, 20000, 32), (bc.BOX_RETESS, 4000, 32),
                                             (bc.TERRAIN, 48, 32)])
def test_host_build_matches_oracle_build(kind, param, bins):
    h = bc.HostScene.workload(kind, param=param, bins=bins, leaf=4)
    o = orc.Scene.from_host(h, bins=bins, leaf=4)
    hb, hl, ho = h.nodes()
    ob, ol, oo = o.nodes()
    assert np.array_equal(hb, ob) and np.array_equal(hl, ol) and np.array_equal(ho, oo)


def test_revision_tracking():
    """TestGPURevisionTracking (renderer_test.go:692-718): Build bumps both, Refit geometry only."""
    s = bc.HostScene.workload(bc.SHOWCASE, bins=16)
    t0, g0 = s.revisions()
    assert t0 > 0 and g0 > 0
    s.translate(1, 1, (0.5, 0, 0))
    s.refit()
    t1, g1 = s.revisions()
    assert t1 == t0 and g1 == g0 + 1
    s.build()
    t2, g2 = s.revisions()
    assert t2 == t1 + 1 and g2 == g1 + 1


def test_refit_matches_rebuild_boxes():
    """TestBVHRefitMatchesRebuild (renderer_test.go:464-493): refit after a move gives boxes that
    contain every primitive, and hits agree with a full rebuild."""
    s = bc.HostScene.workload(bc.STRESS, bins=16)
    s.translate(10, 50, (0.1, 0.2, -0.1))
    s.refit()
    kind, mat, geom, mk, md = s.export()
    box, link, order = s.nodes()
    for i, (l, r, st, c) in enumerate(link):
        if c:
            for p in order[st:st + c]:
                cc, rr = geom[p, :3], geom[p, 3]
                assert np.all(box[i, :3] <= cc - rr + 1e-12) and np.all(box[i, 3:] >= cc + rr - 1e-12)
    rebuilt = orc.Scene(kind, mat, geom, mk, md, bins=16, leaf=4)
    rays = np.array([[9, 5, 8, 1e-4, *(np.array([x, -0.4, -1]) / np.linalg.norm([x, -0.4, -1])), 1e30]
                     for x in np.linspace(-0.8, 0.2, 200)])
    p1, t1 = rebuilt.hit64(rays)
    p2, t2 = rebuilt.brute64(rays)
    assert np.array_equal(p1, p2)


def test_workload_rays_subrange_and_shape():
    full = bc.workload_rays(1, 64, count=1000)
    part = bc.workload_rays(1, 64, first=300, count=100)
    assert np.array_equal(full[300:400], part)
    d = full[:, 4:7]
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6)
    assert np.all(full[:, 3] == np.float32(1e-4)) and np.all(full[:, 7] == np.float32(1e30))
    coh = bc.workload_rays(0, 64, any_hit=True, count=64 * 64)
    assert np.median(coh[:, 7]) > 1.5  # rays entering the unit cube get tmax beyond its entry plane
    inc = bc.workload_rays(1, 64, any_hit=True, count=4096)
    assert inc[:, 7].min() >= 0.05 and inc[:, 7].max() <= 1.0


def test_sampling_levels():
    """TestSamplingLevels (renderer_test.go:198-212): effectiveSPP multiplies SPP by the MSAA
    factor (2, 4 or 8, else 1) and the TSAA frames (TSAA 1-3 -> 2-4), clamped to [1, 4096]."""
    assert bc.effective_spp("Preview", 2, 4, 2) == 24
    assert bc.effective_spp("Preview", 200, 8, 3) == 4096
    assert bc.effective_spp("Preview", 5, 3, 0) == 5
    assert bc.effective_spp("Preview", 5, 2, 1) == 20
    assert bc.effective_spp("Preview", 0, 4, 3) == 1


def test_ultra_realism_preset():
    """TestUltraRealismPreset (renderer_test.go:846-864): the CPU preset keeps the full transport
    and no biased effects; the OpenCL-GPU one is the megakernel-compatible subset."""
    cpu = bc.quality(backend="CPU")
    assert cpu["Quality"] == "Ultra Realism" and (cpu["Width"], cpu["Height"]) == (2560, 1440)
    assert cpu["SPP"] >= 256 and cpu["Bounces"] >= 16 and cpu["MIS"] and cpu["PowerLightSampling"]
    assert cpu["Sampler"] == "Halton"
    assert cpu["HBAO"] == 0 and cpu["HBAOPlus"] == 0 and cpu["FXAA"] == 0 and cpu["FireflyClamp"] == 0
    gpu = bc.quality(backend="OpenCL GPU")
    assert gpu["Backend"] == "OpenCL GPU" and gpu["Sampler"] == "Random"
    assert not gpu["MIS"] and not gpu["PowerLightSampling"] and not gpu["AdaptiveSampling"]
    assert gpu["SPP"] >= 256 and gpu["Bounces"] >= 16 and gpu["Denoise"]
    with pytest.raises(ValueError):
        bc.quality(backend="Vulkan")
    # the other presets as Quality (renderer.go:1064-1117) sets them
    want = {"Draft": (640, 360, 2, 4, 1, 0), "Balanced": (1280, 720, 16, 8, 2, 2), "High": (1600, 900, 48, 10, 2, 2),
            "Ultra": (1920, 1080, 96, 12, 3, 3), "anything else": (960, 540, 8, 6, 1, 1)}
    for name, (w, h, spp, b, fxaa, taa) in want.items():
        q = bc.quality(name)
        assert (q["Width"], q["Height"], q["SPP"], q["Bounces"], q["FXAA"], q["TAA"]) == (w, h, spp, b, fxaa, taa), name
    assert bc.quality("anything else")["Quality"] == "Preview"


def test_median_build_invariants():
    """BVHMode "Median" (buildNodeMedian, renderer.go:506-524): a pre-order tree whose every node
    splits its range at the middle after sorting on the largest centroid axis, leaves of at most
    BVHLeafSize primitives, boxes that contain their children, every primitive in one leaf."""
    s = bc.HostScene.workload(bc.STRESS, bins=16)
    s.build_median(leaf=4)
    box, link, order = s.nodes()
    n = len(order)
    assert sorted(order.tolist()) == list(range(n))
    kind, _, geom, _, _ = s.export()
    cen = np.where((kind == bc.TRIANGLE)[:, None], (geom[:, 0:3] + geom[:, 3:6] + geom[:, 6:9]) / 3.0, geom[:, 0:3])
    seen = np.zeros(n, np.int32)

    def walk(i, start, end):
        l, r, st, c = link[i]
        if l < 0:
            assert st == start and c == end - start and 0 < c <= 4
            seen[start:end] += 1
            return
        assert l == i + 1  # pre-order: the left child follows its parent
        mid = (start + end) // 2
        for ch in (l, r):
            assert np.all(box[ch, :3] >= box[i, :3]) and np.all(box[ch, 3:] <= box[i, 3:])
        c = cen[order[start:end]]
        ext = c.max(axis=0) - c.min(axis=0)
        axis = 1 if ext[1] > ext[0] and ext[1] >= ext[2] else (2 if ext[2] > ext[0] and ext[2] > ext[1] else 0)
        assert cen[order[start:mid], axis].max() <= cen[order[mid:end], axis].min()
        walk(l, start, mid)
        walk(r, mid, end)

    walk(0, 0, n)
    assert np.all(seen == 1)
    assert len(link) <= 2 * n - 1


@pytest.mark.parametrize("kind,param", [(bc.CLOSED_BOX, 0), (bc.SOUP, 20000), (bc.STRESS, 0), (bc.CORNELL, 0),
                                        (bc.TERRAIN, 40), (bc.BOX_RETESS, 3000)])
@pytest.mark.parametrize("leaf", [1, 4, 8, 16])
def test_bvh8_flatten_covers_every_primitive(kind, param, leaf):
    """The BVH8 collapse puts every primitive in exactly one leaf slot, one primitive per slot, its
    record at 8 * node + slot and holes elsewhere, for Go leaves of up to 16 primitives
    (BVHLeafSize 16: leaves of more than eight are split in two before the collapse)."""
    s = bc.HostScene.workload(kind, param=param, bins=16, leaf=leaf)
    nodes, depth, slots, ok = s.flatten_check()
    assert ok and nodes >= 1 and depth >= 1 and slots == s.counts()[0]


def test_bvh8_flatten_far_from_origin():
    """The grid rule far from the scene origin (origin integers |k| < 2^20): every record still
    inside all its enclosing quantised boxes."""
    s = bc.HostScene.workload(bc.SOUP, param=20_000)
    s.translate(0, s.counts()[0], (1.5e5, -2.5e5, 4.0e4))
    s.build(32, 4)
    nodes, depth, slots, ok = s.flatten_check()
    assert ok and slots == s.counts()[0]
