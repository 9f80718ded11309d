"""GPU parity of the BVH8 traversal against the oracle (through the C ABI).

Bar: closest-hit ids and distances bit-exact against the oracle's fp32 restatement of
triHit / sphereHitAccel (same operation order, equal-t resolved to the smaller primitive
index); against the fp64 Go-semantics Scene.Hit at least 99.99% identical ids, mismatches
only where |dt|/t < 1e-5.  Any hit: identical occlusion flags.
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


def soup_rays(n_per_set, width=1024, any_hit=False):
    # strided picks over the full ray sets so the sample spans the whole image / cube
    coh = bc.workload_rays(0, width, any_hit=any_hit, count=width * width)[:: max(1, width * width // n_per_set)]
    inc = bc.workload_rays(1, width, any_hit=any_hit, count=n_per_set)
    return np.concatenate([coh[:n_per_set], inc])


def check_closest(dev, host, rays, fp64_check=True):
    dev.upload(host)
    got = dev.trace(rays)
    o = orc.Scene.from_host(host)
    prim, t = o.hit32(rays)
    bad = np.nonzero((got["prim"] != prim) | (got["t"] != t))[0]
    assert bad.size == 0, f"{bad.size} rays differ from the fp32 oracle, first {bad[:5]}"
    if fp64_check:
        p64, t64 = o.hit64(rays.astype(np.float64))
        mism = np.nonzero(got["prim"] != p64)[0]
        assert mism.size <= 1e-4 * len(rays), f"{mism.size} id mismatches vs fp64"
        kind, _, geom, _, _ = host.export()
        for i in mism:
            # allowed: equal-distance ties at shared edges (|dt|/t < 1e-5), or a ray that grazes
            # an edge of one of the two primitives within single-precision rounding
            if got["prim"][i] >= 0 and p64[i] >= 0 and abs(got["t"][i] - t64[i]) / t64[i] < 1e-5:
                continue
            cands = [p for p in (got["prim"][i], p64[i]) if p >= 0]
            assert any(kind[p] == bc.TRIANGLE and edge_distance(rays[i], geom[p]) < fp32_edge_tolerance(rays[i], geom[p])
                       for p in cands), f"ray {i}: gpu {got['prim'][i]} t={got['t'][i]} vs fp64 {p64[i]} t={t64[i]}"
    return got


def fp32_edge_tolerance(ray, abc):
    """Barycentric uncertainty of the single-precision test: vertices and origin are rounded to
    fp32 (absolute error ~ 2^-24 |coord|), which moves the barycentrics by that error over the
    shortest edge, amplified by 1/sin of the ray's angle to the triangle plane."""
    o, d = ray[:3].astype(np.float64), ray[4:7].astype(np.float64)
    a, b, c = abc[0:3], abc[3:6], abc[6:9]
    e1, e2 = b - a, c - a
    det = abs(e1 @ np.cross(d, e2)) / (np.linalg.norm(e1) * np.linalg.norm(e2) * np.linalg.norm(d))
    scale = max(np.abs(abc).max(), np.abs(o).max())
    edge = min(np.linalg.norm(e1), np.linalg.norm(e2), np.linalg.norm(e2 - e1))
    return 8 * 2.0**-23 * scale / edge / max(det, 1e-12)


def edge_distance(ray, abc):
    """Smallest barycentric distance of the ray's crossing of the triangle plane from an edge."""
    o, d = ray[:3].astype(np.float64), ray[4:7].astype(np.float64)
    a, b, c = abc[0:3], abc[3:6], abc[6:9]
    e1, e2 = b - a, c - a
    h = np.cross(d, e2)
    det = e1 @ h
    if abs(det) < 1e-300:
        return 0.0
    s = o - a
    u = (s @ h) / det
    v = (d @ np.cross(s, e1)) / det
    return float(min(abs(u), abs(v), abs(1 - u - v), abs(1 - u), abs(1 - v)))


def test_soup_closest_and_any(dev):
    print("case: soup 200k, 2x32768 rays", flush=True)
    s = bc.HostScene.workload(bc.SOUP, param=200_000)
    rays = soup_rays(32768)
    got = check_closest(dev, s, rays)
    assert (got["prim"] >= 0).mean() > 0.3
    o = orc.Scene.from_host(s)
    ra = soup_rays(32768, any_hit=True)
    occ = dev.trace(ra, any_hit=True)
    assert np.array_equal(occ, o.any32(ra))
    assert 0.02 < occ.mean() < 0.98


@pytest.mark.parametrize("kind", [bc.STRESS, bc.SHOWCASE, bc.CORNELL, bc.CLOSED_BOX])
def test_builtin_scenes(dev, kind):
    print(f"case: scene kind {kind}", flush=True)
    s = bc.HostScene.workload(kind, bins=16, leaf=4)
    pos = s.camera[:3]
    rng = np.random.default_rng(kind)
    d = rng.normal(size=(20000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.tile(pos, (20000, 1)) + rng.normal(size=(20000, 3)) * 0.5
    rays = np.column_stack([o, np.full(20000, 1e-4), d, np.full(20000, 1e30)]).astype(np.float32)
    check_closest(dev, s, rays)
    rays[:, 7] = rng.random(20000) * 10
    occ = dev.trace(rays, any_hit=True)
    assert np.array_equal(occ, orc.Scene.from_host(s, bins=16).any32(rays))


def test_edge_rays(dev):
    """Axis-parallel directions (zero components), origins on box faces, empty ranges."""
    print("case: edge rays", flush=True)
    s = bc.HostScene.workload(bc.CLOSED_BOX)
    dev.upload(s)
    rows = []
    for o in ([0, 1, 1], [0.5, 0.5, -1.0], [-1, 1, -1], [0.35, 0.0, -0.7], [0, 1.98, -1]):
        for d in ([1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [0.6, 0.8, 0]):
            for tmax in (1e30, 0.5, 1e-4):
                rows.append([*o, 1e-4, *d, tmax])
    rows.append([0, 1, 1, 2.0, 0, 0, -1, 1.0])  # tmin > tmax: no hit
    rays = np.array(rows, np.float32)
    check_closest(dev, s, rays, fp64_check=False)
    occ = dev.trace(rays, any_hit=True)
    assert np.array_equal(occ, orc.Scene.from_host(s).any32(rays))
    assert dev.trace(np.zeros((0, 8), np.float32)).shape == (0,)


def test_full_size_any_matches_closest(dev):
    """1M-triangle soup, 2 x 1M rays: any-hit(tmax) == (closest hit exists with t <= tmax)."""
    print("case: 1M soup property", flush=True)
    s = bc.HostScene.workload(bc.SOUP, param=1_000_000)
    dev.upload(s)
    st = dev.stats()
    assert st["prims"] == 1_000_000 and st["depth8"] >= 2
    for set_id in (0, 1):
        rays = bc.workload_rays(set_id, 1024, any_hit=True, count=1024 * 1024)
        closest = rays.copy()
        closest[:, 7] = 1e30
        h = dev.trace(closest)
        occ = dev.trace(rays, any_hit=True)
        want = (h["prim"] >= 0) & (h["t"] <= rays[:, 7])
        assert np.array_equal(occ.astype(bool), want)


def test_scene_encoding_errors(dev):
    s = bc.HostScene.workload(bc.CLOSED_BOX)
    vp = s.view()
    v = vp.contents
    link = np.ctypeslib.as_array(v.node_link, shape=(v.nnode * 4,))
    saved = link.copy()
    leaf = int(np.nonzero(link[3::4] > 0)[0][0])
    link[4 * leaf + 3] = 17  # validateGPUSceneEncoding: leaf count > 16 (gpu_opencl_windows.go:470-482)
    with pytest.raises(bc.BeamcastError, match="unsupported count"):
        dev.upload_view(vp, 99, 99)
    link[:] = saved
    dev.upload_view(vp, 100, 100)


@pytest.mark.parametrize("n", [1, 63, 1000, 70_001])
def test_refill_schedules(dev, n):
    """Lane refill (a wave hands finished lanes new rays of its run) returns the results of one
    ray per thread, for closest and any hit, for batch sizes that leave partial waves and runs."""
    print(f"case: refill schedules, {n} rays", flush=True)
    s = bc.HostScene.workload(bc.SOUP, param=100_000)
    dev.upload(s)
    rays = np.concatenate([bc.workload_rays(0, 512, count=512 * 512)[::7], bc.workload_rays(1, 512, count=70_000)])
    rays = rays[np.random.default_rng(n).permutation(len(rays))[:n]]
    ra = rays.copy()
    ra[:, 7] = np.random.default_rng(1).random(n) * 0.5
    o = orc.Scene.from_host(s)
    prim, t = o.hit32(rays)
    occ = o.any32(ra)
    for closest, anyh in ((0, 0), (1, 1), (8, 8), (33, 64), (64, 4), (-1, 8)):
        dev.set_trace_refill(closest, anyh)
        got = dev.trace(rays)
        assert np.array_equal(got["prim"], prim) and np.array_equal(got["t"], t), (closest, anyh)
        assert np.array_equal(dev.trace(ra, any_hit=True), occ), (closest, anyh)
    dev.set_trace_refill(-1, 8)


def test_automatic_refill_coherent_and_incoherent(dev):
    """The automatic closest-hit schedule (the default) on batches above the probe's 65,536 rays,
    one coherent (consecutive pinhole primaries) and one incoherent: the oracle's results either way."""
    print("case: automatic closest-hit refill, 70,001 coherent and incoherent rays", flush=True)
    s = bc.HostScene.workload(bc.SOUP, param=100_000)
    dev.upload(s)
    dev.set_trace_refill(-1, 8)
    o = orc.Scene.from_host(s)
    for ray_set in (0, 1):
        rays = bc.workload_rays(ray_set, 512, count=512 * 512)[:70_001]
        prim, t = o.hit32(rays)
        got = dev.trace(rays)
        assert np.array_equal(got["prim"], prim) and np.array_equal(got["t"], t), ray_set


@pytest.mark.parametrize("bits", [2, 3, 4])
def test_ray_order(dev, bits):
    """Binned queries (the batch counting-sorted by origin cell and direction octant, each result
    written back to its ray's index) return the results of the unbinned batch, for closest and
    any hit, with and without lane refill; origins outside the scene box and NaN origins fall into
    border bins."""
    print(f"case: ray order, cell bits {bits}", flush=True)
    s = bc.HostScene.workload(bc.SOUP, param=100_000)
    dev.upload(s)
    n = 65536 + 4099
    rng = np.random.default_rng(bits)
    rays = bc.workload_rays(1, 512, count=n)
    rays[:64, :3] = rng.uniform(-4.0, 5.0, (64, 3)).astype(np.float32)
    rays[64:72, :3] = np.array([0.0, 0.0, 0.0], np.float32)  # a corner of the scene box
    rays[72:80, 0] = np.nan
    ra = rays.copy()
    ra[:, 7] = rng.random(n).astype(np.float32) * 0.5
    ok = slice(80, None)  # the oracle comparison leaves the NaN rows out
    o = orc.Scene.from_host(s)
    prim, t = o.hit32(rays[ok])
    occ = o.any32(ra[ok])
    base = None
    for mode in (0, 1, 2):
        dev.set_trace_order(mode, bits)
        for closest, anyh in ((0, 8), (0, 0), (8, 8), (-1, 8)):
            dev.set_trace_refill(closest, anyh)
            got = dev.trace(rays)
            binned = dev.trace_order_last()
            assert binned[0] == (1 if mode else 0), (mode, binned)
            g_any = dev.trace(ra, any_hit=True)
            assert np.array_equal(got["prim"][ok], prim) and np.array_equal(got["t"][ok], t), (mode, closest)
            assert np.array_equal(g_any[ok], occ), (mode, anyh)
            if base is None:
                base = (got.copy(), g_any.copy())
            assert np.array_equal(got["prim"], base[0]["prim"]) and np.array_equal(got["t"].view(np.int32),
                                                                                    base[0]["t"].view(np.int32))
            assert np.array_equal(g_any, base[1])
    dev.set_trace_refill(-1, 8)
    dev.set_trace_order(0, 4)


def test_ray_order_decision(dev):
    """The automatic mode bins an incoherent batch and leaves a coherent one (pinhole primaries)
    in the caller's order; batches below 65,536 rays run no binning pass."""
    print("case: ray order decision", flush=True)
    s = bc.HostScene.workload(bc.SOUP, param=100_000)
    dev.upload(s)
    dev.set_trace_order(2, 4)
    dev.trace(bc.workload_rays(1, 512, count=1 << 17))
    b, diff = dev.trace_order_last()
    assert b == 1 and diff > 3000
    dev.trace(bc.workload_rays(0, 512, count=1 << 17))
    b, diff = dev.trace_order_last()
    assert b == 0 and 0 <= diff < 200
    dev.trace(bc.workload_rays(1, 512, count=1000))
    assert dev.trace_order_last() == (0, -1)
    dev.set_trace_order(0, 4)
    dev.trace(bc.workload_rays(1, 512, count=1 << 17))
    assert dev.trace_order_last() == (0, -1)
    with pytest.raises(bc.BeamcastError):
        dev.set_trace_order(3, 4)
    with pytest.raises(bc.BeamcastError):
        dev.set_trace_order(1, 5)
    dev.set_trace_order(0, 4)


def test_sah_and_median_agree(dev):
    """TestSAHAndMedianAgree (renderer_test.go:415-441): the Stress scene built as a median tree
    and as a SAH tree gives the same hits for 400 rays from the camera.  Through the BVH8 path the
    hits are bit-identical (equal distances resolve to the smaller primitive index, so the tree
    cannot change a result) and equal to the oracle's fp32 hits; so are rendered images."""
    print("case: SAH vs median tree", flush=True)
    s = bc.HostScene.workload(bc.STRESS, bins=16, leaf=4)
    pos, target = np.array(s.camera[:3]), np.array(s.camera[3:6])
    rng = np.random.default_rng(42)
    off = (rng.random((400, 3)) - 0.5) * np.array([14.0, 5.0, 14.0])
    d = target + off - pos
    rays = np.column_stack([np.tile(pos, (400, 1)), np.full(400, 1e-4), d, np.full(400, 1e30)]).astype(np.float32)
    dev.upload(s)
    sah = dev.trace(rays)
    desc = bc.render_desc(s, 48, 32, 4, 4, rr_depth=3)
    img_sah, _, rays_sah = dev.render(desc)
    s.build_median(leaf=4)
    dev.upload(s)
    med = dev.trace(rays)
    img_med, _, rays_med = dev.render(desc)
    assert (sah["prim"] >= 0).mean() > 0.5
    assert np.array_equal(sah["prim"], med["prim"]) and np.array_equal(sah["t"].view(np.int32), med["t"].view(np.int32))
    prim, t = orc.Scene.from_host(bc.HostScene.workload(bc.STRESS, bins=16, leaf=4), bins=16).hit32(rays)
    assert np.array_equal(med["prim"], prim) and np.array_equal(med["t"], t)
    assert rays_sah == rays_med and np.array_equal(img_sah, img_med)


def test_trace_counts(dev):
    """The counting traversal (bc_trace_counts, the algorithmic bytes of the benchmark's roofline):
    the same results as the timed query, and per-ray counts that obey what the traversal must do --
    the root is visited by every ray and no node twice; a ray that hits tests at least its own
    record and reads the id table for its index; a ray that misses reads no id; steps are at least
    node visits.  A one-triangle scene pins the counts exactly (one node, one record test
    on a hit, none on a miss)."""
    print("case: counting traversal", flush=True)
    one = bc.HostScene()
    m = one.add_material(0, (0.5, 0.5, 0.5))
    one.add_triangle((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), m)
    one.build(16, 4)
    soup = bc.HostScene.workload(bc.SOUP, param=50_000)
    hit_ray = [0.2, 0.2, 1.0, 1e-4, 0.0, 0.0, -1.0, 1e30]
    miss_ray = [5.0, 5.0, 1.0, 1e-4, 0.0, 0.0, -1.0, 1e30]
    for scene, rays in ((one, np.array([hit_ray, miss_ray], np.float32)),
                        (soup, np.concatenate([bc.workload_rays(0, 256, count=256 * 256)[::5],
                                               bc.workload_rays(1, 256, count=20_000)]))):
        dev.upload(scene)
        st = dev.stats()
        n = len(rays)
        d_rays = dev.to_device(np.ascontiguousarray(rays, np.float32))
        for any_hit in (False, True):
            out_b = 4 if any_hit else 8
            d_out, d_ref, d_cnt = dev.alloc(n * out_b), dev.alloc(n * out_b), dev.alloc(n * 16)
            dev.trace_counts(d_rays, n, any_hit, d_out, d_cnt)
            dev.trace_device(d_rays, n, any_hit, d_ref)
            got = dev.download(d_out, np.empty(n * out_b, np.uint8))
            ref = dev.download(d_ref, np.empty(n * out_b, np.uint8))
            cnt = dev.download(d_cnt, np.empty((n, 4), np.uint32)).astype(np.int64)
            for p in (d_out, d_ref, d_cnt):
                dev.free(p)
            assert np.array_equal(got, ref), (scene is one, any_hit)
            nodes, recs, ids, steps = cnt.T
            assert (nodes >= 1).all() and (steps >= nodes).all()
            if any_hit:
                hit = got.view(np.int32) != 0
            else:
                hit = got.view([("t", np.float32), ("prim", np.int32)])["prim"] >= 0
            assert (recs[hit] >= 1).all() and (nodes <= st["nodes8"]).all()
            # the id table is read for a hit's index (closest hit: once at the end, and on exact ties;
            # any hit: once) and only then (no spheres here)
            assert (ids[~hit] == 0).all() and (ids[hit] >= 1).all()
            if scene is one:
                assert nodes.tolist() == [1, 1] and recs.tolist() == [1, 0] and hit.tolist() == [True, False]
                assert ids.tolist() == [1, 0]


def test_soup_far_from_origin(dev):
    """Node grid origins are 21-bit integers on the node's own power-of-two grid; far from the
    scene origin that bound coarsens the grid of small nodes (step at least 2^-20 of the distance).
    A 20k-triangle soup moved to (1.5e5, -2.5e5, 4e4) still traverses exactly (closest and any hit
    against the fp32 oracle), for the host and the device flattening."""
    print("case: soup 20k far from the origin", flush=True)
    s = bc.HostScene.workload(bc.SOUP, param=20_000)
    off = np.array([1.5e5, -2.5e5, 4.0e4])
    s.translate(0, s.counts()[0], tuple(off))
    s.build(32, 4)
    rng = np.random.default_rng(11)
    n = 20_000
    o = off + rng.random((n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.column_stack([o, np.full(n, 1e-4), d, np.full(n, 1e30)]).astype(np.float32)
    ra = rays.copy()
    ra[:, 7] = 0.05
    orc_s = orc.Scene.from_host(s)
    prim, t = orc_s.hit32(rays)
    occ = orc_s.any32(ra)
    assert (prim >= 0).mean() > 0.2 and 0.02 < occ.mean() < 0.98
    for mode in (1, 2):
        dev.set_flatten(mode)
        dev.upload(s)
        got = dev.trace(rays)
        assert np.array_equal(got["prim"], prim) and np.array_equal(got["t"], t), mode
        assert np.array_equal(dev.trace(ra, any_hit=True), occ), mode
    dev.set_flatten(0)


def test_probes(dev):
    """The calibration probes of the traversal's rooflines run, time themselves, and refuse
    buffers they cannot index (bc_line_probe: a power-of-two byte count of at least 64, 16-byte
    aligned)."""
    print("case: VALU and scattered-read probes", flush=True)
    assert dev.valu_probe(64, 16, mix=0) > 0.0 and dev.valu_probe(64, 16, mix=1) > 0.0
    nbytes = 1 << 20
    d = dev.alloc(nbytes)
    try:
        assert dev.line_probe(d, nbytes, 64, 8) > 0.0
        assert dev.line_probe(d, 64, 4, 2) > 0.0
        for bad in (0, 32, nbytes - 64, 3 << 10):
            with pytest.raises(bc.BeamcastError):
                dev.line_probe(d, bad, 64, 8)
        with pytest.raises(bc.BeamcastError):
            dev.line_probe(d + 8, 1 << 10, 4, 2)  # not 16-byte aligned
        with pytest.raises(bc.BeamcastError):
            dev.line_probe(d, nbytes, 0, 8)
    finally:
        dev.free(d)
