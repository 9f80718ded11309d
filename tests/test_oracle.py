"""Pins the CPU oracle against the reference's own known answers and invariants.

The reference (windows_portable/) holds no golden images or radiance vectors; what it does
hold -- the AABB known-answer test, the BVH/traversal invariants of renderer_test.go, the
Halton closed forms, and the published mean path lengths of the GIGA benchmark
(benchmarks/GIGA_SAMPLE_THIS_ENV.txt) -- is checked here, plus closed forms of the
integrator that a dropped term or wrong sign would break.
"""
import math

import numpy as np
import pytest

import beamcast as bc
import oracle as orc


def stress():
    return bc.HostScene.workload(bc.STRESS, bins=16, leaf=4)


# --- AABB.HitAccel: the 3 boxes x 5 rays of TestFastAABBMatchesReference (renderer_test.go:361-413)
BOXES = [[-1, -1, -1, 1, 1, 1], [2, -3, -4, 6, 5, 2], [-10, 0, -2, -2, 0.1, 8]]
RAYS = [[0, 0, 5, 0, 0, -1], [0, 0, 0, 1, 0, 0], [4, 10, 0, 0, -1, 0], [20, 20, 20, 1, 1, 1], [-5, 0.05, 10, 0, 0, -1]]
# worked by hand: slab intervals per axis, zero-direction axes test containment
EXPECT = [[1, 1, 0, 0, 0], [0, 1, 1, 0, 0], [0, 0, 0, 0, 1]]


def test_aabb_known_answers():
    for bi, b in enumerate(BOXES):
        for ri, r in enumerate(RAYS):
            hit, near = orc.aabb_hit(b, r)
            assert hit == bool(EXPECT[bi][ri]), (bi, ri)
    # entry distance of ray 0 into box 0 is 4 (z from 5 to 1)
    assert orc.aabb_hit(BOXES[0], RAYS[0])[1] == 4.0


def test_radical_inverse_closed_forms():
    assert [orc.radical_inverse(n, 2) for n in (1, 2, 3, 4)] == [0.5, 0.25, 0.75, 0.125]
    for n, want in ((1, 1 / 3), (2, 2 / 3), (3, 1 / 9), (4, 4 / 9)):
        assert abs(orc.radical_inverse(n, 3) - want) < 1e-15
    rng = np.random.default_rng(1)
    for n in rng.integers(1, 2**62, 50):
        n = int(n)
        rev = int(format(n, "064b")[::-1], 2)
        assert orc.radical_inverse(n, 2) == float(rev) / 2.0**64


def test_pixel_rotation():
    rx, ry = orc.pixel_rotation(0, 0)
    assert rx == 0.0
    h2 = 0x68BC21EB ^ (0x68BC21EB >> 13)
    assert ry == h2 / 2.0**32
    vals = np.array([orc.pixel_rotation(x, y) for x in range(64) for y in range(64)])
    assert vals.min() >= 0 and vals.max() < 1
    assert abs(vals.mean() - 0.5) < 0.02


def test_power_heuristic():
    assert orc.power_heuristic(0.5, 0.5) == 0.5
    assert orc.power_heuristic(1.0, 0.0) == 1.0
    assert orc.power_heuristic(0.0, 0.0) == 0.0
    for a, b in ((0.3, 2.0), (5.0, 0.1)):
        assert abs(orc.power_heuristic(a, b) + orc.power_heuristic(b, a) - 1) < 1e-15


def test_rng_uniform():
    keys = [orc.lib().orc_path_key(7, p, s) for p in range(64) for s in range(16)]
    u = np.array([orc.lib().orc_rng_float(k, d) for k in keys for d in range(8)])
    assert u.min() >= 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.01
    assert abs(u.var() - 1 / 12) < 0.004


def test_sah_invariants():
    for kind, param in ((bc.STRESS, 0), (bc.SOUP, 5000), (bc.CLOSED_BOX, 0)):
        h = bc.HostScene.workload(kind, param=param, bins=32, leaf=4)
        o = orc.Scene.from_host(h)
        box, link, order = o.nodes()
        n = o.nprim
        assert len(box) <= 2 * n - 1
        assert sorted(order.tolist()) == list(range(n))
        seen = np.zeros(n, int)
        for i, (l, r, st, c) in enumerate(link):
            if c > 0:
                assert c <= 16
                seen[st:st + c] += 1
            else:
                assert l == i + 1, "DFS pre-order: left child follows its parent"
                for ch in (l, r):
                    assert np.all(box[ch, :3] >= box[i, :3]) and np.all(box[ch, 3:] <= box[i, 3:])
        assert np.all(seen == 1)
        assert orc.lib().orc_scene_fallback_sorts(o.h) == 0


def test_sah_brute_force_small():
    """On tiny inputs the chosen split is the cheapest non-empty bin split (renderer.go:574-589)."""
    s = bc.HostScene()
    m = s.add_material(bc.DIFFUSE, (0.5, 0.5, 0.5))
    rng = np.random.default_rng(3)
    for _ in range(12):
        c = rng.random(3) * 4
        s.add_triangle(c, c + rng.random(3) * 0.3, c + rng.random(3) * 0.3, m)
    s.build(bins=8, leaf=2)
    box, link, order = s.nodes()
    kind, mat, geom, _, _ = s.export()
    A, B, C = geom[:, :3], geom[:, 3:6], geom[:, 6:]
    cen = (A + B + C) * (1.0 / 3.0)
    lo = np.minimum(np.minimum(A, B), C) - 1e-6
    hi = np.maximum(np.maximum(A, B), C) + 1e-6

    def area(idx):
        b0, b1 = lo[idx].min(0), hi[idx].max(0)
        e = b1 - b0
        return 2 * (e[0] * e[1] + e[1] * e[2] + e[2] * e[0])

    l, r, _, _ = link[0]
    left = order[link[l, 2]:link[l, 2] + 1] if link[l, 3] else None
    idx = np.arange(12)
    ext = cen.max(0) - cen.min(0)
    ax = int(np.argmax(ext))
    b = np.clip(((cen[:, ax] - cen[:, ax].min()) * (1 / ext[ax]) * 8).astype(int), 0, 7)
    costs = []
    for sp in range(7):
        L, R = idx[b <= sp], idx[b > sp]
        if len(L) and len(R):
            costs.append((area(L) * len(L) + area(R) * len(R), sp))
    best = min(costs)[1]
    nleft = int((b <= best).sum())
    # the left child of the root covers exactly the primitives in bins <= best
    def count(i):
        return link[i, 3] if link[i, 3] else count(link[i, 0]) + count(link[i, 1])
    assert count(l) == nleft and count(r) == 12 - nleft
    del left


def test_escape_table():
    """TestThreadedBVHEscapeTraversal, renderer_test.go:772-795."""
    o = orc.Scene.from_host(stress(), bins=16, leaf=4)
    esc = o.escapes()
    _, link, _ = o.nodes()
    assert esc[0] == -1

    def check(i, want):
        assert esc[i] == want
        l, r, _, c = link[i]
        if c == 0:
            check(l, r)
            check(r, want)

    import sys
    sys.setrecursionlimit(10000)
    check(0, -1)


def camera_rays(scene, nx, ny, dx, dy):
    pos = scene.camera[:3]
    rays = []
    for y in range(ny):
        for x in range(nx):
            d = np.array([(x - nx // 2) * dx, (y - ny // 2) * dy, -1.0])
            d /= np.linalg.norm(d)
            rays.append([*pos, 1e-4, *d, 1e30])
    return np.array(rays)


def test_threaded_equals_stack_closest_hit():
    """TestThreadedBVHClosestHitEquivalence, renderer_test.go:797-844 (47x31 rays, |dt| <= 1e-9)."""
    s = stress()
    o = orc.Scene.from_host(s, bins=16, leaf=4)
    rays = camera_rays(s, 47, 31, 0.065, 0.055)
    p1, t1 = o.hit64(rays)
    p2, t2 = o.hit64(rays, threaded=True)
    assert np.array_equal(p1 >= 0, p2 >= 0)
    assert np.max(np.abs(t1 - t2)) <= 1e-9
    assert (p1 >= 0).mean() > 0.3


def test_occluded_matches_closest_hit():
    """TestOccludedMatchesClosestHit, renderer_test.go:658-671 (3000 rays, random tmax)."""
    s = stress()
    o = orc.Scene.from_host(s, bins=16, leaf=4)
    rng = np.random.default_rng(424242)
    pos, tgt = s.camera[:3], s.camera[3:6]
    target = tgt + (rng.random((3000, 3)) - 0.5) * np.array([24, 8, 24])
    d = target - pos
    tmax = 0.25 + rng.random(3000) * 30
    rays = np.column_stack([np.tile(pos, (3000, 1)), np.full(3000, 1e-4), d, tmax])
    prim, _ = o.hit64(rays)
    occ = o.occluded64(rays)
    assert np.array_equal(occ, (prim >= 0).astype(np.int32))
    r32 = rays.astype(np.float32)
    p32, _ = o.hit32(r32)
    assert np.array_equal(o.any32(r32), (p32 >= 0).astype(np.int32))


def test_traversal_matches_brute_force():
    s = bc.HostScene.workload(bc.SOUP, param=3000)
    o = orc.Scene.from_host(s)
    rays = np.concatenate([bc.workload_rays(0, 128, count=3000), bc.workload_rays(1, 128, count=3000)])
    p, t = o.hit32(rays)
    bp, bt = o.brute32(rays)
    assert np.array_equal(p, bp) and np.array_equal(t, bt)
    q, u = o.hit64(rays.astype(np.float64))
    bq, bu = o.brute64(rays.astype(np.float64))
    assert np.array_equal(q, bq) and np.array_equal(u, bu)
    # fp32 and fp64 agree on ids except at |dt|/t < 1e-5 ties
    diff = p != q
    assert diff.mean() <= 1e-4 or np.all(np.abs(t[diff] - u[diff]) / np.maximum(u[diff], 1e-9) < 1e-5)


def test_traversal_matches_brute_force_far_from_origin():
    """1e5 away from the origin a coordinate's fp32 rounding (up to 0.008) is a sizeable part of a
    small triangle, so the rounded records stick out of boxes built from the fp64 geometry; the
    fp32 traversal must still find exactly what testing every primitive finds."""
    s = bc.HostScene.workload(bc.SOUP, param=20_000)
    off = np.array([1.5e5, -2.5e5, 4.0e4])
    s.translate(0, s.counts()[0], tuple(off))
    s.build(32, 4)
    o = orc.Scene.from_host(s)
    rng = np.random.default_rng(11)
    n = 4000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.column_stack([off + rng.random((n, 3)), np.full(n, 1e-4), d, np.full(n, 1e30)]).astype(np.float32)
    p, t = o.hit32(rays)
    bp, bt = o.brute32(rays)
    assert (p >= 0).mean() > 0.2
    assert np.array_equal(p, bp) and np.array_equal(t, bt)
    ra = rays.copy()
    ra[:, 7] = 0.05
    bq, _ = o.brute32(ra)
    occ = o.any32(ra)
    assert 0.02 < occ.mean() < 0.98 and np.array_equal(occ, (bq >= 0).astype(np.int32))


def test_light_distribution():
    """TestPowerWeightedLightPDF / TestCachedLightDistribution (renderer_test.go:310-329, 563-582)."""
    s = bc.HostScene()
    m1 = s.add_material(bc.EMISSIVE, emission=(10, 10, 10))
    m2 = s.add_material(bc.EMISSIVE, emission=(1, 2, 3))
    d = s.add_material(bc.DIFFUSE, (0.5, 0.5, 0.5))
    s.add_sphere((0, 0, 0), 1.0, m1)
    s.add_sphere((3, 0, 0), 0.5, d)
    s.add_sphere((0, 3, 0), 2.0, m2)
    s.build()
    o = orc.Scene.from_host(s)
    lp, cdf, pdf = o.lights()
    assert lp.tolist() == [0, 2]
    assert cdf[-1] == 1.0 and np.all(np.diff(cdf) > 0)
    assert abs(pdf.sum() - 1) < 1e-12
    w0 = 10 * 4 * math.pi * 1.0
    w1 = (0.2126 * 1 + 0.7152 * 2 + 0.0722 * 3) * 4 * math.pi * 4.0
    assert abs(pdf[0] - w0 / (w0 + w1)) < 1e-12 and pdf[1] == 0


def test_cosine_sampler_distribution():
    """TestCosineSamplerDistribution (renderer_test.go:673-690) plus E[cos] = 2/3."""
    n = np.array([0.2, 0.9, 0.38])
    n /= np.linalg.norm(n)
    out = np.zeros(3)
    cs = []
    for i in range(20000):
        orc.lib().orc_cos_hemisphere64(orc._p(n), orc.lib().orc_path_key(1, i, 0), 21, orc._p(out))
        assert abs(out @ out - 1) <= 2e-12
        c = out @ n
        assert c >= -1e-12
        cs.append(c)
    assert abs(np.mean(cs) - 2 / 3) < 0.01


def sky_sphere_scene(albedo):
    s = bc.HostScene()
    m = s.add_material(bc.DIFFUSE, (albedo, albedo, albedo))
    s.add_sphere((0, 0, -3), 1.0, m)
    s.build()
    s.camera = np.array([0, 0, 0, 0, 0, -3, 0, 1, 0, 40.0])
    return s


@pytest.mark.parametrize("precision", [64, 32])
def test_convex_furnace_closed_form(precision):
    """A convex diffuse sphere under a constant sky cannot see itself: every path that hits it
    returns albedo * sky after one bounce, background pixels return sky (trace, renderer.go:1484-1563)."""
    rho, sky = 0.6, float(np.float32(0.3))  # sky_color is a float field
    s = sky_sphere_scene(rho)
    o = orc.Scene.from_host(s)
    d = orc.desc_from(bc.render_desc(s, 24, 24, 4, 4, rr_depth=8, sky_mode=1, sky_color=(sky, sky, sky)))
    img, _, rays, _ = o.render(d, precision=precision)
    rgb = img[..., :3].reshape(-1, 3)
    segs = img[..., 3].reshape(-1)
    k = segs - 4  # samples that hit the sphere (2 segments) vs escaped (1 segment)
    hit = k == 4
    assert hit.any() and (k == 0).any()
    tol = 1e-12 if precision == 64 else 4e-7
    if precision == 64:
        want = (k * (rho * sky) + (4 - k) * sky) * 0.25
    else:
        want = (k * float(np.float32(0.6) * np.float32(sky)) + (4 - k) * sky) * 0.25
    assert np.all(np.abs(rgb - want[:, None]) <= tol)
    # a single segment cannot escape after the diffuse bounce
    d.bounces = 1
    img1, _, _, _ = o.render(d, precision=precision)
    assert np.all(img1[..., :3].reshape(-1, 3)[hit] == 0)


PUBLISHED = [  # (scene, w, h, spp, bounces, rays, samples) from benchmarks/GIGA_SAMPLE_THIS_ENV.txt
    (bc.SHOWCASE, 640, 360, 4, 6, 1767805, 921600),   # :58-63 (GPU probe on CPU fallback)
    (bc.STRESS, 448, 252, 4, 6, 902298, 451584),      # :4-26 (mean of the four worker runs)
    (bc.SHOWCASE, 640, 360, 2, 4, 872714, 460800),    # :28-32
    (bc.CORNELL, 384, 216, 2, 5, 300707, 165888),     # :40-44
]


@pytest.mark.parametrize("case", PUBLISHED, ids=["showcase6", "stress6", "showcase4", "cornell5"])
def test_published_mean_path_length(case):
    """benchmarkSettings (benchmark.go:71-94): Draft preset, RRDepth 4, SAH 16 bins, leaf 4."""
    kind, w, h, spp, b, rays_pub, samples = case
    s = bc.HostScene.workload(kind, bins=16, leaf=4)
    o = orc.Scene.from_host(s, bins=16, leaf=4)
    d = orc.desc_from(bc.render_desc(s, w, h, spp, b, rr_depth=4, sampler=0, mis=True, power=True,
                                     firefly_clamp=12, seed=20240611))
    _, _, rays, _ = o.render(d, precision=64)
    assert abs(rays / samples / (rays_pub / samples) - 1) < 0.003


def test_mis_is_unbiased():
    """NEE+MIS, NEE without MIS on a light-only-by-NEE setup, and pure BSDF sampling agree in the mean."""
    s = bc.HostScene.workload(bc.CORNELL, bins=16, leaf=4)
    o = orc.Scene.from_host(s, bins=16, leaf=4)
    means = []
    for mis in (True, False):
        d = orc.desc_from(bc.render_desc(s, 48, 32, 64, 6, rr_depth=3, sampler=1, mis=mis, power=True, seed=99))
        img, _, _, _ = o.render(d, precision=64)
        means.append(img[..., :3].mean())
    # with MIS off the Go integrator counts sphere lights twice (renderer.go:1459-1462, 1505-1510)
    assert means[1] > means[0] * 1.02
    d = orc.desc_from(bc.render_desc(s, 48, 32, 256, 6, rr_depth=3, sampler=1, mis=True, power=False, seed=7))
    img2, _, _, _ = o.render(d, precision=64)
    assert abs(img2[..., :3].mean() / means[0] - 1) < 0.03


def sunlit_plane_scene(albedo):
    s = bc.HostScene()
    m = s.add_material(bc.DIFFUSE, (albedo, albedo, albedo))
    h = 60.0
    s.add_triangle((-h, 0, -h), (-h, 0, h), (h, 0, h), m)
    s.add_triangle((-h, 0, -h), (h, 0, h), (h, 0, -h), m)
    s.build()
    s.camera = np.array([0, 3, 3, 0, 0, 0, 0, 1, 0, 40.0])
    return s


@pytest.mark.parametrize("bounces", [1, 3])
def test_ris_plane_closed_form(bounces):
    """Resampled direct lighting on a diffuse plane under a constant sky Ls and a sun E from
    direction s: the expected radiance is albedo * (Ls + E (n.s) / pi) whatever the candidate
    split; the BSDF continuation escapes to the sky, which the resampled estimate already holds,
    so more bounces add nothing."""
    a, ls, e = 0.5, float(np.float32(0.3)), 3.0
    sd = np.array([0.3, 0.8, 0.2]) / np.linalg.norm([0.3, 0.8, 0.2])
    s = sunlit_plane_scene(a)
    o = orc.Scene.from_host(s)
    d = orc.desc_from(bc.render_desc(s, 24, 24, 16, bounces, rr_depth=8, sky_mode=1, sky_color=(ls, ls, ls),
                                     restir=32, sun_dir=sd, sun=(e, e, e)))
    img, _, _, shadow = o.render(d, precision=64)
    want = np.float32(a) * (ls + e * sd[1] / np.pi)
    got = img[..., :3].mean()
    assert abs(got / want - 1) < 0.01
    assert shadow == 24 * 24 * 16  # one visibility ray per diffuse vertex
    # the sun alone (black sky): sky candidates weigh zero but still count in M; the 32 candidates
    # split into 16 sun and 16 sky strata, each sun candidate weighs its target x 32 / 16, so
    # every sample is (16 x 2 / 32) x the sun term: exactly the sun term, in every pixel (to the
    # fp32 rounding of the image and of the stored sun direction)
    d.sky_color[0] = d.sky_color[1] = d.sky_color[2] = 0.0
    img2, _, _, _ = o.render(d, precision=64)
    sun = np.float32(a) * e * sd[1] / np.pi
    assert np.allclose(img2[..., :3], sun, rtol=1e-6, atol=0)
    # three candidates over sun and sky: strata of 2 and 1 (weights x 3/2 and x 3); with a black
    # sky each sample is (2 x 3/2 / 3) x the sun term, again exactly
    d.restir_candidates = 3
    img3, _, _, _ = o.render(d, precision=64)
    assert np.allclose(img3[..., :3], sun, rtol=1e-6, atol=0)
    # one candidate (fewer than the two components): a random component with weight x 2, so a
    # pixel is 2 x the sun term times the fraction of its samples that drew the sun
    d.restir_candidates = 1
    img1, _, _, _ = o.render(d, precision=64)
    frac = img1[..., 0] / (2 * sun)
    assert np.allclose(frac * 16, np.round(frac * 16), atol=1e-4) and abs(frac.mean() - 0.5) < 0.05


def test_ris_matches_one_sample_nee():
    """Resampled direct lighting and NEE+MIS estimate the same image in the mean (closed box:
    quad light, gradient sky through the open side)."""
    s = bc.HostScene.workload(bc.CLOSED_BOX)
    o = orc.Scene.from_host(s, nee_triangles=True)
    means = []
    for restir in (0, 32):
        d = orc.desc_from(bc.render_desc(s, 40, 40, 96, 4, rr_depth=3, sampler=1, restir=restir, seed=5))
        img, _, _, _ = o.render(d, precision=64)
        means.append(img[..., :3].reshape(-1, 3).mean(axis=0))
    assert np.all(np.abs(means[1] / means[0] - 1) < 0.02)


def test_spatial_reuse_without_neighbours_is_ris():
    """Spatial reuse whose neighbours all fall outside the rendered rows (a one-pixel-wide image
    rendered row by row) keeps the pixel's own reservoir: the same image as resampled direct
    lighting alone, although the emitter samples go through the area measure and back."""
    s = bc.HostScene.workload(bc.CLOSED_BOX)
    o = orc.Scene.from_host(s, nee_triangles=True)
    kw = dict(rr_depth=3, sampler=1, restir=16, seed=7)
    for y in range(8):
        a = o.render(orc.desc_from(bc.render_desc(s, 1, 8, 8, 3, **kw)), precision=64, y0=y, rows=1)[0]
        b = o.render(orc.desc_from(bc.render_desc(s, 1, 8, 8, 3, spatial=4, radius=2, **kw)), precision=64, y0=y,
                     rows=1)[0]
        assert np.allclose(a, b, rtol=1e-9, atol=1e-12)


def test_spatial_reuse_on_sunlit_plane():
    """On a diffuse plane under a constant sky and a sun every pixel has the same expected
    radiance albedo (Ls + E (n.s) / pi).  With one candidate (a random component) reusing
    neighbours' reservoirs keeps the mean and cuts the per-pixel spread.  With two candidates
    (one sun and one sky stratum) the estimate has no variance at all: nothing occludes the plane,
    the grey contribution is a fixed multiple of its target, and the weight sum is the same for
    every sky direction (cos theta cancels against the cosine pdf)."""
    a, ls, e = 0.5, float(np.float32(0.3)), 3.0
    sd = np.array([0.3, 0.8, 0.2]) / np.linalg.norm([0.3, 0.8, 0.2])
    s = sunlit_plane_scene(a)
    o = orc.Scene.from_host(s)
    want = np.float32(a) * (ls + e * sd[1] / np.pi)
    spread = {}
    for sp in (0, 16):
        d = orc.desc_from(bc.render_desc(s, 32, 32, 1, 1, sky_mode=1, sky_color=(ls, ls, ls), restir=1, sun_dir=sd,
                                         sun=(e, e, e), spatial=sp, radius=8, seed=11))
        img = o.render(d, precision=64)[0][..., :3]
        assert abs(img.mean() / want - 1) < 0.02
        spread[sp] = img.std() / img.mean()
    assert spread[16] < 0.6 * spread[0]
    d = orc.desc_from(bc.render_desc(s, 32, 32, 1, 1, sky_mode=1, sky_color=(ls, ls, ls), restir=2, sun_dir=sd,
                                     sun=(e, e, e), seed=11))
    img = o.render(d, precision=64)[0][..., :3]
    assert np.allclose(img, want, rtol=1e-5, atol=0)


def test_spatial_reuse_mean_closed_box():
    """The reuse (biased, 1/M normalisation) stays close to the unbiased estimate in the mean."""
    s = bc.HostScene.workload(bc.CLOSED_BOX)
    o = orc.Scene.from_host(s, nee_triangles=True)
    means = []
    for sp in (0, 8):
        d = orc.desc_from(bc.render_desc(s, 40, 40, 64, 4, rr_depth=3, sampler=1, restir=8, seed=5, spatial=sp, radius=8))
        means.append(o.render(d, precision=64)[0][..., :3].reshape(-1, 3).mean(axis=0))
    assert np.all(np.abs(means[1] / means[0] - 1) < 0.03)


# ---------------------------------------------------------------- post stack (oracle_post.c)

def rgba_bytes(word):
    return word & 255, (word >> 8) & 255, (word >> 16) & 255, word >> 24


def test_to_rgba_special_values():
    """toRGBA (renderer.go:2055-2064): identity tone curve and gamma 1 quantise as floor(256 c);
    exposure doubles per stop; gamma 2 is a square root; Reinhard maps 1 -> 1/2 and 3 -> 3/4;
    every curve maps 0 to 0 and saturates at 255; alpha is 0xff, red in the low byte."""
    assert rgba_bytes(orc.to_rgba([0.5, 0.25, 0.75], 0.0, 1.0, 0)) == (128, 64, 192, 255)
    assert rgba_bytes(orc.to_rgba([0.25, 0.125, 0.0], 1.0, 1.0, 0)) == (128, 64, 0, 255)
    assert rgba_bytes(orc.to_rgba([0.25, 0.0625, 1.0], 0.0, 2.0, 0)) == (128, 64, 255, 255)
    assert rgba_bytes(orc.to_rgba([1.0, 3.0, 0.0], 0.0, 1.0, 1)) == (128, 192, 0, 255)
    for hdr in (0, 1, 2, 3):
        assert rgba_bytes(orc.to_rgba([0.0, 0.0, 0.0], 0.0, 2.2, hdr)) == (0, 0, 0, 255)
    for hdr in (0, 2, 3):
        assert rgba_bytes(orc.to_rgba([1e6, 1e6, 1e6], 0.0, 2.2, hdr))[:3] == (255, 255, 255)
    # gamma is floored at 0.1: 0.5 ** 10 * 256 < 1
    assert rgba_bytes(orc.to_rgba([0.5, 0.5, 0.5], 0.0, 0.01, 0))[:3] == (0, 0, 0)
    # monotone in radiance for every curve
    for hdr in (0, 1, 2, 3):
        vals = [rgba_bytes(orc.to_rgba([v, v, v], 0.0, 2.2, hdr))[0] for v in np.linspace(0, 4, 200)]
        assert all(a <= b for a, b in zip(vals, vals[1:]))


def test_decode_oct_normal():
    """decodeOctNormal (gpu_opencl_windows.go:552-562): unit length; the folded corners are -z."""
    for q in [(0, 0), (1023, 1023), (0, 1023), (1023, 0)]:
        assert np.allclose(orc.decode_oct_normal(*q), [0, 0, -1])
    rng = np.random.default_rng(3)
    for qx, qy in rng.integers(0, 1024, size=(200, 2)):
        n = orc.decode_oct_normal(int(qx), int(qy))
        assert abs(np.linalg.norm(n) - 1) < 1e-12
        x, y = qx / 1023 * 2 - 1, qy / 1023 * 2 - 1
        assert (n[2] >= 0) == (abs(x) + abs(y) <= 1)


def guides_uniform(h, w, depth=3.0, valid=None, packed_normal=(511, 511)):
    g = np.zeros((h, w, 3), np.uint32)
    g[..., 0] = np.float32(depth).view(np.uint32)
    g[..., 1] = 0x80000000 | (packed_normal[1] << 10) | packed_normal[0]
    if valid is not None:
        g[..., 1][~valid] = 0
        g[..., 0][~valid] = 0
    return g


def test_atrous_properties():
    """atrousDenoise (renderer.go:2228-2287): a constant image is a fixed point; pixels with and
    without a primary hit never mix; one pass at an interior pixel with equal guides is the
    weighted mean (1.5 a + 8 e b) / (1.5 + 8 e), e = exp(-3.5 |lum(a) - lum(b)|)."""
    h, w = 9, 11
    rgba = np.zeros((h, w, 4), np.float32)
    rgba[..., :3] = (0.3, 0.6, 0.2)
    lin, _ = orc.post(rgba, guides_uniform(h, w), denoise=3, fxaa=0)
    assert np.allclose(lin, np.float32([0.3, 0.6, 0.2]), rtol=1e-12, atol=0)
    valid = np.zeros((h, w), bool)
    valid[:, :5] = True
    rgba[..., :3] = np.where(valid[..., None], 1.0, 2.0)
    lin, _ = orc.post(rgba, guides_uniform(h, w, valid=valid), denoise=2)
    assert np.allclose(lin[valid], 1.0) and np.allclose(lin[~valid], 2.0)
    rgba[..., :3] = 0.2
    rgba[4, 5, :3] = 1.0
    lin, _ = orc.post(rgba, guides_uniform(h, w), denoise=1)
    e = np.exp(-3.5 * 0.8)
    assert np.isclose(lin[4, 5, 0], (1.5 * 1.0 + 8 * e * 0.2) / (1.5 + 8 * e), rtol=1e-12)


def test_fxaa_and_txaa_properties():
    """applyFXAA (renderer.go:2557-2586): borders and low-contrast pixels are untouched; a lone
    bright pixel blends with the mean of its 4 neighbours (level 2: blend 0.5).  TXAA leaves a
    constant image constant."""
    h, w = 7, 8
    rgba = np.full((h, w, 4), 0.1, np.float32)
    lin, _ = orc.post(rgba, fxaa=3, txaa=2)
    assert np.allclose(lin, np.float32(0.1), rtol=1e-12)
    rgba[3, 4, :3] = 2.0
    lin, _ = orc.post(rgba, fxaa=2)
    a, b = 2.0, float(np.float32(0.1))  # the frame is fp32; the stack computes in fp64
    assert np.isclose(lin[3, 4, 0], a * 0.5 + b * 0.5, rtol=1e-12)
    assert np.isclose(lin[2, 4, 0], b * 0.5 + (a + 3 * b) / 4 * 0.5, rtol=1e-12)
    assert np.allclose(lin[0], b) and np.allclose(lin[:, 0], b)


@pytest.mark.parametrize("sampler", [1, 2])
def test_prepared_sample_sequence_bounds(sampler):
    """TestPreparedSampleSequenceBounds (renderer_test.go:610-627): the Halton and R2 camera
    sequences, rotated by the pixel's Cranley-Patterson offset (pixel 17, 29), stay in [0, 1)."""
    for smp in range(64):
        jx, jy = orc.camera_jitter(sampler, 17, 29, smp)
        assert 0.0 <= jx < 1.0 and 0.0 <= jy < 1.0
    # the Halton pair of sample 0 is (1/2, 1/3) plus the rotation, modulo 1
    rx, ry = orc.pixel_rotation(17, 29)
    jx, jy = orc.camera_jitter(1, 17, 29, 0)
    assert jx == pytest.approx((0.5 + rx) % 1.0, abs=1e-15) and jy == pytest.approx((1 / 3 + ry) % 1.0, abs=1e-15)
