"""Device BVH8 refit (bc_scene_refit) after primitives move with the topology kept.

Bar: after Scene.Refit + bc_scene_refit, closest-hit ids and distances bit-exact against the
oracle's fp32 traversal of the moved scene, images bit-identical to those after a fresh
bc_scene_upload of the moved scene (traversal results do not depend on box tightness), node
topology unchanged, and every node's quantised child boxes conservative (each contains its
primitives' exact boxes with the builder's padding) and tight (at most the padding plus one
grid step beyond them, on the smallest power-of-two grid with an origin k 2^e, |k| < 2^20, that
spans the node).
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


def record_boxes(recs, ids):
    """Exact (fp64) boxes of the fp32 records: triangles A, A+E1, A+E2; spheres C +- r."""
    r = recs.astype(np.float64)
    sph = (ids & 0x80000000) != 0
    a, b, c = r[:, 0:3], r[:, 0:3] + r[:, 3:6], r[:, 0:3] + r[:, 6:9]
    lo = np.minimum(np.minimum(a, b), c)
    hi = np.maximum(np.maximum(a, b), c)
    rad = r[:, 3:4]
    lo[sph] = r[sph, 0:3] - rad[sph]
    hi[sph] = r[sph, 0:3] + rad[sph]
    return lo, hi


def node_levels(nodes):
    """Node indices by level, root first (children of a node are contiguous from child_base)."""
    f = bc.Device.node_fields(nodes)
    levels, cur = [], np.array([0])
    seen = np.zeros(len(nodes), bool)
    while cur.size:
        assert not seen[cur].any(), "node reached twice"
        seen[cur] = True
        levels.append(cur)
        cnt = np.array([bin(int(m)).count("1") for m in f["imask"][cur]])
        base = f["child_base"][cur].astype(np.int64)
        cur = np.concatenate([np.arange(b0, b0 + k) for b0, k in zip(base, cnt)]) if cnt.sum() else np.array([], np.int64)
    assert seen.all(), "unreachable nodes"
    return levels


def check_bvh8(nodes, recs, ids, extra=0.0):
    """Conservative and tight quantised boxes; every primitive in exactly one leaf slot, its record at
    8 * node + slot; every other record slot a hole."""
    f = bc.Device.node_fields(nodes)
    assert len(recs) == 8 * len(nodes) and len(ids) == 8 * len(nodes)
    assert not (f["imask"] & f["lmask"]).any(), "a slot is both internal and a leaf"
    rlo, rhi = record_boxes(recs, ids)
    n = len(nodes)
    ulo, uhi = np.full((n, 3), np.inf), np.full((n, 3), -np.inf)
    used = np.zeros(len(recs), np.int32)
    scale = np.ldexp(1.0, f["e"])  # [n, 3]
    org = f["o"]
    assert (np.abs(f["k"]) < 2**20).all()
    for lvl in reversed(node_levels(nodes)):
        slo, shi = np.full((len(lvl), 8, 3), np.inf), np.full((len(lvl), 8, 3), -np.inf)
        nd = nodes[lvl]
        imask, lmask = f["imask"][lvl], f["lmask"][lvl]
        for s in range(8):
            internal = ((imask >> s) & 1) == 1
            leaf = ((lmask >> s) & 1) == 1
            empty = ~internal & ~leaf
            assert (nd["qlo"][empty, :, s] == 255).all() and (nd["qhi"][empty, :, s] == 0).all()
            below = imask & ((1 << s) - 1)
            rank = np.array([bin(int(m)).count("1") for m in below])
            child = f["child_base"][lvl].astype(np.int64) + rank
            slo[internal, s], shi[internal, s] = ulo[child[internal]], uhi[child[internal]]
            r = 8 * lvl.astype(np.int64) + s
            assert (ids[r[~leaf]] == bc.Device.REC_HOLE).all(), "a record in a non-leaf slot"
            r = r[leaf]
            assert (ids[r] != bc.Device.REC_HOLE).all(), "a leaf slot without a record"
            np.add.at(used, r, 1)
            slo[leaf, s], shi[leaf, s] = rlo[r], rhi[r]
        ulo[lvl], uhi[lvl] = slo.min(axis=1), shi.max(axis=1)
        # the node's grid: padded union, origin k 2^e below it, |k| < 2^20, smallest step for these
        nlo, nhi = ulo[lvl], uhi[lvl]
        pad = np.maximum((np.maximum(np.abs(nlo), np.abs(nhi)) + (nhi - nlo)) * 2.0**-20, 1e-20)
        o, sc = org[lvl], scale[lvl]
        assert (o <= nlo - 0.5 * pad).all()
        ext2 = (nhi + pad) - o
        assert (255 * sc >= ext2 - 0.5 * pad).all()
        # the origin snap costs at most one doubling over the span's own step; beyond that only the
        # |k| < 2^20 bound (or the smallest step, 2^-126) may hold the step up
        half_k = np.floor((nlo - pad) / (sc / 2))
        assert ((255 * sc / 4 < ext2 + extra) | (f["e"][lvl] == -126) | (np.abs(half_k) >= 2**20 - 2)).all(), \
            "grid step not minimal"
        for s in range(8):
            present = np.isfinite(slo[:, s, 0])
            qlo = nd["qlo"][:, :, s].astype(np.float64)
            qhi = nd["qhi"][:, :, s].astype(np.float64)
            blo, bhi = o + qlo * sc, o + qhi * sc
            p = present
            assert (blo[p] <= slo[p, s] - 0.5 * pad[p]).all(), "child box not conservative (low side)"
            assert (bhi[p] >= shi[p, s] + 0.5 * pad[p]).all(), "child box not conservative (high side)"
            # a refit takes an internal child's box as the fp32 rounding outward of its union (at most
            # one fp32 ulp, below pad / 8), hence the 1.25 pad
            loose = p[:, None] & ((blo < slo[:, s] - 1.25 * pad - sc - extra) | (bhi > shi[:, s] + 1.25 * pad + sc + extra))
            if loose.any():
                i, a = np.argwhere(loose)[0]
                raise AssertionError(
                    f"child box loose: node {lvl[i]} slot {s} axis {a}: box [{blo[i, a]!r}, {bhi[i, a]!r}] child "
                    f"[{slo[i, s, a]!r}, {shi[i, s, a]!r}] pad {pad[i, a]!r} step {sc[i, a]!r} origin {o[i, a]!r} "
                    f"k {f['k'][lvl[i], a]} e {f['e'][lvl[i], a]} qlo {qlo[i, a]} qhi {qhi[i, a]} "
                    f"internal {(f['imask'][lvl[i]] >> s) & 1} node box [{nlo[i, a]!r}, {nhi[i, a]!r}]")
    assert used.sum() == (ids != bc.Device.REC_HOLE).sum() and (used <= 1).all(), \
        "records not covered exactly once by leaf slots"


def topology(nodes):
    return np.column_stack([nodes["link"], nodes["grid"] >> 29])


def scene_rays(scene, n, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.tile(scene.camera[:3], (n, 1)) + rng.normal(size=(n, 3)) * 0.5
    return np.column_stack([o, np.full(n, 1e-4), d, np.full(n, 1e30)]).astype(np.float32)


def assert_trace_parity(dev, scene, rays, bins=16):
    got = dev.trace(rays)
    prim, t = orc.Scene.from_host(scene, bins=bins).hit32(rays)
    bad = np.nonzero((got["prim"] != prim) | (got["t"] != t))[0]
    assert bad.size == 0, f"{bad.size} rays differ from the oracle after refit, first {bad[:5]}"
    return got


def test_host_builder_boxes(dev):
    """The upload's flattened boxes obey the same rule (the Go boxes add 1e-6 to triangles)."""
    for kind in (bc.SHOWCASE, bc.CORNELL, bc.CLOSED_BOX):
        print(f"case: builder boxes, scene {kind}", flush=True)
        s = bc.HostScene.workload(kind, bins=16, leaf=4)
        dev.upload(s)
        check_bvh8(*dev.scene_arrays(), extra=4e-6)


@pytest.mark.parametrize("kind", [bc.SHOWCASE, bc.CORNELL])
def test_refit_after_move(dev, kind):
    print(f"case: refit after move, scene {kind}", flush=True)
    s = bc.HostScene.workload(kind, bins=16, leaf=4)
    dev.upload(s)
    topo0 = topology(dev.scene_arrays()[0])
    nprim = s.counts()[0]
    s.translate(0, nprim // 3, (0.7, -0.3, 1.1))
    s.translate(nprim // 2, nprim // 5, (-2.5, 0.4, 0.0))
    s.refit()
    dev.refit(s)
    nodes, recs, ids = dev.scene_arrays()
    assert np.array_equal(topology(nodes), topo0), "refit changed the topology"
    check_bvh8(nodes, recs, ids)
    assert_trace_parity(dev, s, scene_rays(s, 20000, kind))
    # same image as after a fresh upload of the moved scene (lights moved with their quads)
    d = bc.render_desc(s, 48, 32, 4, 5, rr_depth=4, sampler=1, mis=True, power=True)
    img, g, rays = dev.render(d, guides=True)
    dev.upload(s)
    img2, g2, rays2 = dev.render(d, guides=True)
    assert rays == rays2 and np.array_equal(img, img2) and np.array_equal(g, g2)


def test_refit_render_parity(dev):
    print("case: refit closed box, render parity vs oracle", flush=True)
    s = bc.HostScene.workload(bc.CLOSED_BOX)
    dev.upload(s, nee_triangles=True)
    nprim = s.counts()[0]
    s.translate(nprim - 12, 12, (0.05, -0.1, 0.08))
    s.refit()
    dev.refit(s, nee_triangles=True)
    d = bc.render_desc(s, 40, 40, 4, 4, rr_depth=3, sampler=1, mis=True, power=True)
    img, g, rays = dev.render(d, guides=True)
    ref, rg, rrays, _ = orc.Scene.from_host(s, nee_triangles=True).render(orc.desc_from(d), precision=32, guides=True)
    assert rays == rrays
    assert np.array_equal(img[..., 3], ref[..., 3])
    err = np.abs(img[..., :3] - ref[..., :3])
    assert not (err > 1e-4 * np.abs(ref[..., :3]) + 1e-6).any(), f"max err {err.max()}"
    assert np.array_equal(g, rg)


def test_refit_soup_large_moves(dev):
    """Half of a 200k-triangle soup moves far outside the scene bounds and back again."""
    print("case: soup 200k, refit far move and back", flush=True)
    s = bc.HostScene.workload(bc.SOUP, param=200_000)
    coh = bc.workload_rays(0, 512, count=512 * 512)[::16]
    inc = bc.workload_rays(1, 512, count=16384)
    rays = np.concatenate([coh, inc])
    dev.upload(s)
    base = dev.trace(rays)
    topo0 = topology(dev.scene_arrays()[0])
    s.translate(0, 100_000, (40.0, -3.0, 7.5))
    s.refit()
    dev.refit(s)
    nodes, recs, ids = dev.scene_arrays()
    assert np.array_equal(topology(nodes), topo0)
    check_bvh8(nodes, recs, ids)
    moved = assert_trace_parity(dev, s, rays)
    assert (moved["prim"] != base["prim"]).any()
    s.translate(0, 100_000, (-40.0, 3.0, -7.5))
    s.refit()
    dev.refit(s)
    back = dev.trace(rays)
    # the vertices came back to within fp64 rounding of the translation; traversal of the
    # restored scene equals the oracle's on the same fp32 records
    assert_trace_parity(dev, s, rays)
    assert (back["prim"] == base["prim"]).mean() > 0.999


def test_refit_errors(dev):
    print("case: refit argument errors", flush=True)
    a = bc.HostScene.workload(bc.CORNELL, bins=16, leaf=4)
    b = bc.HostScene.workload(bc.SHOWCASE, bins=16, leaf=4)
    with bc.Device(0) as fresh:
        with pytest.raises(bc.BeamcastError, match="resident single scene"):
            fresh.refit(a)
    dev.upload(a)
    with pytest.raises(bc.BeamcastError, match="topology changed"):
        dev.refit(b)
    dev.upload_batch([a, a])
    with pytest.raises(bc.BeamcastError, match="resident single scene"):
        dev.refit(a)
    dev.upload(a)
    dev.refit(a)  # an unchanged scene refits to a valid tree
    check_bvh8(*dev.scene_arrays())
