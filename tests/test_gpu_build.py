"""Device SAH build (bc_build_bvh) against the host build of the Go Scene.Build.

Bar: Nodes (Box, Left, Right, Start, Count) and Order identical to the host restatement of
Scene.Build / buildNodeSAH, node for node and index for index, for the workload scenes, several
bin counts and leaf sizes, degenerate inputs (coincident centroids, a single primitive, trees
whose root is a leaf or a one-wave subtree) and the 1M-triangle traversal soup; a scene built
on the device traces exactly as the host-built one.
"""
import numpy as np
import pytest

import beamcast as bc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = bc.Device(0)
    yield d
    d.close()


def same_tree(dev, s, bins, leaf):
    box, cen = s.bounds()
    s.build(bins, leaf)
    hbox, hlink, horder = s.nodes()
    gbox, glink, gorder = dev.build_bvh(box, cen, bins, leaf)
    assert len(gbox) == len(hbox), f"{len(gbox)} nodes vs {len(hbox)} on the host"
    bad = np.nonzero((glink != hlink).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} node links differ, first {bad[:4]}: {glink[bad[:2]]} vs {hlink[bad[:2]]}"
    assert np.array_equal(gorder, horder), f"order differs at {np.nonzero(gorder != horder)[0][:5]}"
    assert np.array_equal(gbox, hbox)
    return len(gbox)


@pytest.mark.parametrize("bins,leaf", [(32, 4), (16, 4), (8, 1), (32, 16), (4, 2)])
def test_soup(dev, bins, leaf):
    print(f"case: soup 200k bins {bins} leaf {leaf}", flush=True)
    s = bc.HostScene.workload(bc.SOUP, param=200_000, bins=bins, leaf=leaf)
    same_tree(dev, s, bins, leaf)


@pytest.mark.parametrize("kind,param", [(bc.CLOSED_BOX, 0), (bc.SHOWCASE, 0), (bc.CORNELL, 0), (bc.STRESS, 0),
                                        (bc.TERRAIN, 255), (bc.BOX_RETESS, 50_000), (bc.BATCH_SCENE, 3)])
def test_workload_scenes(dev, kind, param):
    print(f"case: scene {kind} param {param}", flush=True)
    s = bc.HostScene.workload(kind, param=param, bins=32, leaf=4)
    same_tree(dev, s, 32, 4)
    same_tree(dev, s, 16, 8)


def test_degenerate_and_tiny(dev):
    rng = np.random.default_rng(7)
    # coincident centroids: 600 copies of one triangle (a leaf of 600 by the degenerate-extent
    # rule) among scattered ones, and a sliver axis with equal coordinates
    s = bc.HostScene()
    m = s.add_material(0, (0.5, 0.5, 0.5))
    a, b, c = np.array([0.1, 0.2, 0.3]), np.array([0.2, 0.2, 0.3]), np.array([0.1, 0.3, 0.3])
    for _ in range(600):
        s.add_triangle(a, b, c, m)
    for _ in range(900):
        p = rng.random(3)
        p[2] = 0.5
        s.add_triangle(p, p + [0.01, 0, 0], p + [0, 0.01, 0], m)
    for _ in range(40):
        s.add_sphere(rng.random(3), 0.02, m)
    print("case: coincident centroids", flush=True)
    same_tree(dev, s, 32, 4)
    # tiny trees: a single primitive, a leaf root, a one-wave subtree root, just above it
    for n in (1, 3, 4, 5, 17, 256, 257, 300, 4097):
        print(f"case: {n} primitives", flush=True)
        t = bc.HostScene()
        m = t.add_material(0, (0.5, 0.5, 0.5))
        for _ in range(n):
            p = rng.random(3)
            t.add_triangle(p, p + rng.random(3) * 0.05, p + rng.random(3) * 0.05, m)
        same_tree(dev, t, 32, 4)


def test_soup_1m_and_trace(dev):
    print("case: soup 1M built on the device, traced", flush=True)
    s = bc.HostScene.workload(bc.SOUP, param=1_000_000, bins=32, leaf=4)
    same_tree(dev, s, 32, 4)
    rays = np.concatenate([bc.workload_rays(0, 512, count=512 * 512)[::8], bc.workload_rays(1, 512, count=16384)])
    dev.upload(s)
    host = dev.trace(rays)
    s.build_device(dev, 32, 4)
    dev.upload(s)
    gpu = dev.trace(rays)
    assert np.array_equal(host["prim"], gpu["prim"]) and np.array_equal(host["t"], gpu["t"])


def test_build_errors(dev):
    box = np.zeros((4, 6))
    cen = np.zeros((4, 3))
    box[2, 1] = np.nan
    with pytest.raises(bc.BeamcastError, match="non-finite"):
        dev.build_bvh(box, cen)
