"""BVH8 flattening on the device (bc_set_flatten(ctx, 2)) against the host flattening.

Bar: nodes, records, record ids and statistics byte for byte equal to the host flattening for
every workload scene (so traversal and images are unchanged); the refit level lists it leaves
behind refit correctly (trace parity with the oracle after a move); encoding errors are the same.
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
    d.set_flatten(0)
    d.close()


def both(dev, s, nee=False):
    out = []
    for mode in (1, 2):
        dev.set_flatten(mode)
        dev.upload(s, nee_triangles=nee)
        out.append((dev.stats(), *dev.scene_arrays()))
    dev.set_flatten(0)
    return out


@pytest.mark.parametrize("kind,param", [(bc.SOUP, 200_000), (bc.TERRAIN, 255), (bc.SHOWCASE, 0), (bc.CORNELL, 0),
                                        (bc.STRESS, 0), (bc.CLOSED_BOX, 0), (bc.BOX_RETESS, 50_000), (bc.BATCH_SCENE, 3)])
def test_same_bytes(dev, kind, param):
    print(f"case: flatten scene {kind} param {param}", flush=True)
    s = bc.HostScene.workload(kind, param=param, bins=32, leaf=4)
    (sa, na, ra, ia), (sb, nb, rb, ib) = both(dev, s)
    assert sa == sb
    assert np.array_equal(na.view(np.uint8), nb.view(np.uint8)), "nodes differ"
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32)), "records differ"
    assert np.array_equal(ia, ib)


def test_leaf_sizes_and_tiny(dev):
    for bins, leaf in ((16, 1), (32, 8), (8, 16)):
        print(f"case: soup 30k bins {bins} leaf {leaf}", flush=True)
        s = bc.HostScene.workload(bc.SOUP, param=30_000, bins=bins, leaf=leaf)
        (sa, na, ra, ia), (sb, nb, rb, ib) = both(dev, s)
        assert sa == sb and np.array_equal(na.view(np.uint8), nb.view(np.uint8))
        assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32)) and np.array_equal(ia, ib)
    for n in (1, 2, 3, 9):
        t = bc.HostScene()
        m = t.add_material(0, (0.5, 0.5, 0.5))
        rng = np.random.default_rng(n)
        for _ in range(n):
            p = rng.random(3)
            t.add_triangle(p, p + [0.1, 0, 0], p + [0, 0.1, 0], m)
        t.build(32, 4)
        (sa, na, ra, ia), (sb, nb, rb, ib) = both(dev, t)
        assert sa == sb and np.array_equal(na.view(np.uint8), nb.view(np.uint8)), n


def test_refit_after_device_flatten(dev):
    print("case: refit after the device flatten", flush=True)
    s = bc.HostScene.workload(bc.SOUP, param=100_000)
    dev.set_flatten(2)
    dev.upload(s)
    dev.set_flatten(0)
    s.translate(0, 50_000, (3.0, 0.0, -1.0))
    s.refit()
    dev.refit(s)
    rays = np.concatenate([bc.workload_rays(0, 256, count=256 * 256)[::4], bc.workload_rays(1, 256, count=16384)])
    got = dev.trace(rays)
    prim, t = orc.Scene.from_host(s, bins=16).hit32(rays)
    assert np.array_equal(got["prim"], prim) and np.array_equal(got["t"], t)


def test_encoding_errors_match(dev):
    s = bc.HostScene()
    m = s.add_material(0, (0.5, 0.5, 0.5))
    for _ in range(40):  # one leaf of 40 coincident triangles: not encodable
        s.add_triangle([0, 0, 0], [1, 0, 0], [0, 1, 0], m)
    s.build(32, 4)
    msgs = []
    for mode in (1, 2):
        dev.set_flatten(mode)
        with pytest.raises(bc.BeamcastError) as e:
            dev.upload(s)
        msgs.append(str(e.value))
    dev.set_flatten(0)
    assert msgs[0] == msgs[1] and "unsupported count" in msgs[0]
