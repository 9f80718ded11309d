"""OBJ ingestion straight to a resident device scene (bch_ingest_obj: parallel parse, device SAH
build, upload).

Bar: the ingested scene's tree equals the host build of the same primitives node for node, and
its traversal equals both that of the source scene uploaded from the host build and the oracle's
fp32 closest hit over the ingested primitives (ids and distances bit-exact); a page store appended
and built on the device traces like the host-built scene and like the oracle.
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


def test_obj_to_device(dev, tmp_path):
    print("case: terrain 255 cells through OBJ", flush=True)
    src = bc.HostScene.workload(bc.TERRAIN, param=255, bins=32, leaf=4)
    p = str(tmp_path / "terrain.obj")
    src.write_obj(p)
    rng = np.random.default_rng(3)
    d = rng.normal(size=(30000, 3))
    d[:, 1] = -np.abs(d[:, 1])  # downwards from above the heightfield
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = rng.random((30000, 3)) * [1, 0.1, 1] + [0, 0.25, 0]
    rays = np.column_stack([o, np.full(30000, 1e-4), d, np.full(30000, 1e30)]).astype(np.float32)
    dev.upload(src)
    want = dev.trace(rays)
    got_scene, cam = bc.HostScene.ingest_obj(dev, p, bins=32, leaf=4, threads=8)
    got = dev.trace(rays)
    assert np.array_equal(want["prim"], got["prim"]) and np.array_equal(want["t"], got["t"])
    assert (got["prim"] >= 0).mean() > 0.3
    prim, t = orc.Scene.from_host(got_scene).hit32(rays)
    assert np.array_equal(got["prim"], prim) and np.array_equal(got["t"], t)
    a, b = src.nodes(), got_scene.nodes()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_pages_to_device(dev, tmp_path):
    print("case: page store built on the device", flush=True)
    rng = np.random.default_rng(9)
    n = 50_000
    t = np.zeros(n, bc.PACKED_TRIANGLE)
    t["v0"][:, :3] = rng.random((n, 3))
    t["e1"][:, :3] = rng.normal(size=(n, 3)) * 0.01
    t["e2"][:, :3] = rng.normal(size=(n, 3)) * 0.01
    p = str(tmp_path / "soup.pages")
    bc.pages_write(p, t, per_page=8192)
    host, gpu = bc.HostScene(), bc.HostScene()
    for s in (host, gpu):
        s.add_material(0, (0.6, 0.6, 0.6))
        s.append_pages(bc.Pages(p))
    host.build(32, 4)
    gpu.build_device(dev, 32, 4)
    assert all(np.array_equal(x, y) for x, y in zip(host.nodes(), gpu.nodes()))
    rays = np.concatenate([bc.workload_rays(1, 256, count=20000)])
    dev.upload(host)
    a = dev.trace(rays)
    dev.upload(gpu)
    b = dev.trace(rays)
    assert np.array_equal(a["prim"], b["prim"]) and np.array_equal(a["t"], b["t"])
    prim, t = orc.Scene.from_host(gpu).hit32(rays)
    assert np.array_equal(b["prim"], prim) and np.array_equal(b["t"], t)
    assert (prim >= 0).mean() > 0.05


def test_ingest_errors(dev, tmp_path):
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\n")
    with pytest.raises(ValueError, match="OBJ line 3: index out of range"):
        bc.HostScene.ingest_obj(dev, str(p))
