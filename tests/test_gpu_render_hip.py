"""renderOpenCL's host loop over the HIP backend (bch_render_hip), through the C ABI.

renderOpenCL (gpu_opencl_windows.go:565-723) renders a frame in row chunks, publishes each chunk
through onChunk, polls its cancel flag before every chunk, decodes the compact guides into
Guide{P, N, Albedo, Depth} (:686-716) and keeps the scene resident, refitting it when only the
geometry revision changed (ensureSceneLocked, :484-495).  Checked here:

* chunks: 64 rows at PublishHz 30, each published once in order with its progress, the data
  equal to the returned rows;
* the linear image within 1e-4 of the oracle's fp32 render of the same description, and the ray
  count equal;
* decoded guides against the oracle's fp64 primary hit of sample 0's camera ray, traced
  independently: depth within 1e-4 relative, P within 1e-4 of the hit point, N within the
  octahedral quantisation (sqrt(18)/1023 rad), albedo materialPreviewColor of the hit material,
  valid exactly where the ray hits (bar a few pixels on silhouettes, where fp32 and fp64 differ);
* the scene cache: reuse, refit on a geometry-only change (with the image of the moved scene),
  upload on a topology change or a different scene;
* cancel before the first chunk and between chunks.
"""
import ctypes
import math

import numpy as np
import pytest

import beamcast as bc
import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def runtime():
    yield
    bc.render_hip_release()


def settings(w, h, spp, bounces, denoise=False, progressive=True, hz=30):
    st = bc.settings_quality("Draft")
    st.width, st.height, st.spp, st.bounces = w, h, spp, bounces
    st.msaa, st.tsaa, st.taa, st.txaa = 1, 0, 0, 0
    st.denoise, st.progressive, st.publish_hz = int(denoise), int(progressive), hz
    st.rr_depth, st.firefly_clamp, st.upscale = 3, 0.0, 0
    return st


def oracle_render(scene, st, nee=False):
    d = bc.render_desc_for(scene.camera, st)
    o = orc.Scene.from_host(scene, nee_triangles=nee)
    ref, _, rays, _ = o.render(orc.desc_from(d), precision=32)
    return d, o, ref, rays


def assert_close(lin, ref):
    err = np.abs(lin - ref[..., :3])
    bad = err > 1e-4 * np.abs(ref[..., :3]) + 1e-6
    assert not bad.any(), f"{bad.sum()} values outside 1e-4; max err {err.max()}"


def test_chunks_image_and_guides():
    print("case: render_hip Showcase 64x150 3 spp, guides", flush=True)
    s = bc.HostScene.workload(bc.SHOWCASE, bins=16, leaf=4)
    st = settings(64, 150, 3, 5, denoise=True)
    chunks = []
    out = bc.render_hip(s, s.camera, st, on_chunk=lambda y0, rows, rgb, p: chunks.append((y0, rows, rgb, p)),
                        guides=True)
    assert out["device"]
    assert [(c[0], c[1]) for c in chunks] == [(0, 64), (64, 64), (128, 22)]
    assert [c[3] for c in chunks] == [64 / 150, 128 / 150, 1.0]
    lin = out["linear"]
    for y0, rows, rgb, _ in chunks:
        assert np.array_equal(rgb, lin[y0:y0 + rows])
    d, o, ref, rays = oracle_render(s, st)
    assert out["rays"] == rays
    assert_close(lin, ref)
    # guides against the oracle's own primary hit of sample 0
    w, h = d.width, d.height
    origin, lower, horiz, vert = (np.array(v[:]) for v in (d.origin, d.lower, d.horiz, d.vert))
    ys, xs = np.mgrid[0:h, 0:w]
    jit = np.array([orc.camera_jitter(d.sampler, int(x), int(y), 0, d.seed, w) for y, x in zip(ys.ravel(), xs.ravel())])
    u = (xs.ravel() + jit[:, 0]) / (w - 1)
    v = (h - 1 - ys.ravel() + jit[:, 1]) / (h - 1)
    dirs = lower + np.outer(u, horiz) + np.outer(v, vert) - origin
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    rays64 = np.column_stack([np.tile(origin, (w * h, 1)), np.full(w * h, 1e-4), dirs, np.full(w * h, 1e30)])
    prim, t = o.hit64(rays64)
    g = out["guides"]
    valid = g["valid"] == 1
    hit = prim >= 0
    assert (valid != hit).sum() <= 0.002 * w * h + 2
    kind, mat, geom, mk, md = s.export()
    both = np.nonzero(valid & hit)[0]
    assert len(both) > 0.5 * w * h
    bound = math.sqrt(18.0) / 1023.0 + 1e-6
    mism = 0
    for i in both:
        p_ref = origin + dirs[i] * t[i]
        if kind[prim[i]] == bc.SPHERE:
            n = p_ref - geom[prim[i], 0:3]
        else:
            a, b, c = geom[prim[i], 0:3], geom[prim[i], 3:6], geom[prim[i], 6:9]
            n = np.cross(b - a, c - a)
        n /= np.linalg.norm(n)
        if np.dot(dirs[i], n) >= 0:
            n = -n
        gi = g[i]
        if abs(gi["depth"] - t[i]) > 1e-4 * t[i]:
            mism += 1  # a different primitive on a silhouette or shared edge
            continue
        assert np.linalg.norm(gi["p"] - p_ref) <= 1e-4 * (1 + t[i]), i
        assert math.acos(min(1.0, float(np.dot(gi["n"], n)))) <= bound, i
        m = mat[prim[i]]
        want = (np.clip(md[m, 5:8] / 8, 0, 1) if mk[m] == bc.EMISSIVE
                else (md[m, 0:3] if np.dot(md[m, 0:3], md[m, 0:3]) > 0 else np.full(3, 0.8)))
        assert np.allclose(gi["albedo"], want), i
    assert mism <= 0.002 * len(both) + 2


def test_scene_cache_refit_and_upload():
    print("case: render_hip scene cache", flush=True)
    s = bc.HostScene.workload(bc.STRESS, bins=16, leaf=4)
    st = settings(40, 32, 2, 4, progressive=False)
    first = bc.render_hip(s, s.camera, st)
    assert first["scene_action"] == 2
    again = bc.render_hip(s, s.camera, st)
    assert again["scene_action"] == 0 and np.array_equal(again["linear"], first["linear"])
    # geometry only (TranslateObject + Refit): refit on the device, image of the moved scene
    topo0, geom0 = s.revisions()
    s.translate(0, 40, (0.3, 0.2, -0.1))
    s.refit()
    topo1, geom1 = s.revisions()
    assert topo1 == topo0 and geom1 != geom0
    moved = bc.render_hip(s, s.camera, st)
    assert moved["scene_action"] == 1
    _, _, ref, rays = oracle_render(s, st)
    assert moved["rays"] == rays
    assert_close(moved["linear"], ref)
    assert not np.array_equal(moved["linear"], first["linear"])
    # the refit scene renders like a fresh upload of it, bit for bit
    bc.render_hip_release()
    fresh = bc.render_hip(s, s.camera, st)
    assert fresh["scene_action"] == 2 and np.array_equal(fresh["linear"], moved["linear"])
    # a topology change (Build) uploads
    s.build(16, 4)
    assert bc.render_hip(s, s.camera, st)["scene_action"] == 2
    # a different scene uploads, also one allocated where a freed scene was
    s2 = bc.HostScene.workload(bc.CORNELL, bins=16, leaf=4)
    assert bc.render_hip(s2, s2.camera, st)["scene_action"] == 2
    del s2
    s3 = bc.HostScene.workload(bc.CORNELL, bins=16, leaf=4)
    r3 = bc.render_hip(s3, s3.camera, st)
    assert r3["scene_action"] == 2
    _, _, ref3, _ = oracle_render(s3, st)
    assert_close(r3["linear"], ref3)


def test_cancel_between_chunks():
    print("case: render_hip cancel", flush=True)
    s = bc.HostScene.workload(bc.CORNELL, bins=16, leaf=4)
    st = settings(48, 150, 1, 3)
    flag = ctypes.c_int32(1)
    seen = []
    with pytest.raises(bc.BeamcastError, match="render cancelled"):
        bc.render_hip(s, s.camera, st, cancel=flag, on_chunk=lambda *a: seen.append(a))
    assert seen == []
    flag.value = 0

    def stop(y0, rows, rgb, p):
        seen.append(y0)
        flag.value = 1

    with pytest.raises(bc.BeamcastError, match="render cancelled"):
        bc.render_hip(s, s.camera, st, cancel=flag, on_chunk=stop)
    assert seen == [0]
    flag.value = 0
    done = bc.render_hip(s, s.camera, st, cancel=flag)
    assert done["linear"].shape == (150, 48, 3) and done["linear"].mean() > 0
