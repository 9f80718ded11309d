"""Device post stack (bc_postprocess_device / bc_render_post) against the oracle's fp64
restatement of the renderer's post stack (oracle_post.c), on frames the integrator rendered.

Bar: the device computes in fp32 with the same operation order; the processed linear frame is
within 1e-4 relative (+1e-6 absolute) of the fp64 oracle, and RGBA8 words differ by at most one
step per channel (fp32 rounding near a quantisation boundary) on at most 1% of pixels.
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


@pytest.fixture(scope="module")
def frame(dev):
    s = bc.HostScene.workload(bc.CLOSED_BOX)
    dev.upload(s, nee_triangles=True)
    d = bc.render_desc(s, 96, 64, 16, 4, rr_depth=3, sampler=1)
    img, g, _ = dev.render(d, guides=True)
    return s, d, img, g


def compare(lin_gpu, px_gpu, lin_ref, px_ref):
    err = np.abs(lin_gpu[..., :3] - lin_ref)
    assert not (err > 1e-4 * np.abs(lin_ref) + 1e-6).any(), err.max()
    ch = np.stack([(px_gpu >> s) & 255 for s in (0, 8, 16, 24)], -1).astype(np.int64)
    cr = np.stack([(px_ref >> s) & 255 for s in (0, 8, 16, 24)], -1).astype(np.int64)
    d = np.abs(ch - cr)
    assert d.max() <= 1 and (d.max(-1) > 0).mean() <= 0.01


@pytest.mark.parametrize("denoise,fxaa,txaa,hdr,exposure,gamma",
                         [(3, 2, 0, 2, 0.0, 2.2), (2, 0, 2, 1, 0.0, 2.2), (0, 3, 1, 3, 0.5, 1.8), (0, 0, 0, 0, -1.0, 1.0),
                          (5, 1, 3, 2, 0.25, 2.4)])
def test_post_matches_oracle(dev, frame, denoise, fxaa, txaa, hdr, exposure, gamma):
    print(f"case: post denoise {denoise} fxaa {fxaa} txaa {txaa} hdr {hdr}", flush=True)
    _, d, img, g = frame
    h, w = img.shape[:2]
    post = bc.post_desc(denoise=denoise, fxaa=fxaa, txaa=txaa, hdr=hdr, exposure=exposure, gamma=gamma)
    d_img, d_g = dev.to_device(img), dev.to_device(g)
    d_px, d_lin = dev.alloc(w * h * 4), dev.alloc(w * h * 16)
    dev.postprocess_device(post, w, h, d_img, d_g, d_px, d_lin)
    px = dev.download(d_px, np.empty((h, w), np.uint32))
    lin = dev.download(d_lin, np.empty((h, w, 4), np.float32))
    for p in (d_img, d_g, d_px, d_lin):
        dev.free(p)
    lin_ref, px_ref = orc.post(img, g, denoise=denoise, fxaa=fxaa, txaa=txaa, exposure=exposure, gamma=gamma, hdr=hdr)
    compare(lin, px, lin_ref, px_ref)
    assert np.array_equal(lin[..., 3], img[..., 3])  # path segments ride along unchanged


def test_render_post_whole_frame(dev, frame):
    """bc_render_post = render + post on the device, 4 bytes per pixel back."""
    print("case: render_post whole frame", flush=True)
    s, d, img, g = frame
    dev.upload(s, nee_triangles=True)
    post = bc.post_desc(denoise=3, fxaa=2, txaa=1, hdr=2)
    px, lin, rays = dev.render_post(d, post, want_linear=True)
    assert rays == int(img[..., 3].sum())
    lin_ref, px_ref = orc.post(img, g, denoise=3, fxaa=2, txaa=1, hdr=2)
    compare(lin, px, lin_ref, px_ref)
    px2, none, _ = dev.render_post(d, post)
    assert none is None and np.array_equal(px, px2)


def test_post_argument_errors(dev, frame):
    print("case: post argument errors", flush=True)
    _, d, img, _ = frame
    h, w = img.shape[:2]
    d_img, d_px = dev.to_device(img), dev.alloc(w * h * 4)
    with pytest.raises(bc.BeamcastError):
        dev.postprocess_device(bc.post_desc(denoise=2), w, h, d_img, None, d_px)   # denoise needs guides
    with pytest.raises(bc.BeamcastError):
        dev.postprocess_device(bc.post_desc(), w, h, d_img, None, None)             # no output
    # tiny frames: FXAA is skipped below 3x3, as applyFXAA does
    dev.postprocess_device(bc.post_desc(fxaa=3, txaa=1), 2, 2, d_img, None, d_px)
    dev.free(d_img)
    dev.free(d_px)


def test_post_determinism(dev, frame):
    """TestParallelPostDeterminism (renderer_test.go:720-745): the post stack gives the same bits
    however its work is scheduled.  The reference image pattern (x*13 + y*7 mod 97) with the
    rendered guides goes through a-trous denoise and FXAA twice on the device, interleaved with
    other launches; both results are bit-identical and match the oracle."""
    print("case: post determinism", flush=True)
    _, _, img, g = frame
    h, w = img.shape[:2]
    y, x = np.mgrid[0:h, 0:w]
    v = ((x * 13 + y * 7) % 97) / 97.0
    pat = np.stack([0.1 + 2 * v, 0.2 + v, 0.3 + 0.5 * v, img[..., 3]], -1).astype(np.float32)
    post = bc.post_desc(denoise=2, fxaa=2, hdr=2)
    outs = []
    for _ in range(2):
        d_img, d_g = dev.to_device(pat), dev.to_device(g)
        d_px, d_lin = dev.alloc(w * h * 4), dev.alloc(w * h * 16)
        dev.postprocess_device(post, w, h, d_img, d_g, d_px, d_lin)
        outs.append((dev.download(d_px, np.empty((h, w), np.uint32)), dev.download(d_lin, np.empty((h, w, 4), np.float32))))
        for p in (d_img, d_g, d_px, d_lin):
            dev.free(p)
        dev.trace(bc.workload_rays(1, 64, count=4096))  # other work between the two runs
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1].view(np.int32), outs[1][1].view(np.int32))
    lin_ref, px_ref = orc.post(pat, g, denoise=2, fxaa=2, hdr=2)
    compare(outs[0][1], outs[0][0], lin_ref, px_ref)
