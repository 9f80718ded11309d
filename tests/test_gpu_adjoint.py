"""GPU checks of the differentiable render (path-replay backpropagation, bc_forward/bc_backward).

Gradients of an L2 image loss wrt per-material albedo (Color) and emitter radiance are checked
against central finite differences with common random numbers (eps 1e-3, same seeds), on the
GPU and against the oracle's fp64 CPU path; the bar is 2% relative on parameters whose
gradient is not negligible.  Russian roulette is disabled (RRDepth >= bounces), so the
estimator is a smooth function of the parameters and FD with common random numbers is exact
up to O(eps^2).
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


def setup(dev, w=48, h=48, spp=16, bounces=6, tris=3000):
    s = bc.HostScene.workload(bc.BOX_RETESS, param=tris)
    kind, mat, geom, mk, md_gt = s.export()
    dev.upload(s, nee_triangles=True)
    d = bc.render_desc(s, w, h, spp, bounces, rr_depth=bounces, sampler=1, mis=True, power=True, seed=2024)
    target, _ = dev.forward(d)
    md = md_gt.copy()
    diffuse = mk == bc.DIFFUSE
    md[diffuse, 0:3] = 0.5
    md[mk == bc.EMISSIVE, 5:8] *= 0.5
    dev.update_materials(mk, md)
    return s, d, target, mk, md, (kind, mat, geom)


def loss_of(img, target):
    return float(np.sum((img[..., :3].astype(np.float64) - target[..., :3]) ** 2))


def test_adjoint_matches_finite_differences(dev):
    print("case: adjoint vs FD (GPU)", flush=True)
    s, d, target, mk, md, _ = setup(dev)
    img, _ = dev.forward(d)
    dl = 2 * (img[..., :3] - target[..., :3])
    dcol, dem = dev.backward(d, dl, len(mk))
    # the materials with the largest albedo gradients, plus the light's emission
    order = np.argsort(-np.abs(dcol).max(axis=1))
    params = [("col", int(m), int(np.argmax(np.abs(dcol[m])))) for m in order[:5]]
    light = int(np.nonzero(mk == bc.EMISSIVE)[0][0])
    params += [("em", light, 0), ("em", light, 1)]
    eps = 1e-3
    for kind_, m, ch in params:
        col = 5 + ch if kind_ == "em" else ch
        vals = []
        for sgn in (+1, -1):
            mp = md.copy()
            mp[m, col] += sgn * eps
            dev.update_materials(mk, mp)
            im, _ = dev.forward(d)
            vals.append(loss_of(im, target))
        dev.update_materials(mk, md)
        fd = (vals[0] - vals[1]) / (2 * eps)
        ad = float(dem[m, ch] if kind_ == "em" else dcol[m, ch])
        assert abs(ad - fd) <= 0.02 * abs(fd) + 1e-6, (kind_, m, ch, ad, fd)


def test_adjoint_matches_cpu_reference_gradient(dev):
    """Reference gradient from the CPU path: central FD of the oracle's fp64 integrator (same
    seeds) against the GPU adjoint on a small image."""
    print("case: adjoint vs oracle fp64 FD", flush=True)
    s, d, target, mk, md, (kind, mat, geom) = setup(dev, w=24, h=24, spp=8, bounces=4, tris=600)
    img, _ = dev.forward(d)
    dl = 2 * (img[..., :3] - target[..., :3])
    dcol, dem = dev.backward(d, dl, len(mk))
    m = int(np.argsort(-np.abs(dcol).max(axis=1))[0])
    ch = int(np.argmax(np.abs(dcol[m])))
    eps = 1e-3
    od = orc.desc_from(d)
    vals = []
    for sgn in (+1, -1):
        mp = md.copy()
        mp[m, ch] += sgn * eps
        o = orc.Scene(kind, mat, geom, mk, mp, nee_triangles=True)
        im, _, _, _ = o.render(od, precision=64)
        vals.append(float(np.sum((im[..., :3] - target[..., :3]) ** 2)))
    fd = (vals[0] - vals[1]) / (2 * eps)
    assert abs(dcol[m, ch] - fd) <= 0.02 * abs(fd), (dcol[m, ch], fd)


def test_gradient_descent_reduces_loss(dev):
    """A few steps of the inverse-rendering loop (optimise albedo from 0.5 and light x0.5)."""
    print("case: descent", flush=True)
    s, d, target, mk, md, _ = setup(dev)
    diffuse = mk == bc.DIFFUSE
    light = mk == bc.EMISSIVE
    losses = []
    for it in range(6):
        img, _ = dev.forward(d)
        losses.append(loss_of(img, target))
        dcol, dem = dev.backward(d, 2 * (img[..., :3] - target[..., :3]), len(mk))
        # normalised gradient steps: at most 0.04 in albedo and 8% of the radiance per step
        gc = dcol[diffuse]
        md[diffuse, 0:3] = np.clip(md[diffuse, 0:3] - 0.04 * gc / (np.abs(gc).max() + 1e-30), 0.01, 0.99)
        ge = dem[light]
        md[light, 5:8] = md[light, 5:8] * (1 - 0.08 * ge / (np.abs(ge).max() + 1e-30))
        dev.update_materials(mk, md)
    assert losses[-1] < 0.7 * losses[0], losses


def test_backward_requires_matching_forward(dev):
    s = bc.HostScene.workload(bc.CLOSED_BOX)
    dev.upload(s, nee_triangles=True)
    d = bc.render_desc(s, 16, 16, 2, 3, rr_depth=3)
    dev.forward(d)
    d2 = bc.render_desc(s, 16, 16, 4, 3, rr_depth=3)
    with pytest.raises(bc.BeamcastError, match="bc_forward"):
        dev.backward(d2, np.zeros((16, 16, 3), np.float32), 4)
    d.firefly_clamp = 5.0
    with pytest.raises(bc.BeamcastError, match="firefly"):
        dev.forward(d)
