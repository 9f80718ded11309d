"""bc_render_noalloc (the render whose state lives in the library's fixed scratch array) against
bc_render_device, through the C ABI.

Both run the plain schedule over the same sample indices; the no-allocation call renders in
passes of at most 262,144 paths, so it splits the samples differently, but each sample's path and
the per-pixel sum in sample order do not depend on the pass size: images, guides and ray counts
must be bit-identical.  Two calls in a row, the second smaller than the first, check that the
contents the first call leaves in the shared array do not leak into the second.
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


def both(dev, desc, y0, rows, guides):
    npix = rows * desc.width
    out = {}
    for name in ("device", "noalloc"):
        d_rgba = dev.alloc(npix * 16)
        d_g = dev.alloc(npix * 12) if guides else None
        try:
            fn = dev.render_device if name == "device" else dev.render_noalloc
            rays, shadow = fn(desc, y0, rows, d_rgba, d_g)
            img = dev.download(d_rgba, np.zeros((rows, desc.width, 4), np.float32))
            g = dev.download(d_g, np.zeros((rows, desc.width, 3), np.uint32)) if guides else None
        finally:
            dev.free(d_rgba)
            if d_g:
                dev.free(d_g)
        out[name] = (img, g, rays, shadow)
    return out


def check_equal(out, guides):
    a, b = out["device"], out["noalloc"]
    assert a[2] == b[2] and a[3] == b[3], f"ray counts {a[2:]} vs {b[2:]}"
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)), "images differ"
    if guides:
        assert np.array_equal(a[1], b[1]), "guides differ"
    assert a[0][..., :3].mean() > 0.01


def test_noalloc_equals_render_device_twice(dev):
    # first call: the closed box, full 256 x 256 frame (65,536 pixels: 4 passes of 4 spp)
    print("case: noalloc closed box 256x256 16spp", flush=True)
    box = bc.HostScene.workload(bc.CLOSED_BOX)
    dev.upload(box, nee_triangles=True)
    d = bc.render_desc(box, 256, 256, 16, 4, rr_depth=3, sampler=1, mis=True, power=True)
    check_equal(both(dev, d, 0, 256, True), True)
    # second call, smaller: 24 rows of a 64 x 40 Cornell frame from row 8, other transport
    print("case: noalloc cornell rows 8..32 of 64x40 6spp", flush=True)
    room = bc.HostScene.workload(bc.CORNELL, bins=16, leaf=4)
    dev.upload(room)
    d2 = bc.render_desc(room, 64, 40, 6, 5, rr_depth=2, sampler=0, mis=False, power=True, seed=77)
    check_equal(both(dev, d2, 8, 24, False), False)


def test_noalloc_rejects_what_it_cannot_hold(dev):
    print("case: noalloc argument checks", flush=True)
    box = bc.HostScene.workload(bc.CLOSED_BOX)
    dev.upload(box, nee_triangles=True)
    d_rgba = dev.alloc(16)
    try:
        big = bc.render_desc(box, 1024, 512, 1, 1)  # 524,288 pixels > 262,144
        with pytest.raises(bc.BeamcastError, match="262144 pixels"):
            dev.render_noalloc(big, 0, 512, d_rgba)
        pers = bc.render_desc(box, 16, 16, 1, 1)
        pers.mode = 1
        with pytest.raises(bc.BeamcastError, match="plain schedule"):
            dev.render_noalloc(pers, 0, 16, d_rgba)
    finally:
        dev.free(d_rgba)
