"""Copies between host and device through the C ABI (bc_upload / bc_download).

Copies above 64 KB, in either direction, are staged through two 16 MB pinned buffers that host
threads fill or drain while the other one's DMA runs; scene uploads, refit, the device build,
host-pointer ray queries and frame read-backs all go through the same routines.  Bar: every
byte identical after a round trip, at sizes on both sides of the direct-copy limit, ragged
tails past whole chunks, and back-to-back uploads that reuse both pinned buffers before
anything is read back.
"""
import numpy as np
import pytest

import beamcast as bc

pytestmark = pytest.mark.gpu

MB = 1 << 20
KB = 1 << 10


@pytest.fixture(scope="module")
def dev():
    d = bc.Device(0)
    yield d
    d.close()


def _bytes(n, seed):
    return np.random.default_rng(seed).integers(0, 256, size=n, dtype=np.uint8)


@pytest.mark.parametrize("n", [1, 64 * KB, 64 * KB + 1, 4 * MB + 1, 16 * MB, 16 * MB + 3, 48 * MB + 12345])
def test_round_trip(dev, n):
    print(f"case: upload {n} bytes", flush=True)
    a = _bytes(n, n)
    p = dev.to_device(a)
    try:
        out = np.zeros_like(a)
        dev.download(p, out)
        assert np.array_equal(out, a)
    finally:
        dev.free(p)


def test_back_to_back_uploads(dev):
    """Three uploads before any read-back: later chunks reuse the pinned buffers of earlier
    ones, so each buffer must only be refilled once its transfer has finished."""
    print("case: back-to-back uploads", flush=True)
    arrs = [_bytes(n, 7 + i) for i, n in enumerate((33 * MB + 5, 17 * MB, 5 * MB + 1, 100 * KB))]
    ptrs = [dev.to_device(a) for a in arrs]
    try:
        for a, p in zip(arrs, ptrs):
            out = np.zeros_like(a)
            dev.download(p, out)
            assert np.array_equal(out, a)
    finally:
        for p in ptrs:
            dev.free(p)
