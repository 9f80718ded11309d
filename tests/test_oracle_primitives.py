"""Known answers for the oracle's primitive tests and guide packing.

triHit (renderer.go:684-711) and sphereHitAccel (renderer.go:650-678) are checked on rays
worked out by hand (Moller-Trumbore barycentrics and the quadratic's roots), in both the fp64
restatement of the Go code and the fp32 form the GPU traversal runs.  Every input is a dyadic
rational, so both precisions must give the exact answers; each case is chosen so that a plausible
slip (a wrong sign in s = o - A, a swapped cross product, exclusive instead of inclusive bounds,
the det cut on |det| <= instead of <, the near root skipped) fails at least one of them.
packGuideNormalMat / packJitter (gpu_opencl_windows.go:787-797) are checked against words
packed by hand and by a round trip through the reference's decoder (decodeOctNormal, :552-563,
itself pinned in test_oracle.py).
"""
import math

import numpy as np
import pytest

import oracle as orc

EPS = 2.0 ** -20

# --- triHit ------------------------------------------------------------------------------------
# Unit right triangle A=(0,0,0), E1=(1,0,0), E2=(0,1,0).  A ray from (x, y, 1) along (0,0,-1):
# h = d x E2 = (1,0,0), det = 1, s = (x,y,1), u = x, q = s x E1 = (0,1,-y), v = y, t = E2.q = 1.
T0 = ((0, 0, 0), (1, 0, 0), (0, 1, 0))
DOWN = (0, 0, -1)
TRI_CASES = [
    # (x, y, hit)
    (0.25, 0.25, True),      # interior
    (0.0, 0.5, True),        # u = 0 edge, inclusive
    (0.5, 0.0, True),        # v = 0 edge, inclusive
    (0.5, 0.5, True),        # u + v = 1 edge, inclusive
    (0.0, 0.0, True),        # vertex A
    (1.0, 0.0, True),        # vertex B (u = 1)
    (0.0, 1.0, True),        # vertex C (v = 1)
    (-EPS, 0.5, False),      # u < 0
    (0.5, -EPS, False),      # v < 0
    (0.5 + EPS, 0.5, False),  # u + v > 1
    (1.0 + EPS, 0.0, False),  # u > 1
]


@pytest.mark.parametrize("prec", [64, 32])
def test_tri_barycentric_edges(prec):
    f = orc.tri_hit64 if prec == 64 else orc.tri_test32
    for x, y, want in TRI_CASES:
        hit, t = f(*T0, (x, y, 1.0), DOWN)
        assert hit == want, (prec, x, y)
        if want:
            assert t == 1.0, (prec, x, y, t)


@pytest.mark.parametrize("prec", [64, 32])
def test_tri_tilted_unnormalised(prec):
    """A=(1,1,1), E1=(2,0,0), E2=(0,0,4); o=(2,-3,2), d=(0,2,0) (unnormalised, as Go's camera
    rays are): h=(8,0,0), det=16, s=(1,-4,1), u=1/2, q=(0,2,8), v=1/4, t=32/16=2 (the plane y=1
    is reached after 4 units of y at 2 per unit t)."""
    f = orc.tri_hit64 if prec == 64 else orc.tri_test32
    hit, t = f((1, 1, 1), (2, 0, 0), (0, 0, 4), (2, -3, 2), (0, 2, 0))
    assert hit and t == 2.0
    # the same ray moved to u = 1/2, v = 3/4 (u + v > 1) misses; to u = -1/2 misses
    assert not f((1, 1, 1), (2, 0, 0), (0, 0, 4), (2, -3, 4), (0, 2, 0))[0]
    assert not f((1, 1, 1), (2, 0, 0), (0, 0, 4), (0, -3, 2), (0, 2, 0))[0]


@pytest.mark.parametrize("prec", [64, 32])
def test_tri_back_face_and_behind(prec):
    f = orc.tri_hit64 if prec == 64 else orc.tri_test32
    # from below along +z: h = (-1,0,0), det = -1, f = -1, u = v = 1/4, t = 1 (no back-face cull)
    hit, t = f(*T0, (0.25, 0.25, -1.0), (0, 0, 1))
    assert hit and t == 1.0
    # the triangle behind the origin: t = -1 < tmin
    assert not f(*T0, (0.25, 0.25, 1.0), (0, 0, 1))[0]


@pytest.mark.parametrize("prec", [64, 32])
def test_tri_t_bounds_inclusive(prec):
    f = orc.tri_hit64 if prec == 64 else orc.tri_test32
    o = (0.25, 0.25, 1.0)
    assert f(*T0, o, DOWN, tmin=1.0)[0]
    assert not f(*T0, o, DOWN, tmin=1.0 + EPS)[0]
    assert f(*T0, o, DOWN, tmax=1.0)[0]
    assert not f(*T0, o, DOWN, tmax=1.0 - EPS)[0]


@pytest.mark.parametrize("prec", [64, 32])
def test_tri_det_cut(prec):
    f = orc.tri_hit64 if prec == 64 else orc.tri_test32
    # a ray in the triangle's plane: h = d x E2 = (0,0,1), det = E1.h = 0
    assert not f(*T0, (-1.0, 0.25, 0.0), (1, 0, 0))[0]
    # tiny triangles hit through their interior: E2 = (0,1e-5,0) gives h = (1e-5,0,0) and
    # det = E1.x * 1e-5.  E1.x = 2e-5: det = 2e-10 >= 1e-10 is kept (u = .05, v = .1, t = 1);
    # E1.x = 5e-6: det = 5e-11 < 1e-10 is rejected although u = .2, v = .1 lie inside.
    o = (1e-6, 1e-6, 1.0)
    hit, t = f((0, 0, 0), (2e-5, 0, 0), (0, 1e-5, 0), o, DOWN)
    assert hit and abs(t - 1.0) < 1e-6
    assert not f((0, 0, 0), (5e-6, 0, 0), (0, 1e-5, 0), o, DOWN)[0]


# --- sphereHitAccel ----------------------------------------------------------------------------
# Unit sphere at the origin.  From (0,0,5) along (0,0,-1): hb = oc.d = -5, a = 1, c = 24,
# disc = 1, roots 4 and 6.
@pytest.mark.parametrize("prec", [64, 32])
def test_sphere_roots(prec):
    f = orc.sphere_hit64 if prec == 64 else orc.sphere_test32
    c0 = (0, 0, 0)
    assert f(c0, 1.0, (0, 0, 5), (0, 0, -1)) == (True, 4.0)
    # unnormalised direction: a = 4, hb = -10, c = 24, disc = 100 - 96 = 4, near root (10-2)/4 = 2
    assert f(c0, 1.0, (0, 0, 5), (0, 0, -2)) == (True, 2.0)
    # origin inside: the near root -1 is below tmin, the far root +1 is taken
    assert f(c0, 1.0, (0, 0, 0), (0, 0, 1)) == (True, 1.0)
    # sphere behind the origin: roots -6 and -4
    assert not f(c0, 1.0, (0, 0, 5), (0, 0, 1))[0]
    # off-centre sphere c=(1,2,3), r=2 from (1,2,10) along -z: near root 10-3-2 = 5
    assert f((1, 2, 3), 2.0, (1, 2, 10), (0, 0, -1)) == (True, 5.0)


@pytest.mark.parametrize("prec", [64, 32])
def test_sphere_tangent_and_bounds(prec):
    f = orc.sphere_hit64 if prec == 64 else orc.sphere_test32
    c0 = (0, 0, 0)
    # tangent: the line from (1,0,5) along -z touches at (1,0,0): disc = 0, double root 5
    assert f(c0, 1.0, (1, 0, 5), (0, 0, -1)) == (True, 5.0)
    # just outside: disc < 0 (fp64: 25 - (25 + 2^-19 + 2^-40); fp32: 1 - (1 + 2^-19))
    assert not f(c0, 1.0, (1 + EPS, 0, 5), (0, 0, -1))[0]
    # tmin inclusive on the near root; just above it the far root (6) is returned instead
    assert f(c0, 1.0, (0, 0, 5), (0, 0, -1), tmin=4.0) == (True, 4.0)
    assert f(c0, 1.0, (0, 0, 5), (0, 0, -1), tmin=4.0 + 4 * EPS) == (True, 6.0)
    # tmax inclusive; below the near root nothing is hit
    assert f(c0, 1.0, (0, 0, 5), (0, 0, -1), tmax=4.0) == (True, 4.0)
    assert not f(c0, 1.0, (0, 0, 5), (0, 0, -1), tmax=4.0 - 4 * EPS)[0]
    # a zero direction never hits
    assert not f(c0, 1.0, (0, 0, 5), (0, 0, 0))[0]


# --- guide packing -----------------------------------------------------------------------------
# word = valid bit 31 | material (11 bits) << 20 | qy << 10 | qx, q = rte(clamp(p*.5+.5)*1023)
# where p is the octahedral projection of n/|n|_1 (folded when n.z < 0).
GUIDE_WORDS = [
    ((0, 0, 1), 0, 0x80080200),      # p = (0,0) -> q = rte(511.5) = 512 (ties to even) both
    ((0, 0, -1), 0, 0x800FFFFF),     # folded: p = (copysign(1, +0), copysign(1, +0)) -> 1023, 1023
    ((1, 0, 0), 0, 0x800803FF),      # qx = 1023, qy = 512
    ((-1, 0, 0), 0, 0x80080000),     # qx = 0, qy = 512
    ((0, 1, 0), 0, 0x800FFE00),      # qx = 512, qy = 1023
    ((0, -1, 0), 0, 0x80000200),     # qx = 512, qy = 0
    ((1, 1, 0), 0, 0x800BFEFF),      # p = (.5,.5): rte(767.25) = 767 both
    ((1, 1, -2), 0, 0x800DFF7F),     # p = (.25,.25,-.5) folds to (.75,.75): rte(895.125) = 895
    ((0, 0, 1), 5, 0x80580200),      # material 5 in bits 20-30
    ((0, 0, 1), 2047, 0xFFF80200),   # largest material
    ((0, 0, 1), 5000, 0xFFF80200),   # clamped to 2047
    ((0, 0, 1), -3, 0x80080200),     # clamped to 0
]


def test_pack_guide_known_words():
    for n, m, want in GUIDE_WORDS:
        assert orc.pack_guide(n, m) == want, (n, m, hex(orc.pack_guide(n, m)))


def test_pack_jitter_known_words():
    assert orc.pack_jitter(0.0, 0.0) == 0
    assert orc.pack_jitter(1.0, 1.0) == 0xFFFFFFFF
    assert orc.pack_jitter(0.5, 0.0) == 0x00008000     # rte(32767.5) = 32768 (even) in the low half
    assert orc.pack_jitter(0.0, 0.5) == 0x80000000     # ... and in the high half
    assert orc.pack_jitter(0.25, 0.75) == 0xBFFF4000   # rte(16383.75) = 16384, rte(49151.25) = 49151
    assert orc.pack_jitter(-1.0, 2.0) == 0xFFFF0000    # clamped to [0, 1]


def test_pack_guide_round_trip():
    """decodeOctNormal(pack(n)) returns n within the quantisation: 10,000 directions uniform on
    the sphere plus the axes and the diagonals of the folded (-z) half."""
    rng = np.random.default_rng(11)
    v = rng.normal(size=(10000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    extra = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    extra += [(sx, sy, -1) for sx in (-1, 1) for sy in (-1, 1)]
    extra += [(sx, sy, 1) for sx in (-1, 1) for sy in (-1, 1)]
    v = np.concatenate([v, np.asarray(extra, float) / np.linalg.norm(extra, axis=1, keepdims=True)])
    worst = 0.0
    for i, n in enumerate(v):
        mat = i % 2048
        w = orc.pack_guide(n, mat)
        assert w >> 31 == 1 and (w >> 20) & 2047 == mat
        back = orc.decode_oct_normal(w & 1023, (w >> 10) & 1023)
        worst = max(worst, math.acos(min(1.0, float(np.dot(back, n)))))
    # rounding moves each octahedral coordinate by at most half a step, 1/1023; on the surface
    # |x|+|y|+|z| = 1 that moves the point by at most sqrt(1 + 1 + 2^2)/1023 (z follows x and
    # y), and normalising stretches it by at most 1/|p| <= sqrt(3) (the face centres), so the
    # angle is at most sqrt(18)/1023 = 0.2376 degrees; the sample reaches it to 4 digits
    bound = math.sqrt(18.0) / 1023.0
    assert worst <= bound * (1 + 1e-3), math.degrees(worst)
    assert worst > 0.95 * bound  # the sample comes within 5% of it: the bound is not slack
