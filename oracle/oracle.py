"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker or the timed CPU comparator, never by the product.
The C sources cite the windows_portable/ functions they restate.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_int32, c_int64, c_uint32, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class RenderDesc(ctypes.Structure):
    _fields_ = [
        ("width", c_int32), ("height", c_int32), ("spp", c_int32), ("bounces", c_int32), ("rr_depth", c_int32),
        ("sampler", c_int32), ("mis", c_int32), ("power_lights", c_int32), ("sky_mode", c_int32),
        ("firefly_clamp", c_float), ("seed", c_uint32), ("sky_color", c_float * 3),
        ("origin", c_double * 3), ("lower", c_double * 3), ("horiz", c_double * 3), ("vert", c_double * 3),
        ("restir_candidates", c_int32), ("sun_dir", c_float * 3), ("sun_irradiance", c_float * 3),
        ("restir_spatial", c_int32), ("restir_radius", c_int32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        vp = c_void_p
        sig = {
            "orc_scene_new": (vp, [c_int, vp, vp, c_int, vp, vp, vp, c_int, c_int, c_int]),
            "orc_scene_free": (None, [vp]),
            "orc_scene_nprim": (c_int, [vp]),
            "orc_scene_nnode": (c_int, [vp]),
            "orc_scene_fallback_sorts": (c_int, [vp]),
            "orc_scene_nodes": (None, [vp, vp, vp, vp]),
            "orc_scene_escapes": (None, [vp, vp]),
            "orc_scene_nlights": (c_int, [vp]),
            "orc_scene_lights": (None, [vp, vp, vp, vp]),
            "orc_aabb_hit": (c_int, [vp, vp, c_double, c_double, POINTER(c_double)]),
            "orc_hit64": (None, [vp, c_int64, vp, vp, vp]),
            "orc_hit64_threaded": (None, [vp, c_int64, vp, vp, vp]),
            "orc_occluded64": (None, [vp, c_int64, vp, vp]),
            "orc_brute64": (None, [vp, c_int64, vp, vp, vp]),
            "orc_hit32": (None, [vp, c_int64, vp, vp, vp, vp, vp, c_int]),
            "orc_any32": (None, [vp, c_int64, vp, vp, vp, vp, c_int]),
            "orc_brute32": (None, [vp, c_int64, vp, vp, vp]),
            "orc_time_hit64": (c_double, [vp, c_int64, vp, c_int, c_int, vp]),
            "orc_radical_inverse": (c_double, [ctypes.c_uint64, ctypes.c_uint64]),
            "orc_pixel_rotation": (None, [c_int, c_int, POINTER(c_double), POINTER(c_double)]),
            "orc_camera_jitter": (None, [c_int, c_uint32, c_int, c_int, c_int, c_int, POINTER(c_double), POINTER(c_double)]),
            "orc_power_heuristic": (c_double, [c_double, c_double]),
            "orc_path_key": (c_uint32, [c_uint32, c_uint32, c_uint32]),
            "orc_rng_u32": (c_uint32, [c_uint32, c_uint32]),
            "orc_rng_float": (c_float, [c_uint32, c_uint32]),
            "orc_cos_hemisphere64": (None, [vp, c_uint32, c_uint32, vp]),
            "orc_camera_basis": (None, [vp, vp, vp, c_double, c_double, vp]),
            "orc_render64": (c_int64, [vp, POINTER(RenderDesc), c_int, c_int, vp, vp, POINTER(c_int64), c_int]),
            "orc_render32": (c_int64, [vp, POINTER(RenderDesc), c_int, c_int, vp, vp, POINTER(c_int64), c_int]),
            "orc_post": (c_int, [c_int, c_int, vp, vp, c_int, c_int, c_int, c_double, c_double, c_int, vp, vp]),
            "orc_to_rgba": (c_uint32, [vp, c_double, c_double, c_int]),
            "orc_decode_oct_normal": (None, [c_uint32, c_uint32, vp]),
            "orc_tri_hit64": (c_int, [vp, vp, c_double, c_double, POINTER(c_double)]),
            "orc_sphere_hit64": (c_int, [vp, vp, c_double, c_double, POINTER(c_double)]),
            "orc_tri_test32": (c_int, [vp, vp, vp, c_float, c_float, POINTER(c_float)]),
            "orc_sphere_test32": (c_int, [vp, vp, vp, c_float, c_float, POINTER(c_float)]),
            "orc_pack_guide32": (c_uint32, [vp, c_int]),
            "orc_pack_jitter": (c_uint32, [c_float, c_float]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return c_void_p(a.ctypes.data) if a is not None else None


class Scene:
    """An oracle scene built from Go-style primitives (A,B,C triangles; C,r spheres)."""

    def __init__(self, prim_kind, prim_mat, geom9, mat_kind, mat_data, bins=32, leaf=4, nee_triangles=False):
        self._keep = [np.ascontiguousarray(prim_kind, np.int32), np.ascontiguousarray(prim_mat, np.int32),
                      np.ascontiguousarray(geom9, np.float64), np.ascontiguousarray(mat_kind, np.int32),
                      np.ascontiguousarray(mat_data, np.float64)]
        pk, pm, g, mk, md = self._keep
        self.h = lib().orc_scene_new(len(mk), _p(mk), _p(md), len(pk), _p(pk), _p(pm), _p(g), bins, leaf,
                                     1 if nee_triangles else 0)

    @classmethod
    def from_host(cls, host_scene, bins=32, leaf=4, nee_triangles=False):
        return cls(*host_scene.export(), bins=bins, leaf=leaf, nee_triangles=nee_triangles)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_free(self.h)
            self.h = None

    @property
    def nprim(self):
        return lib().orc_scene_nprim(self.h)

    @property
    def nnode(self):
        return lib().orc_scene_nnode(self.h)

    def nodes(self):
        nn, npr = self.nnode, self.nprim
        box = np.zeros((nn, 6))
        link = np.zeros((nn, 4), np.int32)
        order = np.zeros(npr, np.int32)
        lib().orc_scene_nodes(self.h, _p(box), _p(link), _p(order))
        return box, link, order

    def escapes(self):
        e = np.zeros(self.nnode, np.int32)
        lib().orc_scene_escapes(self.h, _p(e))
        return e

    def lights(self):
        n = lib().orc_scene_nlights(self.h)
        lp = np.zeros(max(n, 1), np.int32)
        cdf = np.zeros(max(n, 1))
        pdf = np.zeros(max(self.nprim, 1))
        lib().orc_scene_lights(self.h, _p(lp), _p(cdf), _p(pdf))
        return lp[:n], cdf[:n], pdf[: self.nprim]

    def hit64(self, rays64, threaded=False):
        rays64 = np.ascontiguousarray(rays64, np.float64)
        n = rays64.shape[0]
        prim = np.zeros(n, np.int32)
        t = np.zeros(n)
        (lib().orc_hit64_threaded if threaded else lib().orc_hit64)(self.h, n, _p(rays64), _p(prim), _p(t))
        return prim, t

    def occluded64(self, rays64):
        rays64 = np.ascontiguousarray(rays64, np.float64)
        occ = np.zeros(rays64.shape[0], np.int32)
        lib().orc_occluded64(self.h, rays64.shape[0], _p(rays64), _p(occ))
        return occ

    def brute64(self, rays64):
        rays64 = np.ascontiguousarray(rays64, np.float64)
        n = rays64.shape[0]
        prim = np.zeros(n, np.int32)
        t = np.zeros(n)
        lib().orc_brute64(self.h, n, _p(rays64), _p(prim), _p(t))
        return prim, t

    def hit32(self, rays32, threads=0, counts=False):
        rays32 = np.ascontiguousarray(rays32, np.float32)
        n = rays32.shape[0]
        prim = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        nv = np.zeros(n, np.uint32) if counts else None
        pt = np.zeros(n, np.uint32) if counts else None
        lib().orc_hit32(self.h, n, _p(rays32), _p(prim), _p(t), _p(nv), _p(pt), threads)
        return (prim, t, nv, pt) if counts else (prim, t)

    def any32(self, rays32, threads=0, counts=False):
        rays32 = np.ascontiguousarray(rays32, np.float32)
        n = rays32.shape[0]
        occ = np.zeros(n, np.int32)
        nv = np.zeros(n, np.uint32) if counts else None
        pt = np.zeros(n, np.uint32) if counts else None
        lib().orc_any32(self.h, n, _p(rays32), _p(occ), _p(nv), _p(pt), threads)
        return (occ, nv, pt) if counts else occ

    def brute32(self, rays32):
        rays32 = np.ascontiguousarray(rays32, np.float32)
        n = rays32.shape[0]
        prim = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        lib().orc_brute32(self.h, n, _p(rays32), _p(prim), _p(t))
        return prim, t

    def time_hit64(self, rays32, any_hit=False, threads=0):
        """Go-semantics fp64 Scene.Hit / Occluded over fp32 rays, all threads; returns seconds."""
        rays32 = np.ascontiguousarray(rays32, np.float32)
        return lib().orc_time_hit64(self.h, rays32.shape[0], _p(rays32), 1 if any_hit else 0, threads, None)

    def render(self, desc: RenderDesc, precision=32, y0=0, rows=None, guides=False, threads=0):
        rows = desc.height - y0 if rows is None else rows
        dt = np.float32 if precision == 32 else np.float64
        rgba = np.zeros((rows, desc.width, 4), dt)
        g = np.zeros((rows, desc.width, 3), np.uint32) if guides else None
        shadow = c_int64()
        f = lib().orc_render32 if precision == 32 else lib().orc_render64
        rays = f(self.h, ctypes.byref(desc), y0, rows, _p(rgba), _p(g), ctypes.byref(shadow), threads)
        return rgba, g, int(rays), int(shadow.value)


def desc_from(d_product):
    """Copies the shared fields of a product RenderDesc into an oracle RenderDesc."""
    d = RenderDesc()
    for name, _ in RenderDesc._fields_:
        v = getattr(d_product, name)
        if hasattr(v, "_length_"):
            arr = getattr(d, name)
            for k in range(v._length_):
                arr[k] = v[k]
        else:
            setattr(d, name, v)
    return d


def radical_inverse(n, base):
    return lib().orc_radical_inverse(n, base)


def pixel_rotation(x, y):
    a, b = c_double(), c_double()
    lib().orc_pixel_rotation(x, y, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def camera_jitter(sampler, x, y, smp, seed=0x9E3779B9, width=1 << 16):
    """Camera jitter (jx, jy) of sample smp at pixel (x, y): sampler 1 Halton, 2 R2, else random."""
    a, b = c_double(), c_double()
    lib().orc_camera_jitter(sampler, seed, width, x, y, smp, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def power_heuristic(a, b):
    return lib().orc_power_heuristic(a, b)


def aabb_hit(box6, ray6, tmin=1e-4, tmax=1e30):
    b = np.asarray(box6, np.float64)
    r = np.asarray(ray6, np.float64)
    near = c_double()
    return bool(lib().orc_aabb_hit(_p(b), _p(r), tmin, tmax, ctypes.byref(near))), near.value


def post(rgba, guides=None, denoise=0, fxaa=0, txaa=0, exposure=0.0, gamma=2.2, hdr=2):
    """Post stack of the renderer on a frame: (linear H x W x 3 float64, RGBA8 H x W uint32)."""
    rgba = np.ascontiguousarray(rgba, np.float32)
    h, w = rgba.shape[:2]
    g = None if guides is None else np.ascontiguousarray(guides, np.uint32)
    lin = np.zeros((h, w, 3))
    pix = np.zeros((h, w), np.uint32)
    if lib().orc_post(w, h, _p(rgba), _p(g), denoise, fxaa, txaa, exposure, gamma, hdr, _p(lin), _p(pix)) != 0:
        raise MemoryError("orc_post")
    return lin, pix


def to_rgba(c, exposure=0.0, gamma=2.2, hdr=2):
    v = np.ascontiguousarray(c, np.float64)
    return int(lib().orc_to_rgba(_p(v), exposure, gamma, hdr))


def decode_oct_normal(qx, qy):
    n = np.zeros(3)
    lib().orc_decode_oct_normal(qx, qy, _p(n))
    return n


def tri_hit64(a, e1, e2, o, d, tmin=1e-4, tmax=1e30):
    """triHit (renderer.go:684-711) in fp64: (hit, t)."""
    g = np.asarray(list(a) + list(e1) + list(e2), np.float64)
    r = np.asarray(list(o) + list(d), np.float64)
    t = c_double()
    return bool(lib().orc_tri_hit64(_p(g), _p(r), tmin, tmax, ctypes.byref(t))), t.value


def sphere_hit64(center, radius, o, d, tmin=1e-4, tmax=1e30):
    """sphereHitAccel (renderer.go:650-678) in fp64: (hit, t)."""
    c = np.asarray(list(center) + [radius], np.float64)
    r = np.asarray(list(o) + list(d), np.float64)
    t = c_double()
    return bool(lib().orc_sphere_hit64(_p(c), _p(r), tmin, tmax, ctypes.byref(t))), t.value


def tri_test32(a, e1, e2, o, d, tmin=1e-4, tmax=1e30):
    """The GPU's fp32 triangle test: (hit, t)."""
    g = np.asarray(list(a) + list(e1) + list(e2), np.float32)
    of = np.asarray(o, np.float32)
    df = np.asarray(d, np.float32)
    t = c_float()
    return bool(lib().orc_tri_test32(_p(g), _p(of), _p(df), tmin, tmax, ctypes.byref(t))), t.value


def sphere_test32(center, radius, o, d, tmin=1e-4, tmax=1e30):
    """The GPU's fp32 sphere test: (hit, t)."""
    c = np.asarray(list(center) + [radius], np.float64)
    of = np.asarray(o, np.float32)
    df = np.asarray(d, np.float32)
    t = c_float()
    return bool(lib().orc_sphere_test32(_p(c), _p(of), _p(df), tmin, tmax, ctypes.byref(t))), t.value


def pack_guide(n, mat):
    """packGuideNormalMat (gpu_opencl_windows.go:787-791) as the integrator writes it."""
    v = np.asarray(n, np.float32)
    return int(lib().orc_pack_guide32(_p(v), int(mat)))


def pack_jitter(ju, jv):
    """packJitter (gpu_opencl_windows.go:792-796)."""
    return int(lib().orc_pack_jitter(float(ju), float(jv)))
