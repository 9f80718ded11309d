"""ctypes bindings of libbeamcast_hip.so (include/beamcast_hip.h, include/beamcast_host.h).

Host-side helper for callers without Go: it exposes the C ABI that the Go renderer's cgo
shim binds (the renderOpenCL / openCLGPUAvailable replacement of
windows_portable/gpu_opencl_windows.go:172,565) and the C++ mirror of the Go Scene.
There is no fallback: if the library is missing or no gfx950 device is present, the
calls raise.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbeamcast_hip.so")

BC_OK = 0

# material kinds (renderer.go:86-91) and primitive kinds (renderer.go:101-104)
DIFFUSE, METAL, GLASS, EMISSIVE = 0, 1, 2, 3
SPHERE, TRIANGLE = 0, 1

# workload kinds of bch_workload_scene
CLOSED_BOX, SOUP, SHOWCASE, CORNELL, STRESS, TERRAIN, BOX_RETESS, BATCH_SCENE = 1, 2, 3, 4, 5, 6, 7, 8


class SceneView(ctypes.Structure):
    _fields_ = [
        ("nnode", c_int32), ("node_box", POINTER(c_double)), ("node_link", POINTER(c_int32)),
        ("nprim", c_int32), ("order", POINTER(c_int32)), ("prim_kind", POINTER(c_int32)),
        ("prim_mat", POINTER(c_int32)), ("prim_geom", POINTER(c_double)),
        ("nmat", c_int32), ("mat_kind", POINTER(c_int32)), ("mat_data", POINTER(c_double)),
        ("nee_triangles", c_int32),
    ]


class PostDesc(ctypes.Structure):
    _fields_ = [
        ("denoise_passes", c_int32), ("fxaa", c_int32), ("txaa", c_int32), ("hdr", c_int32),
        ("exposure", c_float), ("gamma", c_float), ("reserved", c_int32 * 4),
    ]


def post_desc(denoise=0, fxaa=0, txaa=0, hdr=2, exposure=0.0, gamma=2.2):
    p = PostDesc()
    p.denoise_passes, p.fxaa, p.txaa, p.hdr, p.exposure, p.gamma = denoise, fxaa, txaa, hdr, exposure, gamma
    return p


class Settings(ctypes.Structure):
    """bch_settings: the RenderSettings fields of the GPU render path (renderer.go:1027-1062)."""
    _fields_ = [(n, c_int32) for n in ("width", "height", "spp", "bounces", "rr_depth", "sampler", "mis",
                                       "power_lights", "msaa", "tsaa", "taa", "txaa", "hbao", "hbao_plus",
                                       "debug_view", "denoise", "progressive", "publish_hz", "upscale")]
    _fields_ += [("firefly_clamp", c_double)]


class Guide(ctypes.Structure):
    """bch_guide: Guide (renderer.go:726-731) as renderOpenCL decodes it."""
    _fields_ = [("p", c_double * 3), ("n", c_double * 3), ("albedo", c_double * 3), ("depth", c_double),
                ("valid", c_int32), ("pad", c_int32)]


CHUNK_FN = ctypes.CFUNCTYPE(None, c_void_p, c_int32, c_int32, POINTER(c_double), c_double)


class RenderDesc(ctypes.Structure):
    _fields_ = [
        ("width", c_int32), ("height", c_int32), ("spp", c_int32), ("bounces", c_int32),
        ("rr_depth", c_int32), ("sampler", c_int32), ("mis", c_int32), ("power_lights", c_int32),
        ("sky_mode", c_int32), ("firefly_clamp", c_float), ("seed", c_uint32),
        ("sky_color", c_float * 3),
        ("origin", c_double * 3), ("lower", c_double * 3), ("horiz", c_double * 3), ("vert", c_double * 3),
        ("mode", c_int32), ("samples_per_pass", c_int32), ("reserved", c_int32 * 6),
        ("restir_candidates", c_int32), ("sun_dir", c_float * 3), ("sun_irradiance", c_float * 3),
        ("restir_spatial", c_int32), ("restir_radius", c_int32),
    ]


_lib = None


def lib():
    """Loads the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64 = c_void_p, c_int64
        sig = {
            "bc_device_available": (c_int, [c_char_p, c_int]),
            "bc_create": (vp, [c_int]),
            "bc_destroy": (None, [vp]),
            "bc_last_error": (c_char_p, [vp]),
            "bc_device_name": (c_char_p, [vp]),
            "bc_scene_upload": (c_int, [vp, vp, c_uint64, c_uint64]),
            "bc_scene_stats": (c_int, [vp, POINTER(c_int64)]),
            "bc_render": (c_int, [vp, POINTER(RenderDesc), c_int, c_int, vp, vp, POINTER(c_uint64), vp]),
            "bc_render_device": (c_int, [vp, POINTER(RenderDesc), c_int, c_int, vp, vp, POINTER(c_uint64), POINTER(c_uint64)]),
            "bc_render_noalloc": (c_int, [vp, POINTER(RenderDesc), c_int, c_int, vp, vp, POINTER(c_uint64), POINTER(c_uint64)]),
            "bc_trace_counts": (c_int, [vp, vp, ctypes.c_int64, c_int, vp, vp]),
            "bc_valu_probe": (c_int, [vp, c_int, c_int, c_int, POINTER(ctypes.c_float)]),
            "bc_line_probe": (c_int, [vp, vp, c_uint64, c_int, c_int, POINTER(ctypes.c_float)]),
            "bc_render_tuning": (c_int, [vp, POINTER(c_int32)]),
            "bc_batch_upload": (c_int, [vp, c_int, vp]),
            "bc_postprocess_device": (c_int, [vp, POINTER(PostDesc), c_int, c_int, vp, vp, vp, vp]),
            "bc_render_post": (c_int, [vp, POINTER(RenderDesc), POINTER(PostDesc), vp, vp, POINTER(c_uint64)]),
            "bc_batch_render_device": (c_int, [vp, c_int, vp, vp, POINTER(c_uint64), POINTER(c_uint64)]),
            "bc_forward": (c_int, [vp, POINTER(RenderDesc), c_int, c_int, vp, POINTER(c_uint64)]),
            "bc_backward": (c_int, [vp, POINTER(RenderDesc), c_int, c_int, vp, vp, vp]),
            "bc_update_materials": (c_int, [vp, c_int32, vp, vp]),
            "bc_scene_refit": (c_int, [vp, vp, c_uint64]),
            "bc_scene_arrays": (c_int, [vp, vp, vp, vp]),
            "bc_set_flatten": (c_int, [vp, c_int]),
            "bc_set_render_ahead": (c_int, [vp, c_int]),
            "bc_set_trace_refill": (c_int, [vp, c_int, c_int]),
            "bc_set_trace_order": (c_int, [vp, c_int, c_int]),
            "bc_trace_order_last": (c_int, [vp, vp]),
            "bc_trace": (c_int, [vp, vp, i64, c_int, vp]),
            "bc_trace_device": (c_int, [vp, vp, i64, c_int, vp]),
            "bc_alloc": (vp, [vp, c_uint64]),
            "bc_free": (c_int, [vp, vp]),
            "bc_upload": (c_int, [vp, vp, vp, c_uint64]),
            "bc_download": (c_int, [vp, vp, vp, c_uint64]),
            "bc_copy_device": (c_int, [vp, vp, vp, c_uint64]),
            "bc_synchronize": (c_int, [vp]),
            "bc_timer_start": (c_int, [vp]),
            "bc_timer_stop": (c_int, [vp, POINTER(c_float)]),
            "bch_scene_new": (vp, []),
            "bch_scene_free": (None, [vp]),
            "bch_add_material": (c_int, [vp, c_int, POINTER(c_double)]),
            "bch_add_sphere": (None, [vp, POINTER(c_double), c_double, c_int]),
            "bch_add_triangle": (None, [vp, POINTER(c_double), POINTER(c_double), POINTER(c_double), c_int]),
            "bch_build": (None, [vp, c_int, c_int]),
            "bch_build_median": (None, [vp, c_int]),
            "bch_flatten_check": (c_int, [vp, POINTER(c_int64)]),
            "bch_quality": (c_int, [ctypes.c_char_p, ctypes.c_char_p, vp, POINTER(c_double), ctypes.c_char_p, c_int]),
            "bch_effective_spp": (c_int, [ctypes.c_char_p, c_int, c_int, c_int]),
            "bch_refit": (None, [vp]),
            "bch_build_device": (c_int, [vp, vp, c_int, c_int]),
            "bch_load_obj": (vp, [ctypes.c_char_p, vp, vp, c_int, c_int]),
            "bch_write_obj": (c_int, [vp, ctypes.c_char_p]),
            "bch_ingest_obj": (vp, [vp, ctypes.c_char_p, c_int, c_int, c_int, vp, vp, c_int]),
            "bch_pages_write": (c_int, [ctypes.c_char_p, vp, c_int64, ctypes.c_uint32]),
            "bch_pages_open": (vp, [ctypes.c_char_p, vp, c_int]),
            "bch_pages_close": (None, [vp]),
            "bch_pages_count": (c_int64, [vp]),
            "bch_pages_info": (c_int, [vp, c_int64, vp, vp]),
            "bch_pages_load": (c_int, [vp, c_int64, vp]),
            "bch_pages_append": (c_int, [vp, vp, c_int64, c_int64]),
            "bch_export_bounds": (None, [vp, vp, vp]),
            "bc_build_bvh": (c_int, [vp, c_int32, vp, vp, c_int32, c_int32, c_int32, vp, vp, vp, vp]),
            "bch_counts": (None, [vp, POINTER(c_int32)]),
            "bch_revisions": (None, [vp, POINTER(c_uint64)]),
            "bch_view": (POINTER(SceneView), [vp, c_int]),
            "bch_export_prims": (None, [vp, vp, vp, vp]),
            "bch_export_materials": (None, [vp, vp, vp]),
            "bch_translate": (None, [vp, c_int, c_int, POINTER(c_double)]),
            "bch_camera_basis": (None, [POINTER(c_double)] * 3 + [c_double, c_double, POINTER(c_double)]),
            "bch_settings_quality": (c_int, [ctypes.c_char_p, POINTER(Settings)]),
            "bch_render_desc_for": (c_int, [POINTER(c_double), POINTER(Settings), POINTER(RenderDesc)]),
            "bch_render_hip": (c_int, [vp, POINTER(c_double), POINTER(Settings), POINTER(c_int32), CHUNK_FN, vp,
                                       POINTER(c_double), POINTER(Guide), POINTER(ctypes.c_int64), POINTER(c_int32),
                                       ctypes.c_char_p, c_int, ctypes.c_char_p, c_int]),
            "bch_render_hip_release": (None, []),
            "bch_workload_scene": (vp, [c_int, c_uint64, c_int64, c_int, c_int, POINTER(c_double)]),
            "bch_workload_rays": (c_int, [c_int, c_uint64, c_int64, c_int, c_int64, c_int64, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return c_void_p(a.ctypes.data) if a is not None else None


def device_available():
    buf = ctypes.create_string_buffer(256)
    ok = lib().bc_device_available(buf, 256)
    return bool(ok), buf.value.decode()


QUALITY_FIELDS = ("Width", "Height", "SPP", "Bounces", "Denoise", "MIS", "PowerLightSampling", "AdaptiveSampling",
                  "ProgressiveUpdates", "Sampler", "MSAA", "TSAA", "TAA", "TXAA", "FXAA", "HBAO", "HBAOPlus", "RRDepth",
                  "Backend", "BVHBins")
SAMPLERS = ("Random", "Halton", "R2")
BACKENDS = ("Auto", "CPU", "OpenCL GPU")


def quality(name="", backend=None):
    """Quality(name) of the host mirror, or UltraRealismSettings(backend) when backend is given:
    a dict of the RenderSettings fields the presets set (Sampler and Backend as names)."""
    out = np.zeros(20, np.int32)
    ff = c_double()
    label = ctypes.create_string_buffer(64)
    rc = lib().bch_quality(name.encode(), backend.encode() if backend is not None else None, _ptr(out), ctypes.byref(ff),
                           label, 64)
    if rc != 0:
        raise ValueError(f"unknown backend {backend!r}")
    q = {k: int(v) for k, v in zip(QUALITY_FIELDS, out)}
    for k in ("Denoise", "MIS", "PowerLightSampling", "AdaptiveSampling", "ProgressiveUpdates"):
        q[k] = bool(q[k])
    q["Sampler"] = SAMPLERS[q["Sampler"]]
    q["Backend"] = BACKENDS[q["Backend"]]
    q["FireflyClamp"] = ff.value
    q["Quality"] = label.value.decode()
    return q


def effective_spp(name, spp, msaa, tsaa):
    """effectiveSPP of Quality(name) with SPP, MSAA and TSAA replaced."""
    return int(lib().bch_effective_spp(name.encode(), spp, msaa, tsaa))


def settings_quality(name):
    """Quality(name) (renderer.go:1064-1117) as a Settings structure."""
    st = Settings()
    if lib().bch_settings_quality(name.encode(), ctypes.byref(st)) != 0:
        raise ValueError("bch_settings_quality")
    return st


def render_desc_for(camera, settings):
    """The RenderDesc renderOpenCL's host loop renders a frame with."""
    cam = np.ascontiguousarray(camera, np.float64)
    d = RenderDesc()
    if lib().bch_render_desc_for(cam.ctypes.data_as(POINTER(c_double)), ctypes.byref(settings), ctypes.byref(d)) != 0:
        raise ValueError("bch_render_desc_for")
    return d


def render_hip(scene, camera, settings, cancel=None, on_chunk=None, guides=False):
    """renderOpenCL over the HIP backend (bch_render_hip, the process-wide runtime).

    cancel: a ctypes c_int32 polled before every chunk; on_chunk(y0, rows, rgb[rows, W, 3], progress).
    Returns dict(linear[H, W, 3], guides (structured array or None), rays, scene_action, device)."""
    d = render_desc_for(camera, settings)
    w, h = d.width, d.height
    lin = np.zeros((h, w, 3))
    gd = (Guide * (w * h))() if guides else None
    rays = ctypes.c_int64()
    action = c_int32()
    name = ctypes.create_string_buffer(256)
    err = ctypes.create_string_buffer(512)

    def cb(_user, y0, rows, data, progress):
        if on_chunk is not None:
            arr = np.ctypeslib.as_array(data, shape=(rows * w * 3,)).reshape(rows, w, 3).copy()
            on_chunk(int(y0), int(rows), arr, float(progress))

    fn = CHUNK_FN(cb)
    cam = np.ascontiguousarray(camera, np.float64)
    rc = lib().bch_render_hip(scene.h, cam.ctypes.data_as(POINTER(c_double)), ctypes.byref(settings),
                              ctypes.byref(cancel) if cancel is not None else None, fn, None,
                              lin.ctypes.data_as(POINTER(c_double)), gd, ctypes.byref(rays), ctypes.byref(action),
                              name, 256, err, 512)
    if rc != 0:
        raise BeamcastError(f"status {rc}: {err.value.decode()}")
    g = None
    if guides:
        g = np.ctypeslib.as_array(gd).copy()
    return {"linear": lin, "guides": g, "rays": int(rays.value), "scene_action": int(action.value),
            "device": name.value.decode()}


def render_hip_release():
    lib().bch_render_hip_release()


def camera_basis(pos, target, up, fov, aspect):
    out = np.zeros(12)
    arr = [np.asarray(v, dtype=np.float64) for v in (pos, target, up)]
    lib().bch_camera_basis(*[a.ctypes.data_as(POINTER(c_double)) for a in arr], float(fov), float(aspect),
                           out.ctypes.data_as(POINTER(c_double)))
    return out.reshape(4, 3)


class HostScene:
    """The C++ mirror of the Go Scene (renderer.go:225-641)."""

    def __init__(self, handle=None, camera=None):
        self.h = handle if handle is not None else lib().bch_scene_new()
        if not self.h:
            raise ValueError("invalid scene")
        self.camera = camera

    @classmethod
    def load_obj(cls, path, threads=0):
        """LoadOBJ (parallel parse): (scene, camera[10]); not built.  Raises ValueError with
        LoadOBJ's message."""
        cam = np.zeros(10)
        err = ctypes.create_string_buffer(512)
        h = lib().bch_load_obj(os.fsencode(path), cam.ctypes.data_as(POINTER(c_double)), err, 512, threads)
        if not h:
            raise ValueError(err.value.decode())
        return cls(h, cam), cam

    @classmethod
    def ingest_obj(cls, device, path, bins=32, leaf=4, threads=0):
        """OBJ straight to a resident scene on `device` (parallel parse, device SAH build,
        upload): (scene, camera[10])."""
        cam = np.zeros(10)
        err = ctypes.create_string_buffer(512)
        h = lib().bch_ingest_obj(device.ctx, os.fsencode(path), bins, leaf, threads,
                                 cam.ctypes.data_as(POINTER(c_double)), err, 512)
        if not h:
            raise ValueError(err.value.decode())
        return cls(h, cam), cam

    def write_obj(self, path):
        """Writes the triangles as OBJ text (three vertices and one face per triangle, %.17g)."""
        if lib().bch_write_obj(self.h, os.fsencode(path)) != 0:
            raise OSError(f"cannot write {path}")

    def append_pages(self, pages, first=0, count=None):
        count = pages.count() - first if count is None else count
        if lib().bch_pages_append(self.h, pages.h, first, count) != 0:
            raise ValueError("cannot append pages")

    @classmethod
    def workload(cls, kind, seed=0x5EED, param=0, bins=32, leaf=4):
        cam = np.zeros(10)
        h = lib().bch_workload_scene(kind, seed, param, bins, leaf, cam.ctypes.data_as(POINTER(c_double)))
        if not h:
            raise ValueError(f"unknown workload {kind}")
        return cls(h, cam)

    def __del__(self):
        if getattr(self, "h", None):
            lib().bch_scene_free(self.h)
            self.h = None

    def add_material(self, kind, color=(0, 0, 0), rough=0.0, ior=0.0, emission=(0, 0, 0)):
        d = np.array([*color, rough, ior, *emission], dtype=np.float64)
        return lib().bch_add_material(self.h, kind, d.ctypes.data_as(POINTER(c_double)))

    def add_sphere(self, c, r, mat):
        c = np.asarray(c, dtype=np.float64)
        lib().bch_add_sphere(self.h, c.ctypes.data_as(POINTER(c_double)), float(r), mat)

    def add_triangle(self, a, b, c, mat):
        a, b, c = (np.asarray(v, dtype=np.float64) for v in (a, b, c))
        lib().bch_add_triangle(self.h, *(v.ctypes.data_as(POINTER(c_double)) for v in (a, b, c)), mat)

    def build(self, bins=32, leaf=4):
        lib().bch_build(self.h, bins, leaf)

    def flatten_check(self):
        """Host BVH8 flattening without a device: (wide nodes, depth, leaf slots, every primitive in
        exactly one leaf slot)."""
        o = (c_int64 * 4)()
        if lib().bch_flatten_check(self.h, o) != 0:
            raise ValueError("scene view cannot be encoded")
        return int(o[0]), int(o[1]), int(o[2]), bool(o[3])

    def build_median(self, leaf=4):
        """Scene.Build with BVHMode "Median"."""
        lib().bch_build_median(self.h, leaf)

    def refit(self):
        lib().bch_refit(self.h)

    def build_device(self, device, bins=32, leaf=4):
        """Scene.Build with the tree made on the GPU (bc_build_bvh); same Nodes and Order."""
        device._check(lib().bch_build_device(self.h, device.ctx, bins, leaf))

    def bounds(self):
        """(Primitive.Box [n, 6], Primitive.Centroid [n, 3])."""
        n = self.counts()[0]
        box = np.zeros((n, 6))
        cen = np.zeros((n, 3))
        lib().bch_export_bounds(self.h, box.ctypes.data_as(POINTER(c_double)), cen.ctypes.data_as(POINTER(c_double)))
        return box, cen

    def translate(self, start, count, d):
        d = np.asarray(d, dtype=np.float64)
        lib().bch_translate(self.h, start, count, d.ctypes.data_as(POINTER(c_double)))

    def counts(self):
        o = (c_int32 * 3)()
        lib().bch_counts(self.h, o)
        return int(o[0]), int(o[1]), int(o[2])

    def revisions(self):
        o = (c_uint64 * 2)()
        lib().bch_revisions(self.h, o)
        return int(o[0]), int(o[1])

    def view(self, nee_triangles=False):
        return lib().bch_view(self.h, 1 if nee_triangles else 0)

    def nodes(self):
        """(node_box [n,6], node_link [n,4], order [nprim]) copied out of the view."""
        v = self.view().contents
        nn, npr = v.nnode, v.nprim
        box = np.ctypeslib.as_array(v.node_box, shape=(nn * 6,)).reshape(nn, 6).copy()
        link = np.ctypeslib.as_array(v.node_link, shape=(nn * 4,)).reshape(nn, 4).copy()
        order = np.ctypeslib.as_array(v.order, shape=(npr,)).copy()
        return box, link, order

    def export(self):
        """(prim_kind, prim_mat, geom9 [A,B,C or C,r], mat_kind, mat_data8) as numpy arrays."""
        npr, _, nm = self.counts()
        kind = np.zeros(npr, np.int32)
        mat = np.zeros(npr, np.int32)
        geom = np.zeros((npr, 9))
        lib().bch_export_prims(self.h, _ptr(kind), _ptr(mat), _ptr(geom))
        mk = np.zeros(nm, np.int32)
        md = np.zeros((nm, 8))
        lib().bch_export_materials(self.h, _ptr(mk), _ptr(md))
        return kind, mat, geom, mk, md


PACKED_TRIANGLE = np.dtype([("v0", "<f4", 4), ("e1", "<f4", 4), ("e2", "<f4", 4), ("normal", "<f4", 4),
                            ("material", "<u4"), ("pad", "<u4", 3)])


def pages_write(path, tris, per_page=65536):
    """PagedTriangleStore::write of PACKED_TRIANGLE records."""
    tris = np.ascontiguousarray(tris, PACKED_TRIANGLE)
    rc = lib().bch_pages_write(os.fsencode(path), _ptr(tris), len(tris), per_page)
    if rc != 0:
        raise OSError(f"page store write failed ({rc})")


class Pages:
    """PagedTriangleStore (BGXPAGE1) opened for reading."""

    def __init__(self, path):
        err = ctypes.create_string_buffer(256)
        self.h = lib().bch_pages_open(os.fsencode(path), err, 256)
        if not self.h:
            raise ValueError(err.value.decode())

    def close(self):
        if self.h:
            lib().bch_pages_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def count(self):
        return int(lib().bch_pages_count(self.h))

    def info(self, page):
        off, cnt = ctypes.c_uint64(), ctypes.c_uint32()
        if lib().bch_pages_info(self.h, page, ctypes.byref(off), ctypes.byref(cnt)) != 0:
            raise IndexError(page)
        return int(off.value), int(cnt.value)

    def load(self, page):
        out = np.zeros(self.info(page)[1], PACKED_TRIANGLE)
        if lib().bch_pages_load(self.h, page, _ptr(out)) != 0:
            raise ValueError("truncated Beamcast GX page payload")
        return out


def workload_rays(set_id, width, any_hit=False, first=0, count=None, seed=0x5EED):
    """Ray set of the traversal workload: n x 8 float32 (o.xyz, tmin, d.xyz, tmax)."""
    if count is None:
        count = width * width - first
    out = np.empty((count, 8), np.float32)
    rc = lib().bch_workload_rays(set_id, seed, width, 1 if any_hit else 0, first, count, _ptr(out))
    if rc != 0:
        raise ValueError("bad ray-set arguments")
    return out


class BeamcastError(RuntimeError):
    pass


class Device:
    """A bc_ctx: persistent runtime on one gfx950 device."""

    def __init__(self, ordinal=0):
        L = lib()
        self.ctx = L.bc_create(ordinal)
        if not self.ctx:
            raise BeamcastError(L.bc_last_error(None).decode())

    def close(self):
        if self.ctx:
            lib().bc_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != BC_OK:
            raise BeamcastError(f"status {rc}: {lib().bc_last_error(self.ctx).decode()}")

    @property
    def name(self):
        return lib().bc_device_name(self.ctx).decode()

    def upload(self, scene: HostScene, nee_triangles=False, revisions=None):
        topo, geom = revisions if revisions is not None else scene.revisions()
        self._check(lib().bc_scene_upload(self.ctx, scene.view(nee_triangles), topo, geom))

    def refit(self, scene: HostScene, nee_triangles=False, geom_revision=None):
        """Geometry-only update (bc_scene_refit): primitives of the resident scene moved, same
        topology.  Call scene.refit() (Scene.Refit) first so the host BVH matches."""
        geom = geom_revision if geom_revision is not None else scene.revisions()[1]
        self._check(lib().bc_scene_refit(self.ctx, scene.view(nee_triangles), geom))

    def set_trace_refill(self, closest, any_hit):
        """Ray-query schedule: refill a wave's finished lanes once that many are idle (0: off; closest
        -1: chosen per batch by the library's coherence probe, the default)."""
        self._check(lib().bc_set_trace_refill(self.ctx, closest, any_hit))

    def set_trace_order(self, mode, cell_bits=4):
        """Ray binning before queries of 65,536 rays or more: 0 off (default), 1 always, 2 when the
        batch is incoherent; cell_bits 2..4 = log2 of the origin cells per axis."""
        self._check(lib().bc_set_trace_order(self.ctx, mode, cell_bits))

    def trace_order_last(self):
        """(binned, differing sampled pairs of 4096 or -1) for the last ray query."""
        out = np.zeros(2, np.int32)
        self._check(lib().bc_trace_order_last(self.ctx, _ptr(out)))
        return int(out[0]), int(out[1])

    def set_render_ahead(self, on):
        """bc_set_render_ahead: 1 (default) lets bc_render of a row chunk render the following
        chunks with it (one render per block); 0 renders each alone."""
        self._check(lib().bc_set_render_ahead(self.ctx, 1 if on else 0))

    def set_flatten(self, mode):
        """BVH8 flattening of later uploads: 0 automatic, 1 host, 2 device (same bytes)."""
        self._check(lib().bc_set_flatten(self.ctx, mode))

    def upload_view(self, view, topo, geom):
        self._check(lib().bc_scene_upload(self.ctx, view, topo, geom))

    def stats(self):
        o = (c_int64 * 4)()
        self._check(lib().bc_scene_stats(self.ctx, o))
        return {"nodes8": o[0], "depth8": o[1], "prims": o[2], "bytes": o[3]}

    NODE8 = np.dtype([("link", "<u4"), ("grid", "<u4", 3), ("qlo", "u1", (3, 8)), ("qhi", "u1", (3, 8))])
    REC_HOLE = 0xFFFFFFFF

    @staticmethod
    def node_fields(nodes):
        """Decoded bc_node8 fields (bc_layout.h): child_base, imask, lmask (leaf slots), spheres,
        k [n, 3] (origin grid integers), e [n, 3] (grid exponents), o [n, 3] (origin = k 2^e)."""
        link = nodes["link"].astype(np.uint32)
        g = nodes["grid"].astype(np.uint32)
        k = ((g << np.uint32(11)).view(np.int32) >> 11).astype(np.int64)
        e = ((g >> np.uint32(21)) & np.uint32(0xFF)).astype(np.int64) - 127
        lmask = (g[:, 0] >> 29) | ((g[:, 1] >> 29) << 3) | (((g[:, 2] >> 29) & 3) << 6)
        return {"child_base": link & np.uint32(0xFFFFFF), "imask": link >> np.uint32(24), "lmask": lmask,
                "spheres": (g[:, 2] >> 31) != 0, "k": k, "e": e, "o": np.ldexp(k.astype(np.float64), e)}

    def scene_arrays(self):
        """The resident BVH8: (nodes [NODE8], records [8 n, 9] float32, ids [8 n] uint32): eight record
        slots per node, leaf slot s of node i at 8 i + s, holes with id REC_HOLE."""
        st = self.stats()
        nodes = np.zeros(st["nodes8"], self.NODE8)
        recs = np.zeros((8 * st["nodes8"], 9), np.float32)
        ids = np.zeros(8 * st["nodes8"], np.uint32)
        self._check(lib().bc_scene_arrays(self.ctx, _ptr(nodes), _ptr(recs), _ptr(ids)))
        return nodes, recs, ids

    def trace(self, rays, any_hit=False):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = rays.shape[0]
        if any_hit:
            out = np.empty(n, np.int32)
        else:
            out = np.empty(n, dtype=[("t", np.float32), ("prim", np.int32)])
        self._check(lib().bc_trace(self.ctx, _ptr(rays), n, 1 if any_hit else 0, _ptr(out)))
        return out

    # --- device memory / timing helpers
    def alloc(self, nbytes):
        p = lib().bc_alloc(self.ctx, nbytes)
        if not p:
            raise BeamcastError(lib().bc_last_error(self.ctx).decode())
        return p

    def free(self, p):
        self._check(lib().bc_free(self.ctx, p))

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        self._check(lib().bc_upload(self.ctx, p, _ptr(arr), arr.nbytes))
        return p

    def download(self, p, out):
        self._check(lib().bc_download(self.ctx, _ptr(out), p, out.nbytes))
        return out

    def copy_device(self, d_dst, d_src, nbytes):
        self._check(lib().bc_copy_device(self.ctx, d_dst, d_src, nbytes))

    def trace_device(self, d_rays, n, any_hit, d_out):
        self._check(lib().bc_trace_device(self.ctx, d_rays, n, 1 if any_hit else 0, d_out))

    def sync(self):
        self._check(lib().bc_synchronize(self.ctx))

    def timer_start(self):
        self._check(lib().bc_timer_start(self.ctx))

    def timer_stop(self):
        ms = c_float()
        self._check(lib().bc_timer_stop(self.ctx, ctypes.byref(ms)))
        return float(ms.value)

    def render(self, desc: RenderDesc, y0=0, rows=None, guides=False, cancel=None, out=None, guides_out=None):
        """bc_render of rows [y0, y0 + rows) into host memory -> (rgba [rows, W, 4], guides [rows, W, 3]
        or None, path segments).  out / guides_out: C-contiguous float32 / uint32 arrays of those
        shapes to write into (a caller's chunk buffers reused across chunks) instead of new ones."""
        rows = desc.height - y0 if rows is None else rows

        def dest(a, dtype, ch):
            if a is None:
                return np.zeros((rows, desc.width, ch), dtype)
            if a.dtype != dtype or a.shape != (rows, desc.width, ch) or not a.flags.c_contiguous:
                raise ValueError(f"output buffer must be a C-contiguous {np.dtype(dtype).name} array of shape "
                                 f"{(rows, desc.width, ch)}")
            return a

        rgba = dest(out, np.float32, 4)
        g = dest(guides_out, np.uint32, 3) if guides else None
        rays = c_uint64()
        cflag = None
        if cancel is not None:
            cflag = np.array([1 if cancel else 0], np.int32)
        self._check(lib().bc_render(self.ctx, ctypes.byref(desc), y0, rows, _ptr(rgba), _ptr(g) if g is not None else None,
                                    ctypes.byref(rays), _ptr(cflag) if cflag is not None else None))
        return rgba, g, int(rays.value)

    def build_bvh(self, prim_box, prim_centroid, bins=16, leaf=4):
        """bc_build_bvh: the Go SAH tree built on the GPU -> (node_box [m, 6], node_link [m, 4], order [n])."""
        box = np.ascontiguousarray(prim_box, np.float64)
        cen = np.ascontiguousarray(prim_centroid, np.float64)
        n = box.shape[0]
        cap = max(1, 2 * n - 1)
        nbox = np.zeros((cap, 6))
        link = np.zeros((cap, 4), np.int32)
        order = np.zeros(n, np.int32)
        nn = c_int32()
        self._check(lib().bc_build_bvh(self.ctx, n, _ptr(box), _ptr(cen), bins, leaf, cap, _ptr(nbox), _ptr(link),
                                       _ptr(order), ctypes.byref(nn)))
        return nbox[: nn.value], link[: nn.value], order

    def update_materials(self, mat_kind, mat_data):
        mk = np.ascontiguousarray(mat_kind, np.int32)
        md = np.ascontiguousarray(mat_data, np.float64)
        self._check(lib().bc_update_materials(self.ctx, len(mk), _ptr(mk), _ptr(md)))

    def forward(self, desc: RenderDesc, y0=0, rows=None):
        """Differentiable render, forward: image (rows, W, 4) and path segments."""
        rows = desc.height - y0 if rows is None else rows
        rgba = np.zeros((rows, desc.width, 4), np.float32)
        rays = c_uint64()
        self._check(lib().bc_forward(self.ctx, ctypes.byref(desc), y0, rows, _ptr(rgba), ctypes.byref(rays)))
        return rgba, int(rays.value)

    def backward(self, desc: RenderDesc, dloss_dimage, nmat, y0=0, rows=None):
        """Path-replay backward: (d loss / d Color [nmat,3], d loss / d Emission [nmat,3])."""
        rows = desc.height - y0 if rows is None else rows
        dl = np.ascontiguousarray(dloss_dimage, np.float32).reshape(rows, desc.width, 3)
        dc = np.zeros((nmat, 3), np.float32)
        de = np.zeros((nmat, 3), np.float32)
        self._check(lib().bc_backward(self.ctx, ctypes.byref(desc), y0, rows, _ptr(dl), _ptr(dc), _ptr(de)))
        return dc, de

    def tuning(self):
        o = (c_int32 * 8)()
        self._check(lib().bc_render_tuning(self.ctx, o))
        return {"pool": o[0], "iters_per_graph": o[1], "stack": o[2], "windows": o[3], "segment": o[4],
                "material_sort": o[5], "plain": o[6]}

    def render_device(self, desc: RenderDesc, y0, rows, d_rgba, d_guides=None):
        rays, shadow = c_uint64(), c_uint64()
        self._check(lib().bc_render_device(self.ctx, ctypes.byref(desc), y0, rows, d_rgba, d_guides,
                                           ctypes.byref(rays), ctypes.byref(shadow)))
        return int(rays.value), int(shadow.value)

    def trace_counts(self, d_rays, n, any_hit, d_out, d_counts):
        """bc_trace_counts: the query of trace_device plus per-ray (node visits, record tests,
        id reads, 0) as n x 4 uint32 at d_counts (asynchronous)."""
        self._check(lib().bc_trace_counts(self.ctx, d_rays, n, int(any_hit), d_out, d_counts))

    def line_probe(self, d_buf, nbytes, blocks, iters):
        """bc_line_probe: milliseconds of `blocks` x 256 threads each reading `iters` 64-byte blocks at
        hashed positions of the device buffer d_buf (nbytes, a power of two)."""
        ms = ctypes.c_float()
        self._check(lib().bc_line_probe(self.ctx, d_buf, nbytes, blocks, iters, ctypes.byref(ms)))
        return ms.value

    def valu_probe(self, blocks, iters, mix=0):
        """bc_valu_probe: milliseconds of `blocks` x 256 threads running `iters` slab-test visits
        (mix 0) or rounds of eight independent FMAs (mix 1)."""
        ms = ctypes.c_float()
        self._check(lib().bc_valu_probe(self.ctx, mix, blocks, iters, ctypes.byref(ms)))
        return float(ms.value)

    def render_noalloc(self, desc: RenderDesc, y0, rows, d_rgba, d_guides=None):
        """bc_render_noalloc: the plain schedule with its state in the library's fixed scratch array."""
        rays, shadow = c_uint64(), c_uint64()
        self._check(lib().bc_render_noalloc(self.ctx, ctypes.byref(desc), y0, rows, d_rgba, d_guides,
                                            ctypes.byref(rays), ctypes.byref(shadow)))
        return int(rays.value), int(shadow.value)

    def postprocess_device(self, post, width, height, d_rgba, d_guides, d_pixels, d_linear=None):
        self._check(lib().bc_postprocess_device(self.ctx, ctypes.byref(post), width, height, d_rgba, d_guides,
                                                d_pixels, d_linear))

    def render_post(self, desc, post, want_linear=False):
        """Render + device post stack: (RGBA8 H x W uint32, linear H x W x 4 or None, segments)."""
        px = np.zeros((desc.height, desc.width), np.uint32)
        lin = np.zeros((desc.height, desc.width, 4), np.float32) if want_linear else None
        rays = c_uint64()
        self._check(lib().bc_render_post(self.ctx, ctypes.byref(desc), ctypes.byref(post), _ptr(px),
                                         _ptr(lin) if lin is not None else None, ctypes.byref(rays)))
        return px, lin, int(rays.value)

    def upload_batch(self, scenes, nee_triangles=False):
        """Packs several host scenes into one device arena (bc_batch_upload)."""
        views = (SceneView * len(scenes))()
        for i, sc in enumerate(scenes):
            views[i] = sc.view(nee_triangles).contents
        self._check(lib().bc_batch_upload(self.ctx, len(scenes), views))

    def render_batch_device(self, descs, d_rgba):
        """Renders every scene of the resident batch into device buffers; descs[0].mode picks the
        schedule (0: scene after scene, plain; 1: one persistent schedule over all scenes)."""
        n = len(descs)
        da = (RenderDesc * n)(*descs)
        ptrs = (c_void_p * n)(*[c_void_p(p) if not isinstance(p, c_void_p) else p for p in d_rgba])
        rays, shadow = c_uint64(), c_uint64()
        self._check(lib().bc_batch_render_device(self.ctx, n, da, ptrs, ctypes.byref(rays), ctypes.byref(shadow)))
        return int(rays.value), int(shadow.value)

    def render_batch(self, descs):
        """Host-side convenience: list of rgba images (H x W x 4) and the path-segment count."""
        bufs = [self.alloc(d.width * d.height * 16) for d in descs]
        try:
            rays, _ = self.render_batch_device(descs, bufs)
            imgs = [self.download(b, np.empty((d.height, d.width, 4), np.float32)) for b, d in zip(bufs, descs)]
        finally:
            for b in bufs:
                self.free(b)
        return imgs, rays


def render_desc(scene: HostScene, width, height, spp, bounces, rr_depth=4, sampler=1, mis=True, power=True,
                firefly_clamp=0.0, seed=0x9E3779B9, sky_mode=0, sky_color=(0.3, 0.3, 0.3), mode=0,
                samples_per_pass=0, pool=0, iters=0, segment=0, xcd_group=0, material_sort=0, restir=0, sun_dir=(0.0, 1.0, 0.0), sun=(0.0, 0.0, 0.0),
                spatial=0, radius=0, window=0):
    """RenderDesc for a workload scene (camera from the workload, Camera.basis in fp64)."""
    cam = scene.camera
    basis = camera_basis(cam[0:3], cam[3:6], cam[6:9], cam[9], width / height)
    d = RenderDesc()
    d.width, d.height, d.spp, d.bounces, d.rr_depth = width, height, spp, bounces, rr_depth
    d.sampler, d.mis, d.power_lights, d.sky_mode = sampler, int(mis), int(power), sky_mode
    d.firefly_clamp, d.seed = firefly_clamp, seed
    for k in range(3):
        d.sky_color[k] = sky_color[k]
        d.origin[k], d.lower[k], d.horiz[k], d.vert[k] = basis[0][k], basis[1][k], basis[2][k], basis[3][k]
    d.mode, d.samples_per_pass = mode, samples_per_pass
    d.reserved[0], d.reserved[1], d.reserved[2], d.reserved[3] = pool, iters, segment, xcd_group
    d.reserved[4], d.reserved[5] = material_sort, window
    d.restir_candidates = restir
    d.restir_spatial, d.restir_radius = spatial, radius
    n = np.asarray(sun_dir, np.float64)
    n = n / np.linalg.norm(n)
    for k in range(3):
        d.sun_dir[k] = float(n[k])
        d.sun_irradiance[k] = sun[k]
    return d
