"""Pure-Python restatement of the renderer's scene ingestion, for small inputs.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the native loaders, never by the
product.  Restates (reference = windows_portable/ and source/):
  - LoadOBJ                                   renderer.go:2996-3069
      bufio.Scanner lines, strings.TrimSpace / strings.Fields (unicode.IsSpace), "#" comments,
      "v x y z" with strconv.ParseFloat, "f a b c ..." with strconv.Atoi of the text before the
      first "/", negative indices relative to the vertices read so far, fan triangulation,
      "OBJ line N: ..." errors at the first offending line, the camera framing of the bounds
  - Scene.AddTriangle                         renderer.go:257-269 (degenerate drop, padded box)
  - PagedTriangleStore::write / load_page /   gx_stream.cpp:16-55 (BGXPAGE1 layout)
    append_page_to_scene
Deliberate narrowing shared with the native loader: hexadecimal floats with "_" digit
separators are rejected (Go accepts them after a 0x prefix).
"""
from __future__ import annotations

import math
import re
import struct

import numpy as np

# unicode.IsSpace
_SPACE = "\t\n\v\f\r \u0085\u00a0\u1680\u2000\u2001\u2002\u2003\u2004\u2005\u2006\u2007\u2008\u2009\u200a\u2028\u2029\u202f\u205f\u3000"
_FIELDS = re.compile("[^" + re.escape(_SPACE) + "]+")
_DEC = re.compile(r"[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?\Z")
_HEX = re.compile(r"[+-]?0[xX]([0-9a-fA-F]+\.?[0-9a-fA-F]*|\.[0-9a-fA-F]+)[pP][+-]?\d+\Z")
_INF = re.compile(r"[+-]?(inf|infinity)\Z", re.IGNORECASE)
_NAN = re.compile(r"nan\Z", re.IGNORECASE)
_INT = re.compile(r"[+-]?\d+\Z")


def parse_float(tok):
    """strconv.ParseFloat(tok, 64) -> float or None (syntax error or overflow)."""
    if _INF.match(tok):
        return -math.inf if tok[0] == "-" else math.inf
    if _NAN.match(tok):
        return math.nan
    try:
        if _DEC.match(tok):
            v = float(tok)
        elif _HEX.match(tok):
            v = float.fromhex(tok)
        else:
            return None
    except OverflowError:
        return None
    return None if math.isinf(v) else v


def parse_int(tok):
    """strconv.Atoi (64-bit int) -> int or None."""
    if not _INT.match(tok):
        return None
    v = int(tok)
    return v if -(1 << 63) <= v < (1 << 63) else None


def go_min(a, b):
    if math.isnan(a) or math.isnan(b):
        return math.nan
    if a == 0 and b == 0:
        return a if math.copysign(1, a) < 0 else b
    return a if a < b else b


def go_max(a, b):
    if math.isnan(a) or math.isnan(b):
        return math.nan
    if a == 0 and b == 0:
        return b if math.copysign(1, a) < 0 else a
    return a if a > b else b


def add_triangle(prims, a, b, c, mat):
    """Scene.AddTriangle: appends (A, B, C, box_min, box_max, centroid, mat) unless degenerate."""
    e1 = (b[0] - a[0], b[1] - a[1], b[2] - a[2])
    e2 = (c[0] - a[0], c[1] - a[1], c[2] - a[2])
    n = (e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0])
    if n[0] * n[0] + n[1] * n[1] + n[2] * n[2] < 1e-18:
        return
    lo = tuple(go_min(go_min(go_min(math.inf, a[k]), b[k]), c[k]) - 1e-6 for k in range(3))
    hi = tuple(go_max(go_max(go_max(-math.inf, a[k]), b[k]), c[k]) + 1e-6 for k in range(3))
    cen = tuple((a[k] + b[k] + c[k]) * (1.0 / 3.0) for k in range(3))
    prims.append((a, b, c, lo, hi, cen, mat))


def load_obj(data: bytes):
    """LoadOBJ over the file bytes -> (prims, camera[10]) or raises ValueError(message)."""
    text = data.decode("utf-8", "surrogateescape")
    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    prims, verts = [], []
    mn, mx = [math.inf] * 3, [-math.inf] * 3
    for no, line in enumerate(lines, 1):
        if len(line.encode("utf-8", "surrogateescape")) + 1 > 65536:
            raise ValueError("bufio.Scanner: token too long")
        parts = _FIELDS.findall(line)
        if not parts or parts[0].startswith("#"):
            continue
        if parts[0] == "v":
            if len(parts) < 4:
                raise ValueError(f"OBJ line {no}: malformed vertex")
            xyz = [parse_float(t) for t in parts[1:4]]
            if any(v is None for v in xyz):
                raise ValueError(f"OBJ line {no}: invalid vertex")
            verts.append(tuple(xyz))
            mn = [go_min(mn[k], xyz[k]) for k in range(3)]
            mx = [go_max(mx[k], xyz[k]) for k in range(3)]
        elif parts[0] == "f":
            if len(parts) < 4:
                raise ValueError(f"OBJ line {no}: face needs 3 vertices")
            idx = []
            for tok in parts[1:]:
                n = parse_int(tok.split("/")[0])
                if n is None or n == 0:
                    raise ValueError(f"OBJ line {no}: invalid index")
                if n < 0:
                    n = len(verts) + n + 1
                n -= 1
                if n < 0 or n >= len(verts):
                    raise ValueError(f"OBJ line {no}: index out of range")
                idx.append(n)
            for i in range(1, len(idx) - 1):
                add_triangle(prims, verts[idx[0]], verts[idx[i]], verts[idx[i + 1]], 0)
    if not prims:
        raise ValueError("OBJ contains no renderable faces")
    center = [(mn[k] + mx[k]) * 0.5 for k in range(3)]
    d = [mx[k] - mn[k] for k in range(3)]
    rad = max(0.5, math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * 0.5)
    pos = [center[0] + rad * 1.4, center[1] + rad * 0.9, center[2] + rad * 2.6]
    return prims, np.array(pos + center + [0.0, 1.0, 0.0, 42.0])


# --- BGXPAGE1 page store

PAGE_MAGIC = 0x3145474150584742


def pages_write(tris, per_page):
    """PagedTriangleStore::write -> file bytes.  tris: list of (v0, e1, e2, normal, material),
    the vectors as 3 floats."""
    pages = (len(tris) + per_page - 1) // per_page
    out = struct.pack("<QII", PAGE_MAGIC, 1, pages)
    off = 16 + 16 * pages
    for p in range(pages):
        cnt = min(per_page, len(tris) - p * per_page)
        out += struct.pack("<QII", off, cnt, 0)
        off += 80 * cnt
    for v0, e1, e2, nrm, mat in tris:
        for v in (v0, e1, e2, nrm):
            out += struct.pack("<4f", v[0], v[1], v[2], 0.0)
        out += struct.pack("<I12x", mat)
    return out


def pages_append(data: bytes, prims):
    """Every page's append_page_to_scene: AddTriangle(v0, v0 + e1, v0 + e2) with float sums."""
    magic, version, pages = struct.unpack_from("<QII", data, 0)
    if magic != PAGE_MAGIC or version != 1:
        raise ValueError("invalid Beamcast GX page store")
    for p in range(pages):
        off, cnt, _ = struct.unpack_from("<QII", data, 16 + 16 * p)
        for t in range(cnt):
            rec = np.frombuffer(data, np.float32, 16, off + 80 * t).reshape(4, 4)
            mat = struct.unpack_from("<I", data, off + 80 * t + 64)[0]
            v0 = rec[0, :3]
            b = (v0 + rec[1, :3]).astype(np.float32)  # float sums
            c = (v0 + rec[2, :3]).astype(np.float32)
            add_triangle(prims, tuple(float(x) for x in v0), tuple(float(x) for x in b), tuple(float(x) for x in c), mat)
    return prims
