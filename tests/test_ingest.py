"""Scene ingestion on the host: the parallel OBJ loader (bch_load_obj) against the oracle's
restatement of LoadOBJ (renderer.go:2996-3069), and the BGXPAGE1 page store against
PagedTriangleStore (gx_stream.cpp:16-55).

Bar: identical primitives (A, B, C, padded box, centroid; bit-exact doubles), camera and error
text.  The oracle itself is pinned by the reference's TestLoadOBJ known answer
(renderer_test.go:51-63: one triangle), the pyramid asset (source/examples/assets/pyramid.obj,
kept as tests/golden/pyramid.obj: one quad fanned into two triangles plus four triangles, counted
by hand), and strconv.ParseFloat / Atoi cases whose Go results are fixed by the language spec.
"""
import math
import os
import struct

import numpy as np
import pytest

import beamcast as bc
import oracle_obj as oo

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def native(path, threads=0):
    s, cam = bc.HostScene.load_obj(path, threads=threads)
    kind, mat, geom, _, _ = s.export()
    box, cen = s.bounds()
    return kind, mat, geom, box, cen, cam


def assert_same(path, threads=0):
    data = open(path, "rb").read()
    try:
        prims, cam = oo.load_obj(data)
    except ValueError as e:
        with pytest.raises(ValueError) as got:
            native(path, threads)
        assert str(got.value) == str(e)
        return None
    kind, mat, geom, box, cen, ncam = native(path, threads)
    assert len(kind) == len(prims) and (kind == bc.TRIANGLE).all()
    ref = np.array([[*a, *b, *c] for a, b, c, *_ in prims])
    assert np.array_equal(geom, ref, equal_nan=True)
    assert np.array_equal(box, np.array([[*lo, *hi] for _, _, _, lo, hi, _, _ in prims]), equal_nan=True)
    assert np.array_equal(cen, np.array([c for *_, c, _ in prims]), equal_nan=True)
    assert np.array_equal(ncam, cam, equal_nan=True)
    return len(prims)


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text.encode() if isinstance(text, str) else text)
    return str(p)


def test_oracle_known_answers(tmp_path):
    # renderer_test.go:51-63
    prims, _ = oo.load_obj(b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    assert len(prims) == 1
    # the pyramid: quad 1 4 3 2 -> (1,4,3), (1,3,2); then four side triangles
    prims, cam = oo.load_obj(open(os.path.join(GOLDEN, "pyramid.obj"), "rb").read())
    v = [(-0.75, 0.0, -0.75), (0.75, 0.0, -0.75), (0.75, 0.0, 0.75), (-0.75, 0.0, 0.75), (0.0, 1.2, 0.0)]
    want = [(0, 3, 2), (0, 2, 1), (0, 1, 4), (1, 2, 4), (2, 3, 4), (3, 0, 4)]
    assert [(p[0], p[1], p[2]) for p in prims] == [(v[a], v[b], v[c]) for a, b, c in want]
    assert cam[3:6].tolist() == [0.0, 0.6, 0.0] and cam[9] == 42.0
    # ParseFloat / Atoi grammar (Go spec): accepted and rejected spellings
    ok = {"1": 1.0, "-2.5": -2.5, ".5": 0.5, "5.": 5.0, "1e3": 1000.0, "1.e2": 100.0, "+7": 7.0, "0x1p-2": 0.25,
          "0X1.8P1": 3.0, "Inf": math.inf, "-infinity": -math.inf, "1e-400": 0.0}
    for t, want_v in ok.items():
        assert oo.parse_float(t) == want_v, t
    assert math.isnan(oo.parse_float("NaN"))
    for t in ("", "1e", "e5", ".", "1_0", "0x1.8", "+nan", "infi", "1e400", "0x1p5000", "--1", "1..2", "nan(1)"):
        assert oo.parse_float(t) is None, t
    assert oo.parse_int("+12") == 12 and oo.parse_int("-3") == -3 and oo.parse_int("9223372036854775807") is not None
    for t in ("", "+", "1.0", "0x10", "9223372036854775808", " 1"):
        assert oo.parse_int(t) is None, t


def test_reference_cases_native(tmp_path):
    assert assert_same(write(tmp_path, "a.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")) == 1
    assert assert_same(os.path.join(GOLDEN, "pyramid.obj")) == 6


@pytest.mark.parametrize("text", [
    "v 0 0\nf 1 2 3\n",                                     # malformed vertex
    "v 0 0 0\nv 1 x 0\n",                                   # invalid vertex
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2\n",                   # face needs 3 vertices
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 0 3\n",                 # invalid index (zero)
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 a\n",                 # invalid index (text)
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n",                 # index out of range
    "v 0 0 0\nv 1 0 0\nf -3 -2 -1\nv 0 1 0\n",              # relative index before the vertex exists
    "v 0 0 0\nv 1 0 0\nv 2 0 0\nf 1 2 3\n",                 # only a degenerate face
    "# nothing\n\n   \n",                                    # no faces
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\nv 1e400 0 0\n",    # overflowing vertex after a good face
    "v 0 0 0\r\nv 1 0 0\r\nv 0 1 0\r\nf 1/1 2/2/2 3//3\r\n",  # CRLF, slash forms
    "\tv  0\t0 0 \nv 1 0 0 1\nv 0 1 0\n f -3 -2 -1 \nvt 0 0\nvn 0 0 1\no x\ng y\nusemtl m\n",
    "v 0 0 0\nv 1 0 0\nv 0 1\u30000\u00a0\nf\u20031 2 3\n",  # Unicode spaces (unicode.IsSpace)
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nf 1 2 4 3\nf 1 -1 -2\n",  # quad fan, relative indices
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3",                   # no final newline
    "v 0 0 0\nv 1 0 0\nv 0 1 0\n# " + "x" * 70000 + "\nf 1 2 3\n",  # line longer than the scanner buffer
    "v nan 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n",               # NaN vertex: camera goes NaN as math.Min does
])
def test_cases_native_vs_oracle(tmp_path, text):
    assert_same(write(tmp_path, "c.obj", text))


def test_random_files(tmp_path):
    rng = np.random.default_rng(11)
    for k in range(12):
        lines, nv = [], 0
        for _ in range(rng.integers(30, 200)):
            r = rng.random()
            if r < 0.5 or nv < 3:
                x = rng.normal(size=3) * 10.0 ** int(rng.integers(-3, 4))
                fmt = rng.choice(["%.17g", "%.6f", "%g", "%.3e"])
                lines.append("v " + " ".join(fmt % c for c in x))
                nv += 1
            elif r < 0.85:
                n = int(rng.integers(3, 7))
                idx = rng.integers(1, nv + 1, size=n)
                toks = [str(i) if rng.random() < 0.6 else str(i - nv - 1) for i in idx]
                toks = [t + rng.choice(["", "/1", "//2", "/3/4"]) for t in toks]
                lines.append("f " + " ".join(toks))
            elif r < 0.9:
                lines.append("# comment " + str(rng.random()))
            else:
                lines.append(rng.choice(["", "vt 0.5 0.5", "vn 0 1 0", "s off", "g group"]))
        if k % 4 == 3:  # plant an error
            lines.insert(int(rng.integers(1, len(lines))), rng.choice(["v 1 2", "f 1 2", "f 1 2 999999", "v 1 2 z"]))
        assert_same(write(tmp_path, f"r{k}.obj", "\n".join(lines) + "\n"))


def test_parallel_equals_serial(tmp_path):
    """A multi-megabyte file parsed on 1 and on 8 threads: the same primitives; a late planted error
    is reported at its line."""
    s = bc.HostScene.workload(bc.SOUP, param=60_000)
    p = str(tmp_path / "soup.obj")
    s.write_obj(p)
    one = native(p, threads=1)
    many = native(p, threads=8)
    for a, b in zip(one, many):
        assert np.array_equal(a, b)
    kind, mat, geom, _, _ = s.export()
    assert np.array_equal(one[2], geom)
    lines = open(p).read().split("\n")
    lines[200_001] = "f 1 2 0"
    q = str(tmp_path / "bad.obj")
    open(q, "w").write("\n".join(lines))
    with pytest.raises(ValueError, match="OBJ line 200002: invalid index"):
        native(q, threads=8)


def packed(n, seed):
    rng = np.random.default_rng(seed)
    t = np.zeros(n, bc.PACKED_TRIANGLE)
    t["v0"][:, :3] = rng.random((n, 3))
    t["e1"][:, :3] = rng.normal(size=(n, 3)) * 0.01
    t["e2"][:, :3] = rng.normal(size=(n, 3)) * 0.01
    t["normal"][:, :3] = rng.normal(size=(n, 3))
    t["material"] = rng.integers(0, 3, n)
    return t


def test_page_store(tmp_path):
    t = packed(1000, 5)
    p = str(tmp_path / "tri.pages")
    bc.pages_write(p, t, per_page=300)
    data = open(p, "rb").read()
    assert data[:8] == b"BGXPAGE1"
    # same bytes as the oracle's restatement of PagedTriangleStore::write
    tris = [(r["v0"][:3], r["e1"][:3], r["e2"][:3], r["normal"][:3], int(r["material"])) for r in t]
    assert data == oo.pages_write(tris, 300)
    pg = bc.Pages(p)
    assert pg.count() == 4 and [pg.info(k)[1] for k in range(4)] == [300, 300, 300, 100]
    assert pg.info(0)[0] == 16 + 4 * 16 and pg.info(1)[0] == pg.info(0)[0] + 300 * 80
    got = np.concatenate([pg.load(k) for k in range(4)])
    assert np.array_equal(got.view(np.uint8), t.view(np.uint8))
    # append_page_to_scene through AddTriangle, against the oracle
    s = bc.HostScene()
    for _ in range(3):
        s.add_material(0, (0.5, 0.5, 0.5))
    s.append_pages(pg)
    kind, mat, geom, _, _ = s.export()
    box, cen = s.bounds()
    prims = oo.pages_append(data, [])
    assert len(prims) == len(kind) and np.array_equal(mat, [m for *_, m in prims])
    assert np.array_equal(geom, np.array([[*a, *b, *c] for a, b, c, *_ in prims]))
    assert np.array_equal(box, np.array([[*lo, *hi] for _, _, _, lo, hi, _, _ in prims]))
    pg.close()


def test_page_store_errors(tmp_path):
    bad = tmp_path / "bad.pages"
    bad.write_bytes(b"NOTAPAGE" + struct.pack("<II", 1, 0))
    with pytest.raises(ValueError, match="invalid Beamcast GX page store"):
        bc.Pages(str(bad))
    trunc = tmp_path / "trunc.pages"
    trunc.write_bytes(struct.pack("<QII", oo.PAGE_MAGIC, 1, 5) + b"\0" * 20)
    with pytest.raises(ValueError, match="truncated Beamcast GX page table"):
        bc.Pages(str(trunc))
    with pytest.raises(ValueError, match="cannot open page store"):
        bc.Pages(str(tmp_path / "missing.pages"))
    # a table pointing past the end of the file
    short = tmp_path / "short.pages"
    short.write_bytes(struct.pack("<QII", oo.PAGE_MAGIC, 1, 1) + struct.pack("<QII", 32, 10, 0) + b"\0" * 80)
    pg = bc.Pages(str(short))
    with pytest.raises(ValueError, match="truncated"):
        pg.load(0)
    with pytest.raises(IndexError):
        pg.info(3)
