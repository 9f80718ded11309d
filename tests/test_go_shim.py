"""Static checks of the Go cgo shim (go/gpu_hip_linux.go, go/diff_hip_linux.go), which no Go
toolchain here can compile: every C function it calls is declared in include/beamcast_hip.h and
exported by the built library, every field it sets on a C struct is a member of that struct, and
both files carry the build tag that selects them instead of the OpenCL stub.  When the reference
tree is mounted, the Go identifiers the shim takes from the renderer's package exist there too."""
import os
import re

import pytest

import beamcast as bc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GO = os.path.join(ROOT, "beamcast-gx-v1.0-beta-_amd", "go")
FILES = ["gpu_hip_linux.go", "diff_hip_linux.go"]
REF_GO = "/root/reference/windows_portable"


def go_src(name):
    text = open(os.path.join(GO, name)).read()
    return re.sub(r"//[^\n]*", "", text)  # comments may mention names that are not called


def header():
    text = open(os.path.join(ROOT, "include", "beamcast_hip.h")).read()
    return re.sub(r"/\*.*?\*/", "", text, flags=re.S)


def struct_members(name):
    m = re.search(r"typedef struct " + name + r"\s*\{(.*?)\}\s*" + name + ";", header(), flags=re.S)
    assert m, name
    body = m.group(1)
    return set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*(?:\[[^\]]*\])*\s*[;,]", body))


@pytest.mark.parametrize("name", FILES)
def test_build_tag_and_package(name):
    text = open(os.path.join(GO, name)).read()
    assert text.startswith("//go:build linux && cgo && beamcast_hip\n")
    assert re.search(r"^package main$", text, flags=re.M)
    assert 'import "C"' in text


@pytest.mark.parametrize("name", FILES)
def test_c_calls_are_declared_and_exported(name):
    calls = sorted(set(re.findall(r"\bC\.(bc_[a-z0-9_]+)\(", go_src(name))))
    assert calls, name
    decl = set(re.findall(r"\b(bc_[a-z0-9_]+)\s*\(", header()))
    lib = bc.lib()
    for fn in calls:
        assert fn in decl, f"{name} calls C.{fn}, which beamcast_hip.h does not declare"
        assert hasattr(lib, fn), f"{name} calls C.{fn}, which the library does not export"


def test_struct_fields_exist():
    src = go_src("gpu_hip_linux.go")
    desc = struct_members("bc_render_desc")
    fn = src[src.index("func hipRenderDesc"):]
    fn = fn[: fn.index("\n}\n")]
    used = set(re.findall(r"\bd\.([a-z_]+)\b", fn)) | set(re.findall(r"&d\.([a-z_]+)\b", fn))
    assert {"width", "height", "spp", "bounces", "origin"} <= used
    assert used <= desc, f"bc_render_desc has no {sorted(used - desc)}"
    view = struct_members("bc_scene_view")
    lit = src[src.index("C.bc_scene_view{\n") + len("C.bc_scene_view{\n"):]
    lit = lit[: lit.index("}")]
    keys = set(re.findall(r"\b([a-z_]+):", lit))
    assert {"nnode", "nprim", "nmat", "prim_geom"} <= keys
    assert keys <= view, f"bc_scene_view has no {sorted(keys - view)}"


@pytest.mark.skipif(not os.path.isdir(REF_GO), reason="reference tree not mounted")
def test_renderer_identifiers_exist_in_reference():
    """The shim is a file of the renderer's package main: the functions, types and fields it takes
    from there must be defined by the renderer's own sources."""
    ref = ""
    for f in sorted(os.listdir(REF_GO)):
        if f.endswith(".go") and not f.endswith("_test.go"):
            ref += open(os.path.join(REF_GO, f), encoding="utf-8", errors="replace").read()
    funcs = ["internalSize", "effectiveSPP", "temporalWeight", "materialPreviewColor", "iMax", "V"]
    for fn in funcs:
        assert re.search(r"^func " + fn + r"\(", ref, flags=re.M), fn
    assert re.search(r"^func \(\w+ \*?Camera\) basis\(", ref, flags=re.M)
    for ty in ["Scene", "Camera", "RenderSettings", "Material", "Guide", "Vec3"]:
        assert re.search(r"^type " + ty + r"\b", ref, flags=re.M), ty
    for field in ["GPUTopologyRevision", "GPUGeometryRevision", "FireflyClamp", "PowerLightSampling", "RRDepth",
                  "ProgressiveUpdates", "PublishHz", "Emission", "Roughness", "IOR"]:
        assert re.search(r"^\s+" + field + r"\s", ref, flags=re.M), field
    for const in ["SamplerHalton", "SamplerR2", "PrimSphere", "DebugBeauty"]:
        assert re.search(r"\b" + const + r"\b", ref), const
    for method in ["Add", "Sub", "Mul", "Unit"]:
        assert re.search(r"^func \(\w+ Vec3\) " + method + r"\(", ref, flags=re.M), method
    guide = re.search(r"^type Guide struct \{(.*?)^\}", ref, flags=re.M | re.S).group(1)
    for field in ["P", "N", "Albedo", "Depth", "Valid"]:
        assert re.search(r"\b" + field + r"\b", guide), field
    prim = re.search(r"^type Primitive struct \{(.*?)^\}", ref, flags=re.M | re.S).group(1)
    for field in ["Kind", "Material", "Center", "Radius", "A", "E1", "E2"]:
        assert re.search(r"\b" + field + r"\b", prim), field
    node = re.search(r"^type BVHNode struct \{(.*?)^\}", ref, flags=re.M | re.S).group(1)
    for field in ["Box", "Left", "Right", "Start", "Count"]:
        assert re.search(r"\b" + field + r"\b", node), field


# --- the GPU gate (renderer.go:2649) and its widening patch ---------------------------------------

PATCH = os.path.join(GO, "renderer_gate.patch")
GATE_VARS = ["Integrator", "Sampler", "MIS", "PowerLightSampling", "FireflyClamp", "AdaptiveSampling"]


def patch_hunks():
    """(old_start, old_len, new_start, new_len, lines) of each hunk of the patch."""
    text = open(PATCH).read()
    assert "--- a/windows_portable/renderer.go\n+++ b/windows_portable/renderer.go\n" in text
    body = text[text.index("--- a/"):].split("\n")[2:]
    hunks, cur = [], None
    for line in body:
        m = re.match(r"@@ -(\d+),(\d+) \+(\d+),(\d+) @@", line)
        if m:
            cur = tuple(int(g) for g in m.groups()) + ([],)
            hunks.append(cur)
        elif cur is not None and line[:1] in (" ", "-", "+"):
            cur[4].append(line)
    return hunks


def go_gate(expr, full, s):
    """Evaluates a Go gpuCompatible expression for one settings assignment (a translation of its
    &&, ||, ! and comparisons to Python)."""
    py = expr.replace("&&", " and ").replace("||", " or ")
    py = re.sub(r"!(?!=)", " not ", py)
    for c in ("IntegratorPath", "IntegratorRecursive", "SamplerRandom", "SamplerHalton", "SamplerR2"):
        py = re.sub(r"(?<![.\w])" + c + r"\b", repr(c), py)
    for v in GATE_VARS:
        py = py.replace("settings." + v, repr(s[v]))
    py = py.replace("gpuFullTransport", repr(full))
    return eval(py, {"__builtins__": {}})  # noqa: S307 - a fixed expression of this patch


def gate_exprs():
    (hunk,) = patch_hunks()
    old = [ln[1:] for ln in hunk[4] if ln[:1] == "-"]
    new = [ln[1:] for ln in hunk[4] if ln[:1] == "+"]
    assert len(old) == 1
    strip = lambda lines: " ".join(ln.strip() for ln in lines).split(":=", 1)[1].strip()  # noqa: E731
    assert old[0].strip().startswith("gpuCompatible :=") and new[0].strip().startswith("gpuCompatible :=")
    return strip(old), strip(new)


def test_gate_patch_hunk_is_consistent():
    (hunk,) = patch_hunks()
    o0, ol, n0, nl, lines = hunk
    assert o0 == n0 == 2646
    assert ol == sum(1 for ln in lines if ln[:1] in (" ", "-"))
    assert nl == sum(1 for ln in lines if ln[:1] in (" ", "+"))


def test_gate_truth_table():
    """With gpuFullTransport false the widened gate is the reference gate for every combination of
    the settings it reads; with it true only the integrator and adaptive sampling keep a render on
    the CPU."""
    old, new = gate_exprs()
    n = 0
    import itertools
    for integ, samp, mis, power, clamp, adapt in itertools.product(
            ("IntegratorPath", "IntegratorRecursive"), ("SamplerRandom", "SamplerHalton", "SamplerR2"),
            (False, True), (False, True), (0.0, -1.0, 4.0), (False, True)):
        s = dict(Integrator=integ, Sampler=samp, MIS=mis, PowerLightSampling=power, FireflyClamp=clamp,
                 AdaptiveSampling=adapt)
        ref = go_gate(old, None, s)
        assert go_gate(new, False, s) == ref, s
        assert go_gate(new, True, s) == (integ == "IntegratorPath" and not adapt), s
        n += ref
    assert 0 < n < 2 * 3 * 2 * 2 * 3 * 2


def test_full_transport_flag_per_backend():
    """gpuFullTransport is defined once per build: true by the HIP shim, false under exactly the
    complementary build constraint (the OpenCL kernel and the stub), and the shim maps every setting
    the widened gate lets through onto the render description."""
    shim = open(os.path.join(GO, "gpu_hip_linux.go")).read()
    other = open(os.path.join(GO, "gpu_transport_opencl.go")).read()
    assert other.startswith("//go:build !(linux && cgo && beamcast_hip)\n")
    assert re.search(r"^const gpuFullTransport = true$", shim, flags=re.M)
    assert re.search(r"^const gpuFullTransport = false$", other, flags=re.M)
    src = go_src("gpu_hip_linux.go")
    fn = src[src.index("func hipRenderDesc"):]
    fn = fn[: fn.index("\n}\n")]
    for needle in ("case SamplerHalton", "case SamplerR2", "settings.MIS", "settings.PowerLightSampling",
                   "settings.FireflyClamp"):
        assert needle in fn, needle


@pytest.mark.skipif(not os.path.isdir(REF_GO), reason="reference tree not mounted")
def test_gate_patch_applies_to_reference(tmp_path):
    """git apply --check against a copy of the reference renderer, then the applied file differs from
    the original in the gate line only."""
    import shutil
    import subprocess
    dst = tmp_path / "windows_portable"
    dst.mkdir()
    shutil.copy(os.path.join(REF_GO, "renderer.go"), dst / "renderer.go")
    subprocess.run(["git", "apply", "--check", PATCH], cwd=tmp_path, check=True, capture_output=True)
    subprocess.run(["git", "apply", PATCH], cwd=tmp_path, check=True, capture_output=True)
    before = open(os.path.join(REF_GO, "renderer.go")).read().split("\n")
    after = (dst / "renderer.go").read_text().split("\n")
    assert len(after) == len(before) + 1
    assert before[:2648] == after[:2648] and before[2649:] == after[2650:]
    old, new = gate_exprs()
    assert before[2648].strip() == "gpuCompatible := " + old
