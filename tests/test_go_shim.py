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
