"""The C-ABI library loads and exports every symbol include/dpe_hip.h declares;
struct layouts match; host-side logic (result assembly, scene, support maps).
No GPU compute here."""
import ctypes as C
import os
import re

import numpy as np

from dpe_amd import abi, host, scene, support

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = host.load_library()
    header = open(os.path.join(ROOT, "include", "dpe_hip.h")).read()
    declared = set(re.findall(r"\b(dpe_[a-z_]+)\s*\(", header))
    assert declared == set(abi.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.dpe_abi_version() == abi.ABI_VERSION


def test_struct_layouts():
    assert C.sizeof(abi.Camera) == 112  # Camera, main.h:53-62
    assert abi.Camera.K.offset == 0 and abi.Camera.c.offset == 84 and abi.Camera.height.offset == 96
    assert C.sizeof(abi.Params) == 112
    assert abi.Params.seed.offset == 104
    assert C.sizeof(abi.Point) == 16


STRUCTS = {"dpe_camera": abi.Camera, "dpe_params": abi.Params, "dpe_inputs": abi.Inputs, "dpe_outputs": abi.Outputs,
           "dpe_stats": abi.Stats, "dpe_fuse_view": abi.FuseView, "dpe_fuse_params": abi.FuseParams,
           "dpe_point": abi.Point, "dpe_form": abi.Form, "dpe_fuse_tat_params": abi.FuseTatParams}


def test_ctypes_mirror_matches_the_c_header(tmp_path):
    """Every field offset and struct size of the ctypes mirror equals what a C
    compiler lays out from include/dpe_hip.h, and the #define constants agree."""
    import subprocess
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "dpe_hip.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    for const in ("DPE_ABI_VERSION", "DPE_MAX_IMAGES", "DPE_NCC_FAMILIES", "DPE_FAMILY_WEAK", "DPE_TAT_ADVANCED"):
        lines.append(f'printf("const {const} %d\\n", (int)({const}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        a, b, v = line.split()
        got[(a, b)] = int(v)
    for cname, cls in STRUCTS.items():
        assert got[(cname, "size")] == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
    assert got[("const", "DPE_ABI_VERSION")] == abi.ABI_VERSION
    assert got[("const", "DPE_MAX_IMAGES")] == abi.MAX_IMAGES
    assert got[("const", "DPE_NCC_FAMILIES")] == abi.NCC_FAMILIES
    assert got[("const", "DPE_FAMILY_WEAK")] == abi.FAMILY_WEAK
    assert got[("const", "DPE_TAT_ADVANCED")] == abi.TAT_ADVANCED


def test_result_assembly_zeroes_out_of_range_depth():
    out = host.ProblemOutputs(3, 2)
    out.planes[..., 3] = np.array([[1.0, 2.0, 3.0], [0.5, 7.5, 2.5]])
    out.planes[..., 2] = -1.0
    out.pixel_states[:] = abi.STRONG
    r = host.assemble(out, 1.2, 7.2)
    assert r["depth"].tolist() == [[0.0, 2.0, 3.0], [0.0, 0.0, 2.5]]
    assert r["pixel_states"].tolist() == [[abi.UNKNOWN, 1, 1], [abi.UNKNOWN, abi.UNKNOWN, 1]]
    assert np.all(r["normal"][..., 2] == -1.0)  # normals are kept (main.cpp:271-276)


def test_scene_is_seeded_and_in_range():
    a = scene.make_scene(64, 48, 60.0, 3, 20.0, seed=3, supersample=1)
    b = scene.make_scene(64, 48, 60.0, 3, 20.0, seed=3, supersample=1)
    for va, vb in zip(a.views, b.views):
        assert np.array_equal(va.image, vb.image)
        assert np.array_equal(va.depth, vb.depth)
    d = a.views[1].depth
    assert d.min() >= 2.0 and d.max() <= 6.0
    w = a.views[1].weak.mean()
    assert 0.1 < w < 0.3
    # normals face the camera
    v = a.views[1]
    n_cam = v.normal @ v.R.T
    assert np.all(n_cam[..., 2] < 0)


def test_camera_center_in_double():
    K = np.diag([100.0, 100.0, 1.0])
    R = np.eye(3)
    t = np.array([0.1, -0.2, 0.3])
    cam = abi.camera_from_arrays(K, R, t, 10, 10, 1, 2)
    c = np.array(list(cam.c))
    assert np.allclose(np.eye(3) @ c + t, 0, atol=1e-7)  # R c + t = 0


def test_support_map_semantics():
    s = scene.make_scene(80, 64, 75.0, 3, 20.0, seed=4, supersample=1)
    e, lab = support.support_maps(s.views[1].image)
    assert e.dtype == np.uint8 and set(np.unique(e)) <= {0, 255}
    assert lab.dtype == np.int32
    assert set(np.unique(lab)) >= {0}
    assert lab.min() >= -1
