"""The C++ host driver's library (libdpe_host.so) and schedule, on the CPU:
file formats against their reference layouts, the pyramid and nearest rescale
against closed forms / brute force, the support maps against dpe_amd/support.py
bit for bit, and the driver's reference schedule (dry run) against main.cpp's."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from dpe_amd import abi, host, hostio, scene, support

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    return hostio.load()


def test_host_library_exports_every_declared_symbol(lib):
    header = open(os.path.join(ROOT, "include", "dpe_host.h")).read()
    declared = set(re.findall(r"\b(dpe_host_[a-z_]+)\s*\(", header))
    assert declared == set(hostio.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name


@pytest.mark.parametrize("arr", [np.arange(12, dtype=np.float32).reshape(3, 4),
                                 np.arange(6, dtype=np.uint8).reshape(2, 3),
                                 np.arange(-3, 3, dtype=np.int32).reshape(2, 3),
                                 np.arange(18, dtype=np.float32).reshape(2, 3, 3)])
def test_bin_mat_round_trip(lib, tmp_path, arr):
    """WriteBinMat / ReadBinMat layout (DPE.cpp:293-339): int32 version, rows, cols, type, data."""
    p = str(tmp_path / "m.dmb")
    hostio.write_bin_mat(p, arr)
    rows, cols, typ = C.c_int(), C.c_int(), C.c_int()
    out = np.zeros_like(arr)
    assert lib.dpe_host_read_bin_mat(p.encode(), C.byref(rows), C.byref(cols), C.byref(typ), out.ctypes.data,
                                     out.nbytes) == 0
    assert (rows.value, cols.value) == arr.shape[:2]
    assert np.array_equal(out, arr)
    q = str(tmp_path / "n.dmb")
    assert lib.dpe_host_write_bin_mat(q.encode(), arr.shape[0], arr.shape[1], typ.value, arr.ctypes.data) == 0
    assert open(p, "rb").read() == open(q, "rb").read()
    assert np.array_equal(hostio.read_bin_mat(q), arr)
    # too small a buffer and a wrong version are errors, not overruns
    assert lib.dpe_host_read_bin_mat(p.encode(), None, None, None, out.ctypes.data, out.nbytes - 1) == -3
    bad = str(tmp_path / "bad.dmb")
    with open(bad, "wb") as f:
        f.write(np.array([2, 1, 1, 0], np.int32).tobytes() + b"\0")
    assert lib.dpe_host_read_bin_mat(bad.encode(), None, None, None, None, 0) == -2


def test_camera_and_pairs(lib, tmp_path):
    s = scene.make_scene(64, 48, 60.0, 4, 20.0, seed=5, supersample=1)
    scene.write_folder(s, str(tmp_path))
    for i in range(4):
        cam = abi.Camera()
        assert lib.dpe_host_read_camera(str(tmp_path / "cams" / ("%08d_cam.txt" % i)).encode(), C.byref(cam)) == 0
        ref = abi.camera_from_arrays(s.views[i].K, s.views[i].R, s.views[i].t, 64, 48, 2.0, 6.0)
        assert list(cam.K) == list(ref.K) and list(cam.R) == list(ref.R) and list(cam.t) == list(ref.t)
        assert list(cam.c) == list(ref.c)  # c = -R^T t in double
        assert (cam.depth_min, cam.depth_max) == (2.0, 6.0)
    # pair.txt: non-positive scores are dropped (main.cpp:126-131)
    with open(tmp_path / "pair2.txt", "w") as f:
        f.write("2\n7\n3 1 0.5 2 0.0 3 -1\n8\n1 4 2.5\n")
    refs = np.zeros(4, np.int32)
    counts = np.zeros(4, np.int32)
    srcs = np.zeros((4, abi.MAX_IMAGES), np.int32)
    n = lib.dpe_host_read_pairs(str(tmp_path / "pair2.txt").encode(), 4, refs.ctypes.data, counts.ctypes.data,
                                srcs.ctypes.data)
    assert n == 2 and list(refs[:2]) == [7, 8] and list(counts[:2]) == [1, 1]
    assert srcs[0, 0] == 1 and srcs[1, 0] == 4
    # more than 10 sources (ETH3D-style pair files): all positive ones are kept in order
    ids = list(range(10, 30))
    with open(tmp_path / "pair3.txt", "w") as f:
        body = " ".join("%d %.2f" % (i, 30 - i) for i in ids) + " 40 0 41 -2"
        f.write("1" + chr(10) + "3" + chr(10) + "%d %s" % (len(ids) + 2, body) + chr(10))
    n = lib.dpe_host_read_pairs(str(tmp_path / "pair3.txt").encode(), 4, refs.ctypes.data, counts.ctypes.data,
                                srcs.ctypes.data)
    assert n == 1 and counts[0] == 20 and list(srcs[0, :20]) == ids


def test_resize_factor_two_is_box_mean_and_nearest_rescale(lib):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (10, 14)).astype(np.float32)
    out = np.zeros((5, 7), np.float32)
    lib.dpe_host_resize_linear(img.ctypes.data, 14, 10, out.ctypes.data, 7, 5)
    box = img.reshape(5, 2, 7, 2).sum(axis=(1, 3)) * np.float32(0.25)
    assert np.array_equal(out, box)
    # general size: half-pixel centres, clamped borders (INTER_LINEAR), float64 restatement
    out2 = np.zeros((7, 9), np.float32)
    lib.dpe_host_resize_linear(img.ctypes.data, 14, 10, out2.ctypes.data, 9, 7)

    def taps(n, dn):
        f = (np.arange(dn) + 0.5) * (n / dn) - 0.5
        i = np.floor(f).astype(int)
        a = f - i
        a[i < 0] = 0
        i[i < 0] = 0
        a[i >= n - 1] = 0
        i[i >= n - 1] = n - 1
        return i, np.minimum(i + 1, n - 1), a
    xi, xj, xa = taps(14, 9)
    yi, yj, ya = taps(10, 7)
    rows = img[:, xi] * (1 - xa) + img[:, xj] * xa
    ref = rows[yi] * (1 - ya)[:, None] + rows[yj] * ya[:, None]
    assert np.allclose(out2, ref, atol=1e-4)
    # RescaleMatToTargetSize (DPE.cpp:1313-1335): o = int(r / s) with s = target / source in float
    src = np.arange(6 * 5, dtype=np.int32).reshape(6, 5)
    dst = np.zeros((13, 11), np.int32)
    lib.dpe_host_rescale_nearest(src.ctypes.data, 5, 6, dst.ctypes.data, 11, 13, 4)
    sy, sx = np.float32(13 / np.float32(6)), np.float32(11 / np.float32(5))
    for r in range(13):
        for c in range(11):
            orow, ocol = int(np.float32(r) / sy), int(np.float32(c) / sx)
            assert dst[r, c] == src[orow, ocol]


@pytest.mark.parametrize("shape,seed", [((48, 64), 1), ((240, 320), 2), ((37, 53), 3)])
def test_support_maps_match_python(lib, shape, seed):
    s = scene.make_scene(shape[1], shape[0], shape[1] * 0.9, 3, 20.0, seed=seed, supersample=1)
    img = s.views[1].image
    e_c, l_c = hostio.support_maps(lib, img, 0)
    e_p, l_p = support.support_maps(img, 0)
    assert np.array_equal(e_c, e_p)
    assert np.array_equal(l_c, l_p)
    if shape[0] >= 200:  # large enough to hold regions as well as edges
        assert (l_c == 0).any() and (l_c > 0).any() and (e_c == 255).any() and (e_c == 0).any()


def test_ply_layout(lib, tmp_path):
    pts = (abi.Point * 3)()
    for i in range(3):
        pts[i].x, pts[i].y, pts[i].z = i, i + 0.5, -i
        pts[i].b, pts[i].g, pts[i].r = 10 * i, 20 * i, 30 * i
    p = str(tmp_path / "x.ply")
    assert lib.dpe_host_write_ply(p.encode(), pts, 3) == 0
    header, v = hostio.read_ply(p)
    assert header.startswith("ply\nformat binary_little_endian 1.0\nelement vertex 3\n")
    assert "property uchar diffuse_blue" in header
    assert list(v["y"]) == [0.5, 1.5, 2.5] and list(v["r"]) == [0, 30, 60]


def test_driver_dry_run_reference_schedule(tmp_path):
    """ComputeRoundNum and the schedule of main.cpp:390-457 for 1700x900 (3 rounds), 1600x900
    (2 rounds: 800 is not halved) and 320x240 (1 round)."""
    for (W, H), rounds in (((1700, 900), 3), ((1600, 900), 2), ((320, 240), 1)):
        folder = tmp_path / f"s{W}"
        s = scene.make_scene(64, 48, 60.0, 3, 20.0, seed=9, supersample=1)
        scene.write_folder(s, str(folder))
        for i in range(3):  # replace the images by blank ones of the wanted size
            with open(folder / "images" / ("%08d.pgm" % i), "wb") as f:
                f.write(b"P5\n%d %d\n255\n" % (W, H) + bytes(W * H))
        out = subprocess.run([hostio.DRIVER_PATH, str(folder), "--dry-run"], capture_output=True, text=True,
                             check=True).stdout
        calls = [ln for ln in out.splitlines() if ln.startswith("call ")]
        expect = host.reference_schedule(rounds)
        assert len(calls) == len(expect) == 4 * rounds
        for ln, c in zip(calls, expect):
            kv = dict(t.split("=") for t in ln.split()[2:])
            assert int(kv["scale"]) == c["scale"]
            assert kv["state"] == {0: "FIRST_INIT", 1: "REFINE_INIT", 2: "REFINE_ITER"}[c["state"]]
            assert (int(kv["apd"]), int(kv["edge"]), int(kv["geom"]), int(kv["peak"])) == \
                (c["apd"], c["edge"], c["geom"], c["peak"])
            assert np.float32(float(kv["ransac"])) == np.float32(c["ransac"])
            assert int(kv["rotate"]) == c["rotate"] and int(kv["final"]) == int(c["final"])
        assert f"max_scale_size={1 << (rounds - 1)}" in out
        assert "problem 0 ref 0 srcs 1 2" in out
        assert "fusion eth min_consistent=" in out
    for name in ("tat_intermediate", "tat_advanced"):
        out = subprocess.run([hostio.DRIVER_PATH, str(folder), "--dry-run", "--fusion=" + name], capture_output=True,
                             text=True, check=True).stdout
        assert f"fusion {name} min_consistent=" in out


def test_driver_rejects_bad_input(tmp_path):
    r = subprocess.run([hostio.DRIVER_PATH, str(tmp_path)], capture_output=True, text=True)
    assert r.returncode != 0 and "Invalid input folder" in r.stderr
    r = subprocess.run([hostio.DRIVER_PATH, str(tmp_path), "--vis=bad"], capture_output=True, text=True)
    assert r.returncode != 0 and "Invalid --vis value" in r.stderr
