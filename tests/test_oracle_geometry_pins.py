"""Pins of the oracle's geometry and random-hypothesis functions (no GPU), each
against an independent statement of what the function computes:

* ComputeGeomConsistencyCost (DPE.cu:919-957): a float64 evaluation of its
  geometric definition (project into the source, read the source depth at the
  truncated pixel, back-project, reproject, distance capped at 3) on planar
  scenes with ground-truth depth maps; consistent maps give ~0, a hole gives 3.
* ComputeMultiViewInitialCost's mask update (DPE.cu:832-861, unSetBit
  DPE.cu:81-84): the closed form of clearing bits 0..k for an invalid view k.
* GenerateRandomNormal / GeneratePerturbedNormal (DPE.cu:365-428): Marsaglia's
  (1972) disk method and a Z-Y-X Euler rotation evaluated in float64 from the
  same Philox draws (the generator is pinned against rocRAND elsewhere), plus
  the invariants unit length, facing the camera, uniform on the hemisphere and
  the rotation-angle bound.
* TransformNormal / TransformNormal2RefCam (DPE.cu:528-546) and
  GetDepthandNormal (DPE.cu:1944-1959): R^T n and R n in float64, round trips,
  and the depth of a plane built from a known depth.
"""
import ctypes as C

import numpy as np
import pytest
from scipy import stats

from dpe_amd import abi
import oracle_lib


@pytest.fixture(scope="module")
def orc():
    return oracle_lib.load_oracle()


def _rot(axis, a):
    c, s = np.cos(a), np.sin(a)
    if axis == "x":
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    if axis == "y":
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def _cam(R, C_world, W=320, H=240, fx=300.0, fy=280.0, cx=161.5, cy=118.25):
    K = np.array([[fx, 0, cx], [0, fy, cy], [0, 0, 1]], np.float32)
    R = np.asarray(R, np.float32)
    t = (-R.astype(np.float64) @ np.asarray(C_world, np.float64)).astype(np.float32)
    return abi.camera_from_arrays(K, R, t, W, H, 2.0, 6.0)


def _arr(cam):
    K = np.array(cam.K, np.float64).reshape(3, 3)
    R = np.array(cam.R, np.float64).reshape(3, 3)
    t = np.array(cam.t, np.float64)
    c = np.array(cam.c, np.float64)
    return K, R, t, c


def _plane_depth_map(cam, n_w, p0):
    """Ground-truth depth (camera z) of the world plane n_w . (X - p0) = 0 at every pixel centre."""
    K, R, t, c = _arr(cam)
    H, W = cam.height, cam.width
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    rays_c = np.stack([(xs - K[0, 2]) / K[0, 0], (ys - K[1, 2]) / K[1, 1], np.ones_like(xs)], -1)
    rays_w = rays_c @ R  # R^T r
    lam = (n_w @ (p0 - c)) / (rays_w @ n_w)  # X = c + lam * R^T r, camera z = lam
    return lam


def _geom_cost64(ref, src, depth_map, px, py, plane):
    """The geometric-consistency cost from its definition, in float64."""
    Kr, Rr, tr, cr = _arr(ref)
    Ks, Rs, ts, cs = _arr(src)
    n, d = np.asarray(plane[:3], np.float64), float(plane[3])
    ray = np.array([(px - Kr[0, 2]) / Kr[0, 0], (py - Kr[1, 2]) / Kr[1, 1], 1.0])
    depth = -d / (n @ ray)
    Xw = Rr.T @ (depth * ray) + cr
    xs = Ks @ (Rs @ Xw + ts)
    u, v = xs[0] / xs[2], xs[1] / xs[2]
    H, W = depth_map.shape
    ix = min(max(int(np.trunc(u)), 0), W - 1)
    iy = min(max(int(np.trunc(v)), 0), H - 1)
    ds = float(depth_map[iy, ix])
    if ds == 0.0:
        return 3.0, u, v
    X2 = Rs.T @ (ds * np.array([(u - Ks[0, 2]) / Ks[0, 0], (v - Ks[1, 2]) / Ks[1, 1], 1.0])) + cs
    xr = Kr @ (Rr @ X2 + tr)
    return min(3.0, float(np.hypot(px - xr[0] / xr[2], py - xr[1] / xr[2]))), u, v


def _cam_plane(cam, n_w, p0):
    """Camera-frame (n, d) of the world plane through p0 with normal n_w."""
    K, R, t, c = _arr(cam)
    n_c = R @ n_w
    X0 = R @ p0 + t
    return np.array([*n_c, -n_c @ X0], np.float32)


def test_geom_cost_is_zero_on_consistent_depth_and_three_on_a_hole(orc):
    ref = _cam(np.eye(3), [0, 0, 0])
    src = _cam(_rot("y", 0.12) @ _rot("x", -0.05), [0.6, 0.1, 0.2])
    rng = np.random.default_rng(81)
    for trial in range(4):
        n_w = np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), -1.0])
        n_w /= np.linalg.norm(n_w)
        p0 = np.array([0.0, 0.0, rng.uniform(3.0, 5.0)])
        dmap = _plane_depth_map(src, n_w, p0).astype(np.float32)
        plane = _cam_plane(ref, n_w, p0)
        worst = 0.0
        checked = 0
        for _ in range(300):
            px, py = int(rng.integers(20, 300)), int(rng.integers(20, 220))
            exp, u, v = _geom_cost64(ref, src, dmap, px, py, plane)
            if min(abs(u - round(u)), abs(v - round(v))) < 1e-3 or not (0 <= u < 319 and 0 <= v < 239):
                continue  # lookup pixel ambiguous at float precision, or outside the source
            got = orc.oracle_geom_cost(C.byref(ref), C.byref(src), 320, 240, dmap.ctypes.data, px, py,
                                       plane.ctypes.data)
            assert abs(got - exp) < 2e-3, (trial, px, py, got, exp)
            worst = max(worst, got)
            checked += 1
        assert checked > 150
        # the depth read at the truncated pixel is off the exact one by under a pixel's
        # depth change, so a consistent map reprojects within about a pixel
        assert worst < 1.5
        # a hole (depth 0) in the source map costs the maximum, 3
        hole = dmap.copy()
        hole[:] = 0
        assert orc.oracle_geom_cost(C.byref(ref), C.byref(src), 320, 240, hole.ctypes.data, 160, 120,
                                    plane.ctypes.data) == np.float32(3.0)


def test_geom_cost_reads_the_truncated_source_pixel(orc):
    """A fronto-parallel plane seen by a translated source: the consistent map
    gives ~0 everywhere; a map whose depth steps by column gives the cost of
    the truncated column (Q13), not the rounded one, and clamps outside."""
    ref = _cam(np.eye(3), [0, 0, 0])
    src = _cam(np.eye(3), [0.5, 0.0, 0.0])
    n_w, p0 = np.array([0.0, 0.0, -1.0]), np.array([0.0, 0.0, 4.0])
    plane = _cam_plane(ref, n_w, p0)
    flat = np.full((240, 320), 4.0, np.float32)
    for px in range(40, 300, 7):
        got = orc.oracle_geom_cost(C.byref(ref), C.byref(src), 320, 240, flat.ctypes.data, px, 100,
                                   plane.ctypes.data)
        assert got < 1e-3
    step = (4.0 + 0.01 * np.arange(320, dtype=np.float32))[None, :].repeat(240, 0).astype(np.float32)
    for px in range(0, 320, 3):
        exp, u, v = _geom_cost64(ref, src, step, px, 100, plane)
        if abs(u - round(u)) < 1e-3:
            continue
        got = orc.oracle_geom_cost(C.byref(ref), C.byref(src), 320, 240, step.ctypes.data, px, 100,
                                   plane.ctypes.data)
        assert abs(got - exp) < 2e-3, (px, u, got, exp)
        # reading the rounded column instead would differ once u's fraction is >= 0.5
        if 0 <= u < 318 and u - np.floor(u) > 0.6:
            wrong = step.copy()
            wrong[:, :] = step[:, np.minimum(np.arange(320) + 1, 319)]
            alt, _, _ = _geom_cost64(ref, src, wrong, px, 100, plane)
            assert abs(alt - exp) > 1e-2


def test_initial_cost_clears_bits_up_to_an_invalid_view(orc):
    """Q1: an invalid (cost >= 2) prior-selected view k clears bits 0..k, so the
    mask keeps only the bits above the highest invalid selected view, while the
    cost is the float32 mean of the selected valid views in view order."""
    # the case that tells 'clear bits 0..k' from 'clear bit k': mask 0b11011, bit 3's view invalid
    costs = np.array([0.5, 0.25, 0.75, 2.0, 0.125], np.float32)
    sel = C.c_uint32(0b11011)
    cost = orc.oracle_initial_cost(costs.ctypes.data, 6, C.byref(sel))
    assert sel.value == 0b10000
    assert cost == np.float32((np.float32(0.5) + np.float32(0.25) + np.float32(0.125)) / np.float32(3))
    rng = np.random.default_rng(82)
    for _ in range(3000):
        n = int(rng.integers(1, 32))
        costs = rng.uniform(0, 1.99, n).astype(np.float32)
        invalid = rng.random(n) < 0.2
        costs[invalid] = rng.choice(np.array([2.0, 2.0, 3.5, np.inf], np.float32), invalid.sum())
        mask = int(rng.integers(0, 1 << n)) if n < 31 else int(rng.integers(0, 1 << 31))
        sel = C.c_uint32(mask)
        got = orc.oracle_initial_cost(costs.ctypes.data, n + 1, C.byref(sel))
        bits = [k for k in range(n) if mask >> k & 1]
        bad = [k for k in bits if invalid[k]]
        exp_mask = mask & ~((2 << max(bad)) - 1) & 0xFFFFFFFF if bad else mask
        assert sel.value == exp_mask
        good = [k for k in bits if not invalid[k]]
        if good:
            acc = np.float32(0)
            for k in good:
                acc = np.float32(acc + costs[k])
            assert got == np.float32(acc / np.float32(len(good)))
        else:
            assert got == np.float32(2.0)


def _philox_draws(orc, seed, pixel, phase, call, n):
    out = []
    for b in range((n + 3) // 4):
        ctr = (C.c_uint32 * 4)(b, pixel, phase, call)
        key = (C.c_uint32 * 2)(seed & 0xFFFFFFFF, seed >> 32)
        o = (C.c_uint32 * 4)()
        orc.oracle_philox(ctr, key, o)
        out += list(o)
    return out[:n]


def _uniform(x):
    return ((x >> 8) + 1) * 2.0 ** -24


def _view_dir(K, px, py, depth):
    X = np.array([depth * (px - K[0, 2]) / K[0, 0], depth * (py - K[1, 2]) / K[1, 1], depth])
    return X / np.linalg.norm(X)


def test_random_normal_is_marsaglias_method_facing_the_camera(orc):
    cam = _cam(np.eye(3), [0, 0, 0])
    K, _, _, _ = _arr(cam)
    rng = np.random.default_rng(83)
    out = np.zeros(4, np.float32)
    cosines = []
    for i in range(20000):
        px, py = int(rng.integers(0, 320)), int(rng.integers(0, 240))
        depth = float(rng.uniform(1.2, 7.2))
        seed, pixel, phase, call = 0x5EED0002, i, 2, int(rng.integers(0, 4))
        orc.oracle_random_normal(C.byref(cam), px, py, depth, seed, pixel, phase, call, out.ctypes.data)
        # Marsaglia (1972): (q1, q2) uniform in the unit disk by rejection, then
        # (2 q1 sqrt(1-s), 2 q2 sqrt(1-s), 1 - 2 s) is uniform on the sphere
        draws = _philox_draws(orc, seed, pixel, phase, call, 64)
        k = 0
        while True:
            q1 = 2 * _uniform(draws[k]) - 1
            q2 = 2 * _uniform(draws[k + 1]) - 1
            k += 2
            s = q1 * q1 + q2 * q2
            if s < 1:
                break
        n = np.array([2 * q1 * np.sqrt(1 - s), 2 * q2 * np.sqrt(1 - s), 1 - 2 * s])
        v = _view_dir(K, px, py, depth)
        if n @ v > 0:
            n = -n
        n /= np.linalg.norm(n)
        # the float32 sum s = q1^2 + q2^2 is rounded (half an ulp of 1), which
        # sqrt(1 - s) amplifies near the rim of the disk
        np.testing.assert_allclose(out[:3], n, atol=2e-6 + 1.2e-7 / np.sqrt(max(1.0 - s, 1e-12)))
        assert out[3] == 0
        assert abs(np.linalg.norm(out[:3].astype(np.float64)) - 1) < 1e-6
        c = float(out[:3].astype(np.float64) @ v)
        assert c <= 1e-7
        cosines.append(-c)
    # on the hemisphere facing the camera, -n.v is uniform on [0, 1] (Archimedes)
    assert stats.kstest(cosines, "uniform").pvalue > 1e-3


def test_perturbed_normal_is_a_zyx_euler_rotation_within_its_bound(orc):
    cam = _cam(np.eye(3), [0, 0, 0])
    K, _, _, _ = _arr(cam)
    p = np.float32(0.02 * np.pi)
    rng = np.random.default_rng(84)
    out = np.zeros(4, np.float32)
    changed = 0
    for i in range(20000):
        px, py = int(rng.integers(0, 320)), int(rng.integers(0, 240))
        v = _view_dir(K, px, py, 1.0)
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        if n @ v > 0:
            n = -n
        n4 = np.array([*n, 0.25], np.float32)
        seed, pixel, phase, call = 0x5EED0002, i, 8, 1
        orc.oracle_perturbed_normal(C.byref(cam), px, py, n4.ctypes.data, p, seed, pixel, phase, call,
                                    out.ctypes.data)
        a = [(np.float32(_uniform(x)) - np.float32(0.5)) * p for x in _philox_draws(orc, seed, pixel, phase, call, 3)]
        R = _rot("z", float(a[2])) @ _rot("y", float(a[1])) @ _rot("x", float(a[0]))
        m = R @ n4[:3].astype(np.float64)
        exp = n4[:3].astype(np.float64) if m @ v >= 0 else m
        exp = exp / np.linalg.norm(exp)
        np.testing.assert_allclose(out[:3], exp, atol=3e-6)
        assert out[3] == n4[3]  # w passes through
        ang = np.arccos(np.clip(out[:3].astype(np.float64) @ n4[:3], -1, 1))
        assert ang <= 1.5 * p + 1e-5  # |a1| + |a2| + |a3| <= 3 p / 2
        assert out[:3].astype(np.float64) @ v < 1e-6
        changed += ang > 1e-4
    assert changed > 0.95 * 20000


def test_transform_normal_round_trip_and_depth_and_normal(orc):
    rng = np.random.default_rng(85)
    out = np.zeros(4, np.float32)
    back = np.zeros(4, np.float32)
    for _ in range(2000):
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        cam = _cam(R, rng.uniform(-1, 1, 3), fx=float(rng.uniform(200, 900)), fy=float(rng.uniform(200, 900)),
                   cx=float(rng.uniform(100, 200)), cy=float(rng.uniform(80, 160)))
        K, Rf, _, _ = _arr(cam)
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        n4 = np.array([*n, 1.5], np.float32)
        orc.oracle_transform_normal(C.byref(cam), n4.ctypes.data, 1, out.ctypes.data)
        np.testing.assert_allclose(out[:3], Rf.T @ n4[:3], atol=2e-6)
        assert out[3] == n4[3]
        orc.oracle_transform_normal(C.byref(cam), out.ctypes.data, 0, back.ctypes.data)
        np.testing.assert_allclose(back[:3], n4[:3], atol=4e-6)
        np.testing.assert_allclose(back[:3], Rf @ out[:3].astype(np.float64), atol=2e-6)
        # GetDepthandNormal: the camera-frame plane through the point at `depth` on
        # pixel (px, py) gives back that depth and the world normal
        px, py = int(rng.integers(0, 320)), int(rng.integers(0, 240))
        depth = float(rng.uniform(2, 6))
        nc = n if n @ _view_dir(K, px, py, 1.0) < 0 else -n
        X = depth * np.array([(px - K[0, 2]) / K[0, 0], (py - K[1, 2]) / K[1, 1], 1.0])
        if abs(nc @ X / np.linalg.norm(X)) < 0.2:
            continue  # grazing plane: depth ill-conditioned
        plane = np.array([*nc, -nc @ X], np.float32)
        orc.oracle_depth_and_normal(C.byref(cam), px, py, plane.ctypes.data, out.ctypes.data)
        assert abs(out[3] - depth) < 2e-5 * depth
        np.testing.assert_allclose(out[:3], Rf.T @ plane[:3].astype(np.float64), atol=2e-6)
