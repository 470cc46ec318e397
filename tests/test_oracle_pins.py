"""Brute-force pins of the oracle's small building blocks (no GPU): the
insertion sorts, argmin and CDF of view selection (DPE.cu:9-61, 297-311),
FindNearestStrongPoint (DPE.cu:2859-2893) and the strong-pixel median filter
(DPE.cu:1961-2071).  Each check computes the answer another way (library sort,
exhaustive distance search, a neighbour table written from the reference's
conditions) so that a wrong index, bound or tie rule fails it."""
import ctypes as C

import numpy as np
import pytest

from dpe_amd import abi
import oracle_lib


@pytest.fixture(scope="module")
def orc():
    return oracle_lib.load_oracle()


def test_sort_small_is_an_ascending_sort(orc):
    rng = np.random.default_rng(61)
    for n in (0, 1, 2, 5, 17, 32):
        for _ in range(50):
            a = rng.integers(0, 6, n).astype(np.float32) * np.float32(0.25)  # many ties
            got = a.copy()
            orc.oracle_sort_small(got.ctypes.data, n)
            assert np.array_equal(got, np.sort(a))


def test_sort_small_weighted_is_a_stable_sort_by_weight(orc):
    rng = np.random.default_rng(62)
    for n in (1, 3, 9, 31, 64):
        for _ in range(50):
            w = rng.integers(0, 5, n).astype(np.float32)
            xy = rng.integers(-50, 50, (n, 2)).astype(np.int16)
            gw, gxy = w.copy(), xy.copy()
            orc.oracle_sort_small_weighted(gxy.ctypes.data, gw.ctypes.data, n)
            order = np.argsort(w, kind="stable")
            assert np.array_equal(gw, w[order])
            assert np.array_equal(gxy, xy[order])


def test_find_min_cost_index_takes_the_last_minimum(orc):
    rng = np.random.default_rng(63)
    for n in (1, 2, 8, 31):
        for _ in range(100):
            c = rng.integers(0, 4, n).astype(np.float32)
            assert orc.oracle_find_min_cost_index(c.ctypes.data, n) == np.flatnonzero(c == c.min())[-1]


def test_pdf_to_cdf_is_the_normalised_cumulative_sum(orc):
    rng = np.random.default_rng(64)
    for n in (1, 4, 10, 31):
        p = rng.uniform(0.01, 1.0, n).astype(np.float32)
        got = p.copy()
        orc.oracle_pdf_to_cdf(got.ctypes.data, n)
        exp = np.cumsum(p.astype(np.float64)) / p.astype(np.float64).sum()
        assert np.all(np.diff(got) >= 0)
        np.testing.assert_allclose(got, exp, rtol=2e-6, atol=2e-7)


def _nearest_brute(weak, r_max=100):
    H, W = weak.shape
    sy, sx = np.nonzero(weak == abi.STRONG)
    out = np.full((H, W, 2), -1, np.int16)
    for y, x in zip(*np.nonzero(weak == abi.WEAK)):
        if len(sx) == 0:
            continue
        dx, dy = sx - x, sy - y
        cheb = np.maximum(np.abs(dx), np.abs(dy))
        r = cheb.min()
        if r > r_max:
            continue
        cand = np.flatnonzero(cheb == r)
        # the reference scans ring r with x outer, y inner, both ascending
        k = cand[np.lexsort((dy[cand], dx[cand]))[0]]
        out[y, x] = (sx[k], sy[k])
    return out


@pytest.mark.parametrize("seed,strong_frac", [(71, 0.02), (72, 0.2), (73, 0.6)])
def test_nearest_strong_point_is_the_first_on_the_nearest_ring(orc, seed, strong_frac):
    rng = np.random.default_rng(seed)
    H, W = 37, 53
    u = rng.uniform(size=(H, W))
    weak = np.where(u < strong_frac, abi.STRONG, np.where(u < 0.9, abi.WEAK, abi.UNKNOWN)).astype(np.uint8)
    got = np.zeros((H, W, 2), np.int16)
    orc.oracle_nearest_strong(weak.ctypes.data, W, H, got.ctypes.data)
    assert np.array_equal(got, _nearest_brute(weak))


def test_nearest_strong_point_beyond_radius_100_is_none(orc):
    H, W = 5, 240
    weak = np.full((H, W), abi.WEAK, np.uint8)
    weak[2, 0] = abi.STRONG
    got = np.zeros((H, W, 2), np.int16)
    orc.oracle_nearest_strong(weak.ctypes.data, W, H, got.ctypes.data)
    exp = _nearest_brute(weak)
    assert np.array_equal(got, exp)
    assert tuple(got[2, 100]) == (0, 2) and tuple(got[2, 101]) == (-1, -1)


# (dx, dy, condition on (x, y, W, H)) of the 20 neighbours CheckerboardFilterStrong
# reads (DPE.cu:1999-2058); note the y > 2 (not y > 1) guard of the (+-1, -2) pair
FILTER_TAPS = [
    (0, -1, lambda x, y, W, H: y > 0), (0, -3, lambda x, y, W, H: y > 2), (0, -5, lambda x, y, W, H: y > 4),
    (0, 1, lambda x, y, W, H: y < H - 1), (0, 3, lambda x, y, W, H: y < H - 3), (0, 5, lambda x, y, W, H: y < H - 5),
    (-1, 0, lambda x, y, W, H: x > 0), (-3, 0, lambda x, y, W, H: x > 2), (-5, 0, lambda x, y, W, H: x > 4),
    (1, 0, lambda x, y, W, H: x < W - 1), (3, 0, lambda x, y, W, H: x < W - 3), (5, 0, lambda x, y, W, H: x < W - 5),
    (2, -1, lambda x, y, W, H: y > 0 and x < W - 2), (2, 1, lambda x, y, W, H: y < H - 1 and x < W - 2),
    (-2, -1, lambda x, y, W, H: y > 0 and x > 1), (-2, 1, lambda x, y, W, H: y < H - 1 and x > 1),
    (-1, -2, lambda x, y, W, H: x > 0 and y > 2), (1, -2, lambda x, y, W, H: x < W - 1 and y > 2),
    (-1, 2, lambda x, y, W, H: x > 0 and y < H - 2), (1, 2, lambda x, y, W, H: x < W - 1 and y < H - 2),
]


def test_strong_filter_is_the_median_of_its_strong_neighbours(orc):
    rng = np.random.default_rng(81)
    H, W = 23, 29
    for trial in range(6):
        depth = rng.uniform(2.0, 6.0, (H, W)).astype(np.float32)
        costs = rng.uniform(0.0, 1.0, (H, W)).astype(np.float32)
        costs[rng.uniform(size=(H, W)) < 0.1] = 0.0005
        weak = np.where(rng.uniform(size=(H, W)) < 0.7, abi.STRONG, abi.WEAK).astype(np.uint8)
        for y in range(H):
            for x in range(W):
                got = orc.oracle_filter_strong(depth.ctypes.data, costs.ctypes.data, weak.ctypes.data, W, H, x, y)
                if costs[y, x] < np.float32(0.001):
                    assert got == depth[y, x]
                    continue
                v = [depth[y, x]] + [depth[y + dy, x + dx] for dx, dy, ok in FILTER_TAPS
                                     if ok(x, y, W, H) and weak[y + dy, x + dx] == abi.STRONG]
                v = np.sort(np.array(v, np.float32))
                m = len(v) // 2
                exp = v[m] if len(v) % 2 else np.float32((v[m - 1] + v[m]) / np.float32(2))
                assert got == exp, (trial, x, y)


def test_fusion_closed_form_two_coincident_views(orc):
    """RunFusion (DPE.cpp:1387-1548) on two coincident views of a plane: view 0
    keeps one point per pixel (its own back-projection) and masks every pixel of
    view 1, which adds none; a band of rows whose view-1 depth is 5 % off (above
    the 1 % threshold) loses exactly those rows; with min_consistent = 2 (more
    than the one source) nothing is kept."""
    from fuse_util import coincident_pair, fuse_oracle
    W, H = 40, 30
    fs, K = coincident_pair(None, W, H)
    pts = fuse_oracle(orc, fs, abi.fuse_defaults())
    assert len(pts) == W * H
    ys, xs = np.mgrid[0:H, 0:W]
    exp = np.stack([4.0 * (xs - K[0, 2]) / K[0, 0], 4.0 * (ys - K[1, 2]) / K[1, 1], np.full((H, W), 4.0)], -1)
    got = np.stack([pts["x"], pts["y"], pts["z"]], -1).reshape(H, W, 3)  # raster order of view 0
    np.testing.assert_allclose(got, exp, atol=1e-5)
    assert np.all(pts["b"] == 100)
    fs, _ = coincident_pair((slice(10, 17), 1.05), W, H)
    assert len(fuse_oracle(orc, fs, abi.fuse_defaults())) == W * (H - 7)
    fs, _ = coincident_pair(None, W, H)
    assert len(fuse_oracle(orc, fs, abi.fuse_defaults(min_consistent=2))) == 0


def test_fusion_fp64_reprojection_closed_form(orc):
    """The precision form's FP64 fusion geometry (dpe_fuse_params.fp64_reprojection):
    on two coincident views under a rotated, translated camera every pixel of
    view 0 is kept, and its point equals Get3DPointonWorld evaluated in float64
    and rounded to float32, while the fp32 path's points differ from that in
    the last bits (the test fails if either path computes in the other's type)."""
    from fuse_util import coincident_pair, fuse_oracle, rotated_pose, world_points_fp64
    W, H = 40, 30
    R, t = rotated_pose()
    fs, _ = coincident_pair(None, W, H, R, t)
    p64 = fuse_oracle(orc, fs, abi.fuse_defaults(fp64_reprojection=1))
    p32 = fuse_oracle(orc, fs, abi.fuse_defaults())
    assert len(p64) == len(p32) == W * H
    exp = world_points_fp64(fs.views[0].camera, W, H, 4.0)
    got64 = np.stack([p64["x"], p64["y"], p64["z"]], -1)
    got32 = np.stack([p32["x"], p32["y"], p32["z"]], -1)
    assert np.array_equal(got64, exp)
    assert (got32 != exp).any()
    np.testing.assert_allclose(got32, exp, rtol=1e-5, atol=1e-5)
    fs, _ = coincident_pair(None, W, H, R, t)
    bad = abi.fuse_defaults(fp64_reprojection=2)
    cnt = C.c_int64()
    assert orc.oracle_fuse(fs.views, fs.n, C.byref(bad), None, 0, C.byref(cnt)) == abi.ERR_INVALID_ARGUMENT
