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


def _fit(orc, cam, W, H, planes, nbrs, px, py, seed=7):
    nb = np.full((9, 2), -1, np.int16)
    nb[0] = (px, py)
    for k, (x, y) in enumerate(nbrs):
        nb[1 + k] = (x, y)
    out = np.zeros(4, np.float32)
    orc.oracle_fit_plane(C.byref(cam), W, H, np.ascontiguousarray(planes, np.float32).ctypes.data, nb.ctypes.data,
                         px, py, seed, out.ctypes.data)
    return out


def test_ransac_fit_plane_closed_form(orc):
    """RANSACToGetFitPlane (DPE.cu:2895-3128): eight neighbours around the pixel,
    all on one camera-frame plane -> the fit is that plane (normal and offset),
    oriented towards the camera; one neighbour off the plane by 10 % in depth
    does not move it (the triple with the lowest residual sum excludes it);
    neighbours all on one side of the pixel (no triangle contains it) -> no
    fit; fewer than three neighbours -> the pixel's own plane."""
    W, H = 96, 72
    K = np.array([[80.0, 0, 47.5], [0, 80.0, 35.5], [0, 0, 1]])
    cam = abi.camera_from_arrays(K, np.eye(3), np.zeros(3), W, H, 1, 10)
    n = np.array([0.2, -0.1, -1.0])
    n /= np.linalg.norm(n)
    d = -float(n @ np.array([0.0, 0.0, 4.0]))  # the plane n . X + d = 0 through (0, 0, 4)
    planes = np.tile(np.array([n[0], n[1], n[2], d], np.float32), (H, W, 1))
    px, py = 48, 36
    ring = [(px + dx * 12, py + dy * 12) for dx, dy in ((-1, -1), (0, -1), (1, -1), (-1, 0), (1, 0), (-1, 1), (0, 1),
                                                        (1, 1))]
    f = _fit(orc, cam, W, H, planes, ring, px, py)
    assert np.allclose(f[:3], n, atol=1e-4) and abs(f[3] - d) < 1e-4 * abs(d)
    # one neighbour off the plane: 10 % deeper along its ray
    bad = planes.copy()
    bx, by = ring[2]
    bad[by, bx, 3] = d * 1.1  # scales that pixel's depth by 1.1 (depth = -d / (n . K^-1 x))
    f2 = _fit(orc, cam, W, H, bad, ring, px, py)
    assert np.allclose(f2[:3], n, atol=1e-4) and abs(f2[3] - d) < 1e-4 * abs(d)
    # all neighbours to the right of the pixel (not collinear): no triangle contains it
    right = [(px + 8 + (k % 3) * 3, py - 9 + 3 * k) for k in range(8)]
    assert not _fit(orc, cam, W, H, planes, right, px, py).any()
    # two neighbours: the pixel's own plane
    own = planes.copy()
    own[py, px] = (0.0, 0.0, -1.0, 4.0)
    assert np.array_equal(_fit(orc, cam, W, H, own, ring[:2], px, py), own[py, px])


DIRS = [(0, -1), (0, 1), (-1, 0), (1, 0), (-1, -1), (1, 1), (-1, 1), (1, -1)]


def _ray(W, H, dx, dy, k):
    ys, xs = np.mgrid[0:H, 0:W]
    nx, ny = xs + k * dx, ys + k * dy
    inside = (nx >= 0) & (nx < W) & (ny >= 0) & (ny < H)
    return nx, ny, inside


def test_gen_edge_inform_against_vectorised_definitions(orc):
    """GenEdgeInform (DPE.cu:2487-2595) against array formulations of its
    definitions: the first edge pixel along each of the 8 directions (found by
    overwriting from the far end), the edge density of the (2r+1)^2 window
    clipped to the image (a box-filter count) through the sigmoid
    1 / (1 + exp(-25 (density - 0.35))), with use_label the boundary density
    bound_pix / tot_pix in integer division (1 only when the whole window is
    boundary), and for WEAK pixels with a label the last pixel of the same label
    along each direction before a -1 label or the border."""
    import scipy.ndimage as ndi
    rng = np.random.default_rng(41)
    W, H, r = 40, 30, 2
    edge = (rng.uniform(size=(H, W)) < 0.15).astype(np.uint8) * 255
    regions, _ = ndi.label(ndi.uniform_filter(rng.uniform(size=(H, W)), 5) > 0.5)
    label = regions.astype(np.int32)
    label[rng.uniform(size=(H, W)) < 0.05] = -1
    label[(regions == 0) & (rng.uniform(size=(H, W)) < 0.5)] = 0
    label[5:10, 5:10] = 0  # a window that is boundary everywhere
    weak = np.where(rng.uniform(size=(H, W)) < 0.5, abi.WEAK, abi.STRONG).astype(np.uint8)
    for lab in (None, label):
        en = np.zeros((H, W, 8, 2), np.int16)
        cm = np.zeros((H, W), np.float32)
        lb = np.zeros((H, W, 8, 2), np.int16)
        orc.oracle_gen_edge_inform(W, H, edge.ctypes.data, None if lab is None else lab.ctypes.data, weak.ctypes.data, r,
                                   en.ctypes.data, cm.ctypes.data, lb.ctypes.data)
        L = max(W, H)
        for i, (dx, dy) in enumerate(DIRS):
            exp = np.full((H, W, 2), -1, np.int64)
            for k in range(L, 0, -1):
                nx, ny, inside = _ray(W, H, dx, dy, k)
                hit = inside & (edge[np.clip(ny, 0, H - 1), np.clip(nx, 0, W - 1)] != 0)
                exp[hit] = np.stack([nx, ny], -1)[hit]
            assert np.array_equal(en[:, :, i], exp), i
        box = np.ones((2 * r + 1, 2 * r + 1))
        tot = ndi.convolve(np.ones((H, W)), box, mode="constant").astype(np.int64)
        cnt = ndi.convolve((edge != 0).astype(float), box, mode="constant").astype(np.int64)
        density = np.float32(1.0) * cnt.astype(np.float32) / tot.astype(np.float32)
        if lab is not None:
            bnd = ndi.convolve((lab == 0).astype(float), box, mode="constant").astype(np.int64)
            density = np.maximum(density, (bnd // tot).astype(np.float32))
            assert (bnd // tot).max() == 1
        expc = 1.0 / (1.0 + np.exp(-25.0 * (density.astype(np.float64) - 0.35)))
        np.testing.assert_allclose(cm, expc.astype(np.float32), rtol=2e-7, atol=0)
        if lab is None:
            assert (lb == -1).all()
            continue
        sel = (weak == abi.WEAK) & (lab > 0)
        for i, (dx, dy) in enumerate(DIRS):
            exp = np.full((H, W, 2), -1, np.int64)
            alive = np.ones((H, W), bool)
            for k in range(1, L + 1):
                nx, ny, inside = _ray(W, H, dx, dy, k)
                nl = np.where(inside, lab[np.clip(ny, 0, H - 1), np.clip(nx, 0, W - 1)], -1)
                alive &= inside & (nl != -1)
                same = alive & (nl == lab)
                exp[same] = np.stack([nx, ny], -1)[same]
            exp[~sel] = -1
            assert np.array_equal(lb[:, :, i], exp), i


def test_depth_to_weak_peak_analysis_known_answers(orc):
    """DepthToWeak's classification of a 61-sample cost curve (DPE.cu:2699-2750)
    on curves whose answer follows from the stated rules: a single strict
    minimum at the centre is STRONG at cost <= 0.15 (0.15 itself included) and
    WEAK above; a minimum more than weak_peak_radius samples off the centre, or
    with cost > 0.5, is WEAK; with several minima the spread of the others
    above the lowest (sqrt of the sum of squares / (count - 1)) decides at 0.2;
    flat curves, plateaus and minima at the first / last two samples are no
    minima (WEAK)."""
    def state(pc, radius=2):
        a = np.ascontiguousarray(pc, np.float32)
        assert a.shape == (61,)
        return orc.oracle_depth_to_weak_state(a.ctypes.data, radius)

    def valley(at, cost, base=1.0):
        pc = np.full(61, base, np.float32)
        pc[at] = cost
        return pc

    S, Wk = abi.STRONG, abi.WEAK
    assert state(valley(30, 0.1)) == S
    assert state(valley(30, np.float32(0.15))) == S
    assert state(valley(30, np.float32(0.1500001))) == Wk
    assert state(valley(32, 0.1)) == S and state(valley(33, 0.1)) == Wk      # weak_peak_radius 2
    assert state(valley(33, 0.1), radius=3) == S
    assert state(valley(30, 0.6)) == Wk                                      # above 0.5
    assert state(np.full(61, 2.0)) == Wk                                     # no minimum
    pl = np.full(61, 1.0, np.float32)
    pl[30] = pl[31] = 0.1                                                    # plateau: not strict
    assert state(pl) == Wk
    assert state(valley(1, 0.0)) == Wk and state(valley(59, 0.0)) == Wk      # outside samples 2..58
    two = valley(30, 0.1)
    two[45] = 0.9                                                            # spread 0.8 > 0.2
    assert state(two) == S
    two[45] = 0.25                                                           # spread 0.15
    assert state(two) == Wk
    three = valley(30, 0.1)
    three[10], three[50] = 0.5, 0.6                                          # sqrt(0.4^2 + 0.5^2) / 2 = 0.32
    assert state(three) == S
    three[10], three[50] = 0.3, 0.3                                          # sqrt(2 * 0.2^2) / 2 = 0.14
    assert state(three) == Wk
    # the lowest minimum decides the position: a lower one off-centre makes it WEAK
    off = valley(30, 0.1)
    off[45] = 0.05
    assert state(off) == Wk
