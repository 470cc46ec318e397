"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

Bar: >= 99.5 % of pixels with |dd|/d < 1e-3 and normals within
0.5 degrees (the stated tolerance, checked first).  Both sides follow the same
numerics contract and the outputs are bit-identical, which is what the tests
assert beyond the tolerance: every plane, cost, pixel state and selected-view
mask, and every byte of every fused point.
"""
import ctypes as C

import numpy as np
import pytest

from dpe_amd import abi, host, scene
import oracle_lib
from parity_util import bitwise_fractions, depth_agreement, normal_agreement

ALL_BITWISE = dict(planes=1.0, costs=1.0, pixel_states=1.0, selected_views=1.0)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = host.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def orc():
    return oracle_lib.load_oracle()


@pytest.fixture(scope="module")
def c1():
    return scene.config1()


def _ncc_inputs(s, ref, srcs):
    ids = [ref] + list(srcs)
    imgs = [s.views[i].image.astype(np.float32) for i in ids]
    cams = host.scene_cameras(s, ids)
    prm = host.call_params(s.depth_min, s.depth_max, use_APD=0, use_edge=0, use_limit=0, use_label=0)
    return host.ProblemInputs(imgs, cams, prm)


def _gt_planes(s, ref, noise, rng):
    """Camera-frame planes (n, d) from the ground truth of the reference view, optionally perturbed."""
    v = s.views[ref]
    H, W = v.depth.shape
    n = v.normal @ v.R.T  # world -> camera
    depth = v.depth * (1 + noise * rng.standard_normal(v.depth.shape))
    ys, xs = np.mgrid[0:H, 0:W]
    K = v.K
    X = np.stack([depth * (xs - K[0, 2]) / K[0, 0], depth * (ys - K[1, 2]) / K[1, 1], depth], -1)
    d = -np.sum(n * X, -1)
    return np.concatenate([n, d[..., None]], -1).astype(np.float32)


def _random_planes(s, ref, rng):
    v = s.views[ref]
    H, W = v.depth.shape
    n = rng.standard_normal((H, W, 3))
    n[..., 2] = -np.abs(n[..., 2]) - 0.5
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    depth = rng.uniform(1.2, 7.2, (H, W))
    ys, xs = np.mgrid[0:H, 0:W]
    K = v.K
    X = np.stack([depth * (xs - K[0, 2]) / K[0, 0], depth * (ys - K[1, 2]) / K[1, 1], depth], -1)
    d = -np.sum(n * X, -1)
    return np.concatenate([n, d[..., None]], -1).astype(np.float32)


def test_ncc_cost_vectors_match_oracle(ctx, orc, c1):
    """ComputeMultiViewCostVectorOld on GT, perturbed-GT and random planes: bitwise."""
    ref = c1.meta["ref_mid"]
    pin = _ncc_inputs(c1, ref, c1.pairs[ref])
    rng = np.random.default_rng(7)
    planes = np.stack([_gt_planes(c1, ref, 0.0, rng), _gt_planes(c1, ref, 0.01, rng), _random_planes(c1, ref, rng)])
    nh, H, W = planes.shape[0], c1.height, c1.width
    NV = len(c1.pairs[ref])
    g = np.zeros((nh, NV, H, W), np.float32)
    o = np.zeros_like(g)
    ctx.check(ctx.lib.dpe_ncc_cost(ctx.h, C.byref(pin.params), C.byref(pin.struct), nh,
                                   planes.ctypes.data, g.ctypes.data))
    assert orc.oracle_ncc_cost(C.byref(pin.params), C.byref(pin.struct), nh, planes.ctypes.data,
                               o.ctypes.data, 8) == 0
    exact = float((g == o).mean())
    close = float((np.abs(g - o) <= 1e-5).mean())
    print(f"ncc bitwise {exact:.6f} within 1e-5 {close:.6f}")
    assert close >= 0.9999
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    # ground-truth planes give near-zero cost in the best source view
    gt_cost = g[0].min(0)
    assert np.median(gt_cost[~c1.views[ref].weak]) < 0.15
    st = ctx.stats()
    assert st.ncc_hyps == nh * NV * H * W


def _check_counts(st, fam_o, label):
    """The GPU evaluates exactly the oracle's NCC hypotheses and taps, family by
    family (the metric's unit), except the evaluations whose results the
    reference never reads (the oracle's last family), which it skips."""
    fam_g = [(st.family_hyps[f], st.family_taps[f]) for f in range(abi.NCC_FAMILIES)]
    print(f"{label}: NCC (hyps, taps) per family gpu {fam_g} oracle {fam_o}")
    assert sum(h for h, _ in fam_g) == st.ncc_hyps and sum(t for _, t in fam_g) == st.ncc_taps
    assert fam_g == [tuple(x) for x in fam_o[:abi.NCC_FAMILIES]]


def _compare(res_g, res_o, label):
    da = depth_agreement(res_g["depth"], res_o["depth"])
    na = normal_agreement(res_g["normal"], res_o["normal"])
    sa = float((res_g["pixel_states"] == res_o["pixel_states"]).mean())
    va = float((res_g["selected_views"] == res_o["selected_views"]).mean())
    exact = float((res_g["depth"] == res_o["depth"]).mean())
    print(f"{label}: depth {da:.5f} (bitwise {exact:.5f}) normal {na:.5f} states {sa:.5f} views {va:.5f}")
    return da, na, sa, va


def test_config1_first_init_matches_oracle(ctx, orc, c1):
    """320x240 scene, 4 sources: one FIRST_INIT call, 3 checkerboard iterations, FP32."""
    ref = c1.meta["ref_mid"]
    pin = host.first_call_inputs(c1, ref, c1.pairs[ref], max_iterations=3)
    out_g = host.run_hip(ctx, pin)
    fam_o = []
    out_o, _ = oracle_lib.run_oracle(orc, pin, 8, fam_o)
    _check_counts(ctx.stats(), fam_o, "config1")
    dmin, dmax = pin.params.depth_min, pin.params.depth_max
    res_g = host.assemble(out_g, dmin, dmax)
    res_o = host.assemble(out_o, dmin, dmax)
    da, na, sa, va = _compare(res_g, res_o, "config1")
    assert da >= 0.995
    assert na >= 0.995
    assert sa >= 0.99
    assert bitwise_fractions(out_g, out_o) == ALL_BITWISE


def test_config1b_weak_path_matches_oracle(ctx, orc, c1):
    """320x240 scene, 4 sources: a REFINE_ITER call with APD, edge sampling, the weak path and
    geometric consistency on, fed by a FIRST_INIT call of every view."""
    store_g, store_o = host.StateStore(), host.StateStore()
    ref = c1.meta["ref_mid"]
    for i in range(len(c1.views)):
        pin = host.first_call_inputs(c1, i, c1.pairs[i], max_iterations=3, call_index=0)
        dmin, dmax = pin.params.depth_min, pin.params.depth_max
        store_g.put(i, host.assemble(host.run_hip(ctx, pin), dmin, dmax))
        store_o.put(i, host.assemble(oracle_lib.run_oracle(orc, pin, 8)[0], dmin, dmax))
    # the refine call consumes the oracle's first-call state on both sides
    pin = host.refine_call_inputs(c1, ref, c1.pairs[ref], store_o, max_iterations=3, call_index=1)
    weak = int((pin.weak_info == abi.WEAK).sum())
    out_g = host.run_hip(ctx, pin)
    fam_o = []
    out_o, _ = oracle_lib.run_oracle(orc, pin, 8, fam_o)
    _check_counts(ctx.stats(), fam_o, "config1b")
    dmin, dmax = pin.params.depth_min, pin.params.depth_max
    res_g = host.assemble(out_g, dmin, dmax)
    res_o = host.assemble(out_o, dmin, dmax)
    print(f"weak pixels in the refine call: {weak}")
    assert weak > 0
    da, na, sa, va = _compare(res_g, res_o, "config1b")
    assert da >= 0.995
    assert na >= 0.995
    assert bitwise_fractions(out_g, out_o) == ALL_BITWISE


@pytest.mark.parametrize("W,H,n_views,n_src", [(61, 33, 2, 1), (47, 35, 32, 31)])
def test_ragged_sizes_and_source_counts_match_oracle(ctx, orc, W, H, n_views, n_src):
    """Odd widths and heights off every tile size, one source (the fewest a call
    takes) and 31 sources (the most: 32 images, DPE.cpp:944-947).  H = 33 also
    takes the reference's checkerboard launch quirk: its half-height grid
    (H / 2 rounded up to 16 rows, doubled) stops at row 32, so the last row keeps
    its random initialisation on both sides.  FIRST_INIT calls of every view, then
    a REFINE call with the weak path fed by the oracle's states."""
    s = scene.make_scene(W, H, 60.0, n_views, 30.0, seed=0x5EED00F1, n_src=n_src, supersample=1)
    store = host.StateStore()
    ref = n_views // 2
    for i in range(n_views):
        pin = host.first_call_inputs(s, i, s.pairs[i], max_iterations=2, call_index=0)
        assert pin.struct.num_images == n_src + 1
        dmin, dmax = pin.params.depth_min, pin.params.depth_max
        out_o = oracle_lib.run_oracle(orc, pin, 8)[0]
        res_o = host.assemble(out_o, dmin, dmax)
        store.put(i, res_o)
        if i == ref:
            out_g = host.run_hip(ctx, pin)
            res_g = host.assemble(out_g, dmin, dmax)
            da, na, sa, va = _compare(res_g, res_o, f"{W}x{H} first")
            assert da >= 0.995 and na >= 0.995 and sa >= 0.99
            assert bitwise_fractions(out_g, out_o) == ALL_BITWISE
    pin = host.refine_call_inputs(s, ref, s.pairs[ref], store, max_iterations=1, call_index=1)
    assert (pin.weak_info == abi.WEAK).sum() > 0
    out_g = host.run_hip(ctx, pin)
    fam_o = []
    out_o, _ = oracle_lib.run_oracle(orc, pin, 8, fam_o)
    _check_counts(ctx.stats(), fam_o, f"{W}x{H} refine")
    dmin, dmax = pin.params.depth_min, pin.params.depth_max
    da, na, sa, va = _compare(host.assemble(out_g, dmin, dmax), host.assemble(out_o, dmin, dmax), f"{W}x{H} refine")
    assert da >= 0.995 and na >= 0.995
    assert bitwise_fractions(out_g, out_o) == ALL_BITWISE


def test_fusion_matches_oracle(ctx, orc, c1):
    """RunFusion on ground-truth and on perturbed depth maps: the GPU fixpoint
    reproduces the sequential mask semantics, so point counts and points match."""
    from fuse_util import FuseSet, fuse_hip, fuse_oracle
    rng = np.random.default_rng(9)
    n = len(c1.views)
    states = [np.where(v.weak, abi.WEAK, abi.STRONG).astype(np.uint8) for v in c1.views]
    normals = [v.normal for v in c1.views]
    for noise in (0.0, 0.004):
        depths = [v.depth * (1 + noise * rng.standard_normal(v.depth.shape)).astype(np.float32) for v in c1.views]
        fs = FuseSet(c1, depths, normals, states)
        fp = abi.fuse_defaults()
        po = fuse_oracle(orc, fs, fp)
        pg = fuse_hip(ctx, fs, fp)
        print(f"fusion noise {noise}: oracle {len(po)} points, gpu {len(pg)}")
        assert abs(len(pg) - len(po)) <= 0.005 * len(po)
        assert len(pg) == len(po) > 0 and np.array_equal(pg.view(np.uint8), po.view(np.uint8))
        # the 1600x1200 workload's rule: at least 3 consistent views
        fp3 = abi.fuse_defaults(min_consistent=3)
        po3, pg3 = fuse_oracle(orc, fs, fp3), fuse_hip(ctx, fs, fp3)
        print(f"fusion noise {noise}, >= 3 consistent views: oracle {len(po3)} points, gpu {len(pg3)}")
        assert abs(len(pg3) - len(po3)) <= 0.005 * len(po3)
        assert len(pg3) == len(po3) > 0 and np.array_equal(pg3.view(np.uint8), po3.view(np.uint8))


def test_fusion_closed_form_on_gpu(ctx):
    """The GPU fusion on two coincident views of a plane: one point per pixel of
    view 0, none from view 1; a 5 %-off band of view-1 rows drops exactly those
    rows; min_consistent 2 keeps nothing (the closed form the oracle also meets)."""
    from fuse_util import coincident_pair, fuse_hip
    W, H = 40, 30
    fs, _ = coincident_pair(None, W, H)
    assert len(fuse_hip(ctx, fs, abi.fuse_defaults())) == W * H
    fs, _ = coincident_pair((slice(10, 17), 1.05), W, H)
    assert len(fuse_hip(ctx, fs, abi.fuse_defaults())) == W * (H - 7)
    fs, _ = coincident_pair(None, W, H)
    assert len(fuse_hip(ctx, fs, abi.fuse_defaults(min_consistent=2))) == 0


def test_fusion_fp64_reprojection_matches_oracle(ctx, orc, c1):
    """The precision form's FP64 fusion geometry on the GPU equals the oracle's
    FP64 restatement bit for bit (ground-truth and perturbed depths), and on
    coincident views under a rotated camera its points are the float64
    back-projections rounded to float32."""
    from fuse_util import FuseSet, coincident_pair, fuse_hip, fuse_oracle, rotated_pose, world_points_fp64
    rng = np.random.default_rng(19)
    states = [np.where(v.weak, abi.WEAK, abi.STRONG).astype(np.uint8) for v in c1.views]
    normals = [v.normal for v in c1.views]
    for noise in (0.0, 0.004):
        depths = [v.depth * (1 + noise * rng.standard_normal(v.depth.shape)).astype(np.float32) for v in c1.views]
        fs = FuseSet(c1, depths, normals, states)
        fp = abi.fuse_defaults(fp64_reprojection=1)
        po = fuse_oracle(orc, fs, fp)
        pg = fuse_hip(ctx, fs, fp)
        p32 = fuse_hip(ctx, fs, abi.fuse_defaults())
        print(f"fp64 fusion noise {noise}: oracle {len(po)} points, gpu {len(pg)}, fp32 path {len(p32)}")
        assert len(po) > 0 and len(pg) == len(po)
        assert np.array_equal(pg.view(np.uint8), po.view(np.uint8))
        assert abs(len(p32) - len(pg)) <= 0.005 * len(pg)
    W, H = 40, 30
    R, t = rotated_pose()
    fs, _ = coincident_pair(None, W, H, R, t)
    pg = fuse_hip(ctx, fs, abi.fuse_defaults(fp64_reprojection=1))
    assert len(pg) == W * H
    assert np.array_equal(np.stack([pg["x"], pg["y"], pg["z"]], -1), world_points_fp64(fs.views[0].camera, W, H, 4.0))
