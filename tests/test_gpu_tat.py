"""The TAT fusions on the GPU (dpe_fuse_tat through the C-ABI) against the
oracle's sequential restatement (RunFusion_TAT_Intermediate DPE.cpp:1550-1718,
RunFusion_TAT_advanced DPE.cpp:1720-1867).

Bar: bit-identical point lists (xyz and colour, in the reference's order):
the GPU follows the same float arithmetic, and every decision it makes in
parallel reads only masks fixed before the view started plus the per-source
record at the last valid pixel, which is exactly what the sequential loop reads."""
import ctypes as C

import numpy as np
import pytest

from dpe_amd import abi, host, scene
import oracle_lib
from fuse_util import FuseSet, coincident_views, fuse_tat_hip, fuse_tat_oracle

pytestmark = pytest.mark.gpu

INTER, ADV = abi.TAT_INTERMEDIATE, abi.TAT_ADVANCED


@pytest.fixture(scope="module")
def ctx():
    c = host.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def orc():
    return oracle_lib.load_oracle()


def _same(a, b):
    return len(a) == len(b) and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_tat_closed_forms_on_gpu(ctx, orc):
    """The oracle's closed-form scenes (records kept across pixels and masks,
    the accepting k's colour, the angle rule, blocks) give the same points."""
    W, H = 40, 30
    db = float(np.float32(1.0) / np.float32(3500.0))
    blue = (100 + np.arange(H)[:, None] + np.zeros((1, W), np.int64)).astype(np.uint8)
    blk = np.full((H, W), 255, np.uint8)
    blk[5:10] = 127
    blk[10] = 128
    cases = [
        (dict(n=3), None, W * H),
        (dict(n=2), None, 0),
        (dict(n=3, color_b={1: blue}, depth_zero_rows={1: slice(10, 17)}), None, W * H),
        (dict(n=3, depth_zero_rows={1: slice(0, 7)}), None, 2 * W * (H - 7)),
        (dict(n=4, depth_scale={1: 1 + 2.5 * db}, color_b={1: np.full((H, W), 200, np.uint8)}), None, W * H),
        (dict(n=4, depth_scale={1: 1 + 2.5 * db, 2: 1 + 2.5 * db}), None, W * H),
        (dict(n=4, tilt={1: 0.25, 2: 0.25}), None, 0),
        (dict(n=4, tilt={1: 0.2, 2: 0.2}), None, W * H),
        (dict(n=3), [blk, None, None], 75 * W),
        (dict(n=3), [blk, blk, None], 50 * W),
    ]
    for kw, blocks, n_inter in cases:
        kw = dict(kw)
        n = kw.pop("n")
        for variant in (INTER, ADV):
            fs, _ = coincident_views(n, W, H, **kw)
            tp = abi.fuse_tat_defaults(variant)
            pg = fuse_tat_hip(ctx, fs, tp, blocks)
            po = fuse_tat_oracle(orc, fs, tp, blocks)
            print(f"closed form {sorted(kw)} blocks={blocks is not None} variant {variant}: {len(pg)} points")
            assert _same(pg, po), (kw, variant)
            if variant == INTER:
                assert len(pg) == n_inter, kw


@pytest.mark.parametrize("variant", [INTER, ADV])
def test_tat_matches_oracle_on_scenes(ctx, orc, variant):
    """Ground-truth and perturbed depth maps of two seeded scenes, with and
    without block maps: identical point lists."""
    rng = np.random.default_rng(21)
    for s in (scene.config1(), scene.make_scene(480, 360, 450.0, 8, 40.0, seed=0x5EED0077, n_src=5)):
        states = [np.where(v.weak, abi.WEAK, abi.STRONG).astype(np.uint8) for v in s.views]
        normals = [v.normal for v in s.views]
        for noise in (0.0, 0.0005, 0.003):
            depths = [v.depth * (1 + noise * rng.standard_normal(v.depth.shape)).astype(np.float32)
                      for v in s.views]
            # holes, as a filtered depth map has them
            for d in depths:
                d[rng.uniform(size=d.shape) < 0.05] = 0.0
            fs = FuseSet(s, depths, normals, states)
            tp = abi.fuse_tat_defaults(variant)
            for blocks in (None, [np.where(rng.uniform(size=(s.height, s.width)) < 0.3, 0, 255).astype(np.uint8)
                                  if i % 2 == 0 else None for i in range(len(s.views))]):
                pg = fuse_tat_hip(ctx, fs, tp, blocks)
                po = fuse_tat_oracle(orc, fs, tp, blocks)
                print(f"{s.width}x{s.height} noise {noise} blocks={blocks is not None} variant {variant}: "
                      f"oracle {len(po)} points, gpu {len(pg)}")
                assert len(po) > 0
                assert _same(pg, po)


def test_tat_device_memory_and_capacity(ctx):
    """Device pointers in, device points out, and a short output array: the
    count is the full one and the first `capacity` points are written."""
    torch = pytest.importorskip("torch")
    W, H = 40, 30
    fs, _ = coincident_views(3, W, H)
    tp = abi.fuse_tat_defaults(INTER)
    ref = fuse_tat_hip(ctx, fs, tp)
    dev_keep = []
    views = (abi.FuseView * fs.n)()
    for i in range(fs.n):
        v = fs.views[i]
        views[i] = v
        P = W * H
        for name, dt, k in (("depth", np.float32, 1), ("normal", np.float32, 3), ("color", np.uint8, 3)):
            a = np.ctypeslib.as_array(C.cast(getattr(v, name), C.POINTER(C.c_float if dt == np.float32
                                                                             else C.c_uint8)), (P * k,))
            t = torch.from_numpy(a.copy()).to("cuda:0")
            dev_keep.append(t)
            setattr(views[i], name, t.data_ptr())
        views[i].pixel_states = None
    pts = torch.zeros((len(ref), 4), dtype=torch.int32, device="cuda:0")
    cnt = C.c_int64()
    tpd = abi.fuse_tat_defaults(INTER, memory=abi.MEM_DEVICE)
    torch.cuda.synchronize()
    ctx.check(ctx.lib.dpe_fuse_tat(ctx.h, views, fs.n, None, C.byref(tpd), pts.data_ptr(), len(ref), C.byref(cnt)))
    assert cnt.value == len(ref)
    got = pts.cpu().numpy().view(np.uint8).ravel()
    assert np.array_equal(got, ref.view(np.uint8).ravel())
    # host output with room for 100 points only
    short = (abi.Point * 100)()
    ctx.check(ctx.lib.dpe_fuse_tat(ctx.h, fs.views, fs.n, None, C.byref(tp), C.cast(short, C.c_void_p), 100,
                                   C.byref(cnt)))
    assert cnt.value == len(ref)
    assert bytes(short) == ref[:100].tobytes()
    # a bad variant is an error, not a crash
    bad = abi.fuse_tat_defaults(7)
    assert ctx.lib.dpe_fuse_tat(ctx.h, fs.views, fs.n, None, C.byref(bad), None, 0, C.byref(cnt)) == \
        abi.ERR_INVALID_ARGUMENT


def test_fusions_with_31_sources_at_a_ragged_size(ctx, orc):
    """32 views of 47x35 (odd sizes off every tile), each fused against 31
    sources, the most a view can list: all three fusions give the oracle's
    points, with and without holes in the depth maps."""
    from fuse_util import fuse_hip, fuse_oracle
    s = scene.make_scene(47, 35, 60.0, 32, 30.0, seed=0x5EED00F1, n_src=31, supersample=1)
    assert all(len(p) == 31 for p in s.pairs)
    rng = np.random.default_rng(23)
    states = [np.where(v.weak, abi.WEAK, abi.STRONG).astype(np.uint8) for v in s.views]
    normals = [v.normal for v in s.views]
    for noise, holes in ((0.0, 0.0), (0.002, 0.05)):
        depths = [v.depth * (1 + noise * rng.standard_normal(v.depth.shape)).astype(np.float32) for v in s.views]
        for d in depths:
            d[rng.uniform(size=d.shape) < holes] = 0.0
        fs = FuseSet(s, depths, normals, states)
        for fp in (abi.fuse_defaults(), abi.fuse_defaults(min_consistent=3)):
            pg, po = fuse_hip(ctx, fs, fp), fuse_oracle(orc, fs, fp)
            print(f"eth noise {noise}: oracle {len(po)} points, gpu {len(pg)}")
            assert len(po) > 0 and _same(pg, po)
        for variant in (INTER, ADV):
            tp = abi.fuse_tat_defaults(variant)
            pg, po = fuse_tat_hip(ctx, fs, tp), fuse_tat_oracle(orc, fs, tp)
            print(f"tat {variant} noise {noise}: oracle {len(po)} points, gpu {len(pg)}")
            assert len(po) > 0 and _same(pg, po)
