"""Forms beyond the plain path, judged against the synthetic ground truth (their
results are not the reference algorithm's, so the oracle does not apply).

Gauss-Newton form (dpe_form.gauss_newton_steps = 2): on config-1 inputs two
of the refinement's five random hypotheses (the perturbed normal, the random
depth + normal) are replaced by two Gauss-Newton steps on the per-view costs.
Reported: ground-truth accuracy (|dd|/d < 1 % and < 0.2 %) per iteration count
for plain and Gauss-Newton, and the iteration count each needs to reach plain's
4-iteration accuracy.  Bar: within 1 point of plain at every iteration count
(the random draws differ), and ahead of plain at 0.2 % after 2 iterations (the
steps converge faster on the planes they start near)."""
import ctypes as C

import numpy as np
import pytest

from dpe_amd import abi, host, scene

pytestmark = pytest.mark.gpu

ITERS = (1, 2, 3, 4, 6)


def _accuracy(ctx, s, ref, iters, form):
    gt = s.views[ref].depth
    pin = host.first_call_inputs(s, ref, s.pairs[ref], max_iterations=iters)
    out = host.run_hip(ctx, pin, form)
    r = host.assemble(out, pin.params.depth_min, pin.params.depth_max)
    rel = np.abs(r["depth"] - gt) / gt
    return float((rel < 0.01).mean()), float((rel < 0.002).mean()), ctx.stats().gpu_ms


def test_gauss_newton_form_against_ground_truth():
    s = scene.config1()
    ref = s.meta["ref_mid"]
    ctx = host.Context(0)
    try:
        res = {}
        for gn in (0, 2):
            form = abi.Form(gauss_newton_steps=gn)
            res[gn] = [_accuracy(ctx, s, ref, it, form) for it in ITERS]
            print(f"gauss_newton_steps={gn}: " + ", ".join(
                f"{it} it: {a1:.4f} @1% / {a02:.4f} @0.2% ({ms:.1f} ms)" for it, (a1, a02, ms) in zip(ITERS, res[gn])))
        target = res[0][ITERS.index(4)][0]
        need = {gn: next((it for it, r in zip(ITERS, res[gn]) if r[0] >= target), None) for gn in (0, 2)}
        print(f"iterations to reach plain's 4-iteration accuracy {target:.4f} @1%: plain {need[0]}, "
              f"Gauss-Newton {need[2]}")
        for (p1, p02, _), (g1, g02, _) in zip(res[0], res[2]):
            assert g1 >= p1 - 0.01 and g02 >= p02 - 0.01
        assert res[2][ITERS.index(2)][1] > res[0][ITERS.index(2)][1]
        # an invalid form is rejected, not run
        pin = host.first_call_inputs(s, ref, s.pairs[ref], max_iterations=1)
        with pytest.raises(host.DPEError):
            host.run_hip(ctx, pin, abi.Form(gauss_newton_steps=3))
        with pytest.raises(host.DPEError):
            host.run_hip(ctx, pin, abi.Form(precision=3))
    finally:
        ctx.close()


PRECISIONS = {"plain": abi.PRECISION_FP32, "fp16": abi.PRECISION_FP16, "bf16": abi.PRECISION_BF16}


def _cost_inputs(seed, continuous):
    s = scene.make_scene(320, 240, 300.0, 11, 30.0, seed=seed, n_src=10)
    ref = s.meta["ref_mid"]
    ids = [ref] + s.pairs[ref]
    imgs = [s.views[i].image.astype(np.float32) for i in ids]
    if continuous:
        # intensities that are not 8-bit (not fp16-exact): a seeded sub-integer offset
        rng = np.random.default_rng(seed + 1)
        imgs = [(im + rng.uniform(-0.5, 0.5, im.shape)).astype(np.float32) for im in imgs]
    cams = host.scene_cameras(s, ids)
    prm = host.call_params(s.depth_min, s.depth_max, use_APD=0, use_edge=0, use_limit=0, use_label=0)
    return s, ref, imgs, cams, host.ProblemInputs(imgs, cams, prm)


def test_precision_form_cost_error_against_float64():
    """Precision form (dpe_form.precision = FP16 / BF16) against the float64 cost.

    Per-view NCC costs of perturbed ground-truth planes (dpe_ncc_cost_ex) on
    320x240 inputs with 10 source views, 8-bit and continuous intensities,
    compared with the float64 restatement tests/ncc64.py on 10^4 seeded pixels
    (entries where the float64 cost is clamped, floored or the centre projects
    outside are left out).  Reported per form: median, p99 and max absolute
    error, on textured and weak-texture pixels.  Bars: on 8-bit images, where
    the 16-bit quads are exact, both 16-bit forms are below plain at the median
    and the p99 of weak-texture pixels (the one-pass fp32 variance of the
    plain sums cancels there) and within 2x of plain's p99 on textured ones.
    On continuous images the fp16 / bf16 rounding of the texels (2^-11 /
    2^-8 relative) is the error and is reported; bars: fp16 p99 < 2e-3 on
    textured pixels, bf16 p99 < 2e-2."""
    from ncc64 import view_costs64
    ctx = host.Context(0)
    try:
        for continuous in (False, True):
            s, ref, imgs, cams, pin = _cost_inputs(0x5EED0004, continuous)
            H, W = s.height, s.width
            v = s.views[ref]
            rng = np.random.default_rng(31)
            n = v.normal @ v.R.T
            depth = v.depth * (1 + 0.005 * rng.standard_normal(v.depth.shape))
            ys, xs = np.mgrid[0:H, 0:W]
            X = np.stack([depth * (xs - v.K[0, 2]) / v.K[0, 0], depth * (ys - v.K[1, 2]) / v.K[1, 1], depth], -1)
            planes = np.concatenate([n, -np.sum(n * X, -1)[..., None]], -1).astype(np.float32)[None]
            NV = len(imgs) - 1
            ny, nx = np.mgrid[8:H - 8, 8:W - 8]
            cand = np.stack([nx.ravel(), ny.ravel()], -1)
            pix = cand[rng.choice(len(cand), 10000, replace=False)]
            pl = planes[0, pix[:, 1], pix[:, 0]].astype(np.float64)
            c64, sig = view_costs64(pix, pl[:, :3], pl[:, 3], imgs, cams)
            c64 = np.stack(c64)
            keep = np.stack([g[2] & ~g[3] & ~g[4] & ~g[5] for g in sig])
            weak = np.broadcast_to(v.weak[pix[:, 1], pix[:, 0]], keep.shape)
            stats = {}
            for name, prec in PRECISIONS.items():
                form = abi.Form(precision=prec)
                out = np.zeros((1, NV, H, W), np.float32)
                ctx.check(ctx.lib.dpe_ncc_cost_ex(ctx.h, C.byref(pin.params), C.byref(form), C.byref(pin.struct), 1,
                                                  planes.ctypes.data, out.ctypes.data))
                ms = ctx.stats().gpu_ms
                err = np.abs(out[0][:, pix[:, 1], pix[:, 0]].astype(np.float64) - c64)
                e_t, e_w = err[keep & ~weak], err[keep & weak]
                stats[name] = dict(t50=np.median(e_t), t99=np.percentile(e_t, 99), tmax=e_t.max(),
                                   w50=np.median(e_w), w99=np.percentile(e_w, 99), wmax=e_w.max())
                print(f"{'continuous' if continuous else '8-bit'} {name}: textured median {stats[name]['t50']:.2e} "
                      f"p99 {stats[name]['t99']:.2e} max {stats[name]['tmax']:.2e}; weak-texture median "
                      f"{stats[name]['w50']:.2e} p99 {stats[name]['w99']:.2e} max {stats[name]['wmax']:.2e} "
                      f"({ms:.2f} ms)")
            p = stats["plain"]
            if not continuous:
                for name in ("fp16", "bf16"):
                    q = stats[name]
                    assert q["w50"] < p["w50"] and q["w99"] < p["w99"], name
                    assert q["t99"] < 2 * p["t99"], name
            else:
                assert stats["fp16"]["t99"] < 2e-3
                assert stats["bf16"]["t99"] < 2e-2
    finally:
        ctx.close()


def test_precision_form_depth_maps():
    """The precision form through dpe_process_ex on config-1 inputs: ground-truth
    accuracy within 1 point of plain (|dd|/d < 1 % and < 0.2 %), and its call
    time next to plain's."""
    s = scene.config1()
    ref = s.meta["ref_mid"]
    ctx = host.Context(0)
    try:
        res = {}
        for name, prec in PRECISIONS.items():
            res[name] = _accuracy(ctx, s, ref, 4, abi.Form(precision=prec))
            print(f"{name}: {res[name][0]:.4f} @1% / {res[name][1]:.4f} @0.2% ({res[name][2]:.1f} ms)")
        # with the Gauss-Newton form (its steps use the plain sums on the 16-bit quads)
        res["fp16+gn2"] = _accuracy(ctx, s, ref, 4, abi.Form(gauss_newton_steps=2, precision=abi.PRECISION_FP16))
        print(f"fp16 + 2 Gauss-Newton steps: {res['fp16+gn2'][0]:.4f} @1% / {res['fp16+gn2'][1]:.4f} @0.2% "
              f"({res['fp16+gn2'][2]:.1f} ms)")
        for name in ("fp16", "bf16", "fp16+gn2"):
            assert res[name][0] >= res["plain"][0] - 0.01 and res[name][1] >= res["plain"][1] - 0.01, name
    finally:
        ctx.close()


def test_precision_form_depth_error_against_fp64_oracle():
    """Depth error of the plain path and the precision form against the FP64
    oracle (oracle_process_fp64: the same PatchMatch with the multi-view cost --
    homography, projection, bilinear sample, weights, window sums, NCC -- in
    double; states, random streams and decisions as in the fp32 oracle), on
    config-1 inputs (320x240, 4 sources, FIRST_INIT with 4 iterations).

    Reported per form: the fraction of pixels whose depth lies within 1e-4,
    1e-3 and 1e-2 (relative) of the FP64 oracle's, the p90 relative error and
    the ground-truth accuracy.  PatchMatch is a random search: an argmin or
    threshold decision flips on a cost difference of rounding size, and the
    pixel then follows another sample path, so a form departs from the FP64
    depth map on a fraction of pixels rather than by a uniform small error.
    The plain GPU path equals the fp32 oracle bit for bit (asserted), so its
    figure is the fp32 arithmetic's own.  Measured (one MI355X): plain 77.6 /
    86.6 / 91.9 % within 1e-4 / 1e-3 / 1e-2; fp16 and bf16 (the 8-bit texels
    are exact in both) 98.1 / 99.2 / 99.7 %.  Bars: each 16-bit form within
    1e-3 of the FP64 depth on >= 97 % of pixels and on 5 points more than
    plain, and within 1 point of the FP64 oracle's ground-truth accuracy."""
    import oracle_lib
    from parity_util import bitwise_fractions, depth_agreement
    s = scene.config1()
    ref = s.meta["ref_mid"]
    gt = s.views[ref].depth
    pin = host.first_call_inputs(s, ref, s.pairs[ref], max_iterations=4)
    dmin, dmax = pin.params.depth_min, pin.params.depth_max
    orc = oracle_lib.load_oracle()
    out64, _ = oracle_lib.run_oracle(orc, pin, 16, fp64=True)
    out32, _ = oracle_lib.run_oracle(orc, pin, 16)
    d64 = host.assemble(out64, dmin, dmax)["depth"].astype(np.float64)

    def report(name, d):
        d = np.asarray(d, np.float64)
        agree = {tol: depth_agreement(d, d64, tol) for tol in (1e-4, 1e-3, 1e-2)}
        with np.errstate(divide="ignore", invalid="ignore"):
            rel = np.where(d64 > 0, np.abs(d - d64) / d64, 0.0)
        acc = float((np.abs(d - gt) / gt < 0.01).mean())
        print(f"{name}: vs FP64 oracle within 1e-4 {agree[1e-4]:.4f}, 1e-3 {agree[1e-3]:.4f}, 1e-2 "
              f"{agree[1e-2]:.4f}, p90 rel {np.quantile(rel, 0.9):.2e}; ground truth within 1 % {acc:.4f}")
        return agree, acc

    _, acc64 = report("FP64 oracle", d64)
    res = {}
    ctx = host.Context(0)
    try:
        for name, prec in PRECISIONS.items():
            out = host.run_hip(ctx, pin, abi.Form(precision=prec))
            if name == "plain":
                assert bitwise_fractions(out, out32) == dict(planes=1.0, costs=1.0, pixel_states=1.0,
                                                             selected_views=1.0)
            res[name] = report(name, host.assemble(out, dmin, dmax)["depth"])
    finally:
        ctx.close()
    for name in ("fp16", "bf16"):
        assert res[name][0][1e-3] >= 0.97 and res[name][0][1e-3] >= res["plain"][0][1e-3] + 0.05, name
        assert res[name][1] >= acc64 - 0.01, name


def _same(a, b):
    """Bit-identical outputs (NaN payloads included)."""
    return all(np.array_equal(getattr(a, n).view(np.uint8), getattr(b, n).view(np.uint8))
               for n in ("planes", "costs", "pixel_states", "selected_views"))


def test_autotuned_tiles_are_bit_identical():
    """Autotuner (dpe_form.autotune = 1): the first calls of a problem class
    time the five tiles of the shared-patch NCC launches (one tile per
    half-iteration) and report tile -1; once every tile has two launches of
    each kernel timed, the fastest are kept and reported.  Every call's planes,
    costs, states and selected views equal the plain call's bit for bit."""
    s = scene.config1()
    ref = s.meta["ref_mid"]
    ctx = host.Context(0)
    try:
        pin = host.first_call_inputs(s, ref, s.pairs[ref], max_iterations=3)
        plain = host.run_hip(ctx, pin)
        seen = []
        for call in range(4):
            out = host.run_hip(ctx, pin, abi.Form(autotune=1))
            st = ctx.stats()
            seen.append((st.tile_full, st.tile_weighted))
            assert _same(out, plain), call
        print(f"tiles per call (full, weighted): {seen}")
        assert seen[0] == (-1, -1)
        assert all(0 <= t <= 4 for t in seen[-1])
        assert seen[-1] == seen[-2]  # decided once, then kept
    finally:
        ctx.close()




def test_graph_form_replays_are_bit_identical():
    """Graph form (dpe_form.graph = 1): the schedule is captured and launched as
    one hipGraph; the context updates its executable graph in place for the
    next call of the same shape.  Graph calls alternate with eager calls and
    with a problem of another size (which rebuilds the graph); every graph
    call equals the eager call of the same problem bit for bit."""
    c1 = scene.config1()
    small = scene.make_scene(160, 120, 150.0, 4, 20.0, seed=0x5EED0031)
    ref = c1.meta["ref_mid"]
    store = host.StateStore()
    ctx = host.Context(0)
    try:
        for i in range(len(c1.views)):
            pin = host.first_call_inputs(c1, i, c1.pairs[i], max_iterations=3, call_index=0)
            store.put(i, host.assemble(host.run_hip(ctx, pin), pin.params.depth_min, pin.params.depth_max))
        problems = [host.first_call_inputs(c1, ref, c1.pairs[ref], max_iterations=3),
                    host.refine_call_inputs(c1, ref, c1.pairs[ref], store, max_iterations=2, call_index=1),
                    host.first_call_inputs(small, 1, small.pairs[1], max_iterations=2)]
        eager = [host.run_hip(ctx, p) for p in problems]
        graph = abi.Form(graph=1)
        for rep in range(3):
            for k, p in enumerate(problems):
                out = host.run_hip(ctx, p, graph)
                assert ctx.stats().graph == 1
                assert _same(out, eager[k]), (rep, k)
                # eager work between the replays
                assert _same(host.run_hip(ctx, p), eager[k])
    finally:
        ctx.close()


def test_batched_form_is_bit_identical_to_single_calls():
    """Batched form (dpe_batch_process): 9 problems of three image sizes and 3-6
    source views (FIRST_INIT) plus a REFINE_ITER call with the weak path, on 3
    lanes with the autotuner and the graph form; run twice (the first batch
    samples tiles, the second uses the decided tiles and graphs).  Every output
    equals the problem's single plain call bit for bit."""
    c1 = scene.config1()
    scenes = [scene.make_scene(160, 120, 150.0, 7, 25.0, seed=0x5EED0041),
              scene.make_scene(224, 168, 210.0, 7, 25.0, seed=0x5EED0042), c1]
    ctx = host.Context(0)
    try:
        pins = []
        for si, sc in enumerate(scenes):
            for k, ref in enumerate((1, 2, 3)):
                nsrc = 3 + (si + k) % 4
                pins.append(host.first_call_inputs(sc, ref, sc.pairs[ref][:nsrc], max_iterations=2))
        store = host.StateStore()
        for i in range(len(c1.views)):
            pin = host.first_call_inputs(c1, i, c1.pairs[i], max_iterations=2, call_index=0)
            store.put(i, host.assemble(host.run_hip(ctx, pin), pin.params.depth_min, pin.params.depth_max))
        ref = c1.meta["ref_mid"]
        pins.append(host.refine_call_inputs(c1, ref, c1.pairs[ref], store, max_iterations=2, call_index=1))
        single = [host.run_hip(ctx, p) for p in pins]
    finally:
        ctx.close()
    batch = host.Batch(0, lanes=3)
    try:
        form = abi.Form(autotune=1, graph=1)
        for rep in range(2):
            outs, stats, wall = batch.run(pins, form)
            print(f"batch {rep}: {len(pins)} problems in {wall:.1f} ms; tiles "
                  f"{[(s.tile_full, s.tile_weighted) for s in stats]}; graph {[s.graph for s in stats]}")
            for k, (o, ref_out) in enumerate(zip(outs, single)):
                assert _same(o, ref_out), (rep, k)
            assert all(s.ncc_hyps > 0 for s in stats)
    finally:
        batch.close()


def test_batched_form_at_the_batch_workload_sizes():
    """The batched form at the batch workload's image sizes: one seeded
    problem each of 640x480, 1024x768, 1600x1200 and 2048x1536 with 5-10
    source views, FIRST_INIT with 8 iterations, in the combined form (fp16
    gathers, 2 Gauss-Newton steps, autotuner, graph) on 4 lanes, run twice
    (tiles sampled, then decided).  Every output equals the same problem's
    single call in the same form bit for bit, and the plain single call's
    depth agrees with ground truth about as well."""
    rng = np.random.default_rng(0x5EED00C6)
    sizes = [(640, 480), (1024, 768), (1600, 1200), (2048, 1536)]
    pins, gts = [], []
    for k, (w, h) in enumerate(sizes):
        sc = scene.make_scene(w, h, 1500.0 * w / 1600, 11, 60.0, seed=0x5EED0060 + k, n_src=10, supersample=1)
        ref = int(rng.integers(0, len(sc.views)))
        nsrc = int(rng.integers(5, 11))
        pins.append(host.first_call_inputs(sc, ref, sc.pairs[ref][:nsrc], max_iterations=8))
        gts.append(sc.views[ref].depth)
        print(f"problem {k}: {w}x{h}, reference {ref}, {nsrc} sources", flush=True)
    form = abi.Form(precision=abi.PRECISION_FP16, gauss_newton_steps=2, autotune=1, graph=1)
    ctx = host.Context(0)
    try:
        single = [host.run_hip(ctx, p, form) for p in pins]
        plain = [host.run_hip(ctx, p) for p in pins]
    finally:
        ctx.close()
    for k, p in enumerate(pins):
        dmin, dmax = p.params.depth_min, p.params.depth_max
        acc = [float((np.abs(host.assemble(o, dmin, dmax)["depth"] - gts[k]) / gts[k] < 0.01).mean())
               for o in (single[k], plain[k])]
        print(f"problem {k}: within 1 % of ground truth: combined form {acc[0]:.4f}, plain {acc[1]:.4f}")
        assert acc[0] > acc[1] - 0.05
    batch = host.Batch(0, lanes=4)
    try:
        for rep in range(2):
            outs, stats, wall = batch.run(pins, form)
            print(f"batch {rep}: {len(pins)} problems in {wall:.1f} ms; tiles "
                  f"{[(s.tile_full, s.tile_weighted) for s in stats]}; graph {[s.graph for s in stats]}")
            for k, (o, ref_out) in enumerate(zip(outs, single)):
                assert _same(o, ref_out), (rep, k)
    finally:
        batch.close()
