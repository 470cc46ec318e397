"""Benchmark of the MI355X PatchMatch + fusion path on the 1600x1200, 10-source workload.

One step = the whole job on the synthetic scene: 20 reference views of
1600x1200 with their 10 nearest sources each; per view a FIRST_INIT call with 8
checkerboard iterations, then per view a REFINE_ITER call with geometric
consistency, the weak-texture path and edge-adaptive sampling (2 iterations),
then fusion into a point cloud (1 % relative depth, >= 3 consistent views).
Images, edge/label maps and all state stay in HBM; the timed region is K steps
bracketed by device synchronisation.

Prints one JSON line: value = NCC plane-hypotheses per second of the whole job
(a plane-hypothesis = one (pixel, plane, source view) bilateral NCC window),
plus seconds per depth map, accuracy against ground truth, the roofline of the
NCC kernels and the CPU oracle baseline on a bounded sample.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dpe-mvs-fast_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# VALU issue peak of MI355X without packed fp32: the published 157.3 TFLOP/s fp32
# vector peak counts packed FMAs (2 fp32 FMAs per lane per instruction, 2 flops
# each); one unpacked wave64 instruction is 64 lane operations, so the issue
# peak is 157.3e12 / 2 / 2 / 64 = 614.5e9 wave instructions per second (256 CUs x
# 4 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction).
VALU_ISSUE_PEAK = 157.3e12 / 2 / 2 / 64
# the WEAK family is the weak half-iteration: k_weak_patch, k_weak_cand, k_weak_select,
# k_weak_terms (twice), k_weak_mid, k_weak_final
FAMILY_NAMES = ["k_ncc_full<6>", "k_ncc_weighted<6,REFINE>", "k_ncc_weighted<6,POST>", "weak half-iteration"]
# HBM traffic of one k_ncc_full<6> launch on this workload, from two rocprofv3
# --pmc passes (FETCH_SIZE, WRITE_SIZE) over `python bench.py --no-cpu-baseline`,
# each scaled by the counter's reading for a 256 MiB device copy
# (`python bench.py --copy-calibration`: FETCH_SIZE x2.000, WRITE_SIZE x1.000);
# the runs and the arithmetic are in profiles/ncc_kernels_1600x1200_20views_hbm_traffic.md.
# Counters cannot be read inside a timed run.
NCC_FULL_TRAFFIC_BYTES = 6.8665e9
# VALU wave instructions of one k_ncc_full<6> launch on this workload (SQ_INSTS_VALU
# summed over the step's launches / launches, one rocprofv3 --pmc pass over
# `python bench.py --no-cpu-baseline`; profiles/ncc_kernels_1600x1200_20views_valu_issue.md).
NCC_FULL_VALU_INSTS = 1.3723e9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--views", type=int, default=20)
    ap.add_argument("--iters", type=int, default=8)
    ap.add_argument("--geom-iters", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", choices=["plain", "fp16", "bf16"], default="plain",
                    help="dpe_form.precision of every call (plain = the reference's fp32 arithmetic)")
    ap.add_argument("--gauss-newton", type=int, default=0, help="dpe_form.gauss_newton_steps of every call")
    ap.add_argument("--workload", choices=["c2", "batch", "fusion"], default="c2",
                    help="c2: the 1600x1200 20-view job (default); batch: 32 independent FIRST_INIT problems of "
                         "uneven size and source count, sequential plain calls vs the batched form; fusion: the "
                         "three fusions (RunFusion, TAT intermediate, TAT advanced) on 20 resident 1600x1200 "
                         "depth maps")
    ap.add_argument("--lanes", type=int, default=4, help="lanes of the batched form (--workload batch)")
    ap.add_argument("--io", choices=["device", "host"], default="device",
                    help="device: inputs, state and outputs resident in HBM (the measured value); host: every "
                         "call takes and returns host buffers (the PCIe-inclusive rate, reported beside it)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device-to-device 3 times (the known-size kernel that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


def copy_calibration():
    import torch
    n = 256 * 1024 * 1024 // 4
    a = torch.ones(n, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {n * 4} bytes 3 times", flush=True)


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        batch_bench(args)
        return
    if args.workload == "fusion":
        fusion_bench(args)
        return
    if args.io == "host":
        host_io_bench(args)
        return
    import numpy as np
    import torch
    from dpe_amd import abi, host, scene, support

    dev = torch.device("cuda:0")
    t0 = time.time()
    sc = scene.make_scene(1600, 1200, 1500.0, args.views, 60.0, seed=0x5EED0001, n_src=10, supersample=1)
    print(f"[bench] scene ready in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    V, H, W = args.views, sc.height, sc.width
    P = H * W
    # ---- resident inputs (HBM)
    imgs = [torch.from_numpy(v.image.astype(np.float32)).to(dev) for v in sc.views]
    edges, labels = [], []
    for v in sc.views:
        e, l = support.support_maps(v.image)
        edges.append(torch.from_numpy(e).to(dev))
        labels.append(torch.from_numpy(l).to(dev))
    colors = [torch.from_numpy(np.repeat(v.image[..., None], 3, -1)).to(dev) for v in sc.views]
    cams = host.scene_cameras(sc, list(range(V)))
    # ---- per-view state (HBM)
    planes = [torch.zeros((H, W, 4), dtype=torch.float32, device=dev) for _ in range(V)]
    costs = [torch.zeros((H, W), dtype=torch.float32, device=dev) for _ in range(V)]
    states = [torch.zeros((H, W), dtype=torch.uint8, device=dev) for _ in range(V)]
    selected = [torch.zeros((H, W), dtype=torch.int32, device=dev) for _ in range(V)]
    depth = [torch.zeros((H, W), dtype=torch.float32, device=dev) for _ in range(V)]
    normal = [torch.zeros((H, W, 3), dtype=torch.float32, device=dev) for _ in range(V)]
    prior = [torch.zeros((H, W, 4), dtype=torch.float32, device=dev) for _ in range(V)]
    torch.cuda.synchronize()
    ctx = host.Context(0)
    lib = ctx.lib
    print(f"[bench] inputs resident after {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    def ptr(t):
        return t.data_ptr()

    form = abi.Form(gauss_newton_steps=args.gauss_newton,
                    precision={"plain": abi.PRECISION_FP32, "fp16": abi.PRECISION_FP16,
                               "bf16": abi.PRECISION_BF16}[args.precision])

    def assemble(planes_v, depth_v, states_v, normal_v, dmin, dmax):
        """Result assembly (main.cpp:264-280) on the device."""
        d = planes_v[..., 3]
        bad = (d < dmin) | (d > dmax)
        depth_v.copy_(torch.where(bad, torch.zeros_like(d), d))
        states_v.masked_fill_(bad, abi.UNKNOWN)
        normal_v.copy_(planes_v[..., :3])

    def call(v, first):
        srcs = sc.pairs[v]
        ids = [v] + srcs
        if first:
            prm = host.call_params(sc.depth_min, sc.depth_max, max_iterations=args.iters, state=abi.FIRST_INIT,
                                   use_APD=0, use_edge=0, geom_consistency=0, weak_peak_radius=6, call_index=0)
        else:
            prm = host.call_params(sc.depth_min, sc.depth_max, max_iterations=args.geom_iters,
                                   state=abi.REFINE_ITER, use_APD=1, use_edge=1, geom_consistency=1,
                                   weak_peak_radius=4, call_index=1, ransac_threshold=0.01 - 0.00125, rotate_time=2)
        prm.num_images = len(ids)
        s = abi.Inputs()
        s.width, s.height, s.num_images, s.memory = W, H, len(ids), abi.MEM_DEVICE
        for k, i in enumerate(ids):
            s.images[k] = ptr(imgs[i])
            s.cameras[k] = cams[i]
            if not first:
                s.geom_depths[k] = ptr(depth[i])
        if not first:
            torch.cat([normal[v], depth[v][..., None]], -1, out=prior[v])
            s.weak_info = ptr(states[v])
            s.prior_planes = ptr(prior[v])
            s.prior_selected = ptr(selected[v])
        s.edges = ptr(edges[v])
        s.edges_low = ptr(edges[v])
        s.low_width, s.low_height = W, H
        s.labels = ptr(labels[v])
        out = abi.Outputs(ptr(planes[v]), ptr(costs[v]), ptr(states[v]), ptr(selected[v]))
        torch.cuda.synchronize()
        ctx.check(lib.dpe_process_ex(ctx.h, C.byref(prm), C.byref(form), C.byref(s), C.byref(out)))
        st = ctx.stats()
        assemble(planes[v], depth[v], states[v], normal[v], prm.depth_min, prm.depth_max)
        return st

    # torch loads the code object of each of its kernels at the kernel's first
    # launch (20-90 ms apiece here): run the assembly's and the prior's torch
    # operations once on scratch tensors of the same shapes before timing
    scratch = [torch.zeros_like(t) for t in (planes[0], depth[0], states[0], normal[0], prior[0])]
    assemble(*scratch[:4], 1.0, 2.0)
    torch.cat([scratch[3], scratch[1][..., None]], -1, out=scratch[4])
    del scratch
    torch.cuda.synchronize()

    fviews = (abi.FuseView * V)()
    for i in range(V):
        fv = fviews[i]
        fv.width, fv.height = W, H
        fv.camera = cams[i]
        fv.depth, fv.normal, fv.pixel_states, fv.color = ptr(depth[i]), ptr(normal[i]), ptr(states[i]), ptr(colors[i])
        fv.num_src = len(sc.pairs[i])
        for j, s_ in enumerate(sc.pairs[i]):
            fv.src_index[j] = s_
    # the precision form also takes fusion's reprojection geometry to FP64
    fparams = abi.fuse_defaults(min_consistent=3, memory=abi.MEM_DEVICE,
                                fp64_reprojection=int(args.precision != "plain"))

    def step():
        tot = dict(hyps=0, taps=0, gpu_ms=0.0, fam_ms=[0.0] * 4, fam_hyps=[0] * 4, fam_taps=[0] * 4,
                   fam_launches=[0] * 4)
        for first in (True, False):
            for v in range(V):
                st = call(v, first)
                tot["hyps"] += st.ncc_hyps
                tot["taps"] += st.ncc_taps
                tot["gpu_ms"] += st.gpu_ms
                tot["quad_bytes"] = st.quad_bytes
                for f in range(4):
                    tot["fam_ms"][f] += st.family_ms[f]
                    tot["fam_hyps"][f] += st.family_hyps[f]
                    tot["fam_taps"][f] += st.family_taps[f]
                    tot["fam_launches"][f] += st.family_launches[f]
            print(f"[bench] {'first' if first else 'geometric'} calls done", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        n = C.c_int64()
        ctx.check(lib.dpe_fuse(ctx.h, fviews, V, C.byref(fparams), None, 0, C.byref(n)))
        tot["points"] = n.value
        return tot

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    totals = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t1
    ms_per_step = elapsed * 1e3 / args.steps
    hyps = sum(t["hyps"] for t in totals)
    taps = sum(t["taps"] for t in totals)
    fam = {k: [sum(t[k][f] for t in totals) for f in range(4)]
           for k in ("fam_ms", "fam_hyps", "fam_taps", "fam_launches")}
    # accuracy of the final depth maps against ground truth, overall and by
    # what the ground truth says each pixel can be resolved from
    acc = []
    for v in range(V):
        gt = torch.from_numpy(sc.views[v].depth).to(dev)
        rel = (depth[v] - gt).abs() / gt
        acc.append(float((rel < 0.01).float().mean()))
    breakdown = accuracy_breakdown(sc, depth, dev)
    # roofline of the dominant kernel, k_ncc_full<6> (candidate cost vectors):
    # algorithmic bytes per launch / its average launch time, both from the
    # hipEvent pairs and exact tap counters of the timed steps.  Algorithmic
    # bytes per tap: one reference sample and one bilinear source quad at the
    # gather precision, 20 B with fp32 quads (4 + 16), 10 B with 16-bit quads
    # (2 + 8), which the plain path uses on 8-bit images (same samples).
    bytes_per_tap = 20 if totals[-1]["quad_bytes"] == 16 else 10
    full_launches = max(fam["fam_launches"][0], 1)
    full_avg_ms = fam["fam_ms"][0] / full_launches
    full_bytes = fam["fam_taps"][0] * bytes_per_tap / full_launches
    achieved = full_bytes / (full_avg_ms * 1e-3) / 1e9
    # what binds it: the counted DRAM traffic is a quarter of the HBM peak, the
    # VALU issue rate is near its peak
    dram_frac = NCC_FULL_TRAFFIC_BYTES / (full_avg_ms * 1e-3) / (HBM_PEAK_GBS * 1e9)
    valu_frac = (None if NCC_FULL_VALU_INSTS is None
                 else NCC_FULL_VALU_INSTS / (full_avg_ms * 1e-3) / VALU_ISSUE_PEAK)
    families = {FAMILY_NAMES[f]: {"ms": fam["fam_ms"][f], "launches": fam["fam_launches"][f],
                                  "hyps": fam["fam_hyps"][f], "taps": fam["fam_taps"][f],
                                  "hyp_per_s": fam["fam_hyps"][f] / max(fam["fam_ms"][f] * 1e-3, 1e-12)}
                for f in range(4)}
    result = {
        "metric": "NCC plane-hyp/s + s/depth-map @1600x1200,10 src; %px |Δd|/d<1% vs plain",
        "value": hyps / elapsed,
        "unit": "NCC plane-hyp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded 8-plane scene, value-noise texture, 20% weak texture)",
        "config": {"workload": f"{V} reference views 1600x1200, 10 sources each: FIRST_INIT {args.iters} iterations, "
                               f"REFINE_ITER+geometric consistency {args.geom_iters} iterations, fusion "
                               f"(1% depth, >=3 consistent views)",
                   "views": V, "width": W, "height": H, "sources": 10},
        "s_per_depth_map": elapsed / args.steps / V,
        "accuracy_px_lt_1pct_vs_gt": float(np.mean(acc)),
        "accuracy_breakdown": breakdown,
        "fused_points": totals[-1]["points"],
        "ncc_hyps_per_step": hyps // args.steps,
        "ncc_taps_per_step": taps // args.steps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": NCC_FULL_TRAFFIC_BYTES,
                     "kernel": "k_ncc_full<6>", "launch_avg_ms": full_avg_ms, "bytes_per_launch": full_bytes,
                     "bytes_per_tap": bytes_per_tap, "launches": fam["fam_launches"][0],
                     "dram_frac": dram_frac, "binding": "valu_issue", "valu_frac": valu_frac,
                     "valu_insts_per_launch": NCC_FULL_VALU_INSTS,
                     "valu_issue_peak_per_s": VALU_ISSUE_PEAK},
        "form": {"precision": args.precision, "gauss_newton_steps": args.gauss_newton},
        "ncc_families": families,
    }
    if not args.no_cpu_baseline:
        def state(i):
            return (depth[i].cpu().numpy(), normal[i].cpu().numpy(), states[i].cpu().numpy(),
                    selected[i].cpu().numpy().view(np.uint32))
        result["cpu_baseline"] = cpu_baseline(sc, args.cpu_threads, state)
    print(json.dumps(result), flush=True)
    ctx.close()


def accuracy_breakdown(sc, depth, dev):
    """Where the pixels that miss the 1 % ground-truth bar are.  Every pixel of
    every reference view falls in one class, in this order: the 6-px border
    that DepthToWeak marks UNKNOWN (DPE.cu:2605-2611); seen by fewer than 2 of
    its sources (out of view or occluded: its ground-truth point projects
    outside the source or lands over 1 % in front of the source's ground-truth
    depth at the nearest pixel); on a weak-texture patch of the scene; textured
    interior.  Per class: share of all pixels and share within 1 %.
    "resolvable" = the textured interior seen by >= 2 sources."""
    import numpy as np
    import torch
    classes = ["border", "seen_by_lt2_sources", "weak_texture", "textured_interior"]
    tot = {c: [0, 0] for c in classes}
    for v, view in enumerate(sc.views):
        H, W = view.depth.shape
        gt = torch.from_numpy(view.depth.astype(np.float64)).to(dev)
        K = torch.from_numpy(view.K.astype(np.float64)).to(dev)
        R = torch.from_numpy(view.R.astype(np.float64)).to(dev)
        t = torch.from_numpy(view.t.astype(np.float64)).to(dev)
        ys, xs = torch.meshgrid(torch.arange(H, device=dev, dtype=torch.float64),
                                torch.arange(W, device=dev, dtype=torch.float64), indexing="ij")
        Xc = torch.stack([(xs - K[0, 2]) / K[0, 0] * gt, (ys - K[1, 2]) / K[1, 1] * gt, gt], -1)
        Xw = (Xc - t) @ R  # R^T (Xc - t)
        seen = torch.zeros((H, W), dtype=torch.int32, device=dev)
        for j in sc.pairs[v]:
            sv = sc.views[j]
            Ks = torch.from_numpy(sv.K.astype(np.float64)).to(dev)
            Rs = torch.from_numpy(sv.R.astype(np.float64)).to(dev)
            ts = torch.from_numpy(sv.t.astype(np.float64)).to(dev)
            Xs = Xw @ Rs.T + ts
            z = Xs[..., 2]
            u = Ks[0, 0] * Xs[..., 0] / z + Ks[0, 2]
            w = Ks[1, 1] * Xs[..., 1] / z + Ks[1, 2]
            inb = (z > 0) & (u >= 0) & (u <= W - 1) & (w >= 0) & (w <= H - 1)
            ui = u.round().clamp(0, W - 1).long()
            wi = w.round().clamp(0, H - 1).long()
            gs = torch.from_numpy(sv.depth.astype(np.float64)).to(dev)[wi, ui]
            seen += (inb & ((z - gs).abs() / z < 0.01)).int()
        good = ((depth[v].double() - gt).abs() / gt < 0.01)
        border = torch.zeros((H, W), dtype=torch.bool, device=dev)
        border[:6, :] = border[-6:, :] = True
        border[:, :6] = border[:, -6:] = True
        weak = torch.from_numpy(view.weak).to(dev)
        rest = ~border
        masks = {"border": border, "seen_by_lt2_sources": rest & (seen < 2)}
        rest = rest & (seen >= 2)
        masks["weak_texture"] = rest & weak
        masks["textured_interior"] = rest & ~weak
        for c in classes:
            tot[c][0] += int(masks[c].sum())
            tot[c][1] += int((masks[c] & good).sum())
    n = sum(a for a, _ in tot.values())
    out = {c: {"share_of_pixels": a / n, "within_1pct": b / max(a, 1)} for c, (a, b) in tot.items()}
    out["resolvable_within_1pct"] = out["textured_interior"]["within_1pct"]
    return out


def host_io_bench(args):
    """The 1600x1200 job with host buffers at the boundary: every call uploads
    its images, support maps, prior state and geometric depths and downloads
    its outputs (dpe_inputs.memory = DPE_MEM_HOST), the result assembly and the
    state hand-off run on the host, and fusion reads host views.  Reports the
    PCIe-inclusive NCC plane-hyp/s of the whole job."""
    import numpy as np
    from dpe_amd import abi, host, scene, support
    t0 = time.time()
    sc = scene.make_scene(1600, 1200, 1500.0, args.views, 60.0, seed=0x5EED0001, n_src=10, supersample=1)
    print(f"[bench] scene ready in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    V, H, W = args.views, sc.height, sc.width
    imgs = [np.ascontiguousarray(v.image, np.float32) for v in sc.views]
    sup = [support.support_maps(v.image) for v in sc.views]
    cams = host.scene_cameras(sc, list(range(V)))
    depth = [np.zeros((H, W), np.float32) for _ in range(V)]
    normal = [np.zeros((H, W, 3), np.float32) for _ in range(V)]
    states = [np.zeros((H, W), np.uint8) for _ in range(V)]
    selected = [np.zeros((H, W), np.uint32) for _ in range(V)]
    ctx = host.Context(0)

    def call(v, first):
        ids = [v] + sc.pairs[v]
        if first:
            prm = host.call_params(sc.depth_min, sc.depth_max, max_iterations=args.iters, state=abi.FIRST_INIT,
                                   use_APD=0, use_edge=0, geom_consistency=0, weak_peak_radius=6, call_index=0)
            pin = host.ProblemInputs([imgs[i] for i in ids], [cams[i] for i in ids], prm, edges=sup[v][0],
                                     edges_low=sup[v][0], labels=sup[v][1])
        else:
            prm = host.call_params(sc.depth_min, sc.depth_max, max_iterations=args.geom_iters,
                                   state=abi.REFINE_ITER, use_APD=1, use_edge=1, geom_consistency=1,
                                   weak_peak_radius=4, call_index=1, ransac_threshold=0.01 - 0.00125, rotate_time=2)
            prior = np.concatenate([normal[v], depth[v][..., None]], -1)
            pin = host.ProblemInputs([imgs[i] for i in ids], [cams[i] for i in ids], prm,
                                     geom_depths=[depth[i] for i in ids], weak_info=states[v], prior_planes=prior,
                                     prior_selected=selected[v], edges=sup[v][0], edges_low=sup[v][0],
                                     labels=sup[v][1])
        out = host.run_hip(ctx, pin)
        r = host.assemble(out, prm.depth_min, prm.depth_max)
        depth[v], normal[v], states[v], selected[v] = r["depth"], r["normal"], r["pixel_states"], out.selected_views
        return ctx.stats().ncc_hyps

    def step():
        hyps = 0
        for first in (True, False):
            for v in range(V):
                hyps += call(v, first)
        from fuse_util import FuseSet
        fs = FuseSet(sc, depth, normal, states)
        n = C.c_int64()
        ctx.check(ctx.lib.dpe_fuse(ctx.h, fs.views, fs.n, C.byref(abi.fuse_defaults(min_consistent=3)), None, 0,
                                   C.byref(n)))
        return hyps, n.value

    for _ in range(args.warmup):
        step()
    t1 = time.perf_counter()
    res = [step() for _ in range(args.steps)]
    elapsed = time.perf_counter() - t1
    hyps = sum(h for h, _ in res)
    acc = float(np.mean([((np.abs(depth[v] - sc.views[v].depth) / sc.views[v].depth) < 0.01).mean()
                         for v in range(V)]))
    ctx.close()
    print(json.dumps({
        "metric": "NCC plane-hyp/s + s/depth-map @1600x1200,10 src, host buffers at the boundary (PCIe-inclusive)",
        "value": hyps / elapsed, "unit": "NCC plane-hyp/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic (seeded 8-plane scene, value-noise texture, 20% weak texture)",
        "config": {"workload": f"{V} reference views 1600x1200, 10 sources each, host I/O", "views": V},
        "s_per_depth_map": elapsed / args.steps / V, "accuracy_px_lt_1pct_vs_gt": acc,
        "fused_points": res[-1][1]}), flush=True)


def batch_bench(args):
    """32 FIRST_INIT problems (8 checkerboard iterations) with sizes drawn
    uniformly from 640x480, 1024x768, 1600x1200 and 2048x1536 and 5-10 source
    views each (seeded), inputs and outputs resident in HBM.  Timed: the
    problems as sequential plain calls on one context, then as one
    dpe_batch_process call (autotuner + graph form, `--lanes` lanes) after an
    untimed batch that lets the autotuner decide.  The outputs of both are
    compared bit for bit."""
    import numpy as np
    import torch
    from dpe_amd import abi, host, scene
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(0x5EED00C5)
    sizes = [(640, 480), (1024, 768), (1600, 1200), (2048, 1536)]
    t0 = time.time()
    scenes = [scene.make_scene(w, h, 1500.0 * w / 1600, 12, 60.0, seed=0x5EED0050 + k, n_src=10, supersample=1)
              for k, (w, h) in enumerate(sizes)]
    print(f"[bench] scenes ready in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    keep = []

    def on_device(a):
        t = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        keep.append(t)
        return t.data_ptr()

    pins, structs, outs, meta = [], [], [], []
    for i in range(32):
        k = int(rng.integers(0, 4))
        sc = scenes[k]
        ref = int(rng.integers(0, len(sc.views)))
        nsrc = int(rng.integers(5, 11))
        pin = host.first_call_inputs(sc, ref, sc.pairs[ref][:nsrc], max_iterations=8)
        s = abi.Inputs()
        C.memmove(C.addressof(s), C.addressof(pin.struct), C.sizeof(abi.Inputs))
        s.memory = abi.MEM_DEVICE
        for j, im in enumerate(pin.images):
            s.images[j] = on_device(im)
        s.edges = on_device(pin.edges)
        s.edges_low = on_device(pin.edges_low)
        s.labels = on_device(pin.labels)
        W, H = pin.width, pin.height
        o = [torch.zeros((H, W, 4), dtype=torch.float32, device=dev), torch.zeros((H, W), dtype=torch.float32, device=dev),
             torch.zeros((H, W), dtype=torch.uint8, device=dev), torch.zeros((H, W), dtype=torch.int32, device=dev)]
        pins.append(pin)
        structs.append(s)
        outs.append(o)
        meta.append((k, ref, nsrc))
    torch.cuda.synchronize()

    def outputs_struct(o):
        return abi.Outputs(o[0].data_ptr(), o[1].data_ptr(), o[2].data_ptr(), o[3].data_ptr())

    ctx = host.Context(0)
    # sequential plain calls (one warm-up call first: the device pool grows to the largest problem)
    ctx.check(ctx.lib.dpe_process(ctx.h, C.byref(pins[0].params), C.byref(structs[0]),
                                  C.byref(outputs_struct(outs[0]))))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    hyps_seq = 0
    for i in range(32):
        ctx.check(ctx.lib.dpe_process(ctx.h, C.byref(pins[i].params), C.byref(structs[i]),
                                      C.byref(outputs_struct(outs[i]))))
        hyps_seq += ctx.stats().ncc_hyps
    torch.cuda.synchronize()
    seq_s = time.perf_counter() - t1
    ref_out = [[t.clone() for t in o] for o in outs]
    ctx.close()
    print(f"[bench] sequential: {seq_s:.2f}s", file=sys.stderr, flush=True)
    # batched form
    b = host.Batch(0, lanes=args.lanes)
    form = abi.Form(autotune=1, graph=1)
    params = (abi.Params * 32)(*[p.params for p in pins])
    inputs = (abi.Inputs * 32)(*structs)
    outputs = (abi.Outputs * 32)(*[outputs_struct(o) for o in outs])
    stats = (abi.Stats * 32)()
    wall = C.c_float()
    for rep in range(2):
        for o in outs:
            for t in o:
                t.zero_()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        st = b.lib.dpe_batch_process(b.h, 32, params, C.byref(form), inputs, outputs, stats, C.byref(wall))
        torch.cuda.synchronize()
        bat_s = time.perf_counter() - t2
        if st != abi.OK:
            raise host.DPEError(st, b.lib.dpe_batch_last_error(b.h).decode())
        print(f"[bench] batch {rep}: {bat_s:.2f}s", file=sys.stderr, flush=True)
    hyps_bat = sum(stats[i].ncc_hyps for i in range(32))
    tiles = sorted({(stats[i].tile_full, stats[i].tile_weighted) for i in range(32)})
    b.close()
    same = all(torch.equal(a.view(torch.uint8), r.view(torch.uint8)) for o, ro in zip(outs, ref_out)
               for a, r in zip(o, ro))
    acc = []
    for (k, ref, nsrc), o in zip(meta, outs):
        gt = torch.from_numpy(scenes[k].views[ref].depth).to(dev)
        d = o[0][..., 3]
        acc.append(float(((d - gt).abs() / gt < 0.01).float().mean()))
    result = {
        "metric": "NCC plane-hyp/s + s/depth-map, batched form vs sequential plain calls",
        "value": hyps_bat / bat_s,
        "unit": "NCC plane-hyp/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": bat_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded 8-plane scenes, one per size)",
        "config": {"workload": "32 FIRST_INIT problems (8 iterations), sizes uniform in 640x480/1024x768/"
                               "1600x1200/2048x1536, 5-10 source views", "lanes": args.lanes,
                   "problems": [{"size": f"{sizes[k][0]}x{sizes[k][1]}", "sources": n} for k, _, n in meta]},
        "depth_maps_per_s": {"sequential": 32 / seq_s, "batched": 32 / bat_s},
        "hyp_per_s": {"sequential": hyps_seq / seq_s, "batched": hyps_bat / bat_s},
        "speedup": seq_s / bat_s,
        "bit_identical": same,
        "tiles": tiles,
        "accuracy_px_lt_1pct_vs_gt": float(np.mean(acc)),
    }
    print(json.dumps(result), flush=True)


def fusion_bench(args):
    """The fusions alone: 20 reference views of 1600x1200 with 10 sources each,
    ground-truth depth with 0.05 % noise and 5 % holes, ground-truth normals and
    pixel states, colours; everything resident in HBM, points written to HBM.
    Timed per fusion: `--steps` calls after `--warmup` (at least one) untimed
    ones.  CPU baseline: the oracle's sequential restatement fusing reference
    view 0 against its 10 sources (the other views keep their data but list no
    sources), one thread, as the reference runs it."""
    import numpy as np
    import torch
    import oracle_lib
    from dpe_amd import abi, host, scene
    from fuse_util import FuseSet
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(0x5EED00F5)
    t0 = time.time()
    sc = scene.make_scene(1600, 1200, 1500.0, args.views, 60.0, seed=0x5EED0001, n_src=10, supersample=1)
    V, H, W = args.views, sc.height, sc.width
    depths = []
    for v in sc.views:
        d = (v.depth * (1 + 0.0005 * rng.standard_normal(v.depth.shape))).astype(np.float32)
        d[rng.uniform(size=d.shape) < 0.05] = 0.0
        depths.append(d)
    states = [np.where(v.weak, abi.WEAK, abi.STRONG).astype(np.uint8) for v in sc.views]
    fs = FuseSet(sc, depths, [v.normal for v in sc.views], states)
    print(f"[bench] fusion inputs ready in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    keep = []
    views = (abi.FuseView * V)()
    for i in range(V):
        C.memmove(C.addressof(views[i]), C.addressof(fs.views[i]), C.sizeof(abi.FuseView))
        P = W * H
        for name, ctype, k in (("depth", C.c_float, 1), ("normal", C.c_float, 3), ("pixel_states", C.c_uint8, 1),
                               ("color", C.c_uint8, 3)):
            a = np.ctypeslib.as_array(C.cast(getattr(fs.views[i], name), C.POINTER(ctype)), (P * k,))
            t = torch.from_numpy(a.copy()).to(dev)
            keep.append(t)
            setattr(views[i], name, t.data_ptr())
    points = torch.empty((V * H * W, 4), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    ctx = host.Context(0)
    lib = ctx.lib
    n = C.c_int64()
    fp = abi.fuse_defaults(min_consistent=3, memory=abi.MEM_DEVICE)
    runs = {
        "RunFusion (min_consistent 3)": lambda: lib.dpe_fuse(ctx.h, views, V, C.byref(fp), points.data_ptr(),
                                                            V * H * W, C.byref(n)),
    }
    for name, variant in (("TAT intermediate", abi.TAT_INTERMEDIATE), ("TAT advanced", abi.TAT_ADVANCED)):
        tp = abi.fuse_tat_defaults(variant, memory=abi.MEM_DEVICE)
        runs[name] = (lambda tp=tp: lib.dpe_fuse_tat(ctx.h, views, V, None, C.byref(tp), points.data_ptr(),
                                                     V * H * W, C.byref(n)))
    res = {}
    for name, fn in runs.items():
        for _ in range(max(args.warmup, 1)):
            ctx.check(fn())
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            ctx.check(fn())
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t1) / args.steps
        res[name] = {"ms": dt * 1e3, "ref_pixels_per_s": V * H * W / dt, "points": n.value}
        print(f"[bench] {name}: {dt * 1e3:.1f} ms, {n.value} points", file=sys.stderr, flush=True)
    ctx.close()
    cpu = None
    if not args.no_cpu_baseline:
        orc = oracle_lib.load_oracle()
        from fuse_util import fuse_oracle, fuse_tat_oracle
        for i in range(1, V):
            fs.views[i].num_src = 0
        cpu = {}
        t2 = time.perf_counter()
        fuse_oracle(orc, fs, abi.fuse_defaults(min_consistent=3))
        cpu["RunFusion (min_consistent 3)"] = H * W / (time.perf_counter() - t2)
        t2 = time.perf_counter()
        fuse_tat_oracle(orc, fs, abi.fuse_tat_defaults(abi.TAT_INTERMEDIATE))
        cpu["TAT intermediate"] = H * W / (time.perf_counter() - t2)
        cpu = {"value": cpu["TAT intermediate"], "unit": "reference pixels/s", "cores": 1, "kind": "port",
               "sample": "oracle fusing reference view 0 (1600x1200) against its 10 sources; per fusion: " +
                         ", ".join(f"{k} {v:.3g} px/s" for k, v in cpu.items())}
    result = {
        "metric": "reference pixels fused per second (TAT intermediate)",
        "value": res["TAT intermediate"]["ref_pixels_per_s"],
        "unit": "reference pixels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(args.warmup, 1),
        "ms_per_step": res["TAT intermediate"]["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded 8-plane scene; ground-truth depth + 0.05 % noise, 5 % holes)",
        "config": {"workload": f"fusion of {V} depth maps of {W}x{H}, 10 sources each"},
        "fusions": res,
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sc, threads, state):
    """The CPU oracle on a bounded sample of the same workload: both process
    calls of one reference view (FIRST_INIT with 8 iterations, then REFINE_ITER
    with geometric consistency, the weak path and edge sampling, 2 iterations)
    on the 800x600 centre crop of reference view 10 and its 10 sources
    (principal points shifted), timed with `threads` host threads; the REFINE
    call reads the GPU's final states of the same crop (prior plane, pixel
    states, selected views, the 11 views' depths).  s_per_depth_map scales the
    two calls' time by the full view's pixel count (x4).  Beside it, one
    FIRST_INIT call with 2 iterations on a 240x180 crop with one thread.
    `state(i)` returns view i's (depth, normal, pixel states, selected views)."""
    import numpy as np
    import oracle_lib
    from dpe_amd import abi, host, support
    nthreads = threads or min(16, os.cpu_count() or 1)
    ref = min(10, len(sc.views) - 1)
    ids = [ref] + sc.pairs[ref]
    lib = oracle_lib.load_oracle()

    def crop_inputs(cw, ch):
        x0, y0 = (sc.width - cw) // 2, (sc.height - ch) // 2
        imgs, cams = [], []
        for i in ids:
            v = sc.views[i]
            imgs.append(v.image[y0:y0 + ch, x0:x0 + cw].astype(np.float32))
            K = v.K.astype(np.float32).copy()
            K[0, 2] -= x0
            K[1, 2] -= y0
            cams.append(abi.camera_from_arrays(K, v.R, v.t, cw, ch, sc.depth_min, sc.depth_max))
        return x0, y0, imgs, cams

    def timed(pin, nt):
        fam = []
        t = time.perf_counter()
        oracle_lib.run_oracle(lib, pin, nt, fam)
        dt = time.perf_counter() - t
        # the GPU's unit: every family but the evaluations the reference never reads
        return sum(h for h, _ in fam[:abi.NCC_FAMILIES]), dt

    cw, ch = 800, 600
    x0, y0, imgs, cams = crop_inputs(cw, ch)
    edges, labels = support.support_maps(sc.views[ref].image[y0:y0 + ch, x0:x0 + cw])
    prm = host.call_params(sc.depth_min, sc.depth_max, max_iterations=8, state=abi.FIRST_INIT, use_APD=0,
                           use_edge=0, geom_consistency=0, weak_peak_radius=6, call_index=0)
    h1, t1 = timed(host.ProblemInputs(imgs, cams, prm, edges=edges, edges_low=edges, labels=labels), nthreads)
    crop = (slice(y0, y0 + ch), slice(x0, x0 + cw))
    st = {i: state(i) for i in ids}
    depth, normal, states, selected = st[ref]
    prm = host.call_params(sc.depth_min, sc.depth_max, max_iterations=2, state=abi.REFINE_ITER, use_APD=1,
                           use_edge=1, geom_consistency=1, weak_peak_radius=4, call_index=1,
                           ransac_threshold=0.01 - 0.00125, rotate_time=2)
    prior = np.concatenate([normal[crop], depth[crop][..., None]], -1)
    pin = host.ProblemInputs(imgs, cams, prm, geom_depths=[st[i][0][crop] for i in ids], weak_info=states[crop],
                             prior_planes=prior, prior_selected=selected[crop], edges=edges, edges_low=edges,
                             labels=labels)
    h2, t2 = timed(pin, nthreads)
    _, _, imgs1, cams1 = crop_inputs(240, 180)
    prm1 = host.call_params(sc.depth_min, sc.depth_max, max_iterations=2, state=abi.FIRST_INIT, use_APD=0,
                            use_edge=0, use_limit=0, use_label=0, geom_consistency=0)
    hyps1, dt1 = timed(host.ProblemInputs(imgs1, cams1, prm1), 1)
    full = sc.width * sc.height / (cw * ch)
    return {"value": (h1 + h2) / (t1 + t2), "unit": "NCC plane-hyp/s", "cores": nthreads, "kind": "port",
            "sample": f"oracle, both calls of reference view {ref} (FIRST_INIT 8 iterations: {h1} NCC-hyps in "
                      f"{t1:.1f}s; geometric REFINE_ITER 2 iterations from the GPU's states: {h2} in {t2:.1f}s) on "
                      f"its {cw}x{ch} centre crop + 10 sources, {nthreads} threads",
            "s_per_depth_map": (t1 + t2) * full,
            "s_per_depth_map_note": f"the sample's two calls scaled by the view's pixel count (x{full:g})",
            "cpu_model": _cpu_model(), "host_cpus": os.cpu_count(),
            "single_thread": {"value": hyps1 / dt1, "unit": "NCC plane-hyp/s",
                              "sample": f"FIRST_INIT call, 2 iterations, 240x180 centre crop, {hyps1} NCC-hyps in "
                                        f"{dt1:.1f}s"}}


if __name__ == "__main__":
    main()
