"""The C++ host driver end to end on the GPU: it reads a scene folder (pair.txt,
cams, PGM images), runs the reference schedule (main.cpp:390-457) through
libdpe_hip.so, writes the checkpoints and the fused PLY.  Its final state must
equal, bit for bit, the same schedule run through the Python host API, and its
point cloud the Python-driven dpe_fuse on that state."""
import os
import subprocess

import numpy as np
import pytest

from dpe_amd import abi, host, hostio, scene
from fuse_util import FuseSet, fuse_hip

pytestmark = pytest.mark.gpu


def test_driver_matches_python_schedule(tmp_path):
    s = scene.make_scene(160, 120, 150.0, 4, 24.0, seed=0x5EED0003, supersample=1)
    folder = tmp_path / "scene"
    scene.write_folder(s, str(folder))
    out = tmp_path / "out"
    print("driver: 4 views 160x120, reference schedule (1 round, 4 calls)", flush=True)
    r = subprocess.run([hostio.DRIVER_PATH, str(folder), "0", "--output=" + str(out), "--checkpoint=final"],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    ctx = host.Context(0)
    try:
        store = host.run_schedule_scale1(ctx, s, host.reference_schedule(1))
        for i in range(4):
            f = out / "views" / ("%08d" % i)
            st = store.get(i)
            assert np.array_equal(hostio.read_bin_mat(str(f / "depths.dmb")), st["depth"]), i
            assert np.array_equal(hostio.read_bin_mat(str(f / "normals.dmb")), st["normal"]), i
            assert np.array_equal(hostio.read_bin_mat(str(f / "weak.bin")), st["pixel_states"]), i
            assert np.array_equal(hostio.read_bin_mat(str(f / "selected_views.bin")).view(np.uint32),
                                  st["selected_views"]), i
        fs = FuseSet(s, [store.get(i)["depth"] for i in range(4)], [store.get(i)["normal"] for i in range(4)],
                     [store.get(i)["pixel_states"] for i in range(4)])
        pts = fuse_hip(ctx, fs, abi.fuse_defaults())
    finally:
        ctx.close()
    header, verts = hostio.read_ply(str(out / "DPE.ply"))
    print(f"driver fused {len(verts)} points, python {len(pts)}")
    assert len(verts) == len(pts) > 0
    assert np.array_equal(verts["x"], pts["x"]) and np.array_equal(verts["b"], pts["b"])


def test_driver_with_reference_segmentation(tmp_path):
    """--support=reference: the support maps come from the EdgeSegment
    restatement (Canny edges, Roberts + Hough + Connect labels); the run
    completes, writes maps of the call's size and a point cloud whose depths
    agree with ground truth about as well as with this project's detector."""
    s = scene.make_scene(320, 240, 300.0, 4, 24.0, seed=0x5EED0005, supersample=1)
    folder = tmp_path / "scene"
    scene.write_folder(s, str(folder))
    # a REFINE_ITER call with APD, edge sampling and the weak path consumes the maps
    sched = tmp_path / "refine.sched"
    sched.write_text("call scale=1 state=FIRST_INIT iters=3 apd=0 edge=0 geom=0 peak=6\n"
                     "call scale=1 state=REFINE_ITER iters=2 apd=1 edge=1 geom=1 peak=4 ransac=0.00875 rotate=2\n"
                     "fusion min_consistent=1\n")
    acc, pts = {}, {}
    for mode in ("own", "reference"):
        out = tmp_path / ("out_" + mode)
        print(f"driver --support={mode}: 4 views 320x240", flush=True)
        r = subprocess.run([hostio.DRIVER_PATH, str(folder), "0", "--output=" + str(out), "--checkpoint=final",
                            "--schedule=" + str(sched), "--support=" + mode], capture_output=True, text=True,
                           timeout=300)
        print(r.stdout[-1500:], r.stderr[-1500:])
        assert r.returncode == 0
        labels = hostio.read_bin_mat(str(out / "views" / "00000001" / "labels_0.dmb"))
        edges = hostio.read_bin_mat(str(out / "views" / "00000001" / "edges_0.dmb"))
        assert labels.shape == edges.shape == (240, 320)
        good = []
        for i in range(4):
            d = hostio.read_bin_mat(str(out / "views" / ("%08d" % i) / "depths.dmb"))
            gt = s.views[i].depth
            good.append(float((np.abs(d - gt) / gt < 0.01).mean()))
        acc[mode] = float(np.mean(good))
        _, verts = hostio.read_ply(str(out / "DPE.ply"))
        pts[mode] = len(verts)
        print(f"--support={mode}: {len(verts)} points, {acc[mode]:.4f} of pixels within 1 % of ground truth")
        assert len(verts) > 0
    assert acc["reference"] >= acc["own"] - 0.05
    assert pts["reference"] != pts["own"]  # the maps are consumed: the runs differ


def test_driver_reads_jpeg_folders(tmp_path):
    """images/%08d.jpg input: the driver decodes the JPEGs itself (as
    cv::imread does through libjpeg) and its final state equals the Python
    schedule run on PIL's (libjpeg-turbo's) decode of the same files."""
    import copy
    from PIL import Image
    s = scene.make_scene(160, 120, 150.0, 4, 24.0, seed=0x5EED0006, supersample=1)
    folder = tmp_path / "scene"
    scene.write_folder(s, str(folder))
    s2 = copy.deepcopy(s)
    for i, v in enumerate(s.views):
        pgm = folder / "images" / ("%08d.pgm" % i)
        jpg = folder / "images" / ("%08d.jpg" % i)
        col = np.stack([v.image, np.roll(v.image, 2, 1), 255 - v.image], -1).astype(np.uint8)
        Image.fromarray(col).save(jpg, quality=92, subsampling=2)
        pgm.unlink()
        im = Image.open(jpg)
        im.draft("L", im.size)
        s2.views[i].image = np.array(im.convert("L") if im.mode != "L" else im)
    out = tmp_path / "out"
    r = subprocess.run([hostio.DRIVER_PATH, str(folder), "0", "--output=" + str(out), "--checkpoint=final"],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    ctx = host.Context(0)
    try:
        store = host.run_schedule_scale1(ctx, s2, host.reference_schedule(1))
    finally:
        ctx.close()
    for i in range(4):
        f = out / "views" / ("%08d" % i)
        assert np.array_equal(hostio.read_bin_mat(str(f / "depths.dmb")), store.get(i)["depth"]), i
    _, verts = hostio.read_ply(str(out / "DPE.ply"))
    assert len(verts) > 0


def test_driver_tat_fusions_with_blocks(tmp_path):
    """--fusion=tat_intermediate / tat_advanced with a blocks/ folder: the PLY
    equals dpe_fuse_tat on the driver's own final state and on the block masks
    decoded from blocks/mask_<id>.jpg; a mask smaller than the depth map is an
    error."""
    from PIL import Image
    from fuse_util import fuse_tat_hip
    s = scene.make_scene(160, 120, 150.0, 4, 24.0, seed=0x5EED0008, supersample=1)
    folder = tmp_path / "scene"
    scene.write_folder(s, str(folder))
    (folder / "blocks").mkdir()
    yy, xx = np.mgrid[0:120, 0:160]
    for i in range(4):
        m = np.where((xx // 20 + yy // 20 + i) % 3 == 0, 0, 255).astype(np.uint8)
        Image.fromarray(m).save(folder / "blocks" / ("mask_%d.jpg" % i), quality=90)
    sched = tmp_path / "short.sched"
    sched.write_text("call scale=1 state=FIRST_INIT iters=3 apd=0 edge=0 geom=0 peak=6\n")
    hl = hostio.load()
    blocks = [hostio.read_jpeg(hl, folder / "blocks" / ("mask_%d.jpg" % i)) for i in range(4)]
    assert 0 < np.mean([(b < 128).mean() for b in blocks]) < 0.6
    for name, variant in (("tat_intermediate", abi.TAT_INTERMEDIATE), ("tat_advanced", abi.TAT_ADVANCED)):
        out = tmp_path / ("out_" + name)
        print(f"driver --fusion={name}: 4 views 160x120 with blocks", flush=True)
        r = subprocess.run([hostio.DRIVER_PATH, str(folder), "0", "--output=" + str(out), "--checkpoint=final",
                            "--schedule=" + str(sched), "--fusion=" + name], capture_output=True, text=True,
                           timeout=300)
        print(r.stdout[-1500:], r.stderr[-1500:])
        assert r.returncode == 0
        views = [out / "views" / ("%08d" % i) for i in range(4)]
        fs = FuseSet(s, [hostio.read_bin_mat(str(v / "depths.dmb")) for v in views],
                     [hostio.read_bin_mat(str(v / "normals.dmb")) for v in views],
                     [hostio.read_bin_mat(str(v / "weak.bin")) for v in views])
        ctx = host.Context(0)
        try:
            pts = fuse_tat_hip(ctx, fs, abi.fuse_tat_defaults(variant), blocks)
        finally:
            ctx.close()
        _, verts = hostio.read_ply(str(out / "DPE.ply"))
        print(f"--fusion={name}: driver {len(verts)} points, python {len(pts)}")
        assert len(verts) == len(pts) > 0
        for k in ("x", "y", "z", "b", "g", "r"):
            assert np.array_equal(verts[k], pts[k]), k
    Image.fromarray(np.zeros((100, 160), np.uint8)).save(folder / "blocks" / "mask_2.jpg")
    r = subprocess.run([hostio.DRIVER_PATH, str(folder), "0", "--output=" + str(tmp_path / "bad"),
                        "--schedule=" + str(sched), "--fusion=tat_advanced"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode != 0 and "Block mask smaller than the depth map" in r.stderr
