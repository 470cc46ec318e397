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
