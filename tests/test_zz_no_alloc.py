"""dpe_ncc_cost_noalloc, the entry point that allocates no device memory, against
dpe_ncc_cost on the same device-resident inputs: bit-identical cost vectors and
counts on two calls in a row, the second with a smaller problem (fewer pixels,
sources and hypotheses), so the scratch array's leftovers from the first call
must not leak into the second.  Problems beyond the scratch array and host
inputs are refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from dpe_amd import abi, host, scene

pytestmark = pytest.mark.gpu


def _device_problem(torch, s, ref, srcs, W, H, nh, seed):
    """Device-resident inputs of a cost-vector call over the top-left W x H crop
    of the scene's views (pixel coordinates unchanged, so the cameras only take
    the crop's size): images, random camera-frame planes and two cost arrays."""
    dev = torch.device("cuda:0")
    ids = [ref] + list(srcs)
    imgs = [torch.from_numpy(np.ascontiguousarray(s.views[i].image[:H, :W], np.float32)).to(dev) for i in ids]
    cams = host.scene_cameras(s, ids)
    for c in cams:
        c.width, c.height = W, H
    prm = host.call_params(s.depth_min, s.depth_max, use_APD=0, use_edge=0, use_limit=0, use_label=0)
    prm.num_images = len(ids)
    st = abi.Inputs()
    st.width, st.height, st.num_images, st.memory = W, H, len(ids), abi.MEM_DEVICE
    for k, im in enumerate(imgs):
        st.images[k] = im.data_ptr()
        st.cameras[k] = cams[k]
    rng = np.random.default_rng(seed)
    n = rng.standard_normal((nh, H, W, 3))
    n[..., 2] = -np.abs(n[..., 2]) - 0.5
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    d = rng.uniform(2.0, 6.0, (nh, H, W, 1))
    planes = torch.from_numpy(np.concatenate([n, d], -1).astype(np.float32)).to(dev)
    NV = len(srcs)
    a = torch.full((nh, NV, H, W), -1.0, dtype=torch.float32, device=dev)
    b = torch.full((nh, NV, H, W), -2.0, dtype=torch.float32, device=dev)
    return prm, st, imgs, planes, a, b


def test_noalloc_cost_vectors_equal_the_pooled_call():
    torch = pytest.importorskip("torch")
    c1 = scene.config1()
    ctx = host.Context(0)
    keep = []
    try:
        cases = [(c1.meta["ref_mid"], c1.pairs[c1.meta["ref_mid"]], 320, 240, 3, 1),
                 (1, c1.pairs[1][:2], 160, 120, 2, 2)]
        for ref, srcs, W, H, nh, seed in cases:
            print(f"noalloc case {W}x{H}, {len(srcs)} sources, {nh} hypotheses", flush=True)
            prm, st, imgs, planes, a, b = _device_problem(torch, c1, ref, srcs, W, H, nh, seed)
            keep.append((imgs, planes, a, b))
            torch.cuda.synchronize()
            ctx.check(ctx.lib.dpe_ncc_cost(ctx.h, C.byref(prm), C.byref(st), nh, planes.data_ptr(), a.data_ptr()))
            sa = ctx.stats()
            ctx.check(ctx.lib.dpe_ncc_cost_noalloc(ctx.h, C.byref(prm), C.byref(st), nh, planes.data_ptr(),
                                                   b.data_ptr()))
            sb = ctx.stats()
            ga, gb = a.cpu().numpy(), b.cpu().numpy()
            assert (ga >= 0).all() and (ga <= 2).all()
            assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32))
            assert sa.ncc_hyps == sb.ncc_hyps == nh * len(srcs) * W * H
            assert sa.ncc_taps == sb.ncc_taps
        # the pooled call still works after the scratch calls
        ref, srcs, W, H, nh, seed = cases[0]
        prm, st, imgs, planes, a, b = _device_problem(torch, c1, ref, srcs, W, H, nh, seed)
        ctx.check(ctx.lib.dpe_ncc_cost(ctx.h, C.byref(prm), C.byref(st), nh, planes.data_ptr(), a.data_ptr()))
        assert np.array_equal(a.cpu().numpy().view(np.uint32), keep[0][2].cpu().numpy().view(np.uint32))
    finally:
        ctx.close()


def test_noalloc_refuses_oversized_and_host_problems():
    torch = pytest.importorskip("torch")
    ctx = host.Context(0)
    try:
        dev = torch.device("cuda:0")
        W, H, N = 2048, 1536, 11
        imgs = [torch.zeros((H, W), dtype=torch.float32, device=dev) for _ in range(N)]
        cam = abi.camera_from_arrays(np.array([[1500, 0, W / 2], [0, 1500, H / 2], [0, 0, 1]], np.float32),
                                     np.eye(3, dtype=np.float32), np.zeros(3, np.float32), W, H, 2.0, 6.0)
        prm = host.call_params(2.0, 6.0, use_APD=0, use_edge=0, use_limit=0, use_label=0)
        prm.num_images = N
        st = abi.Inputs()
        st.width, st.height, st.num_images, st.memory = W, H, N, abi.MEM_DEVICE
        for k in range(N):
            st.images[k] = imgs[k].data_ptr()
            st.cameras[k] = cam
        planes = torch.zeros((1, H, W, 4), dtype=torch.float32, device=dev)
        costs = torch.zeros((1, N - 1, H, W), dtype=torch.float32, device=dev)
        rc = ctx.lib.dpe_ncc_cost_noalloc(ctx.h, C.byref(prm), C.byref(st), 1, planes.data_ptr(), costs.data_ptr())
        assert rc == abi.ERR_UNSUPPORTED
        assert b"bytes of scratch" in ctx.lib.dpe_last_error(ctx.h)
        st.memory = abi.MEM_HOST
        rc = ctx.lib.dpe_ncc_cost_noalloc(ctx.h, C.byref(prm), C.byref(st), 1, planes.data_ptr(), costs.data_ptr())
        assert rc == abi.ERR_INVALID_ARGUMENT
    finally:
        ctx.close()
