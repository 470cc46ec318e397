"""Fusion inputs for tests: dpe_fuse_view structs over host arrays."""
import ctypes as C

import numpy as np

from dpe_amd import abi, host


class FuseSet:
    def __init__(self, scene, depths, normals, states, pairs=None):
        self.keep = []
        n = len(scene.views)
        self.views = (abi.FuseView * n)()
        cams = host.scene_cameras(scene, list(range(n)))
        pairs = pairs or scene.pairs
        for i, v in enumerate(scene.views):
            d = np.ascontiguousarray(depths[i], np.float32)
            nm = np.ascontiguousarray(normals[i], np.float32)
            st = np.ascontiguousarray(states[i], np.uint8)
            col = np.ascontiguousarray(np.repeat(v.image[..., None], 3, -1), np.uint8)
            self.keep += [d, nm, st, col]
            fv = self.views[i]
            fv.width, fv.height = scene.width, scene.height
            fv.camera = cams[i]
            fv.depth = d.ctypes.data
            fv.normal = nm.ctypes.data
            fv.pixel_states = st.ctypes.data
            fv.color = col.ctypes.data
            fv.num_src = len(pairs[i])
            for j, s in enumerate(pairs[i]):
                fv.src_index[j] = s
        self.n = n


POINT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("b", "u1"), ("g", "u1"), ("r", "u1"),
                        ("pad", "u1")])


def _blocks_arg(fs, blocks):
    """NULL, or an array of per-view u8 block pointers (None entries = no block)."""
    if blocks is None:
        return None
    arr = (C.c_void_p * fs.n)()
    for i, b in enumerate(blocks):
        if b is not None:
            b = np.ascontiguousarray(b, np.uint8)
            fs.keep.append(b)
            arr[i] = b.ctypes.data
    fs.keep.append(arr)
    return C.cast(arr, C.c_void_p)


def fuse_tat_oracle(orc, fs, tp, blocks=None):
    bl = _blocks_arg(fs, blocks)
    cnt = C.c_int64()
    assert orc.oracle_fuse_tat(fs.views, fs.n, bl, C.byref(tp), None, 0, C.byref(cnt)) == 0
    pts = (abi.Point * max(cnt.value, 1))()
    assert orc.oracle_fuse_tat(fs.views, fs.n, bl, C.byref(tp), pts, cnt.value, C.byref(cnt)) == 0
    return np.frombuffer(pts, dtype=POINT_DTYPE)[:cnt.value]


def fuse_tat_hip(ctx, fs, tp, blocks=None):
    bl = _blocks_arg(fs, blocks)
    cnt = C.c_int64()
    ctx.check(ctx.lib.dpe_fuse_tat(ctx.h, fs.views, fs.n, bl, C.byref(tp), None, 0, C.byref(cnt)))
    pts = (abi.Point * max(cnt.value, 1))()
    ctx.check(ctx.lib.dpe_fuse_tat(ctx.h, fs.views, fs.n, bl, C.byref(tp), C.cast(pts, C.c_void_p), cnt.value,
                                   C.byref(cnt)))
    return np.frombuffer(pts, dtype=POINT_DTYPE)[:cnt.value]


def fuse_oracle(orc, fs, fp, capacity=None):
    """The oracle's RunFusion points; with `capacity`, one pass that writes up to
    that many points (a second pass only if there are more)."""
    cnt = C.c_int64()
    if capacity is None:
        assert orc.oracle_fuse(fs.views, fs.n, C.byref(fp), None, 0, C.byref(cnt)) == 0
        capacity = cnt.value
    pts = (abi.Point * max(capacity, 1))()
    assert orc.oracle_fuse(fs.views, fs.n, C.byref(fp), pts, capacity, C.byref(cnt)) == 0
    if cnt.value > capacity:
        return fuse_oracle(orc, fs, fp, cnt.value)
    return np.frombuffer(pts, dtype=np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("b", "u1"), ("g", "u1"),
                                              ("r", "u1"), ("pad", "u1")]))[:cnt.value]


def fuse_hip(ctx, fs, fp):
    cnt = C.c_int64()
    ctx.check(ctx.lib.dpe_fuse(ctx.h, fs.views, fs.n, C.byref(fp), None, 0, C.byref(cnt)))
    pts = (abi.Point * max(cnt.value, 1))()
    ctx.check(ctx.lib.dpe_fuse(ctx.h, fs.views, fs.n, C.byref(fp), C.cast(pts, C.c_void_p), cnt.value,
                               C.byref(cnt)))
    return np.frombuffer(pts, dtype=np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("b", "u1"), ("g", "u1"),
                                              ("r", "u1"), ("pad", "u1")]))[:cnt.value]


class _Views:
    pass


def coincident_pair(depth1_scale_rows=None, W=40, H=30, R=None, t=None):
    """Two views with the same camera looking at the plane z = 4 in camera
    coordinates (all pixels strong, normals (0, 0, -1)): every pixel of view 0
    reprojects onto itself in view 1.  depth1_scale_rows: (row slice, factor)
    applied to view 1's depth; R, t: the shared pose (default identity).
    Returns (views holder for fuse_oracle / fuse_hip, K)."""
    K = np.array([[50.0, 0, (W - 1) / 2], [0, 50.0, (H - 1) / 2], [0, 0, 1]])
    cam = abi.camera_from_arrays(K, np.eye(3) if R is None else R, np.zeros(3) if t is None else t, W, H, 2, 6)
    fs = _Views()
    fs.keep = []
    fs.views = (abi.FuseView * 2)()
    fs.n = 2
    for i in range(2):
        d = np.full((H, W), 4.0, np.float32)
        if i == 1 and depth1_scale_rows is not None:
            rows, f = depth1_scale_rows
            d[rows] *= np.float32(f)
        nrm = np.zeros((H, W, 3), np.float32)
        nrm[..., 2] = -1.0
        if R is not None:
            nrm = (nrm.reshape(-1, 3) @ np.asarray(R, np.float64)).astype(np.float32).reshape(H, W, 3)  # world frame
        st = np.full((H, W), abi.STRONG, np.uint8)
        col = np.full((H, W, 3), 100 + i, np.uint8)
        fs.keep += [d, nrm, st, col]
        v = fs.views[i]
        v.width, v.height, v.camera = W, H, cam
        v.depth, v.normal, v.pixel_states, v.color = d.ctypes.data, nrm.ctypes.data, st.ctypes.data, col.ctypes.data
        v.num_src = 1
        v.src_index[0] = 1 - i
    return fs, K


def rotated_pose():
    """A fixed non-trivial rotation (about an oblique axis) and translation."""
    ax = np.array([0.3, -0.8, 0.52])
    ax /= np.linalg.norm(ax)
    a = 0.7
    Kx = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    R = np.eye(3) + np.sin(a) * Kx + (1 - np.cos(a)) * Kx @ Kx
    return R, np.array([0.37, -1.21, 2.9])


def world_points_fp64(cam, W, H, depth):
    """Get3DPointonWorld of every pixel evaluated in float64 from the camera's
    float32 entries (the FP64 form's arithmetic up to operation order), rounded
    to float32 as the point cloud stores it.  (H*W, 3)."""
    K = np.array(cam.K[:], np.float64).reshape(3, 3)
    R = np.array(cam.R[:], np.float64).reshape(3, 3)
    t = np.array(cam.t[:], np.float64)
    ys, xs = np.mgrid[0:H, 0:W]
    Xc = np.stack([depth * (xs - K[0, 2]) / K[0, 0], depth * (ys - K[1, 2]) / K[1, 1],
                   np.full((H, W), depth)], -1).reshape(-1, 3)
    return (Xc @ R - R.T @ t).astype(np.float32)


def coincident_views(n, W=40, H=30, depth_scale=None, tilt=None, color_b=None, depth_zero_rows=None):
    """n views with one camera looking at the plane z = 4, each listing all the
    others (ascending) as sources; normals (0, 0, -1), colour 100 + i.
    depth_scale: {view: factor} on the depth map; tilt: {view: angle} rotates
    that view's normals about y; color_b: {view: (H, W) u8} blue channel;
    depth_zero_rows: {view: row slice} with depth 0 (holes)."""
    K = np.array([[50.0, 0, (W - 1) / 2], [0, 50.0, (H - 1) / 2], [0, 0, 1]])
    cam = abi.camera_from_arrays(K, np.eye(3), np.zeros(3), W, H, 2, 6)
    fs = _Views()
    fs.keep = []
    fs.views = (abi.FuseView * n)()
    fs.n = n
    for i in range(n):
        d = np.full((H, W), 4.0, np.float32)
        if depth_scale and i in depth_scale:
            d *= np.float32(depth_scale[i])
        if depth_zero_rows and i in depth_zero_rows:
            d[depth_zero_rows[i]] = 0.0
        nrm = np.zeros((H, W, 3), np.float32)
        th = (tilt or {}).get(i, 0.0)
        nrm[..., 0] = np.float32(np.sin(th))
        nrm[..., 2] = np.float32(-np.cos(th))
        st = np.full((H, W), abi.STRONG, np.uint8)
        col = np.full((H, W, 3), 100 + i, np.uint8)
        if color_b and i in color_b:
            col[..., 0] = color_b[i]
        fs.keep += [d, nrm, st, col]
        v = fs.views[i]
        v.width, v.height, v.camera = W, H, cam
        v.depth, v.normal, v.pixel_states, v.color = d.ctypes.data, nrm.ctypes.data, st.ctypes.data, col.ctypes.data
        srcs = [j for j in range(n) if j != i]
        v.num_src = len(srcs)
        for j, s in enumerate(srcs):
            v.src_index[j] = s
    return fs, K
