"""Host side above the C-ABI: the Python mirror of the reference's per-image
process call and schedule.

  * `ProblemInputs` gathers what DPE::InuputInitialization and
    SupportInitialization read (DPE.cpp:934-1069, 1187-1214): float images at the
    call's scale, cameras scaled to it (SceneCache::GetCamera, DPE.cpp:739-760),
    the depth range x0.6 / x1.2 (DPE.cpp:953-954), geometric depths, pixel
    states, the prior plane state and the edge/label maps.
  * `process` is ProcessProblem (main.cpp:241-319): one dpe_process call plus
    the result assembly (depth zeroed and the pixel marked UNKNOWN outside the
    depth range, normal kept; main.cpp:264-280).
  * `StateStore` is the in-memory hand-off between calls (DPE.cpp:687-703).
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "libdpe_hip.so")


class DPEError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{abi.STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def load_library(path=LIB_PATH):
    """Load libdpe_hip.so; fails loudly when it has not been built."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it with __graft_entry__.build()")
    return abi.declare(C.CDLL(path))


class Context:
    """A dpe_ctx: one GPU, one stream, a device buffer pool."""

    def __init__(self, device=0, lib=None):
        self.lib = lib or load_library()
        h = C.c_void_p()
        st = self.lib.dpe_create(device, C.byref(h))
        if st != abi.OK:
            raise DPEError(st, self.lib.dpe_last_error(None).decode())
        self.h = h

    def close(self):
        if self.h:
            self.lib.dpe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, st):
        if st != abi.OK:
            raise DPEError(st, self.lib.dpe_last_error(self.h).decode())

    def stats(self):
        s = abi.Stats()
        self.check(self.lib.dpe_last_stats(self.h, C.byref(s)))
        return s


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p).value if a is not None else None


class Batch:
    """A dpe_batch: `lanes` contexts on one GPU processing independent problems
    concurrently (the batched form)."""

    def __init__(self, device=0, lanes=4, lib=None):
        self.lib = lib or load_library()
        h = C.c_void_p()
        st = self.lib.dpe_batch_create(device, lanes, C.byref(h))
        if st != abi.OK:
            raise DPEError(st, self.lib.dpe_batch_last_error(None).decode())
        self.h = h

    def close(self):
        if self.h:
            self.lib.dpe_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, pins, form=None):
        """Process the ProblemInputs `pins`; returns (outputs, stats, wall ms)."""
        n = len(pins)
        params = (abi.Params * n)(*[p.params for p in pins])
        inputs = (abi.Inputs * n)(*[p.struct for p in pins])
        outs = [ProblemOutputs(p.width, p.height) for p in pins]
        outputs = (abi.Outputs * n)(*[o.struct for o in outs])
        stats = (abi.Stats * n)()
        wall = C.c_float()
        st = self.lib.dpe_batch_process(self.h, n, params, C.byref(form) if form is not None else None, inputs,
                                        outputs, stats, C.byref(wall))
        if st != abi.OK:
            raise DPEError(st, self.lib.dpe_batch_last_error(self.h).decode())
        return outs, list(stats), wall.value


class ProblemInputs:
    """Host arrays of one process call plus the dpe_inputs struct pointing at them."""

    def __init__(self, images, cameras, params, geom_depths=None, weak_info=None, prior_planes=None,
                 prior_selected=None, edges=None, edges_low=None, labels=None):
        self.images = [np.ascontiguousarray(im, np.float32) for im in images]
        H, W = self.images[0].shape
        self.width, self.height = W, H
        self.cameras = cameras
        self.params = params
        self.params.num_images = len(self.images)
        self.geom_depths = None if geom_depths is None else [np.ascontiguousarray(g, np.float32) for g in geom_depths]
        self.weak_info = None if weak_info is None else np.ascontiguousarray(weak_info, np.uint8)
        self.prior_planes = None if prior_planes is None else np.ascontiguousarray(prior_planes, np.float32)
        self.prior_selected = None if prior_selected is None else np.ascontiguousarray(prior_selected, np.uint32)
        self.edges = None if edges is None else np.ascontiguousarray(edges, np.uint8)
        self.edges_low = None if edges_low is None else np.ascontiguousarray(edges_low, np.uint8)
        self.labels = None if labels is None else np.ascontiguousarray(labels, np.int32)
        self.struct = self._make()

    def _make(self):
        s = abi.Inputs()
        s.width, s.height, s.num_images, s.memory = self.width, self.height, len(self.images), abi.MEM_HOST
        for i, im in enumerate(self.images):
            s.images[i] = _ptr(im)
            s.cameras[i] = self.cameras[i]
            if self.geom_depths is not None:
                s.geom_depths[i] = _ptr(self.geom_depths[i])
        s.weak_info = _ptr(self.weak_info)
        s.prior_planes = _ptr(self.prior_planes)
        s.prior_selected = _ptr(self.prior_selected)
        s.edges = _ptr(self.edges)
        s.edges_low = _ptr(self.edges_low)
        if self.edges_low is not None:
            s.low_height, s.low_width = self.edges_low.shape
        s.labels = _ptr(self.labels)
        return s


class ProblemOutputs:
    def __init__(self, width, height):
        self.planes = np.zeros((height, width, 4), np.float32)
        self.costs = np.zeros((height, width), np.float32)
        self.pixel_states = np.zeros((height, width), np.uint8)
        self.selected_views = np.zeros((height, width), np.uint32)
        self.struct = abi.Outputs(_ptr(self.planes), _ptr(self.costs), _ptr(self.pixel_states),
                                  _ptr(self.selected_views))


def run_hip(ctx, pin, form=None):
    """One dpe_process (dpe_process_ex with a form) call on the GPU; returns ProblemOutputs."""
    out = ProblemOutputs(pin.width, pin.height)
    if form is None:
        ctx.check(ctx.lib.dpe_process(ctx.h, C.byref(pin.params), C.byref(pin.struct), C.byref(out.struct)))
    else:
        ctx.check(ctx.lib.dpe_process_ex(ctx.h, C.byref(pin.params), C.cast(C.byref(form), C.c_void_p),
                                         C.byref(pin.struct), C.byref(out.struct)))
    return out


def assemble(out, depth_min, depth_max):
    """Result assembly of ProcessProblem (main.cpp:264-280)."""
    depth = out.planes[..., 3].copy()
    states = out.pixel_states.copy()
    bad = (depth < depth_min) | (depth > depth_max)
    depth[bad] = 0.0
    states[bad] = abi.UNKNOWN
    normal = out.planes[..., :3].copy()
    return dict(depth=depth, normal=normal, pixel_states=states, selected_views=out.selected_views.copy(),
                costs=out.costs.copy())


class StateStore:
    """In-memory per-view state between calls (StateStore, DPE.cpp:687-703)."""

    def __init__(self):
        self.states = {}

    def put(self, image_id, result):
        self.states[image_id] = result

    def get(self, image_id):
        return self.states.get(image_id)


def scene_cameras(scene, ids, width=None, height=None):
    """Cameras of scene views, K scaled to width x height (SceneCache::GetCamera,
    DPE.cpp:739-760); c = -R^T t in double (ReadCamera, DPE.cpp:364-367)."""
    cams = []
    for i in ids:
        v = scene.views[i]
        W = width or scene.width
        H = height or scene.height
        K = v.K.astype(np.float32).copy()
        sx = np.float32(W / np.float32(scene.width))
        sy = np.float32(H / np.float32(scene.height))
        K[0, 0] *= sx
        K[0, 2] *= sx
        K[1, 1] *= sy
        K[1, 2] *= sy
        cams.append(abi.camera_from_arrays(K, v.R, v.t, W, H, scene.depth_min, scene.depth_max))
    return cams


def call_params(scene_dmin, scene_dmax, **kw):
    """Params of one call with the depth range of DPE.cpp:953-954."""
    p = abi.default_params(**kw)
    p.depth_min = float(np.float32(scene_dmin) * np.float32(0.6))
    p.depth_max = float(np.float32(scene_dmax) * np.float32(1.2))
    return p


def first_call_inputs(scene, ref, srcs, max_iterations=3, seed=0x5EED0002, call_index=0, support=None):
    """FIRST_INIT call as the reference's round 0 runs it (main.cpp:398-414):
    APD and edge sampling off, the weak path inactive."""
    ids = [ref] + list(srcs)
    imgs = [scene.views[i].image.astype(np.float32) for i in ids]
    cams = scene_cameras(scene, ids)
    prm = call_params(scene.depth_min, scene.depth_max, max_iterations=max_iterations, state=abi.FIRST_INIT,
                      use_APD=0, use_edge=0, geom_consistency=0, weak_peak_radius=6, seed=seed,
                      call_index=call_index)
    if support is None:
        from . import support as sup
        edges, labels = sup.support_maps(scene.views[ref].image)
    else:
        edges, labels = support
    return ProblemInputs(imgs, cams, prm, edges=edges, edges_low=edges, labels=labels)


def refine_call_inputs(scene, ref, srcs, store, max_iterations=3, seed=0x5EED0002, call_index=1, support=None,
                       state=abi.REFINE_ITER, weak_peak_radius=4, ransac_threshold=0.01 - 0.00125, rotate_time=2):
    """A REFINE call with APD, edge sampling and geometric consistency on
    (main.cpp:430-445 for round >= 1), fed from the state of earlier calls."""
    ids = [ref] + list(srcs)
    imgs = [scene.views[i].image.astype(np.float32) for i in ids]
    cams = scene_cameras(scene, ids)
    prm = call_params(scene.depth_min, scene.depth_max, max_iterations=max_iterations, state=state, use_APD=1,
                      use_edge=1, geom_consistency=1, weak_peak_radius=weak_peak_radius, seed=seed,
                      call_index=call_index, ransac_threshold=ransac_threshold, rotate_time=rotate_time)
    st = store.get(ref)
    planes = np.concatenate([st["normal"], st["depth"][..., None]], -1).astype(np.float32)
    geom = [store.get(i)["depth"] if store.get(i) is not None else np.zeros_like(st["depth"]) for i in ids]
    if support is None:
        from . import support as sup
        edges, labels = sup.support_maps(scene.views[ref].image)
    else:
        edges, labels = support
    return ProblemInputs(imgs, cams, prm, geom_depths=geom, weak_info=st["pixel_states"], prior_planes=planes,
                         prior_selected=st["selected_views"], edges=edges, edges_low=edges, labels=labels)


def reference_schedule(round_num):
    """The call groups of main.cpp:390-457 (one INIT call, then 3 REFINE_ITER
    calls per round, coarse to fine), as dicts; the C++ driver builds the same
    list (dpe_driver --dry-run prints it)."""
    calls = []
    for i in range(round_num):
        scale = 1 << (round_num - 1 - i)
        calls.append(dict(scale=scale, state=abi.FIRST_INIT if i == 0 else abi.REFINE_INIT, iters=3,
                          apd=0 if i == 0 else 1, edge=0 if i == 0 else 1, geom=0, peak=6,
                          ransac=0.005 if i == 0 else 0.01 - i * 0.00125, rotate=4 if i == 0 else min(1 << i, 4),
                          final=False))
        for j in range(3):
            calls.append(dict(scale=scale, state=abi.REFINE_ITER, iters=3, apd=0 if i == 0 else 1,
                              edge=0 if i == 0 else 1, geom=1, peak=max(4 - 2 * j, 2), ransac=0.01 - i * 0.00125,
                              rotate=min(1 << i, 4), final=(i == round_num - 1 and j == 2)))
    return calls


def scaled_size(width, height, scale):
    """Image size at a pyramid scale, as SceneCache::GetGrayFloatImage rounds it
    (round(W / s) x round(H / s), DPE.cpp:720-737; the driver's lround of the
    float product)."""
    inv = np.float32(1.0) / np.float32(scale)
    return int(np.floor(np.float32(width) * inv + 0.5)), int(np.floor(np.float32(height) * inv + 0.5))


class ScaledScene:
    """Float images, cameras and support maps of a synthetic scene at every
    pyramid scale, built as the C++ driver builds them: INTER_LINEAR reduction
    of the 8-bit image (dpe_host_resize_linear; factor 2 = 2x2 box mean),
    cameras with K scaled to the level (SceneCache::GetCamera, DPE.cpp:739-760),
    and support maps of the level's image rounded back to 8 bits
    (GetProblemEdges, main.cpp:166-221)."""

    def __init__(self, scene, hostlib=None):
        self.scene = scene
        self.hostlib = hostlib
        self._img = {}
        self._maps = {}

    def size(self, scale):
        return scaled_size(self.scene.width, self.scene.height, scale)

    def image(self, i, scale):
        key = (i, scale)
        if key not in self._img:
            full = self.scene.views[i].image.astype(np.float32)
            if scale == 1:
                self._img[key] = full
            else:
                w, h = self.size(scale)
                out = np.zeros((h, w), np.float32)
                self.hostlib.dpe_host_resize_linear(full.ctypes.data, self.scene.width, self.scene.height,
                                                    out.ctypes.data, w, h)
                self._img[key] = out
        return self._img[key]

    def cameras(self, ids, scale):
        w, h = self.size(scale)
        return scene_cameras(self.scene, ids, w, h)

    def support(self, i, scale):
        key = (i, scale)
        if key not in self._maps:
            from . import support as sup
            u8 = np.clip(np.rint(self.image(i, scale)), 0, 255).astype(np.uint8)
            self._maps[key] = sup.support_maps(u8)
        return self._maps[key]


def rescale_nearest(a, width, height):
    """RescaleMatToTargetSize (DPE.cpp:1313-1335): nearest, o = int(r / s) with
    s = target / source in float; identity at the same size."""
    h0, w0 = a.shape[:2]
    if (w0, h0) == (width, height):
        return a
    sx = np.float32(width) / np.float32(w0)
    sy = np.float32(height) / np.float32(h0)
    rows = (np.arange(height, dtype=np.float32) / sy).astype(np.int64)
    cols = (np.arange(width, dtype=np.float32) / sx).astype(np.int64)
    out = np.zeros((height, width) + a.shape[2:], a.dtype)
    ok_r = rows < h0
    ok_c = cols < w0
    out[np.ix_(ok_r, ok_c)] = a[np.ix_(rows[ok_r], cols[ok_c])]
    return out


def schedule_call_inputs(ss, c, call_index, ref, srcs, store, max_scale_size):
    """ProblemInputs of one call of a schedule at the call's scale: what the C++
    driver's ProcessProblem hands to dpe_process (main.cpp:241-319 with
    InuputInitialization and SupportInitialization, DPE.cpp:934-1069,
    1187-1214).  Stored states of another size are brought to the call's size
    with RescaleMatToTargetSize; the low-resolution edge map is the one of the
    coarsest scale (high_res_img)."""
    sc = ss.scene
    scale = c["scale"]
    W, H = ss.size(scale)
    ids = [ref] + list(srcs)
    imgs = [ss.image(i, scale) for i in ids]
    cams = ss.cameras(ids, scale)
    prm = call_params(sc.depth_min, sc.depth_max, max_iterations=c["iters"], state=c["state"], use_APD=c["apd"],
                      use_edge=c["edge"], geom_consistency=c["geom"], weak_peak_radius=c["peak"],
                      ransac_threshold=np.float32(c["ransac"]), rotate_time=c["rotate"], call_index=call_index,
                      max_scale_size=max_scale_size, scale_size=scale)
    kw = {}
    if c["geom"]:
        zero = np.zeros((H, W), np.float32)
        kw["geom_depths"] = [rescale_nearest(store.get(i)["depth"], W, H) if store.get(i) is not None else zero
                             for i in ids]
    st = store.get(ref)
    if c["apd"]:
        kw["weak_info"] = rescale_nearest(st["pixel_states"], W, H)
    if c["state"] != abi.FIRST_INIT:
        depth = rescale_nearest(st["depth"], W, H)
        normal = rescale_nearest(st["normal"], W, H)
        kw["prior_planes"] = np.concatenate([normal, depth[..., None]], -1).astype(np.float32)
        kw["prior_selected"] = rescale_nearest(st["selected_views"], W, H)
    edges, labels = ss.support(ref, scale)
    edges_low, _ = ss.support(ref, max_scale_size)
    return ProblemInputs(imgs, cams, prm, edges=edges, edges_low=edges_low, labels=labels, **kw)


def run_schedule(runner, scene, calls, views=None, hostlib=None, on_call=None):
    """Python mirror of the C++ driver for schedules over a synthetic scene (any
    scales): every call group over every view in order, state hand-off through
    a StateStore, call_index = call group index, max_scale_size = the largest
    scale of the schedule.  `runner(pin)` returns the ProblemOutputs of one call
    (the GPU: `lambda pin: run_hip(ctx, pin)`); `on_call(group, ref, pin, out)`
    sees every call.  Scales above 1 need libdpe_host (`hostlib`) for the
    pyramid.  Returns the StateStore."""
    views = list(range(len(scene.views))) if views is None else list(views)
    ss = ScaledScene(scene, hostlib)
    max_scale_size = max(c["scale"] for c in calls)
    store = StateStore()
    for it, c in enumerate(calls):
        for ref in views:
            srcs = [j for j in scene.pairs[ref] if j in views]
            pin = schedule_call_inputs(ss, c, it, ref, srcs, store, max_scale_size)
            out = runner(pin)
            if on_call is not None:
                on_call(it, ref, pin, out)
            store.put(ref, assemble(out, pin.params.depth_min, pin.params.depth_max))
    return store


def run_schedule_scale1(ctx, scene, calls, views=None):
    """run_schedule on the GPU for schedules whose calls are all at scale 1."""
    assert all(c["scale"] == 1 for c in calls)
    return run_schedule(lambda pin: run_hip(ctx, pin), scene, calls, views)
