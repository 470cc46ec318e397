"""Error paths of the C-ABI (the reference exits the process on these: CudaSafeCall
DPE.cpp:646-679, "Can't process so much images" DPE.cpp:944-947, missing weak.bin
DPE.cpp:995-998).  Every bad call must return its status with a message from
dpe_last_error, touch no buffer it was not meant to, and leave the context usable:
the valid call after all of them equals the same call on a fresh context bit for bit.
All failures here are detected on the host before any kernel runs."""
import ctypes as C

import numpy as np
import pytest

from dpe_amd import abi, host, scene
from fuse_util import coincident_pair

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    s = scene.make_scene(64, 48, 60.0, 3, 20.0, seed=0x5EED00E1, supersample=1)
    return s, host.first_call_inputs(s, 0, s.pairs[0], max_iterations=1)


def _call(ctx, prm, inp, out, form=None):
    if form is None:
        return ctx.lib.dpe_process(ctx.h, C.byref(prm), C.byref(inp), C.byref(out.struct) if out else None)
    return ctx.lib.dpe_process_ex(ctx.h, C.byref(prm), C.cast(C.byref(form), C.c_void_p), C.byref(inp),
                                  C.byref(out.struct))


def _expect(ctx, st, status, text):
    assert st == status, (st, status, text)
    msg = ctx.lib.dpe_last_error(ctx.h).decode()
    assert text in msg, (text, msg)


def test_create_errors():
    lib = host.load_library()
    h = C.c_void_p()
    assert lib.dpe_create(99, C.byref(h)) == abi.ERR_NO_DEVICE
    assert "out of range" in lib.dpe_last_error(None).decode()
    assert lib.dpe_create(0, None) == abi.ERR_INVALID_ARGUMENT
    b = C.c_void_p()
    for lanes in (0, abi.MAX_LANES + 1):
        assert lib.dpe_batch_create(0, lanes, C.byref(b)) == abi.ERR_INVALID_ARGUMENT


def test_process_errors_leave_the_context_usable(small):
    s, pin = small
    ctx = host.Context(0)
    try:
        out = host.ProblemOutputs(pin.width, pin.height)

        def variant(params=None, inputs=None):
            prm = abi.Params.from_buffer_copy(pin.params)
            inp = abi.Inputs.from_buffer_copy(pin.struct)
            for k, v in (params or {}).items():
                setattr(prm, k, v)
            for k, v in (inputs or {}).items():
                setattr(inp, k, v)
            return prm, inp

        assert ctx.lib.dpe_process(ctx.h, None, C.byref(pin.struct), C.byref(out.struct)) == abi.ERR_INVALID_ARGUMENT
        prm, inp = variant()
        _expect(ctx, _call(ctx, prm, inp, None), abi.ERR_INVALID_ARGUMENT, "null output")
        prm, inp = variant(inputs={"num_images": 33})
        _expect(ctx, _call(ctx, prm, inp, out), abi.ERR_TOO_MANY_IMAGES, "Can't process so much images")
        prm, inp = variant(inputs={"num_images": 1})
        _expect(ctx, _call(ctx, prm, inp, out), abi.ERR_INVALID_ARGUMENT, "at least one source")
        prm, inp = variant(inputs={"width": 70000})
        _expect(ctx, _call(ctx, prm, inp, out), abi.ERR_UNSUPPORTED, "larger than 65536")
        prm, inp = variant(inputs={"memory": 5})
        _expect(ctx, _call(ctx, prm, inp, out), abi.ERR_INVALID_ARGUMENT, "bad memory kind")
        prm, inp = variant()
        inp.images[1] = None
        _expect(ctx, _call(ctx, prm, inp, out), abi.ERR_MISSING_INPUT, "image 1 is null")
        prm, inp = variant(params={"state": 7})
        _expect(ctx, _call(ctx, prm, inp, out), abi.ERR_INVALID_ARGUMENT, "bad run state")
        prm, inp = variant(params={"state": abi.REFINE_ITER})
        _expect(ctx, _call(ctx, prm, inp, out), abi.ERR_MISSING_INPUT, "prior planes")
        prm, inp = variant(params={"use_APD": 1}, inputs={"weak_info": None})
        _expect(ctx, _call(ctx, prm, inp, out), abi.ERR_MISSING_INPUT, "weak info")
        prm, inp = variant(params={"use_edge": 1}, inputs={"edges": None})
        _expect(ctx, _call(ctx, prm, inp, out), abi.ERR_MISSING_INPUT, "edge maps")
        prm, inp = variant(params={"use_label": 1}, inputs={"labels": None})
        _expect(ctx, _call(ctx, prm, inp, out), abi.ERR_MISSING_INPUT, "label map")
        prm, inp = variant(params={"geom_consistency": 1})
        _expect(ctx, _call(ctx, prm, inp, out), abi.ERR_MISSING_INPUT, "depth map")
        prm, inp = variant(params={"rotate_time": 0})
        _expect(ctx, _call(ctx, prm, inp, out), abi.ERR_INVALID_ARGUMENT, "rotate_time")
        prm, inp = variant(params={"strong_increment": 0})
        _expect(ctx, _call(ctx, prm, inp, out), abi.ERR_INVALID_ARGUMENT, "bad window parameters")
        prm, inp = variant(params={"strong_radius": 12})
        _expect(ctx, _call(ctx, prm, inp, out), abi.ERR_UNSUPPORTED, "wider than 11 taps")
        prm, inp = variant()
        inp.cameras[2].width = 65
        _expect(ctx, _call(ctx, prm, inp, out), abi.ERR_INVALID_ARGUMENT, "camera 2 size")
        for field, value, text in (("precision", 9, "unknown precision"), ("gauss_newton_steps", 3, "gauss_newton_steps"),
                                   ("autotune", 2, "autotune"), ("graph", 2, "graph")):
            prm, inp = variant()
            form = abi.Form()
            setattr(form, field, value)
            _expect(ctx, _call(ctx, prm, inp, out, form), abi.ERR_INVALID_ARGUMENT, text)
        form = abi.Form()
        form.reserved[1] = 1
        _expect(ctx, _call(ctx, *variant(), out, form), abi.ERR_INVALID_ARGUMENT, "reserved")
        # the context still works and gives what a fresh one gives
        res = host.run_hip(ctx, pin)
    finally:
        ctx.close()
    fresh = host.Context(0)
    try:
        ref = host.run_hip(fresh, pin)
    finally:
        fresh.close()
    assert np.array_equal(res.planes.view(np.uint8), ref.planes.view(np.uint8))
    assert np.array_equal(res.selected_views, ref.selected_views)


def test_cost_and_fusion_errors(small):
    s, pin = small
    ctx = host.Context(0)
    try:
        costs = np.zeros(4, np.float32)
        planes = np.zeros((1, pin.height, pin.width, 4), np.float32)
        st = ctx.lib.dpe_ncc_cost(ctx.h, C.byref(pin.params), C.byref(pin.struct), -1, planes.ctypes.data,
                                  costs.ctypes.data)
        _expect(ctx, st, abi.ERR_INVALID_ARGUMENT, "bad hypotheses")
        fs, _ = coincident_pair(None, 16, 12)
        cnt = C.c_int64()
        fp = abi.fuse_defaults()
        assert ctx.lib.dpe_fuse(ctx.h, fs.views, 0, C.byref(fp), None, 0, C.byref(cnt)) == abi.ERR_INVALID_ARGUMENT
        fs.views[0].src_index[0] = 0  # its own source
        _expect(ctx, ctx.lib.dpe_fuse(ctx.h, fs.views, fs.n, C.byref(fp), None, 0, C.byref(cnt)),
                abi.ERR_INVALID_ARGUMENT, "bad source index")
        fs.views[0].src_index[0] = 1
        bad = abi.fuse_defaults(memory=3)
        _expect(ctx, ctx.lib.dpe_fuse(ctx.h, fs.views, fs.n, C.byref(bad), None, 0, C.byref(cnt)),
                abi.ERR_INVALID_ARGUMENT, "bad memory kind")
        fs.views[1].depth = None
        _expect(ctx, ctx.lib.dpe_fuse(ctx.h, fs.views, fs.n, C.byref(fp), None, 0, C.byref(cnt)),
                abi.ERR_INVALID_ARGUMENT, "is incomplete")
        fs, _ = coincident_pair(None, 16, 12)
        ctx.check(ctx.lib.dpe_fuse(ctx.h, fs.views, fs.n, C.byref(fp), None, 0, C.byref(cnt)))
        assert cnt.value == 16 * 12
    finally:
        ctx.close()
