"""Convergence pins of the composed PatchMatch in the oracle (no GPU): on a scene
whose answer is known -- one textured plane filling every view, small baseline,
no occlusion -- the propagation, view selection and refinement
(CheckerboardPropagationStrong, RefinementStrong: DPE.cu:1069-1122, 1218-1670)
must find the plane at every interior pixel, DepthToWeak (DPE.cu:2597-2751) must
call those pixels STRONG (one clear cost minimum at the current depth), and
LocalRefine (DPE.cu:2753-2839) must leave converged depths where they are.  A
REFINE call with geometric consistency, the strong-pixel filter and the weak
path switched on must keep them.  A dropped term, a wrong sign or a transposed
index in any of these steps breaks convergence on this scene.

The GPU path equals the oracle bit for bit (tests/test_gpu_parity.py), so these
pins hold for it as well."""
import numpy as np
import pytest

from dpe_amd import abi, host, scene
import oracle_lib


@pytest.fixture(scope="module")
def orc():
    return oracle_lib.load_oracle()


@pytest.fixture(scope="module")
def plane_run(orc):
    s = scene.make_scene(128, 96, 120.0, 5, 20.0, seed=0x5EED00C1, n_planes=1, weak_fraction=0.0, supersample=1)
    store = host.StateStore()
    for ref in range(len(s.views)):
        pin = host.first_call_inputs(s, ref, s.pairs[ref], max_iterations=6)
        out, _ = oracle_lib.run_oracle(orc, pin, 8)
        store.put(ref, host.assemble(out, pin.params.depth_min, pin.params.depth_max))
    interior = np.zeros((s.height, s.width), bool)
    interior[8:-8, 8:-8] = True
    return s, store, interior


def test_first_call_finds_the_plane(plane_run):
    """Views 1-3 have sources on both sides; the end views 0 and 4 see part of
    the plane from one side only and converge a little less."""
    s, store, interior = plane_run
    for ref in range(len(s.views)):
        r = store.get(ref)
        gt = s.views[ref].depth
        rel = np.abs(r["depth"] - gt) / gt
        inner = 0 < ref < len(s.views) - 1
        assert (rel[interior] < 0.01).mean() >= (0.995 if inner else 0.95), ref
        assert (r["pixel_states"][interior] == abi.STRONG).mean() >= (0.995 if inner else 0.99), ref
        # the normals are the plane's within 10 degrees at most pixels (a depth
        # error of 1 % tolerates more normal error than that)
        cosang = (r["normal"] * s.views[ref].normal).sum(-1)
        assert (cosang[interior] > np.cos(np.radians(10))).mean() >= 0.85, ref


def test_refine_call_keeps_the_plane(orc, plane_run):
    s, store, interior = plane_run
    ref = 2
    pin = host.refine_call_inputs(s, ref, s.pairs[ref], store, max_iterations=2, call_index=1)
    assert pin.params.geom_consistency == 1 and pin.params.use_APD == 1
    out, _ = oracle_lib.run_oracle(orc, pin, 8)
    res = host.assemble(out, pin.params.depth_min, pin.params.depth_max)
    gt = s.views[ref].depth
    rel = np.abs(res["depth"] - gt) / gt
    assert (rel[interior] < 0.01).mean() >= 0.999
    assert (res["pixel_states"][interior] == abi.STRONG).mean() >= 0.999


def test_weak_path_recovers_forced_weak_pixels(orc, plane_run):
    """The weak-texture path (GenNeighbours DPE.cu:2107-2467, RANSACToGetFitPlane
    DPE.cu:2895-3128, CheckerboardPropagationWeak / RefinementWeak
    DPE.cu:1672-1902, 1124-1216): interior pixels of the converged plane marked
    WEAK with their depth 30 % off (the strong propagation skips WEAK pixels,
    and LocalRefine's +-5 disparity steps cannot bridge that) get the plane back
    from their strong neighbours in one REFINE iteration."""
    s, store, interior = plane_run
    ref = 2
    rng = np.random.default_rng(5)
    inner = np.zeros_like(interior)
    inner[12:-12, 12:-12] = True
    forced = inner & (rng.uniform(size=inner.shape) < 0.2)
    st = {k: v.copy() for k, v in store.get(ref).items()}
    st["pixel_states"][forced] = abi.WEAK
    st["depth"][forced] *= np.float32(1.3)
    store2 = host.StateStore()
    for i in range(len(s.views)):
        store2.put(i, st if i == ref else store.get(i))
    pin = host.refine_call_inputs(s, ref, s.pairs[ref], store2, max_iterations=1, call_index=1)
    assert (pin.weak_info[forced] == abi.WEAK).all()
    out, _ = oracle_lib.run_oracle(orc, pin, 8)
    res = host.assemble(out, pin.params.depth_min, pin.params.depth_max)
    gt = s.views[ref].depth
    rel = np.abs(res["depth"] - gt) / gt
    assert forced.sum() > 1000
    assert (np.abs(st["depth"] - gt)[forced] / gt[forced] > 0.25).all()
    assert (rel[forced] < 0.01).mean() >= 0.99
