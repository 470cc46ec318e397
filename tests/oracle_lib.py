"""Loader of the CPU oracle (oracle/liboracle_dpe.so) for tests and the bench's
CPU baseline.  The oracle is the checker, never the thing measured or shipped."""
import ctypes as C
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle_dpe.so")


def build_oracle():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
    return ORACLE_LIB


def load_oracle():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "dpe-mvs-fast_amd"))
    from dpe_amd import abi
    if not os.path.exists(ORACLE_LIB):
        build_oracle()
    lib = C.CDLL(ORACLE_LIB)
    P = C.POINTER
    lib.oracle_process.argtypes = [P(abi.Params), P(abi.Inputs), P(abi.Outputs), C.c_int, P(C.c_int64)]
    lib.oracle_process.restype = C.c_int
    lib.oracle_process_fp64.argtypes = [P(abi.Params), P(abi.Inputs), P(abi.Outputs), C.c_int, P(C.c_int64)]
    lib.oracle_process_fp64.restype = C.c_int
    lib.oracle_ncc_cost.argtypes = [P(abi.Params), P(abi.Inputs), C.c_int, C.c_void_p, C.c_void_p, C.c_int]
    lib.oracle_ncc_cost.restype = C.c_int
    lib.oracle_ncc_cost_fp64.argtypes = lib.oracle_ncc_cost.argtypes
    lib.oracle_ncc_cost_fp64.restype = C.c_int
    lib.oracle_fuse.argtypes = [P(abi.FuseView), C.c_int, P(abi.FuseParams), C.c_void_p, C.c_int64, P(C.c_int64)]
    lib.oracle_fuse.restype = C.c_int
    lib.oracle_fuse_tat.argtypes = [P(abi.FuseView), C.c_int, C.c_void_p, P(abi.FuseTatParams), C.c_void_p,
                                    C.c_int64, P(C.c_int64)]
    lib.oracle_fuse_tat.restype = C.c_int
    lib.oracle_expf.argtypes = [C.c_float]
    lib.oracle_expf.restype = C.c_float
    lib.oracle_exp.argtypes = [C.c_double]
    lib.oracle_exp.restype = C.c_double
    lib.oracle_acosf.argtypes = [C.c_float]
    lib.oracle_acosf.restype = C.c_float
    lib.oracle_sincosf.argtypes = [C.c_float, P(C.c_float), P(C.c_float)]
    lib.oracle_sincosf.restype = None
    lib.oracle_philox.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
    lib.oracle_philox.restype = None
    lib.oracle_bilinear.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_float]
    lib.oracle_bilinear.restype = C.c_float
    lib.oracle_homography.argtypes = [P(abi.Camera), P(abi.Camera), C.c_void_p, C.c_void_p]
    lib.oracle_homography.restype = None
    lib.oracle_point_in_triangle.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    lib.oracle_point_in_triangle.restype = C.c_int
    lib.oracle_sort_small.argtypes = [C.c_void_p, C.c_int]
    lib.oracle_sort_small.restype = None
    lib.oracle_sort_small_weighted.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    lib.oracle_sort_small_weighted.restype = None
    lib.oracle_find_min_cost_index.argtypes = [C.c_void_p, C.c_int]
    lib.oracle_find_min_cost_index.restype = C.c_int
    lib.oracle_pdf_to_cdf.argtypes = [C.c_void_p, C.c_int]
    lib.oracle_pdf_to_cdf.restype = None
    lib.oracle_nearest_strong.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    lib.oracle_nearest_strong.restype = None
    lib.oracle_filter_strong.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.oracle_filter_strong.restype = C.c_float
    lib.oracle_fit_plane.argtypes = [P(abi.Camera), C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                     C.c_uint64, C.c_void_p]
    lib.oracle_fit_plane.restype = None
    lib.oracle_gen_edge_inform.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                           C.c_void_p, C.c_void_p]
    lib.oracle_gen_edge_inform.restype = None
    lib.oracle_depth_to_weak_state.argtypes = [C.c_void_p, C.c_int]
    lib.oracle_depth_to_weak_state.restype = C.c_uint8
    lib.oracle_geom_cost.argtypes = [P(abi.Camera), P(abi.Camera), C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                     C.c_void_p]
    lib.oracle_geom_cost.restype = C.c_float
    lib.oracle_geom_cost_fp64.argtypes = lib.oracle_geom_cost.argtypes
    lib.oracle_geom_cost_fp64.restype = C.c_float
    lib.oracle_initial_cost.argtypes = [C.c_void_p, C.c_int, P(C.c_uint32)]
    lib.oracle_initial_cost.restype = C.c_float
    lib.oracle_random_normal.argtypes = [P(abi.Camera), C.c_int, C.c_int, C.c_float, C.c_uint64, C.c_uint32,
                                         C.c_uint32, C.c_uint32, C.c_void_p]
    lib.oracle_random_normal.restype = None
    lib.oracle_perturbed_normal.argtypes = [P(abi.Camera), C.c_int, C.c_int, C.c_void_p, C.c_float, C.c_uint64,
                                            C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
    lib.oracle_perturbed_normal.restype = None
    lib.oracle_transform_normal.argtypes = [P(abi.Camera), C.c_void_p, C.c_int, C.c_void_p]
    lib.oracle_transform_normal.restype = None
    lib.oracle_depth_and_normal.argtypes = [P(abi.Camera), C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    lib.oracle_depth_and_normal.restype = None
    lib.oracle_bresenham.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                     C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
    lib.oracle_bresenham.restype = C.c_int
    lib.oracle_direction_search.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    lib.oracle_direction_search.restype = C.c_int
    return lib


ORACLE_FAMILIES = 5  # the GPU's four NCC families + evaluations the reference never reads


def run_oracle(lib, pin, nthreads=8, families=None, fp64=False):
    """One oracle process call on ProblemInputs; returns (ProblemOutputs, (hyps, taps)).
    If `families` is a list, it receives (hyps, taps) of each NCC family.
    fp64: the FP64 cost path (oracle_process_fp64)."""
    from dpe_amd import host
    out = host.ProblemOutputs(pin.width, pin.height)
    cnt = (C.c_int64 * (2 + 2 * ORACLE_FAMILIES))()
    fn = lib.oracle_process_fp64 if fp64 else lib.oracle_process
    st = fn(C.byref(pin.params), C.byref(pin.struct), C.byref(out.struct), nthreads, cnt)
    if st != 0:
        raise RuntimeError(f"oracle_process failed with status {st}")
    if families is not None:
        families[:] = [(cnt[2 + 2 * f], cnt[3 + 2 * f]) for f in range(ORACLE_FAMILIES)]
    return out, (cnt[0], cnt[1])
