"""ctypes mirror of include/dpe_hip.h (the C-ABI of libdpe_hip.so).

The structs are byte-identical to the C declarations; `dpe_camera` is the
reference's 112-byte `Camera` (main.h:53-62) and `dpe_params` mirrors
`PatchMatchParams` (main.h:81-109) with bools widened to int32.
"""
import ctypes as C

MAX_IMAGES = 32
NEIGHBOUR_NUM = 9

# dpe_status
OK, ERR_INVALID_ARGUMENT, ERR_TOO_MANY_IMAGES, ERR_MISSING_INPUT, ERR_HIP, ERR_NO_DEVICE, ERR_UNSUPPORTED, ERR_INTERNAL = range(8)
STATUS_NAMES = {0: "DPE_OK", 1: "DPE_ERR_INVALID_ARGUMENT", 2: "DPE_ERR_TOO_MANY_IMAGES",
                3: "DPE_ERR_MISSING_INPUT", 4: "DPE_ERR_HIP", 5: "DPE_ERR_NO_DEVICE", 6: "DPE_ERR_UNSUPPORTED",
                7: "DPE_ERR_INTERNAL"}
# RunState (main.h:69-73), PixelState (main.h:75-79)
FIRST_INIT, REFINE_INIT, REFINE_ITER = 0, 1, 2
WEAK, STRONG, UNKNOWN = 0, 1, 2
MEM_HOST, MEM_DEVICE = 0, 1


class Camera(C.Structure):
    _fields_ = [("K", C.c_float * 9), ("R", C.c_float * 9), ("t", C.c_float * 3), ("c", C.c_float * 3),
                ("height", C.c_int32), ("width", C.c_int32), ("depth_min", C.c_float), ("depth_max", C.c_float)]


class Params(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("num_images", C.c_int32), ("sigma_spatial", C.c_float),
                ("sigma_color", C.c_float), ("top_k", C.c_int32), ("depth_min", C.c_float),
                ("depth_max", C.c_float), ("geom_consistency", C.c_int32), ("strong_radius", C.c_int32),
                ("strong_increment", C.c_int32), ("weak_radius", C.c_int32), ("weak_increment", C.c_int32),
                ("use_APD", C.c_int32), ("use_edge", C.c_int32), ("use_limit", C.c_int32),
                ("use_label", C.c_int32), ("use_radius", C.c_int32), ("high_res_img", C.c_int32),
                ("max_scale_size", C.c_int32), ("scale_size", C.c_int32), ("weak_peak_radius", C.c_int32),
                ("rotate_time", C.c_int32), ("ransac_threshold", C.c_float), ("geom_factor", C.c_float),
                ("state", C.c_int32), ("call_index", C.c_int32), ("seed", C.c_uint64)]


class Inputs(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("num_images", C.c_int32), ("memory", C.c_int32),
                ("images", C.c_void_p * MAX_IMAGES), ("cameras", Camera * MAX_IMAGES),
                ("geom_depths", C.c_void_p * MAX_IMAGES), ("weak_info", C.c_void_p),
                ("prior_planes", C.c_void_p), ("prior_selected", C.c_void_p), ("edges", C.c_void_p),
                ("edges_low", C.c_void_p), ("low_width", C.c_int32), ("low_height", C.c_int32),
                ("labels", C.c_void_p)]


class Outputs(C.Structure):
    _fields_ = [("planes", C.c_void_p), ("costs", C.c_void_p), ("pixel_states", C.c_void_p),
                ("selected_views", C.c_void_p)]


NCC_FAMILIES = 4
FAMILY_FULL, FAMILY_REFINE, FAMILY_POST, FAMILY_WEAK = range(NCC_FAMILIES)
ABI_VERSION = 11


class Stats(C.Structure):
    _fields_ = [("gpu_ms", C.c_float), ("ncc_ms", C.c_float), ("ncc_hyps", C.c_int64), ("ncc_taps", C.c_int64),
                ("ncc_launches", C.c_int32), ("family_ms", C.c_float * NCC_FAMILIES),
                ("family_hyps", C.c_int64 * NCC_FAMILIES), ("family_taps", C.c_int64 * NCC_FAMILIES),
                ("family_launches", C.c_int32 * NCC_FAMILIES), ("quad_bytes", C.c_int32),
                ("tile_full", C.c_int32), ("tile_weighted", C.c_int32), ("graph", C.c_int32)]


PRECISION_FP32, PRECISION_FP16, PRECISION_BF16 = 0, 1, 2


class Form(C.Structure):
    _fields_ = [("gauss_newton_steps", C.c_int32), ("precision", C.c_int32), ("autotune", C.c_int32),
                ("graph", C.c_int32), ("reserved", C.c_int32 * 4)]


class FuseView(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("camera", Camera), ("depth", C.c_void_p),
                ("normal", C.c_void_p), ("pixel_states", C.c_void_p), ("color", C.c_void_p),
                ("num_src", C.c_int32), ("src_index", C.c_int32 * MAX_IMAGES)]


class FuseParams(C.Structure):
    _fields_ = [("max_reproj_px", C.c_float), ("max_rel_depth", C.c_float), ("max_angle_rad", C.c_float),
                ("strong_factor", C.c_float), ("weak_factor", C.c_float), ("min_consistent", C.c_int32),
                ("memory", C.c_int32), ("fp64_reprojection", C.c_int32)]


TAT_INTERMEDIATE, TAT_ADVANCED = 0, 1


class FuseTatParams(C.Structure):
    _fields_ = [("variant", C.c_int32), ("dist_base", C.c_float), ("depth_base", C.c_float),
                ("angle_base", C.c_float), ("angle_grad", C.c_float), ("memory", C.c_int32)]


class Point(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("b", C.c_uint8), ("g", C.c_uint8),
                ("r", C.c_uint8), ("pad", C.c_uint8)]


assert C.sizeof(Camera) == 112
assert C.sizeof(Params) == 112


def default_params(**kw):
    """PatchMatchParams defaults (main.h:81-109)."""
    p = Params(max_iterations=3, num_images=5, sigma_spatial=5.0, sigma_color=3.0, top_k=4, depth_min=0.0,
               depth_max=1.0, geom_consistency=0, strong_radius=5, strong_increment=2, weak_radius=5,
               weak_increment=5, use_APD=1, use_edge=1, use_limit=1, use_label=1, use_radius=1, high_res_img=1,
               max_scale_size=1, scale_size=1, weak_peak_radius=2, rotate_time=4, ransac_threshold=0.005,
               geom_factor=0.2, state=FIRST_INIT, call_index=0, seed=0x5EED0002)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def fuse_defaults(**kw):
    """RunFusion thresholds (DPE.cpp:1508, 1518-1519)."""
    f = FuseParams(max_reproj_px=2.0, max_rel_depth=0.01, max_angle_rad=0.174533, strong_factor=0.3,
                   weak_factor=0.45, min_consistent=1, memory=MEM_HOST)
    for k, v in kw.items():
        setattr(f, k, v)
    return f


def fuse_tat_defaults(variant, **kw):
    """RunFusion_TAT_Intermediate / RunFusion_TAT_advanced constants (DPE.cpp:1555-1559, 1725-1726),
    with 1.0f / 3500.0f and 1.0f / 3000.0f evaluated in float as the reference does."""
    import numpy as np
    depth_base = np.float32(1.0) / np.float32(3500.0 if variant == TAT_INTERMEDIATE else 3000.0)
    f = FuseTatParams(variant=variant, dist_base=0.25, depth_base=float(depth_base),
                      angle_base=0.06981317007977318, angle_grad=0.05235987755982988, memory=MEM_HOST)
    for k, v in kw.items():
        setattr(f, k, v)
    return f


def camera_from_arrays(K, R, t, width, height, dmin, dmax):
    """Camera with c = -R^T t computed in double, as ReadCamera does (DPE.cpp:364-367)."""
    import numpy as np
    cam = Camera()
    K = np.asarray(K, np.float32).reshape(9)
    R = np.asarray(R, np.float32).reshape(9)
    t = np.asarray(t, np.float32).reshape(3)
    for i in range(9):
        cam.K[i] = float(K[i])
        cam.R[i] = float(R[i])
    for i in range(3):
        cam.t[i] = float(t[i])
    for j in range(3):
        c = float(R[0 + j]) * float(t[0]) + float(R[3 + j]) * float(t[1]) + float(R[6 + j]) * float(t[2])
        cam.c[j] = float(np.float32(-c))
    cam.width, cam.height = int(width), int(height)
    cam.depth_min, cam.depth_max = float(dmin), float(dmax)
    return cam


def declare(lib):
    """Attach argtypes/restype of the C-ABI entry points to a loaded libdpe_hip."""
    P = C.POINTER
    lib.dpe_create.argtypes = [C.c_int, P(C.c_void_p)]
    lib.dpe_create.restype = C.c_int
    lib.dpe_destroy.argtypes = [C.c_void_p]
    lib.dpe_destroy.restype = None
    lib.dpe_last_error.argtypes = [C.c_void_p]
    lib.dpe_last_error.restype = C.c_char_p
    lib.dpe_abi_version.argtypes = []
    lib.dpe_abi_version.restype = C.c_int
    lib.dpe_process.argtypes = [C.c_void_p, P(Params), P(Inputs), P(Outputs)]
    lib.dpe_process.restype = C.c_int
    lib.dpe_process_ex.argtypes = [C.c_void_p, P(Params), C.c_void_p, P(Inputs), P(Outputs)]
    lib.dpe_process_ex.restype = C.c_int
    lib.dpe_last_stats.argtypes = [C.c_void_p, P(Stats)]
    lib.dpe_last_stats.restype = C.c_int
    lib.dpe_ncc_cost.argtypes = [C.c_void_p, P(Params), P(Inputs), C.c_int, C.c_void_p, C.c_void_p]
    lib.dpe_ncc_cost.restype = C.c_int
    lib.dpe_ncc_cost_ex.argtypes = [C.c_void_p, P(Params), C.c_void_p, P(Inputs), C.c_int, C.c_void_p, C.c_void_p]
    lib.dpe_ncc_cost_ex.restype = C.c_int
    lib.dpe_ncc_cost_noalloc.argtypes = [C.c_void_p, P(Params), P(Inputs), C.c_int, C.c_void_p, C.c_void_p]
    lib.dpe_ncc_cost_noalloc.restype = C.c_int
    lib.dpe_ncc_cost_adjoint.argtypes = [C.c_void_p, P(Params), P(Inputs), C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p]
    lib.dpe_ncc_cost_adjoint.restype = C.c_int
    lib.dpe_fuse.argtypes = [C.c_void_p, P(FuseView), C.c_int, P(FuseParams), C.c_void_p, C.c_int64,
                             P(C.c_int64)]
    lib.dpe_fuse.restype = C.c_int
    lib.dpe_fuse_tat.argtypes = [C.c_void_p, P(FuseView), C.c_int, C.c_void_p, P(FuseTatParams), C.c_void_p,
                                 C.c_int64, P(C.c_int64)]
    lib.dpe_fuse_tat.restype = C.c_int
    lib.dpe_batch_create.argtypes = [C.c_int, C.c_int, P(C.c_void_p)]
    lib.dpe_batch_create.restype = C.c_int
    lib.dpe_batch_destroy.argtypes = [C.c_void_p]
    lib.dpe_batch_destroy.restype = None
    lib.dpe_batch_last_error.argtypes = [C.c_void_p]
    lib.dpe_batch_last_error.restype = C.c_char_p
    lib.dpe_batch_process.argtypes = [C.c_void_p, C.c_int, P(Params), C.c_void_p, P(Inputs), P(Outputs),
                                      C.c_void_p, P(C.c_float)]
    lib.dpe_batch_process.restype = C.c_int
    return lib


# Every symbol include/dpe_hip.h declares.
EXPORTS = ["dpe_create", "dpe_destroy", "dpe_last_error", "dpe_abi_version", "dpe_process", "dpe_process_ex",
           "dpe_last_stats", "dpe_ncc_cost", "dpe_ncc_cost_ex", "dpe_ncc_cost_noalloc", "dpe_ncc_cost_adjoint", "dpe_fuse",
           "dpe_fuse_tat", "dpe_batch_create", "dpe_batch_destroy", "dpe_batch_last_error", "dpe_batch_process"]
MAX_LANES = 16
