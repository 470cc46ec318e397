"""ctypes mirror of include/dpe_host.h (libdpe_host.so: the C++ driver's file
formats, image pyramid and support maps; CPU only) and the driver binary."""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
HOST_LIB_PATH = os.path.join(PKG, "libdpe_host.so")
DRIVER_PATH = os.path.join(PKG, "dpe_driver")

CV_8UC1, CV_32SC1, CV_32FC1, CV_32FC3 = 0, 4, 5, 21
_NP = {CV_8UC1: (np.uint8, 1), CV_32SC1: (np.int32, 1), CV_32FC1: (np.float32, 1), CV_32FC3: (np.float32, 3)}

# Every symbol include/dpe_host.h declares.
EXPORTS = ["dpe_host_type_size", "dpe_host_write_bin_mat", "dpe_host_read_bin_mat", "dpe_host_read_camera",
           "dpe_host_read_pairs", "dpe_host_resize_linear", "dpe_host_rescale_nearest", "dpe_host_support_maps",
           "dpe_host_write_ply", "dpe_host_edge_segment", "dpe_host_roberts", "dpe_host_connect", "dpe_host_canny",
           "dpe_host_hough_lines_p", "dpe_host_read_jpeg"]


def load(path=HOST_LIB_PATH):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(path)
    P = C.POINTER
    lib.dpe_host_type_size.argtypes = [C.c_int]
    lib.dpe_host_type_size.restype = C.c_int
    lib.dpe_host_write_bin_mat.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
    lib.dpe_host_write_bin_mat.restype = C.c_int
    lib.dpe_host_read_bin_mat.argtypes = [C.c_char_p, P(C.c_int), P(C.c_int), P(C.c_int), C.c_void_p, C.c_size_t]
    lib.dpe_host_read_bin_mat.restype = C.c_int
    lib.dpe_host_read_camera.argtypes = [C.c_char_p, P(abi.Camera)]
    lib.dpe_host_read_camera.restype = C.c_int
    lib.dpe_host_read_pairs.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.dpe_host_read_pairs.restype = C.c_int
    lib.dpe_host_resize_linear.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int]
    lib.dpe_host_resize_linear.restype = None
    lib.dpe_host_rescale_nearest.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int]
    lib.dpe_host_rescale_nearest.restype = None
    lib.dpe_host_support_maps.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    lib.dpe_host_support_maps.restype = None
    lib.dpe_host_write_ply.argtypes = [C.c_char_p, C.c_void_p, C.c_int64]
    lib.dpe_host_write_ply.restype = C.c_int
    lib.dpe_host_edge_segment.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_void_p, C.c_void_p, P(C.c_int), P(C.c_int)]
    lib.dpe_host_edge_segment.restype = C.c_int
    lib.dpe_host_roberts.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    lib.dpe_host_roberts.restype = None
    lib.dpe_host_connect.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
    lib.dpe_host_connect.restype = C.c_int
    lib.dpe_host_canny.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p]
    lib.dpe_host_canny.restype = None
    lib.dpe_host_hough_lines_p.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                           C.c_int, C.c_void_p, C.c_int]
    lib.dpe_host_hough_lines_p.restype = C.c_int
    lib.dpe_host_read_jpeg.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_size_t, P(C.c_int), P(C.c_int)]
    lib.dpe_host_read_jpeg.restype = C.c_int
    return lib


def read_jpeg(lib, path, color=False):
    """Decode a JPEG as the driver does: grayscale (H, W) u8, or colour (H, W, 3) BGR."""
    w, h = C.c_int(), C.c_int()
    st = lib.dpe_host_read_jpeg(str(path).encode(), int(color), None, 0, C.byref(w), C.byref(h))
    if st != 0:
        raise ValueError(f"cannot decode {path}: status {st}")
    out = np.zeros((h.value, w.value, 3) if color else (h.value, w.value), np.uint8)
    st = lib.dpe_host_read_jpeg(str(path).encode(), int(color), out.ctypes.data, out.nbytes, C.byref(w), C.byref(h))
    if st != 0:
        raise ValueError(f"cannot decode {path}: status {st}")
    return out


def edge_segment(lib, image, scale, mode, use_canny, high_res=False):
    """EdgeSegment restatement: mode 0 -> edges (u8, image size), mode 1 -> labels
    (i32 at round(size / 2^scale))."""
    image = np.ascontiguousarray(image, np.uint8)
    h, w = image.shape
    f = 1.0 / (1 << scale) if mode == 1 else 1.0
    oh, ow = int(round(h * f)), int(round(w * f))
    edges = np.zeros((oh, ow), np.uint8)
    labels = np.zeros((oh, ow), np.int32)
    gw, gh = C.c_int(), C.c_int()
    st = lib.dpe_host_edge_segment(scale, image.ctypes.data, w, h, mode, int(use_canny), int(high_res),
                                   edges.ctypes.data, labels.ctypes.data, C.byref(gw), C.byref(gh))
    if st != 0:
        raise ValueError("dpe_host_edge_segment rejected its arguments")
    assert (gw.value, gh.value) == (ow, oh)
    return edges if mode == 0 else labels


def read_bin_mat(path):
    """A .dmb / .bin matrix (ReadBinMat layout) as a NumPy array, read in Python."""
    with open(path, "rb") as f:
        version, rows, cols, typ = np.frombuffer(f.read(16), np.int32)
        if version != 1:
            raise ValueError(f"{path}: version {version}")
        dt, ch = _NP[int(typ)]
        data = np.frombuffer(f.read(), dt)
    shape = (rows, cols) if ch == 1 else (rows, cols, ch)
    return data.reshape(shape).copy()


def write_bin_mat(path, arr):
    """Write a NumPy array with the WriteBinMat layout."""
    arr = np.ascontiguousarray(arr)
    if arr.dtype == np.uint8:
        typ = CV_8UC1
    elif arr.dtype == np.int32 or arr.dtype == np.uint32:
        typ = CV_32SC1
    elif arr.dtype == np.float32 and arr.ndim == 3 and arr.shape[2] == 3:
        typ = CV_32FC3
    elif arr.dtype == np.float32:
        typ = CV_32FC1
    else:
        raise ValueError(f"unsupported dtype {arr.dtype}")
    with open(path, "wb") as f:
        f.write(np.array([1, arr.shape[0], arr.shape[1], typ], np.int32).tobytes())
        f.write(arr.tobytes())


def support_maps(lib, image_u8, scale=0):
    img = np.ascontiguousarray(image_u8, np.uint8)
    H, W = img.shape
    edges = np.zeros((H, W), np.uint8)
    labels = np.zeros((H, W), np.int32)
    lib.dpe_host_support_maps(img.ctypes.data, W, H, scale, edges.ctypes.data, labels.ctypes.data)
    return edges, labels


def read_ply(path):
    """Vertices of a binary PLY written by ExportPointCloud's layout."""
    with open(path, "rb") as f:
        data = f.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    header = data[:end].decode()
    n = int([ln for ln in header.splitlines() if ln.startswith("element vertex")][0].split()[-1])
    rec = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("b", "u1"), ("g", "u1"), ("r", "u1")])
    return header, np.frombuffer(data[end:end + n * 15], rec)
