"""Edge and label maps consumed by the weak-texture path.

The reference computes them on the CPU with OpenCV (`EdgeSegment`,
DPE.cpp:136-291: Canny for the edge map, Roberts + HoughLinesP + connected
components for the label map).  OpenCV is absent, so this module is the
project's own documented detector; it keeps the reference's map semantics and
file formats, not its pixel values (parity with OpenCV is unpinned):
  * edge map: uint8, 255 on edges, 0 elsewhere.  Sobel gradient magnitude of
    the image above the order statistic k = int(edge_quantile * (n - 1)) of the
    n magnitudes (at least 16), with the one-pixel border copy rule of
    DPE.cpp:262-273.
  * label map: int32, 0 on edge pixels, -1 on regions of at most
    weak_tex_num = rows*cols / (1024 << scale << scale) pixels (DPE.cpp:140,
    282-284), region index > 0 elsewhere.  Edges for the labels come from a
    Roberts cross (DPE.cpp:9-25) on a 4x box-downsampled image, threshold
    robthr = 4 (DPE.cpp:139), upsampled by nearest neighbour; regions are
    4-connected components of non-edge pixels (Connect, DPE.cpp:28-134).
The C++ host driver restates both maps (dpe-mvs-fast_amd/host/dpe_hostio.cpp)
and a CPU test checks that the two agree bit for bit.
"""
import numpy as np
from scipy import ndimage


def _border_rule(e):
    # DPE.cpp:262-273: first/last column and row copy the "not edge" state inwards
    e = e.copy()
    e[:, 0] = np.where(e[:, 1] == 0, 0, e[:, 0])
    e[:, -1] = np.where(e[:, -2] == 0, 0, e[:, -1])
    e[0, :] = np.where(e[1, :] == 0, 0, e[0, :])
    e[-1, :] = np.where(e[-2, :] == 0, 0, e[-1, :])
    return e


def edge_map(image_u8, edge_quantile=0.85):
    img = image_u8.astype(np.float32)
    gx = ndimage.sobel(img, axis=1, mode="nearest")
    gy = ndimage.sobel(img, axis=0, mode="nearest")
    mag = np.hypot(gx, gy)
    flat = mag.ravel()
    k = int(edge_quantile * (flat.size - 1))
    thr = max(float(np.partition(flat, k)[k]), 16.0)
    e = np.where(mag > thr, 255, 0).astype(np.uint8)
    return _border_rule(e)


def label_map(image_u8, scale=0, robthr=4):
    H, W = image_u8.shape
    h4, w4 = max(1, H // 4), max(1, W // 4)
    small = image_u8[:h4 * 4, :w4 * 4].astype(np.float32).reshape(h4, 4, w4, 4).mean(axis=(1, 3))
    s = np.rint(small).astype(np.int32)
    t1 = np.zeros_like(s)
    t2 = np.zeros_like(s)
    t1[:-1, :-1] = s[:-1, :-1] - s[1:, 1:]
    t2[:-1, :-1] = s[1:, :-1] - s[:-1, 1:]
    rob = np.sqrt((t1 * t1 + t2 * t2).astype(np.float32)).astype(np.int32)
    rob[0, :] = rob[-1, :] = 50
    rob[:, 0] = rob[:, -1] = 50
    edges_small = rob > robthr
    yi = np.minimum(np.arange(H) * h4 // H, h4 - 1)
    xi = np.minimum(np.arange(W) * w4 // W, w4 - 1)
    edges = edges_small[yi][:, xi]
    e8 = _border_rule(np.where(edges, 255, 0).astype(np.uint8))
    lab, n = ndimage.label(e8 == 0, structure=[[0, 1, 0], [1, 1, 1], [0, 1, 0]])
    lab = lab.astype(np.int32)
    counts = np.bincount(lab.ravel(), minlength=n + 1)
    weak_tex_num = int(1.0 * H * W / (1024 << scale << scale))
    small_region = (counts[lab] <= weak_tex_num) & (lab != 0)
    lab = np.where(small_region, -1, lab).astype(np.int32)
    return lab


def support_maps(image_u8, scale=0):
    """(edges, labels) of one reference image at one scale."""
    return edge_map(image_u8), label_map(image_u8, scale)
