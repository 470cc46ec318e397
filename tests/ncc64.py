"""float64 (or complex128) NumPy restatement of the per-view bilateral NCC cost
ComputeBilateralNCCOld (DPE.cu:696-782) for the 11x11 step-2 window, used as
the exact-arithmetic reference of the GPU tests (adjoint gradients, precision
form).  Tap offsets, bilateral weights, bilinear clamp-to-edge sampling, the
variance floor and the [0, 2] clamp follow the reference; the sums are taken in
double, so this is the cost the reference's formula gives without rounding."""
import warnings

import numpy as np

OFFS = np.array([-5, -3, -1, 1, 3, 5], np.float64)


def cam64(c):
    K = np.array(c.K, np.float64).reshape(3, 3)
    R = np.array(c.R, np.float64).reshape(3, 3)
    t = np.array(c.t, np.float64)
    return K, R, t


def bilinear(im, u, v):
    Hh, Ww = im.shape
    fx = np.floor(u.real)
    fy = np.floor(v.real)
    a = u - fx
    b = v - fy
    ix = np.clip(fx, -1, Ww - 1).astype(np.int64)
    iy = np.clip(fy, -1, Hh - 1).astype(np.int64)
    x0, x1 = np.clip(ix, 0, Ww - 1), np.clip(ix + 1, 0, Ww - 1)
    y0, y1 = np.clip(iy, 0, Hh - 1), np.clip(iy + 1, 0, Hh - 1)
    top = im[y0, x0] + a * (im[y0, x1] - im[y0, x0])
    bot = im[y1, x0] + a * (im[y1, x1] - im[y1, x0])
    return top + b * (bot - top), (ix, iy)


def view_costs64(pix, n, d, imgs, cams, sigma_spatial=5.0, sigma_color=3.0):
    """Per-view costs of the camera-frame planes (n, d) (n . X + d = 0) at the
    pixels pix [(x, y)]; returns ([cost per source view], [state signature per
    view: tap cells, centre validity, variance floor, clamps])."""
    warnings.filterwarnings("ignore", category=RuntimeWarning)  # degenerate views give 0/0, masked below
    x = pix[:, 0].astype(np.float64)
    y = pix[:, 1].astype(np.float64)
    nu = n / d[:, None]
    oi, oj = np.meshgrid(OFFS, OFFS, indexing="ij")
    oi, oj = oi.ravel(), oj.ravel()
    X = x[:, None] + oi[None]
    Y = y[:, None] + oj[None]
    ref = imgs[0].astype(np.float64)
    Hh, Ww = ref.shape
    r = ref[np.clip(Y, 0, Hh - 1).astype(int), np.clip(X, 0, Ww - 1).astype(int)]
    rc = ref[pix[:, 1], pix[:, 0]]
    w = np.exp(-np.sqrt(oi ** 2 + oj ** 2)[None] / (2 * sigma_spatial ** 2)
               - np.abs(r - rc[:, None]) / (2 * sigma_color ** 2))
    Wt = w.sum(1)
    mu_r = (w * r).sum(1) / Wt
    var_r = (w * r * r).sum(1) / Wt - mu_r ** 2
    Kr, Rr_, tr_ = cam64(cams[0])
    Kr_inv = np.linalg.inv(Kr)
    Cr = -Rr_.T @ tr_
    costs, sig = [], []
    for k in range(1, len(imgs)):
        Ks, Rs, ts = cam64(cams[k])
        Cs = -Rs.T @ ts
        Rrel = Rs @ Rr_.T
        trel = Rs @ (Cr - Cs)
        Gm = Rrel[None] - trel[None, :, None] * nu[:, None, :]
        Hm = Ks[None] @ Gm @ Kr_inv[None]
        hc = Hm @ np.stack([x, y, np.ones_like(x)], -1)[..., None]
        qcx, qcy = hc[:, 0, 0] / hc[:, 2, 0], hc[:, 1, 0] / hc[:, 2, 0]
        valid = (qcx.real >= 0) & (qcx.real < Ww) & (qcy.real >= 0) & (qcy.real < Hh)
        u = Hm[:, 0, 0, None] * X + Hm[:, 0, 1, None] * Y + Hm[:, 0, 2, None]
        v = Hm[:, 1, 0, None] * X + Hm[:, 1, 1, None] * Y + Hm[:, 1, 2, None]
        ww = Hm[:, 2, 0, None] * X + Hm[:, 2, 1, None] * Y + Hm[:, 2, 2, None]
        src = imgs[k].astype(np.float64)
        sv, cells = bilinear(src, u / ww, v / ww)
        mu_s = (w * sv).sum(1) / Wt
        var_s = (w * sv * sv).sum(1) / Wt - mu_s ** 2
        cov = (w * r * sv).sum(1) / Wt - mu_r * mu_s
        c = 1 - cov / np.sqrt(var_r * var_s)
        floor = (var_r < 1e-5) | (var_s.real < 1e-5)
        lo, hi = c.real <= 0, c.real >= 2
        c = np.where(lo, 0.0, np.where(hi, 2.0, c))
        c = np.where(floor | ~valid, 2.0, c)
        costs.append(c)
        sig.append((cells[0], cells[1], valid, floor, lo, hi))
    return costs, sig


def plane_from_tpd(pix, tpd, Kr):
    """(theta, phi, depth) -> camera-frame (n, d) with n = (sin t cos p,
    sin t sin p, -cos t) and d = -depth n . K^-1 [x, y, 1]."""
    x = pix[:, 0].astype(np.float64)
    y = pix[:, 1].astype(np.float64)
    th, ph, D = tpd[:, 0], tpd[:, 1], tpd[:, 2]
    n = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), -np.cos(th)], -1)
    ray = np.stack([(x - Kr[0, 2]) / Kr[0, 0], (y - Kr[1, 2]) / Kr[1, 1], np.ones_like(x)], -1)
    return n, -D * np.sum(n * ray, -1)


def agg_cost64(pix, tpd, imgs, cams):
    """Mean over the source views of the cost of (theta, phi, depth) planes;
    returns the cost and the per-view state signatures."""
    n, d = plane_from_tpd(pix, tpd, cam64(cams[0])[0])
    costs, sig = view_costs64(pix, n, d, imgs, cams)
    return np.mean(costs, 0), sig
