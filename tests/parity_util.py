"""Comparison helpers for the HIP path against the oracle."""
import numpy as np


def depth_agreement(d_gpu, d_ref, tol=1e-3):
    """Fraction of pixels with |d_gpu - d_ref| / d_ref < tol (both zero counts as agreement)."""
    a = np.asarray(d_gpu, np.float64)
    b = np.asarray(d_ref, np.float64)
    both_zero = (a == 0) & (b == 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.abs(a - b) / np.abs(b)
    ok = both_zero | (rel < tol)
    return float(ok.mean())


def normal_agreement(n_gpu, n_ref, max_deg=0.5):
    a = np.asarray(n_gpu, np.float64).reshape(-1, 3)
    b = np.asarray(n_ref, np.float64).reshape(-1, 3)
    dot = np.sum(a * b, -1) / np.maximum(np.linalg.norm(a, axis=-1) * np.linalg.norm(b, axis=-1), 1e-30)
    ang = np.degrees(np.arccos(np.clip(dot, -1, 1)))
    same = np.all(a == b, -1)
    return float((same | (ang <= max_deg)).mean())
