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


def bitwise_fractions(out_g, out_o):
    """Fractions of pixels whose raw outputs are bit-identical between two
    ProblemOutputs (planes: all four floats; costs; pixel states; selected
    views), compared as integers so that NaNs compare by their bits."""
    def frac(a, b, axis=None):
        a = np.ascontiguousarray(a)
        b = np.ascontiguousarray(b)
        if a.dtype == np.float32:
            a, b = a.view(np.uint32), b.view(np.uint32)
        eq = a == b
        if axis is not None:
            eq = eq.all(axis)
        return float(eq.mean())
    return dict(planes=frac(out_g.planes, out_o.planes, -1), costs=frac(out_g.costs, out_o.costs),
                pixel_states=frac(out_g.pixel_states, out_o.pixel_states),
                selected_views=frac(out_g.selected_views, out_o.selected_views))
