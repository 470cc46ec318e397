"""Pins of the oracle's weak-pixel neighbour search (no GPU):

* BresenhamLine (DPE.cu:162-248) against Bresenham's integer line algorithm in
  its textbook error-term form (Bresenham 1965; the all-octant integer variant
  of Foley/van Dam and Rosetta Code: err = (dx > dy ? dx : -dy) / 2, x steps
  while e2 > -dx, y steps while e2 < dy) on every endpoint pair of small grids.
  The reference's walk differs from it in three documented ways, each checked
  on its own: it starts one step after the start point and takes one step past
  the end point; for steep segments (dy > dx) it starts its error at +dy/2
  instead of -dy/2, so its first step is diagonal and a vertical segment is
  walked one column to the side (sx = -1 when the columns are equal); it caps
  each walk at round(max(h, w) / 60) steps for high-resolution images, maps the
  endpoints to the low-resolution map by round(x * low_w / W) and reads a pixel
  past the right border as the first pixel of the next row.  The boolean result
  is an edge on either walk, and false when an endpoint is an edge at full
  resolution.
* The ray search of GenNeighbours (DPE.cu:2170-2222) against a brute-force
  statement of it evaluated in float64 with fixed random draws: directions at
  45 / rotate_time degree steps from the 8 base directions, radii 2, 4, 8, ...
  (min(2r, r + 25)), 4 jittered probes per radius with the unsigned shift of
  Q5, a probe's nearest strong pixel when it is not STRONG, acceptance within
  half the rotation angle and across no edge.
"""
import ctypes as C
import math

import numpy as np
import pytest

from dpe_amd import abi
import oracle_lib


@pytest.fixture(scope="module")
def orc():
    return oracle_lib.load_oracle()


def textbook_line(x0, y0, x1, y1, err0=None):
    """Pixels of Bresenham's line from (x0, y0) to (x1, y1), both included
    (integer error-term form).  err0 overrides the initial error term."""
    dx, sx = abs(x1 - x0), (1 if x0 < x1 else -1)
    dy, sy = abs(y1 - y0), (1 if y0 < y1 else -1)
    err = (dx if dx > dy else -dy) // 2 if dx > dy else -(dy // 2)
    if err0 is not None:
        err = err0
    pts = [(x0, y0)]
    while (x0, y0) != (x1, y1):
        e2 = err
        if e2 > -dx:
            err -= dy
            x0 += sx
        if e2 < dy:
            err += dx
            y0 += sy
        pts.append((x0, y0))
        if len(pts) > 4 * (dx + dy + 2):
            break
    return pts


def textbook_steps(x0, y0, x1, y1, err0, n):
    """The first n pixels after (x0, y0) of the textbook error-term recurrence
    for the segment to (x1, y1), started from error err0 and not stopped at the
    end point."""
    dx, sx = abs(x1 - x0), (1 if x0 < x1 else -1)
    dy, sy = abs(y1 - y0), (1 if y0 < y1 else -1)
    err, pts = err0, []
    for _ in range(n):
        e2 = err
        if e2 > -dx:
            err -= dy
            x0 += sx
        if e2 < dy:
            err += dx
            y0 += sy
        pts.append((x0, y0))
    return pts


def _bres(orc, W, H, edge, lw, lh, low, high_res, a, b):
    """(result, [walk B -> A, walk A -> B]) with each walk the (x, y) it tested."""
    trace = np.zeros(8 * (lw + lh) + 64, np.int64)
    steps = np.zeros(2, np.int32)
    r = orc.oracle_bresenham(W, H, edge.ctypes.data, lw, lh, low.ctypes.data, high_res, a[0], a[1], b[0], b[1],
                             trace.ctypes.data, len(trace), steps.ctypes.data)
    n0, n1 = int(steps[0]), int(steps[1])
    xy = [tuple(int(v) for v in trace[2 * k:2 * k + 2]) for k in range(n0 + n1)]
    return r, [xy[:n0], xy[n0:]]


def test_bresenham_walks_gentle_segments_as_the_textbook_line(orc):
    W, H = 11, 9
    zero = np.zeros((H, W), np.uint8)
    for ax in range(W):
        for ay in range(H):
            for bx in range(W):
                for by in range(H):
                    if (ax, ay) == (bx, by):
                        continue
                    dx, dy = abs(ax - bx), abs(ay - by)
                    r, walks = _bres(orc, W, H, zero, W, H, zero, 0, (ax, ay), (bx, by))
                    assert r == 0
                    for (sx_, sy_), (ex, ey), walk in (((bx, by), (ax, ay), walks[0]),
                                                       ((ax, ay), (bx, by), walks[1])):
                        line = textbook_line(sx_, sy_, ex, ey)
                        got = walk
                        if dx >= dy:
                            # the textbook pixels after the start, then one step past the end
                            assert got[:-1] == line[1:], ((sx_, sy_), (ex, ey), got, line)
                            assert len(got) == len(line)
                            last = got[-1]
                            assert max(abs(last[0] - ex), abs(last[1] - ey)) == 1
                        else:
                            # steep: the same recurrence started from +dy/2; one row per step,
                            # ending one row past its end, and led by up to one column in the
                            # x direction (the textbook line stays within half a column)
                            quirk = textbook_steps(sx_, sy_, ex, ey, dy // 2, len(got))
                            assert all(abs(got[k][1] - got[k - 1][1]) == 1 for k in range(1, len(got)))
                            assert got[-1][1] == ey + (1 if ey > sy_ else -1)
                            sxs = 1 if sx_ < ex else -1
                            for (x, y) in got[:-1]:
                                ideal = sx_ + (ex - sx_) * (y - sy_) / (ey - sy_)
                                assert -0.5 - 1e-9 <= sxs * (x - ideal) <= 1.5 + 1e-9
                            if dy // 2 > -dx:  # all steep segments but the one-row vertical step
                                assert got[0] == (sx_ + (1 if sx_ < ex else -1), sy_ + (1 if sy_ < ey else -1))
                            assert got == quirk


def test_bresenham_vertical_segments_walk_one_column_aside(orc):
    W, H = 9, 12
    zero = np.zeros((H, W), np.uint8)
    r, walks = _bres(orc, W, H, zero, W, H, zero, 0, (4, 2), (4, 9))
    # B -> A from row 9 up to one row past row 2; A -> B from row 2 down to one past row 9; both in column 3
    assert walks[0] == [(3, y) for y in range(8, 0, -1)]
    assert walks[1] == [(3, y) for y in range(3, 11)]


def test_bresenham_result_scaling_cap_and_row_wrap(orc):
    rng = np.random.default_rng(91)
    # full resolution 40 x 30 walked on a 20 x 15 map: endpoints at round(x / 2)
    W, H, lw, lh = 40, 30, 20, 15
    for _ in range(400):
        a = (int(rng.integers(0, W)), int(rng.integers(0, H)))
        b = (int(rng.integers(0, W)), int(rng.integers(0, H)))
        edge = (rng.random((H, W)) < 0.05).astype(np.uint8) * 255
        low = (rng.random((lh, lw)) < 0.08).astype(np.uint8) * 255
        zero_low = np.zeros_like(low)
        r0, walks = _bres(orc, W, H, np.zeros_like(edge), lw, lh, zero_low, 0, a, b)
        sa = [min(int(math.floor(np.float32(a[0]) * np.float32(0.5) + 0.5)), lw - 1),
              min(int(math.floor(np.float32(a[1]) * np.float32(0.5) + 0.5)), lh - 1)]
        sb = [min(int(math.floor(np.float32(b[0]) * np.float32(0.5) + 0.5)), lw - 1),
              min(int(math.floor(np.float32(b[1]) * np.float32(0.5) + 0.5)), lh - 1)]
        if sa != sb and abs(sa[0] - sb[0]) >= abs(sa[1] - sb[1]):
            line = textbook_line(sb[0], sb[1], sa[0], sa[1])
            assert walks[0][:-1] == line[1:]
        # the result: an edge pixel on either walk, read at linear index x + y * low_w
        # (0 outside the map), unless an endpoint is an edge at full resolution
        tested = [x + y * lw for w in walks for (x, y) in w if 0 <= x + y * lw < lw * lh]
        exp = (not edge[a[1], a[0]] and not edge[b[1], b[0]]) and any(low.ravel()[tested])
        r, _ = _bres(orc, W, H, edge, lw, lh, low, 0, a, b)
        assert r == int(exp)
    # high_res_img: each walk stops after round(max(h, w) / 60) steps (128 / 60 -> 2)
    W = H = 128
    zero = np.zeros((H, W), np.uint8)
    _, walks = _bres(orc, W, H, zero, W, H, zero, 1, (10, 10), (100, 40))
    assert len(walks[0]) == len(walks[1]) == 2
    # a gentle walk that ends on the right border steps to x = W, which reads the
    # first pixel of the next row
    W, H = 12, 6
    low = np.zeros((H, W), np.uint8)
    low[3, 0] = 255
    r, walks = _bres(orc, W, H, np.zeros_like(low), W, H, low, 0, (3, 2), (11, 2))
    assert r == 1 and walks[1][-1] == (12, 2)


# ---------------------------------------------------------------- direction search

DIRS = [(-1, -1), (-1, 0), (-1, 1), (0, -1), (0, 1), (1, -1), (1, 0), (1, 1)]


def _shift(sign_draw, mag_draw, shift_range):
    """((sign ? 1 : -1) * magnitude) % shift_range in unsigned 32-bit arithmetic (Q5)."""
    factor = 1 if sign_draw % 2 == 0 else 0xFFFFFFFF
    return (factor * mag_draw) % (1 << 32) % shift_range


def brute_force_search(W, H, weak, nearest, px, py, rotate_time, draws, crosses):
    """Expected slots (index 4 * base + rotation -> (x, y)) of the ray search, or
    'ambiguous' where a float64 value lands within 1e-4 of a decision boundary
    without being on it exactly (an integer pixel coordinate, the image border,
    the angle threshold), where float32 rounding could decide either way.
    Directions along the axes are exact (cos and sin 0 or +-1).""" 
    angle = 45.0 / rotate_time
    thr = math.cos(math.radians(angle / 2))
    shift_range = max(int(math.tan(math.radians(angle / 2)) * 20), 1)
    slots = {}
    probe = 0
    for odi, (ox, oy) in enumerate(DIRS):
        base = math.atan2(oy, ox)
        for rot in range(rotate_time):
            th = base + math.radians(angle) * rot
            od = tuple(float(round(v)) if abs(v - round(v)) < 1e-12 else v for v in (math.cos(th), math.sin(th)))
            di = 4 * odi + rot
            found = None
            ambiguous = False
            radius = 2
            while radius <= 4096 and found is None:  # MAX_SEARCH_RADIUS, main.h
                tx, ty = px + od[0] * radius, py + od[1] * radius
                if any(0 < abs(v) < 1e-4 for v in (tx, ty, tx - W, ty - H)):
                    ambiguous = True
                if tx < 0 or ty < 0 or tx >= W or ty >= H:
                    break
                for _ in range(4):
                    d = [draws[(4 * probe + k) % len(draws)] for k in range(4)]
                    probe += 1
                    rx = _shift(d[0], d[1], shift_range)
                    ry = _shift(d[2], d[3], shift_range)
                    vx, vy = od[0] * 20 + rx, od[1] * 20 + ry
                    nrm = math.hypot(vx, vy)
                    fx, fy = px + vx / nrm * radius, py + vy / nrm * radius
                    if any(0 < abs(v - round(v)) < 1e-4 for v in (fx, fy)):
                        ambiguous = True
                    nx, ny = int(fx), int(fy)  # toward zero
                    if nx < 6 or ny < 6 or nx >= W - 6 or ny >= H - 6:
                        continue
                    if weak[ny, nx] != abi.STRONG:
                        nx, ny = int(nearest[ny, nx, 0]), int(nearest[ny, nx, 1])
                        if nx == -1 or ny == -1:
                            continue
                    tdx, tdy = nx - px, ny - py
                    tn = math.hypot(tdx, tdy)
                    ca = (tdx * od[0] + tdy * od[1]) / tn if tn > 0 else 0.0
                    if abs(ca - thr) < 1e-4:
                        ambiguous = True
                    if ca > thr and not crosses((px, py), (nx, ny)):
                        found = (nx, ny)
                        break
                radius = min(radius * 2, radius + 25)
            slots[di] = "ambiguous" if ambiguous else found
    return slots


def _search(orc, W, H, weak, nearest, edge, rotate_time, edge_limit, px, py, draws):
    pts = np.zeros((64, 2), np.int16)
    valid = np.zeros(64, np.uint8)
    d = np.asarray(draws, np.uint32)
    n = orc.oracle_direction_search(W, H, weak.ctypes.data, nearest.ctypes.data, edge.ctypes.data, edge.ctypes.data,
                                    W, H, 0, rotate_time, edge_limit, px, py, d.ctypes.data, len(d),
                                    pts.ctypes.data, valid.ctypes.data)
    return n, pts, valid


def _nearest(orc, weak):
    H, W = weak.shape
    out = np.zeros((H, W, 2), np.int16)
    orc.oracle_nearest_strong(np.ascontiguousarray(weak).ctypes.data, W, H, out.ctypes.data)
    return out


@pytest.mark.parametrize("rotate_time", [1, 2, 4])
def test_direction_search_matches_brute_force(orc, rotate_time):
    W, H = 96, 80
    rng = np.random.default_rng(92 + rotate_time)
    edge0 = np.zeros((H, W), np.uint8)
    cases = 0
    compared = 0
    for trial in range(12):
        weak = np.full((H, W), abi.STRONG, np.uint8)
        px, py = int(rng.integers(20, W - 20)), int(rng.integers(20, H - 20))
        if trial % 3 != 0:
            # a weak disc around the centre and scattered weak / unknown pixels
            ys, xs = np.mgrid[0:H, 0:W]
            weak[(xs - px) ** 2 + (ys - py) ** 2 <= int(rng.integers(3, 15)) ** 2] = abi.WEAK
            weak[rng.random((H, W)) < 0.15] = abi.UNKNOWN
        nearest = _nearest(orc, weak)
        draws = [0, 0, 0, 0] if trial % 2 == 0 else rng.integers(0, 1 << 32, 4 * int(rng.integers(1, 9))).tolist()
        edge = edge0
        edge_limit = trial % 4 == 3
        if edge_limit:
            edge = edge0.copy()
            edge[:, px + int(rng.integers(4, 12))] = 255  # a wall to the right of the centre

        def crosses(a, b):
            if not edge_limit:
                return False
            r, _ = _bres(orc, W, H, edge, W, H, edge, 0, a, b)
            return r == 1

        exp = brute_force_search(W, H, weak, nearest, px, py, rotate_time, draws, crosses)
        n, pts, valid = _search(orc, W, H, weak, nearest, edge, rotate_time, int(edge_limit), px, py, draws)
        for di, e in exp.items():
            if e == "ambiguous":
                continue
            compared += 1
            if e is None:
                assert valid[di] == 0, (trial, di)
            else:
                assert valid[di] == 1 and tuple(pts[di]) == e, (trial, di, tuple(pts[di]), e)
        assert n == int(valid.sum())
        assert all(valid[4 * k + r] == 0 for k in range(8) for r in range(rotate_time, 4))
        cases += 1
    assert compared >= 0.8 * cases * 8 * rotate_time
