"""The reference's support-map segmentation restated without OpenCV
(dpe_host_edge_segment and its parts, DPE.cpp:9-291; no GPU).

OpenCV is absent here, so agreement with an OpenCV build is parity unpinned;
the parts are pinned by closed forms and an independent library instead:
Roberts by its formula (with the reference's 8-bit wrap), Connect by
scipy.ndimage.label on unions of rectangles (the reference's link overwrite
can split a component in noise, which the restatement keeps: it never merges
two components), Canny by step edges (one-pixel-wide response at the step,
nothing below the low threshold), HoughLinesP by drawn segments, and
EdgeSegment by the shapes of its outputs on a synthetic scene image."""
import numpy as np
import pytest
import scipy.ndimage as ndi

from dpe_amd import hostio, scene


@pytest.fixture(scope="module")
def lib():
    return hostio.load()


def test_roberts_closed_form(lib):
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (17, 23)).astype(np.uint8)
    out = np.zeros_like(img)
    lib.dpe_host_roberts(img.ctypes.data, 23, 17, out.ctypes.data)
    i = img.astype(np.int64)
    t1 = i[1:-2, 1:-2] - i[2:-1, 2:-1]
    t2 = i[2:-1, 1:-2] - i[1:-2, 2:-1]
    exp = (np.floor(np.sqrt(t1 * t1 + t2 * t2)).astype(np.int64) % 256).astype(np.uint8)
    assert np.array_equal(out[1:-2, 1:-2], exp)
    # border pixels (first/last row and column) use (50, 50) -> 70
    assert np.all(out[0] == 70) and np.all(out[-1] == 70) and np.all(out[:, 0] == 70) and np.all(out[:, -1] == 70)


def _connect(lib, b):
    h, w = b.shape
    lab = np.zeros((h, w), np.int32)
    cnt = np.zeros(h * w + 1, np.int32)
    n = lib.dpe_host_connect(b.ctypes.data, w, h, lab.ctypes.data, cnt.ctypes.data, h * w + 1)
    return lab, cnt[:n], n


def test_connect_equals_scipy_on_unions_of_rectangles(lib):
    rng = np.random.default_rng(2)
    for _ in range(100):
        h, w = rng.integers(5, 50, 2)
        b = np.full((h, w), 255, np.uint8)
        for _ in range(rng.integers(1, 6)):
            y0, x0 = rng.integers(0, h), rng.integers(0, w)
            b[y0:y0 + rng.integers(1, 12), x0:x0 + rng.integers(1, 12)] = 0
        lab, cnt, n = _connect(lib, b)
        ref, nr = ndi.label(b == 0)
        assert np.array_equal(lab, ref) and n == nr + 1
        assert np.array_equal(cnt, np.bincount(ref.ravel(), minlength=n))


def test_connect_never_merges_components(lib):
    rng = np.random.default_rng(3)
    for _ in range(100):
        h, w = rng.integers(3, 40, 2)
        b = np.where(rng.uniform(size=(h, w)) < rng.uniform(0.2, 0.6), 255, 0).astype(np.uint8)
        lab, cnt, n = _connect(lib, b)
        ref, _ = ndi.label(b == 0)
        assert np.all((lab == 0) == (b == 255))
        for k in range(1, n):
            assert len(np.unique(ref[lab == k])) == 1  # one true component per label
        # labels are numbered in raster order of their first pixel
        first = [np.flatnonzero(lab.ravel() == k)[0] for k in range(1, n)]
        assert first == sorted(first)


def _canny(lib, img, lo, hi):
    h, w = img.shape
    out = np.zeros_like(img)
    lib.dpe_host_canny(np.ascontiguousarray(img).ctypes.data, w, h, lo, hi, out.ctypes.data)
    return out


def test_canny_step_edges(lib):
    img = np.full((30, 40), 50, np.uint8)
    img[:, 20:] = 200
    e = _canny(lib, img, 50, 100)
    cols = np.unique(np.nonzero(e[2:-2])[1])
    assert len(cols) == 1 and cols[0] in (19, 20)  # one pixel wide, at the step
    assert np.all(e[2:-2, cols[0]] == 255)
    # the transposed image gives the transposed response
    assert np.array_equal(_canny(lib, np.ascontiguousarray(img.T), 50, 100), e.T)
    # a step whose gradient (4 * 10 = 40) stays below the low threshold gives nothing
    weak = np.full((30, 40), 50, np.uint8)
    weak[:, 20:] = 60
    assert not _canny(lib, weak, 50, 100).any()
    # hysteresis: a faint continuation of a strong edge is kept, an isolated faint one is not
    # (a step whose contrast falls from 60 to 15 down the image: gradient 240 -> 60)
    img2 = np.full((40, 40), 100, np.uint8)
    img2[:, 20:] += np.linspace(60, 15, 40).astype(np.uint8)[:, None]
    e2 = _canny(lib, img2, 50, 100)
    assert e2[2:-2].any(axis=1).all()
    img3 = np.full((40, 40), 100, np.uint8)
    img3[:, 20:] = 115
    assert not _canny(lib, img3, 50, 100).any()


def test_hough_lines_p_finds_drawn_segments(lib):
    img = np.zeros((80, 100), np.uint8)
    img[40, 10:90] = 255          # horizontal, 80 px
    for t in range(50):           # diagonal, 50 px
        img[10 + t, 5 + t] = 255
    lines = np.zeros((64, 4), np.int32)
    n = lib.dpe_host_hough_lines_p(img.ctypes.data, 100, 80, 1.0, np.pi / 180, 30, 30, 3, lines.ctypes.data, 64)
    segs = lines[:n]
    assert n >= 2
    horiz = [s for s in segs if s[1] == 40 and s[3] == 40]
    assert horiz and min(min(s[0], s[2]) for s in horiz) <= 11 and max(max(s[0], s[2]) for s in horiz) >= 88
    diag = [s for s in segs if abs((s[3] - s[1]) - (s[2] - s[0])) <= 2 and abs(s[2] - s[0]) >= 30]
    assert diag
    empty = np.zeros((80, 100), np.uint8)
    assert lib.dpe_host_hough_lines_p(empty.ctypes.data, 100, 80, 1.0, np.pi / 180, 30, 30, 3, None, 0) == 0


def test_edge_segment_outputs_on_a_scene_image(lib):
    s = scene.make_scene(320, 240, 300.0, 3, 20.0, seed=0x5EED0091)
    img = s.views[1].image
    edges = hostio.edge_segment(lib, img, 0, 0, True)
    assert edges.shape == img.shape and set(np.unique(edges)) <= {0, 255}
    assert 0.005 < (edges == 255).mean() < 0.4
    labels = hostio.edge_segment(lib, img, 1, 1, False)
    assert labels.shape == (120, 160)
    weak_tex_num = int(240 * 320 / (1024 << 1 << 1))
    ids, counts = np.unique(labels[labels > 0], return_counts=True)
    assert len(ids) >= 1 and np.all(counts > weak_tex_num)  # small regions became -1
    assert labels.min() >= -1
    # the weak-texture patches (flat, 20 % of the area) lie mostly inside labelled regions
    weak = s.views[1].weak[::2, ::2]
    assert (labels[weak] > 0).mean() > 0.5
