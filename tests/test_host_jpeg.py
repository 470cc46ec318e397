"""The driver's JPEG decoder (dpe_host_read_jpeg; cv::imread's role at
DPE.cpp:713 and 1422) against libjpeg-turbo as PIL links it (no GPU): bit for
bit on grayscale files, on the Y component of colour files (what
IMREAD_GRAYSCALE returns) and on the colour output with 4:4:4, 4:2:2 and 4:2:0
chroma, at sizes that are and are not multiples of the MCU, with standard and
optimised Huffman tables and with restart intervals.  Progressive files are
rejected with a status, not decoded wrongly."""
import numpy as np
import pytest

from dpe_amd import hostio, scene

PIL = pytest.importorskip("PIL.Image")


@pytest.fixture(scope="module")
def lib():
    return hostio.load()


@pytest.fixture(scope="module")
def image():
    return scene.make_scene(160, 120, 150.0, 3, 20.0, seed=0x5EED00A1, supersample=1).views[1].image


def _y_of(path):
    im = PIL.open(path)
    im.draft("L", im.size)  # decode to the Y component, as IMREAD_GRAYSCALE does
    return np.array(im if im.mode == "L" else im.convert("L"))


@pytest.mark.parametrize("size", [(120, 160), (29, 37), (8, 8), (1, 17)])
@pytest.mark.parametrize("quality", [60, 95])
def test_grayscale_files_bit_exact(lib, image, tmp_path, size, quality):
    h, w = size
    g = np.ascontiguousarray(image[:h, :w])
    for opt in (False, True):
        p = tmp_path / "g.jpg"
        PIL.fromarray(g).save(p, quality=quality, optimize=opt)
        assert np.array_equal(hostio.read_jpeg(lib, p), np.array(PIL.open(p)))


@pytest.mark.parametrize("size", [(120, 160), (29, 37), (17, 3)])
@pytest.mark.parametrize("subsampling", [0, 1, 2])
def test_colour_files_bit_exact(lib, image, tmp_path, size, subsampling):
    h, w = size
    g = image[:h, :w]
    col = np.ascontiguousarray(np.stack([g, np.roll(g, 3, 1), 255 - g], -1).astype(np.uint8))
    p = tmp_path / "c.jpg"
    PIL.fromarray(col).save(p, quality=90, subsampling=subsampling)
    assert np.array_equal(hostio.read_jpeg(lib, p), _y_of(p))
    bgr = np.array(PIL.open(p).convert("RGB"))[..., ::-1]
    assert np.array_equal(hostio.read_jpeg(lib, p, color=True), bgr)


def test_restart_intervals(lib, image, tmp_path):
    g = np.ascontiguousarray(image)
    col = np.ascontiguousarray(np.stack([g, g // 2, 255 - g], -1).astype(np.uint8))
    p = tmp_path / "r.jpg"
    try:
        PIL.fromarray(col).save(p, quality=85, restart_marker_blocks=3)
    except TypeError:
        pytest.skip("this PIL cannot write restart markers")
    assert b"\xff\xdd" in p.read_bytes()  # a DRI segment
    assert np.array_equal(hostio.read_jpeg(lib, p), _y_of(p))
    assert np.array_equal(hostio.read_jpeg(lib, p, color=True), np.array(PIL.open(p).convert("RGB"))[..., ::-1])


def test_progressive_and_broken_files_are_rejected(lib, image, tmp_path):
    p = tmp_path / "p.jpg"
    PIL.fromarray(np.ascontiguousarray(image)).save(p, quality=80, progressive=True)
    with pytest.raises(ValueError):
        hostio.read_jpeg(lib, p)
    bad = tmp_path / "bad.jpg"
    bad.write_bytes(b"\xff\xd8\xff\xe0\x00\x10JFIF")
    with pytest.raises(ValueError):
        hostio.read_jpeg(lib, bad)
    with pytest.raises(ValueError):
        hostio.read_jpeg(lib, tmp_path / "missing.jpg")
