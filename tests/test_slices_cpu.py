"""Depth-plane slice images on the CPU: the oracle's TransientImage::save
restatement (FB/TransientImage.cpp:27-45, FB/Image.h:45-177) pinned by known
answers and a float64 bilinear check, and the library's PNG writer checked by
decoding its files."""
import ctypes as C

import numpy as np
import pytest

import oracle
from fasterbackprojection_amd import _lib
from pngio import chunks, read_rgba


def test_gray_bytes_and_flip_known_answers():
    # 2 x 2 pixels (p = r * W + c), one plane; no resize.
    act = np.array([[0.5], [1.0], [0.999], [-1.0]], np.float32)
    img = oracle.depth_slices_rgba(act, 2, 2)[0]
    # 0.5 * 255 = 127.5 -> 127; 1.0 -> 255; 0.999 * 255 = 254.745 -> 254;
    # negative -> 0.  Row 0 of the activation is the bottom row of the image.
    assert img[1, :, 0].tolist() == [127, 255]
    assert img[0, :, 0].tolist() == [254, 0]
    assert (img[..., 0] == img[..., 1]).all() and (img[..., 0] == img[..., 2]).all()
    assert (img[..., 3] == 255).all()


def test_gray_bytes_saturate_and_nan():
    act = np.array([[2.0, np.nan, 1.0 / 255.0, 0.0]], np.float32).T.reshape(4, 1)
    img = oracle.depth_slices_rgba(act.reshape(4, 1), 4, 1)[0]
    assert img[0, :, 0].tolist() == [255, 0, 1, 0]


def test_plane_order_is_depth_fastest():
    # act[p][d]: plane d takes element d of every pixel row.
    act = np.array([[0.0, 1.0], [1.0, 0.0]], np.float32)  # W = 2, H = 1, D = 2
    img = oracle.depth_slices_rgba(act, 2, 1)
    assert img[0, 0, :, 0].tolist() == [0, 255]
    assert img[1, 0, :, 0].tolist() == [255, 0]


def test_resize_half_steps_known_answer():
    # W = H = 3 -> 5: ratio 0.5 exactly, samples at 0, .5, 1, 1.5, 2 on both
    # axes.  Rows of the flipped image are identical, so only x matters:
    # values 0, 100, 201 -> 0, 50, 100, 150 (150.5 truncated), 201.
    row = np.array([0.0, 100.0, 201.0], np.float32) / 255.0
    act = np.tile(row, 3).reshape(9, 1)
    img = oracle.depth_slices_rgba(act, 3, 3, 5, 5)[0]
    gray = oracle.depth_slices_rgba(act, 3, 3)[0][0, :, 0]
    assert gray.tolist() == [0, 100, 201]
    for y in range(5):
        assert img[y, :, 0].tolist() == [0, 50, 100, 150, 201]
    assert (img[..., 3] == 255).all()


@pytest.mark.parametrize("W,H,OW,OH", [(16, 16, 256, 256), (12, 20, 48, 80), (7, 5, 28, 20)])
def test_resize_matches_float64_bilinear(W, H, OW, OH):
    rng = np.random.default_rng(W * 1000 + OW)
    act = rng.random((W * H, 2), dtype=np.float32)
    img = oracle.depth_slices_rgba(act, W, H, OW, OH)
    # The unresized images (gray bytes and flip pinned above).
    base = oracle.depth_slices_rgba(act, W, H)
    for d in range(2):
        g = base[d, :, :, 0].astype(np.float64)
        xs = np.arange(OW) * (W - 1) / (OW - 1)
        ys = np.arange(OH) * (H - 1) / (OH - 1)
        x0 = np.floor(xs).astype(int)
        y0 = np.floor(ys).astype(int)
        x1 = np.minimum(x0 + 1, W - 1)
        y1 = np.minimum(y0 + 1, H - 1)
        dx = (xs - x0)[None, :]
        dy = (ys - y0)[:, None]
        v = ((1 - dx) * (1 - dy) * g[np.ix_(y0, x0)] + dx * (1 - dy) * g[np.ix_(y0, x1)]
             + (1 - dx) * dy * g[np.ix_(y1, x0)] + dx * dy * g[np.ix_(y1, x1)])
        got = img[d, :, :, 0].astype(np.int64)
        clear = np.abs(v - np.round(v)) > 1e-3
        # Away from integers the FP32 expression truncates to the same byte.
        assert (got[clear] == np.floor(v[clear])).all()
        # Next to an integer it may land on either side.
        near = ~clear
        assert (np.abs(got[near] - np.round(v[near])) <= 1).all()
        assert np.isin(img[d, :, :, 3], [254, 255]).all()


def _write_png(path, rgba):
    lib = _lib.load()
    a = np.ascontiguousarray(rgba, np.uint8)
    return lib.fbp_write_png_rgba(str(path).encode(), a.ctypes.data_as(C.c_void_p),
                                  a.shape[1], a.shape[0])


def test_png_writer_round_trip(tmp_path):
    rgba = np.random.default_rng(3).integers(0, 256, (23, 37, 4), dtype=np.uint8)
    p = tmp_path / "x.png"
    assert _write_png(p, rgba) == 0
    kinds = [k for k, _ in chunks(p.read_bytes())]
    assert kinds[0] == b"IHDR" and kinds[-1] == b"IEND" and b"IDAT" in kinds
    assert (read_rgba(p) == rgba).all()


def test_png_writer_errors(tmp_path):
    lib = _lib.load()
    rgba = np.zeros((2, 2, 4), np.uint8)
    assert lib.fbp_write_png_rgba(str(tmp_path / "z.png").encode(),
                                  rgba.ctypes.data_as(C.c_void_p), 0, 2) == 1
    assert _write_png(tmp_path / "missing" / "dir" / "z.png", rgba) == 7
