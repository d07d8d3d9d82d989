"""Pins the refiner oracle (CPU): parameter count, identity behaviour of the
convolution, SSIM(x, x) = 1 and its zero gradient at the optimum, and the
finite-difference helper on a function with a known gradient."""
import numpy as np

from fasterbackprojection_amd import _lib
from oracle import refiner as orf


def test_param_count_matches_library():
    assert orf.param_count() == _lib.load().fbp_refiner_param_count()


def test_conv_identity_and_shift():
    rng = np.random.default_rng(0)
    x = rng.random((8, 8, 8, 2))
    w = np.zeros((27, 2, 2))
    w[13] = np.eye(2)                     # centre tap
    np.testing.assert_allclose(orf.conv(x, w, np.zeros(2), 3), x)
    w = np.zeros((27, 1, 1))
    w[14] = 1.0                           # tap (dz, dy, dx) = (0, 0, +1): out[x] = in[x + 1]
    out = orf.conv(x[..., :1], w, np.zeros(1), 3)
    np.testing.assert_allclose(out[:, :, :-1, 0], x[:, :, 1:, 0])
    assert np.all(out[:, :, -1, 0] == 0)


def test_pool_up_adjoint():
    rng = np.random.default_rng(1)
    a = rng.random((8, 8, 8, 3))
    b = rng.random((4, 4, 4, 3))
    # <pool(a), b> = <a, up(b) / 8>
    assert abs((orf.pool(a) * b).sum() - (a * orf.up(b)).sum() / 8) < 1e-12


def test_ssim_identity():
    rng = np.random.default_rng(2)
    x = rng.random((10, 10, 10))
    assert abs(orf.ssim_mean(x, x) - 1.0) < 1e-12
    assert orf.ssim_mean(x, 1 - x) < 0.5


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, 3.14159265], np.float32)
    r = orf.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0          # tie to even
    assert r[2] == 1.0078125                    # rounds up
    assert abs(r[3] - 3.140625) < 1e-12


def test_forward_shapes_and_normalisation():
    p = np.random.default_rng(3).normal(size=orf.param_count()) * 0.05
    v = np.random.default_rng(4).random((16, 16, 16))
    a = orf.forward(p, v)
    b = orf.forward(p, 7.5 * v)                 # input is max-normalised
    assert a.shape == (16, 16, 16)
    np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-12)


def test_zero_network_is_identity_on_normalised_input():
    """y = net(x) + x: with every parameter zero the output is the input
    divided by its largest magnitude."""
    v = np.random.default_rng(6).random((16, 16, 16)) * 4 - 3
    y = orf.forward(np.zeros(orf.param_count()), v)
    np.testing.assert_allclose(y, v / np.abs(v).max(), rtol=0, atol=1e-15)
