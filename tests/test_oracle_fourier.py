"""Pins the frequency-domain oracles (filter_H, LoGFFT, LCT) with closed forms,
the reference's literal construction steps and known answers (CPU only)."""
import numpy as np
import pytest

import oracle
from oracle import lct as olct


@pytest.mark.parametrize("M", [2, 4, 8, 16, 32, 64])
def test_resampling_closed_form_matches_reference_steps(M):
    ref_mtx, ref_mtxi = olct.resampling_operator_reference(M)
    mtx, mtxi = olct.resampling_operator(M)
    np.testing.assert_allclose(mtx.toarray(), ref_mtx, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(mtxi.toarray(), ref_mtxi, rtol=1e-12, atol=1e-15)
    # mtxi is the transpose (python/lct.py:397 and the column halving)
    np.testing.assert_array_equal(mtxi.toarray(), mtx.toarray().T)


def test_resampling_rows_are_averages():
    """Row r averages 1/sqrt(x) over x in [rM+1, (r+1)M]: row sums are known."""
    M = 32
    mtx, _ = olct.resampling_operator(M)
    x = np.arange(1, M * M + 1)
    expect = (1 / np.sqrt(x)).reshape(M, M).mean(axis=1)
    np.testing.assert_allclose(np.asarray(mtx.sum(axis=1)).ravel(), expect, rtol=1e-12)


def test_psf_invariants():
    U, V, slope = 8, 16, 0.3
    psf = olct.define_psf(U, V, slope)
    assert psf.shape == (2 * V, 2 * U, 2 * U)
    assert abs(np.sum(psf ** 2) - 1) < 1e-12          # unit L2 norm
    vals = np.unique(psf)
    assert len(vals) == 2 and vals[0] == 0            # binary mask, scaled
    ones = psf > 0
    assert np.all(ones.sum(axis=0) >= 1)              # every column has its argmin
    # un-roll and check the argmin against a direct search on one column
    un = np.roll(psf, (0, -U, -U), axis=(0, 1, 2))
    x = np.linspace(-1, 1, 2 * U)
    z = np.linspace(0, 2, 2 * V)
    yi, xi = 3, 11
    c = (4 * slope) ** 2 * (x[xi] ** 2 + x[yi] ** 2)
    assert un[np.argmin(np.abs(c - z)), yi, xi] > 0


@pytest.mark.parametrize("frac", [0.6, 0.8])
def test_lct_point_scatterer_known_answer(frac):
    """A point in front of the wall reconstructs at its (x, y) pixel and at
    the depth index z / (range / 2) * (M - 1) (N = 64, M = 256; coarser
    grids undersample the blur cone)."""
    N, M = 64, 256
    width, bin_res, c = 0.5, 16e-12, 3e8
    rng_max = M * c * bin_res
    xs = np.linspace(-width, width, N)
    x0, y0, z0 = xs[40], xs[22], frac * rng_max / 2
    data = np.zeros((N, N, M))
    for i in range(N):
        for j in range(N):
            r = np.sqrt((xs[i] - x0) ** 2 + (xs[j] - y0) ** 2 + z0 ** 2)
            t = int(round(2 * r / (c * bin_res)))
            if t < M:
                data[i, j, t] = 1.0
    vol = olct.lct(data, width, rng_max, snr=8.0)
    t, y, x = np.unravel_index(np.argmax(vol), vol.shape)
    assert abs(int(x) - 40) <= 1 and abs(int(y) - 22) <= 1
    assert abs(t - z0 / (rng_max / 2) * (M - 1)) <= 2


def test_filter_h_lengths_follow_reference_rule():
    # wl_mean = 10 dt, sigma = wl_mean / sqrt(2): t6 = round(42.43) = 42 -> pad 41
    assert oracle.filter_h_nt_pad(1024, 0.0048, 0.048) == 1024 + 82
    assert oracle.filter_h_nt_pad(256, 0.0096, 0.096) == 256 + 82
    # odd t6 is made even: sigma = 1.25 dt -> round(7.5) = 8 (even) ; sigma = 1.1 dt -> 7 -> 8
    assert oracle.filter_h_nt_pad(100, 1.0, 10.0, 1.1) == 100 + 2 * 7


def test_filter_h_tone_and_dc_response():
    """K = (Gaussian envelope, unit mass) x exp(i 2 pi t / wl), centred at
    t_max / 2 and ifftshifted by N / 2.  A real tone at 1/wl keeps only its
    positive-frequency half: Re(out) / nt_pad = 0.5 cos(w0 (n + N/2)) (the
    N/2 is the carrier phase left by the shift); DC is rejected."""
    dt, wl = 1.0, 10.0
    nt = 400
    N = oracle.filter_h_nt_pad(nt, dt, wl)
    w0 = 2 * np.pi / wl
    n = np.arange(nt)
    tone = np.cos(w0 * n)[None, :]
    out = oracle.filter_h(tone, dt, wl, border=0)
    mid = slice(100, 300)
    expect = 0.5 * np.cos(w0 * (n + N // 2))
    assert np.max(np.abs(out[0, mid] / N - expect[mid])) < 1e-6
    # DC leaks only through the envelope's spectrum at w0: exp(-(w0 sigma)^2 / 2)
    dc = oracle.filter_h(np.ones((1, nt)), dt, wl, border=1)
    sigma = wl / np.sqrt(2)
    assert np.max(np.abs(dc[0, mid])) / N <= 1.01 * np.exp(-(w0 * sigma) ** 2 / 2)


def test_filter_h_linear_and_shift_invariant():
    rng = np.random.default_rng(0)
    # integer inputs: exact in float32, so linearity holds to FP64 rounding
    a = rng.integers(0, 100, (2, 64)).astype(np.float64)
    b = rng.integers(0, 100, (2, 64)).astype(np.float64)
    fa = oracle.filter_h(a, 1.0, 8.0)
    fb = oracle.filter_h(b, 1.0, 8.0)
    np.testing.assert_allclose(oracle.filter_h(2 * a + b, 1.0, 8.0), 2 * fa + fb, rtol=1e-10, atol=1e-7)
    imp = np.zeros((1, 64))
    imp[0, 20] = 1
    imp2 = np.zeros((1, 64))
    imp2[0, 25] = 1
    f1 = oracle.filter_h(imp, 1.0, 8.0)[0]
    f2 = oracle.filter_h(imp2, 1.0, 8.0)[0]
    np.testing.assert_allclose(f1[10:40], f2[15:45], atol=1e-9)


def test_log_fft_known_answers():
    """Constant -> 0; a cosine along x with k cycles -> V * lambda(k/N) * cosine."""
    v = np.full((8, 8, 8), 3.0)
    assert np.max(np.abs(oracle.log_fft(v, 1.0))) < 1e-9
    N, k, s = 16, 3, 1.2
    x = np.arange(N)
    v = np.broadcast_to(np.cos(2 * np.pi * k * x / N)[None, None, :], (N, N, N)).copy()
    f = k / N
    lam = -4 * np.pi ** 2 * f ** 2 * np.exp(-2 * np.pi ** 2 * s ** 2 * f ** 2)
    np.testing.assert_allclose(oracle.log_fft(v, s), v * lam * v.size, rtol=1e-9, atol=1e-9)


def test_lct_downsample_known_answers():
    """python/lct.py:98-105: each round sums bins (2t, 2t+1) and halves
    z_trim with Python's round, which takes halves to the even neighbour
    (600 -> 300 -> 150 -> 75 -> 38; 5 -> 2; 7 -> 4; 3 -> 2 -> 1)."""
    rect = np.arange(2 * 3 * 8, dtype=np.float64).reshape(2, 3, 8)
    r1, z1 = olct.downsample(rect, 5, 1)
    np.testing.assert_array_equal(r1, rect[:, :, 0::2] + rect[:, :, 1::2])
    assert r1[1, 2, 3] == rect[1, 2, 6] + rect[1, 2, 7] and z1 == 2
    r2, z2 = olct.downsample(rect, 3, 2)
    np.testing.assert_array_equal(r2, rect.reshape(2, 3, 2, 4).sum(-1))
    assert z2 == 1
    assert olct.downsample(rect, 7, 1)[1] == 4
    assert olct.downsample(np.zeros((1, 1, 64)), 600, 4)[1] == 38
    assert olct.downsample(rect, 600, 0)[1] == 600 and olct.downsample(rect, 600, 0)[0] is rect


def test_lct_downsample_point_scatterer_keeps_its_depth():
    """A point scatterer rendered at 8 ps and downsampled once reconstructs
    at its (x, y) pixel and at the depth index of the 16 ps grid,
    z / (range / 2) * (M/2 - 1) (N = 64, as in the undownsampled case)."""
    N, M = 64, 512
    width, bin_res, c = 0.5, 8e-12, 3e8
    rng_max = M * c * bin_res
    xs = np.linspace(-width, width, N)
    x0, y0, z0 = xs[40], xs[22], 0.7 * rng_max / 2
    data = np.zeros((N, N, M))
    for i in range(N):
        for j in range(N):
            r = np.sqrt((xs[i] - x0) ** 2 + (xs[j] - y0) ** 2 + z0 ** 2)
            t = int(round(2 * r / (c * bin_res)))
            if t < M:
                data[i, j, t] = 1.0
    vol = olct.lct(data, width, rng_max, snr=8.0, K=1)
    assert vol.shape == (M // 2, N, N)
    t, y, x = np.unravel_index(np.argmax(vol), vol.shape)
    assert abs(int(x) - 40) <= 1 and abs(int(y) - 22) <= 1
    assert abs(t - z0 / (rng_max / 2) * (M // 2 - 1)) <= 2
