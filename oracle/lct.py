"""NumPy/SciPy restatement of the confocal LCT (python/lct.py).

TEST INFRASTRUCTURE ONLY (checker and CPU baseline).  Follows
cnlos_reconstruction (python/lct.py:8-292), define_psf (:295-347) and
resampling_operator (:350-419).  resampling_operator_reference() performs the
reference's dense construction step by step (small M only);
resampling_operator() is the closed form of the same matrix, sparse, used
for large M.  The two are checked against each other in the tests.
"""
from __future__ import annotations

import numpy as np
import scipy.fft
import scipy.sparse as sp


def define_psf(U: int, V: int, slope: float) -> np.ndarray:
    """python/lct.py:295-347: binary argmin mask of |(4 slope)^2 (x^2+y^2) - z|
    along z, normalised, rolled by (0, U, U).  Shape (2V, 2U, 2U)."""
    x = np.linspace(-1, 1, 2 * U)
    y = np.linspace(-1, 1, 2 * U)
    z = np.linspace(0, 2, 2 * V)
    gz, gy, gx = np.meshgrid(z, y, x, indexing="ij")
    psf = np.abs(((4.0 * slope) ** 2) * (gx ** 2 + gy ** 2) - gz)
    psf = (psf == psf.min(axis=0, keepdims=True)).astype(np.float64)
    psf = psf / psf[:, U - 1, U - 1].sum()
    psf = psf / np.linalg.norm(psf.ravel())
    return np.roll(psf, shift=(0, U, U), axis=(0, 1, 2))


def resampling_operator_reference(M: int):
    """python/lct.py:350-419 literally (dense, M^2 x M start): small M only."""
    mtx = np.zeros((M * M, M))
    i = np.arange(M * M)
    col = np.clip(np.ceil(np.sqrt(i + 1)) - 1, 0, M - 1).astype(int)
    mtx[i, col] = 1.0
    mtx = mtx * (1.0 / np.sqrt(np.arange(1, M * M + 1)))[:, None]
    mtxi = mtx.T
    for _ in range(round(np.log2(M))):
        mtx = 0.5 * (mtx[0::2, :] + mtx[1::2, :])
        mtxi = 0.5 * (mtxi[:, 0::2] + mtxi[:, 1::2])
    return mtx, mtxi


def resampling_operator(M: int):
    """Closed form of the same operator: mtx[r, j] = (1/M) sum over
    x in [rM+1, (r+1)M] with ceil(sqrt(x)) = j+1 of 1/sqrt(x); mtxi = mtx^T.
    Returned as scipy CSR matrices (M x M)."""
    x = np.arange(1, M * M + 1, dtype=np.float64)
    r = (np.arange(M * M) // M)
    col = np.clip(np.ceil(np.sqrt(x)) - 1, 0, M - 1).astype(np.int64)
    vals = 1.0 / np.sqrt(x) / M
    mtx = sp.coo_matrix((vals, (r, col)), shape=(M, M)).tocsr()
    mtx.sum_duplicates()
    return mtx, mtx.T.tocsr()


def inverse_filter(N: int, M: int, width: float, range_max: float, snr: float,
                   is_backprop: bool = False, workers: int = -1) -> np.ndarray:
    """python/lct.py:112-120: invpsf = conj(F) / (|F|^2 + 1/snr), F = fftn(psf)."""
    psf = define_psf(N, M, width / range_max)
    f = scipy.fft.fftn(psf, workers=workers)
    if is_backprop:
        return np.conj(f)
    return np.conj(f) / (np.abs(f) ** 2 + 1.0 / snr)


def downsample(rect: np.ndarray, z_trim: int, K: int):
    """python/lct.py:98-105: K rounds of M //= 2, rect = rect[:, :, 0::2] +
    rect[:, :, 1::2] and z_trim = round(z_trim / 2) (Python's round, halves
    to even).  Returns (rect, z_trim)."""
    rect = np.asarray(rect)
    for _ in range(K):
        rect = rect[:, :, 0::2] + rect[:, :, 1::2]
        z_trim = round(z_trim / 2)
    return rect, z_trim


def lct(data: np.ndarray, width: float, range_max: float, snr: float = 0.8,
        is_diffuse: bool = False, is_backprop: bool = False, z_trim: int = 0,
        invpsf: np.ndarray | None = None, workers: int = -1, K: int = 0) -> np.ndarray:
    """python/lct.py:98-211 on rect_data (N, N, m): downsampled K times
    (downsample(); range_max stays the caller's pre-downsampling value, as
    lct.py:95 computes it before the loop), then steps 1-4.
    Returns vol (M, N, N) = max(Re(mtxi @ crop(ifftn(fftn(pad) * invpsf))), 0),
    M = m / 2^K."""
    data, z_trim = downsample(np.asarray(data, np.float64), z_trim, K)
    data = np.array(data, np.float64, copy=True)
    N, _, M = data.shape
    data[:, :, :z_trim] = 0
    if invpsf is None:
        invpsf = inverse_filter(N, M, width, range_max, snr, is_backprop, workers)
    mtx, mtxi = resampling_operator(M)
    d = np.transpose(data, (2, 1, 0))                     # (M, N, N)
    gz = np.linspace(0, 1, M)[:, None, None]
    d = d * (gz ** 4 if is_diffuse else gz ** 2)
    t = (mtx @ d.reshape(M, N * N)).reshape(M, N, N)
    tdata = np.zeros((2 * M, 2 * N, 2 * N), np.complex128)
    tdata[:M, :N, :N] = t
    tvol = scipy.fft.ifftn(scipy.fft.fftn(tdata, workers=workers) * invpsf, workers=workers)
    tvol = tvol[:M, :N, :N]
    vol = (mtxi @ tvol.reshape(M, N * N)).reshape(M, N, N)
    return np.maximum(np.real(vol), 0)
