"""CPU oracle of the NLOS reconstruction hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's CPU-baseline leg.  The product
(fasterbackprojection_amd) never imports it.

* fbp_oracle.c (built into liboracle.so by the Makefile next to it) restates
  the reference's backprojection, filter and filter_H formulas: bins by the
  FP32 IEEE expression, weights and sums in FP64.
* lct.py restates python/lct.py in NumPy/SciPy.

Every function names the reference file:line it follows.  How each is
pinned (closed forms, invariants, brute force, known answers) is in the
tests that exercise it; the reference ships no golden vectors for this path.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_lib = None
_lib_path = LIB_PATH


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB_PATH


def use_native() -> Path:
    """Switches every later oracle call to an -O3 -march=native build made
    now, for the host this process runs on, in a fresh temporary directory
    (never shipped with the tree).  It computes the same values as the
    portable build (no fast-math, no contraction); for CPU baseline timing."""
    import shutil
    import tempfile
    global _lib, _lib_path
    out = Path(tempfile.mkdtemp(prefix="oracle_native_"))
    shutil.copy(HERE / "fbp_oracle.c", out / "fbp_oracle.c")
    shutil.copy(HERE / "Makefile", out / "Makefile")
    subprocess.run(["make", "-s", "-C", str(out), "liboracle_native.so"], check=True)
    _lib, _lib_path = None, out / "liboracle_native.so"
    load()
    return _lib_path


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        src = HERE / "fbp_oracle.c"
        if _lib_path == LIB_PATH and (not LIB_PATH.exists() or
                                      LIB_PATH.stat().st_mtime < src.stat().st_mtime):
            build()
        _lib = C.CDLL(str(_lib_path))
        P = C.c_void_p
        _lib.oracle_recinfo_size.restype = C.c_int
        _lib.oracle_confocal_bin.restype = C.c_int64
        _lib.oracle_confocal_bin.argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        _lib.oracle_confocal_distf.restype = C.c_float
        _lib.oracle_confocal_distf.argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        _lib.oracle_confocal_voxel.argtypes = [P, P, C.c_int, P, P]
        _lib.oracle_exhaustive_voxel.argtypes = [P, P, P, P, P]
        _lib.oracle_depth.argtypes = [P, P, P, P, P, C.c_uint32, C.c_int, P]
        _lib.oracle_confocal_bins.argtypes = [P, P, C.c_uint32, P, P]
        _lib.oracle_exhaustive_bins.argtypes = [P, P, P, C.c_uint32, C.c_uint32, P, P]
        _lib.oracle_depth_bins.argtypes = [P, P, P, P, C.c_uint32, C.c_int, C.c_uint32,
                                           C.c_uint32, P]
        _lib.oracle_mis.argtypes = [P, P, P, P, P, P, P, C.c_uint32, C.c_uint32, C.c_int,
                                    C.c_int, P]
        _lib.oracle_confocal_voxel_scatter.argtypes = [P, P, C.c_int, P, P]
        _lib.oracle_scatter_bins.restype = C.c_int
        _lib.oracle_scatter_bins.argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                             P]
        _lib.oracle_alias_tables.restype = C.c_int
        _lib.oracle_alias_tables.argtypes = [P, C.c_uint32, P, P]
        _lib.oracle_set_threads.argtypes = [C.c_int]
        _lib.oracle_max_threads.restype = C.c_int
        _lib.oracle_depth_slices_rgba.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32,
                                                  C.c_uint32, C.c_uint32, P]
        _lib.oracle_laplacian.argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int]
        _lib.oracle_log_kernel.argtypes = [C.c_int, C.c_double, P]
        _lib.oracle_log_filter.argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, P]
        _lib.oracle_normalize.argtypes = [P, C.c_int64]
        _lib.oracle_filter_h_nt_pad.restype = C.c_uint32
        _lib.oracle_filter_h_nt_pad.argtypes = [C.c_uint32, C.c_float, C.c_float, C.c_float]
        _lib.oracle_filter_h.argtypes = [P, C.c_uint64, C.c_uint32, C.c_float, C.c_float,
                                         C.c_float, C.c_int, P]
        assert _lib.oracle_recinfo_size() == 148
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _res(info):
    r = info.voxel_resolution
    return (r.z, r.y, r.x)


def confocal_voxel(info, I, laser, layout: int = 0) -> np.ndarray:
    """backprojectConfocalVoxel (FB/transient_reconstruction.cuh:93-136); FP64 [Z][Y][X]."""
    I, laser = _f32(I), _f32(laser)
    out = np.empty(_res(info), np.float64)
    load().oracle_confocal_voxel(C.byref(info), _p(I), layout, _p(laser), _p(out))
    return out


def confocal_bin(info, laser, l, vx, vy, z) -> int:
    laser = _f32(laser)
    return int(load().oracle_confocal_bin(C.byref(info), _p(laser), l, vx, vy, z))


def confocal_distf(info, laser, l, vx, vy, z) -> float:
    laser = _f32(laser)
    return float(load().oracle_confocal_distf(C.byref(info), _p(laser), l, vx, vy, z))


def confocal_bins(info, laser, l):
    """Bin (-1 outside [0, T)) and FP32 dist of every voxel for scan point l
    (FB/transient_reconstruction.cuh:102-130): int64 and float32 [Z][Y][X]."""
    laser = _f32(laser)
    bins = np.empty(_res(info), np.int64)
    dist = np.empty(_res(info), np.float32)
    load().oracle_confocal_bins(C.byref(info), _p(laser), int(l), _p(bins), _p(dist))
    return bins, dist


def exhaustive_bins(info, laser, sensor, l, s):
    """Bin and FP32 dist of every voxel for the pair (l, s)
    (FB/transient_reconstruction.cuh:150-161)."""
    laser, sensor = _f32(laser), _f32(sensor)
    bins = np.empty(_res(info), np.int64)
    dist = np.empty(_res(info), np.float32)
    load().oracle_exhaustive_bins(C.byref(info), _p(laser), _p(sensor), int(l), int(s), _p(bins),
                                  _p(dist))
    return bins, dist


def depth_bins(info, laser, depths, l, sensor=None, s=0, exhaustive=False):
    """Depth-plane bins [c][d] of one lit scan point l (or pair (l, s))
    (FB/transient_reconstruction.cuh:41-91); -1 outside [0, T)."""
    laser = _f32(laser)
    sensor = laser if sensor is None else _f32(sensor)
    depths = np.ascontiguousarray(depths, np.float64)
    bins = np.empty((info.num_laser_targets, len(depths)), np.int64)
    load().oracle_depth_bins(C.byref(info), _p(laser), _p(sensor), _p(depths), len(depths),
                             int(exhaustive), int(l), int(s), _p(bins))
    return bins


def exhaustive_voxel(info, I, laser, sensor) -> np.ndarray:
    """backprojectExhaustiveVoxel (FB/transient_reconstruction.cuh:138-166), all pairs."""
    I, laser, sensor = _f32(I), _f32(laser), _f32(sensor)
    out = np.empty(_res(info), np.float64)
    load().oracle_exhaustive_voxel(C.byref(info), _p(I), _p(laser), _p(sensor), _p(out))
    return out


def depth(info, I, laser, depths, sensor=None, exhaustive=False) -> np.ndarray:
    """backprojectConfocal / backprojectExhaustive (FB/transient_reconstruction.cuh:41-91)."""
    I, laser = _f32(I), _f32(laser)
    sensor = laser if sensor is None else _f32(sensor)
    depths = np.ascontiguousarray(depths, np.float64)
    out = np.empty((info.num_laser_targets, len(depths)), np.float64)
    load().oracle_depth(C.byref(info), _p(I), _p(laser), _p(sensor), _p(depths), len(depths),
                        int(exhaustive), _p(out))
    return out


def alias_tables(cdf):
    """Laser::buildAliasTables (FB/Laser.cu:253-315).  Returns (alias, prob) or raises."""
    cdf = _f32(cdf)
    n = len(cdf)
    alias = np.empty(n, np.uint32)
    prob = np.empty(n, np.float32)
    rc = load().oracle_alias_tables(_p(cdf), n, _p(alias), _p(prob))
    if rc != 0:
        raise ValueError("invalid probabilities")
    return alias, prob


def mis_cdf(I) -> np.ndarray:
    """Scan-point CDF: prefix sums (FP64) of per-scan-point totals of I > 1e-5,
    min-max normalised, as float32 (FB/Laser.cu:547-574 over the intensity)."""
    I = np.asarray(I, np.float32)
    tot = np.where(I > np.float32(1e-5), I, 0).astype(np.float64).sum(axis=1)
    pre = np.cumsum(tot)
    mn, mx = pre.min(), pre.max()
    return ((pre - mn) / (mx - mn)).astype(np.float32) if mx > mn else np.zeros_like(pre, np.float32)


def mis(info, I, laser, cdf, alias, prob, noise, num_samples, mode=0, scatter=False) -> np.ndarray:
    """backprojectConfocalVoxelMIS (FB/transient_reconstruction.cuh:194-260);
    scatter: its ACCUMULATE_VOXEL_SCATTERING branch (:235-250)."""
    I, laser, cdf, prob, noise = map(_f32, (I, laser, cdf, prob, noise))
    alias = np.ascontiguousarray(alias, np.uint32)
    out = np.empty(_res(info), np.float64)
    load().oracle_mis(C.byref(info), _p(I), _p(laser), _p(cdf), _p(alias), _p(prob), _p(noise),
                      len(noise), num_samples, mode, int(bool(scatter)), _p(out))
    return out


def confocal_voxel_scatter(info, I, laser, layout: int = 0) -> np.ndarray:
    """ACCUMULATE_VOXEL_SCATTERING branch of backprojectConfocalVoxel
    (FB/transient_reconstruction.cuh:112-126); FP64 [Z][Y][X]."""
    I, laser = _f32(I), _f32(laser)
    out = np.empty(_res(info), np.float64)
    load().oracle_confocal_voxel_scatter(C.byref(info), _p(I), layout, _p(laser), _p(out))
    return out


def scatter_bins(info, laser, l, vx, vy, z) -> list:
    """Bins visited by the voxel-scattering walk of one (voxel, scan point)."""
    laser = _f32(laser)
    buf = np.empty(int(info.num_time_bins) + 2, np.int64)
    n = load().oracle_scatter_bins(C.byref(info), _p(laser), l, vx, vy, z, _p(buf))
    return buf[:n].tolist()


def mis_tiled_samples(I, noise, num_samples, tiles):
    """Scan-point samples of the tile-shared MIS mode (no reference
    counterpart; restatement of fbp_mis_params FBP_MIS_TILED): per tile t and
    sample i, LCG state t*S + i (s = s*747796405 + 2891336453 mod 2^32,
    u = noise[s % n]); u0 < 1/2 picks the first l whose inclusive FP64 prefix
    of the row sums R exceeds u1 * sum R, else l = int(u1 * L); weight
    1 / (S * (R_l / sum R + 1/L) / 2).  Returns (l[tiles][S], w[tiles][S])."""
    I = np.asarray(I, np.float32)
    noise = np.asarray(noise, np.float32)
    L = I.shape[0]
    R = np.where(I > np.float32(1e-5), I, 0).astype(np.float64).sum(axis=1)
    cum = np.cumsum(R)
    total = cum[-1]
    n = len(noise)
    ls = np.empty((tiles, num_samples), np.int64)
    ws = np.empty((tiles, num_samples), np.float64)
    for t in range(tiles):
        for i in range(num_samples):
            state = (t * num_samples + i) & 0xFFFFFFFF
            state = (state * 747796405 + 2891336453) & 0xFFFFFFFF
            u0 = noise[state % n]
            state = (state * 747796405 + 2891336453) & 0xFFFFFFFF
            u1 = noise[state % n]
            if total > 0 and u0 < np.float32(0.5):
                l = int(np.searchsorted(cum, float(u1) * total, side="right"))
                l = min(l, L - 1)
            else:
                l = min(int(np.float32(u1) * np.float32(L)), L - 1)
            pmix = 0.5 * (R[l] / total + 1.0 / L) if total > 0 else 1.0 / L
            ls[t, i] = l
            ws[t, i] = np.float32(1.0 / (pmix * num_samples))
    return ls, ws


def mis_tiled(info, I, laser, noise, num_samples, scatter=False) -> np.ndarray:
    """Tile-shared MIS backprojection: every 16^3 tile (vx, vy, z; x fastest)
    gathers only its sampled scan points, each with its weight, through the
    confocal_voxel restatement (the per-tile weighted transient sum_i w_i I[l_i]
    thresholded first; the voxel-scattering restatement when scatter).
    FP64 [Z][Y][X]."""
    I = np.asarray(I, np.float32)
    Z, Y, X = _res(info)
    tx, ty, tz = -(-X // 16), -(-Y // 16), -(-Z // 16)
    ls, ws = mis_tiled_samples(I, noise, num_samples, tx * ty * tz)
    Ithr = np.where(I > np.float32(1e-5), I, np.float32(0))
    out = np.zeros((Z, Y, X))
    for t in range(tx * ty * tz):
        bx, by, bz = t % tx, (t // tx) % ty, t // (tx * ty)
        wl = np.zeros(I.shape[0])
        np.add.at(wl, ls[t], ws[t])
        It = (Ithr.astype(np.float64) * wl[:, None]).astype(np.float32)
        vol = (confocal_voxel_scatter if scatter else confocal_voxel)(info, It, laser)
        sl = (slice(16 * bz, 16 * bz + 16), slice(16 * by, 16 * by + 16), slice(16 * bx, 16 * bx + 16))
        out[sl] = vol[sl]
    return out


def set_threads(n: int) -> None:
    """OpenMP threads of the oracle's loops."""
    load().oracle_set_threads(int(n))


def max_threads() -> int:
    return int(load().oracle_max_threads())


def depth_slices_rgba(act, width, height, out_width=None, out_height=None) -> np.ndarray:
    """TransientImage::save images of every depth plane
    (FB/TransientImage.cpp:27-45, FB/Image.h:45-177): act [W*H][D] float32 ->
    uint8 [D][OH][OW][4]."""
    act = _f32(act)
    D = act.shape[1]
    ow, oh = int(out_width or width), int(out_height or height)
    out = np.empty((D, oh, ow, 4), np.uint8)
    load().oracle_depth_slices_rgba(_p(act), width, height, D, ow, oh, _p(out))
    return out


def laplacian(v, k=5) -> np.ndarray:
    """laplacianFilter (FB/transient_postprocessing.cuh:21-50)."""
    v = _f32(v)
    Z, Y, X = v.shape
    out = np.empty(v.shape, np.float64)
    load().oracle_laplacian(_p(v), _p(out), X, Y, Z, k)
    return out


def log_kernel(k=5, sigma=1.0) -> np.ndarray:
    """calculateLaplacianKernel (FB/PostprocessingFilters.cu:102-175) in FP64, [kz][ky][kx]."""
    w = np.empty(k ** 3, np.float64)
    load().oracle_log_kernel(k, float(sigma), _p(w))
    return w.reshape(k, k, k)


def log_filter(v, k=5, sigma=1.0, weights=None) -> np.ndarray:
    """LoGFilter (FB/transient_postprocessing.cuh:52-83), clamp-to-edge."""
    v = _f32(v)
    Z, Y, X = v.shape
    w = np.ascontiguousarray(log_kernel(k, sigma) if weights is None else weights, np.float64)
    out = np.empty(v.shape, np.float64)
    load().oracle_log_filter(_p(v), _p(out), X, Y, Z, k, _p(w))
    return out


def normalize(v) -> np.ndarray:
    """normalizeReconstruction after min/max (FB/Laser.cu:317-339)."""
    v = np.array(v, np.float64, copy=True)
    load().oracle_normalize(_p(v), v.size)
    return v


def filter_h_nt_pad(nt, dt, wl_mean, wl_sigma=0.0) -> int:
    return int(load().oracle_filter_h_nt_pad(nt, dt, wl_mean, wl_sigma))


def filter_h(data, dt, wl_mean, wl_sigma=0.0, border=0) -> np.ndarray:
    """Laser::filter_H_cuda (FB/Laser.cu:128-240): unnormalised, FP64, direct convolution."""
    data = _f32(data)
    nt = data.shape[-1]
    batch = data.size // nt
    out = np.empty(data.shape, np.float64)
    load().oracle_filter_h(_p(data), batch, nt, dt, wl_mean, wl_sigma, border, _p(out))
    return out


def log_fft(v, sigma=1.0) -> np.ndarray:
    """LoGFFT::compute (FB/PostprocessingFilters.cu:36-77) in FP64 with NumPy's
    FFT: R2C, times -4 pi^2 |f|^2 exp(-2 pi^2 sigma^2 |f|^2) (f in cycles per
    sample), C2R without the 1/V normalisation (the reference keeps it off)."""
    v = np.asarray(v, np.float64)
    Z, Y, X = v.shape
    fz = np.fft.fftfreq(Z)[:, None, None]
    fy = np.fft.fftfreq(Y)[None, :, None]
    fx = np.fft.rfftfreq(X)[None, None, :]
    k2 = fx ** 2 + fy ** 2 + fz ** 2
    spec = -4 * np.pi ** 2 * k2 * np.exp(-2 * np.pi ** 2 * sigma ** 2 * k2)
    return np.fft.irfftn(np.fft.rfftn(v) * spec, s=v.shape) * v.size
