"""Python mirror of the reconstruction entry points, over the C ABI.

torch tensors are used only as device memory and for the current HIP stream;
every computation runs in libfbp.so's HIP kernels.  Names follow the
reference's call surface (FB/Laser.cuh:41-47, FB/PostprocessingFilters.h,
FB/LCT.h, python/lct.py).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import (CONFOCAL, EXHAUSTIVE, LAYOUT_TIME_FASTEST, LAYOUT_TIME_MAJOR, MIS_REFERENCE,
                   MIS_TILED, MIS_UNBIASED, STORAGE_BF16, STORAGE_F16, STORAGE_F32, FbpError, LctParams,
                   MisParams, ReconstructionInfo, Transient, UVec3, Vec3)

__all__ = [
    "Context", "Geometry", "make_info", "CONFOCAL", "EXHAUSTIVE", "STORAGE_F32", "STORAGE_F16",
    "STORAGE_BF16", "LAYOUT_TIME_FASTEST", "LAYOUT_TIME_MAJOR", "MIS_REFERENCE", "MIS_UNBIASED", "MIS_TILED",
    "FbpError",
]


def _ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("tensor must live on the GPU")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return t.data_ptr()


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


_TWO_BYTE = (torch.int16, torch.uint16, torch.float16, torch.bfloat16)


def _need(t, name: str, dtypes, numel: int | None = None, shape: tuple | None = None,
          min_numel: int | None = None) -> None:
    """Raises ValueError unless tensor t has one of dtypes and the given
    element count / shape.  Runs before any pointer reaches the C ABI, so a
    short or mistyped buffer is refused on the host instead of being read or
    written past its end by a kernel."""
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor")
    dtypes = dtypes if isinstance(dtypes, tuple) else (dtypes,)
    if t.dtype not in dtypes:
        raise ValueError(f"{name} must be {' or '.join(map(str, dtypes))}, got {t.dtype}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name} must hold {numel} elements, got {t.numel()}")
    if min_numel is not None and t.numel() < min_numel:
        raise ValueError(f"{name} must hold at least {min_numel} elements, got {t.numel()}")


def _need_points(t, name: str, n: int) -> None:
    """float32 (n, 3) points (fbp_vec3[n])."""
    _need(t, name, torch.float32)
    if t.numel() != 3 * n or t.shape[-1] != 3:
        raise ValueError(f"{name} must be float32 ({n}, 3), got {tuple(t.shape)}")


def _need_transient(info, intensity, storage: int, scales, exhaustive: bool) -> None:
    """The transient holds L*T (confocal) or L*S*T (exhaustive) samples of the
    storage type; F16/BF16 carry one float32 scale per time series."""
    rows = info.num_laser_targets * (info.num_sensor_targets if exhaustive else 1)
    n = rows * info.num_time_bins
    if storage == STORAGE_F32:
        _need(intensity, "intensity", torch.float32, numel=n)
    elif storage in (STORAGE_F16, STORAGE_BF16):
        _need(intensity, "intensity", _TWO_BYTE, numel=n)
        _need(scales, "scales", torch.float32, numel=rows)
    else:
        raise ValueError(f"unknown storage {storage}")


@dataclass
class Geometry:
    """Host description of a capture, the fields NLosData::toGpu copies into
    ReconstructionInfo (FB/NLosData.cpp:131-159)."""
    laser_targets: np.ndarray            # (L, 3) float32
    sensor_targets: np.ndarray           # (S, 3) float32 (== laser for confocal)
    num_time_bins: int
    time_step: float                     # metres per bin
    hidden_min: tuple
    hidden_max: tuple
    voxel_resolution: tuple              # (X, Y, Z)
    confocal: bool = True
    time_offset: float = 0.0
    discard_first_last_bounces: int = 1
    laser_position: tuple = (0.0, 0.0, 0.0)
    sensor_position: tuple = (0.0, 0.0, 0.0)
    wall_normal: tuple = (0.0, 1.0, 0.0)
    wall_size: tuple = (2.0, 2.0)


def make_info(g: Geometry) -> ReconstructionInfo:
    info = ReconstructionInfo()
    f32 = np.float32
    hmin = np.asarray(g.hidden_min, f32)
    hmax = np.asarray(g.hidden_max, f32)
    res = np.asarray(g.voxel_resolution, np.uint32)
    size = (hmax - hmin).astype(f32)
    vs = (size / res.astype(f32)).astype(f32)
    info.sensor_position = Vec3(*map(float, np.asarray(g.sensor_position, f32)))
    info.capture_system = CONFOCAL if g.confocal else EXHAUSTIVE
    info.laser_position = Vec3(*map(float, np.asarray(g.laser_position, f32)))
    info.time_offset = float(f32(g.time_offset))
    info.num_laser_targets = len(g.laser_targets)
    info.num_sensor_targets = len(g.sensor_targets)
    info.num_time_bins = int(g.num_time_bins)
    info.time_step = float(f32(g.time_step))
    info.hidden_volume_min = Vec3(*map(float, hmin))
    info.discard_first_last_bounces = int(g.discard_first_last_bounces)
    info.hidden_volume_max = Vec3(*map(float, hmax))
    info.hidden_volume_size = Vec3(*map(float, size))
    info.hidden_volume_voxel_size = Vec3(*map(float, vs))
    wx, wz = (f32(g.wall_size[0]), f32(g.wall_size[1]))
    info.relay_wall_min_position = Vec3(float(-wx / f32(2)), 0.0, float(-wz / f32(2)))
    info.relay_wall_size = Vec3(float(wx), 0.0, float(wz))
    info.relay_wall_normal = Vec3(*map(float, np.asarray(g.wall_normal, f32)))
    info.voxel_resolution = UVec3(*map(int, res))
    return info


class Context:
    """Owns an fbp_context (scratch memory, FFT plans) on one device."""

    def __init__(self, device: int | None = None):
        self.lib = _lib.load()
        if device is None:
            device = torch.cuda.current_device()
        h = C.c_void_p()
        st = self.lib.fbp_context_create(int(device), C.byref(h))
        if st != 0:
            raise FbpError(st, "fbp_context_create failed")
        self.handle = h
        self.device = int(device)

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.fbp_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, st: int) -> None:
        if st != 0:
            msg = self.lib.fbp_context_last_error(self.handle).decode()
            raise FbpError(st, msg)

    def set_aligned_path(self, enabled: bool) -> None:
        """Allow (default) or forbid the aligned-grid confocal path."""
        self.check(self.lib.fbp_context_set_option(self.handle, _lib.OPTION_CONFOCAL_ALIGNED,
                                                   int(bool(enabled))))

    def set_voxel_scattering(self, enabled: bool) -> None:
        """Voxel-scattering variant of the confocal voxel and MIS
        backprojections (the reference's ACCUMULATE_VOXEL_SCATTERING switch,
        FB/transient_reconstruction.cuh:112-126, 235-250): every bin a voxel
        reaches adds its sample.  Off by default."""
        self.check(self.lib.fbp_context_set_option(self.handle, _lib.OPTION_VOXEL_SCATTERING,
                                                   int(bool(enabled))))

    def set_exhaustive_laser_groups(self, enabled: bool) -> None:
        """Allow (default) or forbid the laser-group kernel of the exhaustive
        voxel backprojection (sensor distances shared by eight lasers; taken
        for linear-weight captures with at least 8 lasers and 64 sensors)."""
        self.check(self.lib.fbp_context_set_option(self.handle,
                                                   _lib.OPTION_EXHAUSTIVE_LASER_GROUPS,
                                                   int(bool(enabled))))

    def option(self, option: int) -> int:
        v = C.c_int32()
        self.check(self.lib.fbp_context_get_option(self.handle, option, C.byref(v)))
        return int(v.value)

    def last_confocal_path_aligned(self) -> bool:
        """True when the last confocal voxel backprojection took the
        aligned-grid path (synchronises)."""
        v = C.c_int32()
        self.check(self.lib.fbp_context_last_confocal_path(self.handle, C.byref(v)))
        return bool(v.value)

    def last_exhaustive_laser_groups(self) -> bool:
        """True when the last exhaustive voxel backprojection ran the
        laser-group kernel (host-side record)."""
        v = C.c_int32()
        self.check(self.lib.fbp_context_last_exhaustive_kernel(self.handle, C.byref(v)))
        return bool(v.value)

    # ------------------------------------------------------------ helpers
    def _dev(self) -> torch.device:
        return torch.device("cuda", self.device)

    @staticmethod
    def _transient(data: torch.Tensor, storage: int, layout: int,
                   scales: torch.Tensor | None) -> Transient:
        return Transient(_ptr(data), storage, layout, _ptr(scales))

    def _volume(self, info: ReconstructionInfo, out: torch.Tensor | None) -> torch.Tensor:
        r = info.voxel_resolution
        shape = (r.z, r.y, r.x)
        if out is None:
            return torch.empty(shape, dtype=torch.float32, device=self._dev())
        _need(out, "out", torch.float32, shape=shape)
        return out

    # ------------------------------------------------------ backprojection
    def backproject_confocal(self, info, intensity, laser_targets, *, storage=STORAGE_F32,
                             layout=LAYOUT_TIME_FASTEST, scales=None, out=None):
        """backprojectConfocalVoxel (FB/transient_reconstruction.cuh:93-136)."""
        _need_transient(info, intensity, storage, scales, exhaustive=False)
        _need_points(laser_targets, "laser_targets", info.num_laser_targets)
        vol = self._volume(info, out)
        tr = self._transient(intensity, storage, layout, scales)
        self.check(self.lib.fbp_backproject_confocal_voxel(
            self.handle, C.byref(info), C.byref(tr), _ptr(laser_targets), _ptr(vol), _stream()))
        return vol

    def backproject_confocal_splat(self, info, intensity, laser_targets, *,
                                   layout=LAYOUT_TIME_FASTEST, out=None):
        """Scatter form (fbp_backproject_confocal_splat): every lit sample
        splatted into the voxels of its shell, in place of
        DrawEllipsoids::solveBackprojection (FB/DrawEllipsoids.cpp:332-440).
        Same volume as backproject_confocal; F32 storage."""
        _need_transient(info, intensity, STORAGE_F32, None, exhaustive=False)
        _need_points(laser_targets, "laser_targets", info.num_laser_targets)
        vol = self._volume(info, out)
        tr = self._transient(intensity, STORAGE_F32, layout, None)
        self.check(self.lib.fbp_backproject_confocal_splat(
            self.handle, C.byref(info), C.byref(tr), _ptr(laser_targets), _ptr(vol), _stream()))
        return vol

    def backproject_exhaustive(self, info, intensity, laser_targets, sensor_targets, *,
                               storage=STORAGE_F32, scales=None, out=None):
        """backprojectExhaustiveVoxel (FB/transient_reconstruction.cuh:138-166)."""
        _need_transient(info, intensity, storage, scales, exhaustive=True)
        _need_points(laser_targets, "laser_targets", info.num_laser_targets)
        _need_points(sensor_targets, "sensor_targets", info.num_sensor_targets)
        vol = self._volume(info, out)
        tr = self._transient(intensity, storage, LAYOUT_TIME_FASTEST, scales)
        self.check(self.lib.fbp_backproject_exhaustive_voxel(
            self.handle, C.byref(info), C.byref(tr), _ptr(laser_targets), _ptr(sensor_targets),
            _ptr(vol), _stream()))
        return vol

    def backproject_exhaustive_noalloc(self, info, intensity, laser_targets, sensor_targets,
                                       out, *, storage=STORAGE_F32, scales=None):
        """fbp_backproject_exhaustive_voxel_noalloc: the exhaustive
        backprojection with its partial sums in the library's fixed per-GPU
        array; `out` must be given (nothing is allocated)."""
        _need_transient(info, intensity, storage, scales, exhaustive=True)
        _need_points(laser_targets, "laser_targets", info.num_laser_targets)
        _need_points(sensor_targets, "sensor_targets", info.num_sensor_targets)
        r = info.voxel_resolution
        _need(out, "out", torch.float32, shape=(r.z, r.y, r.x))
        tr = self._transient(intensity, storage, LAYOUT_TIME_FASTEST, scales)
        self.check(self.lib.fbp_backproject_exhaustive_voxel_noalloc(
            self.handle, C.byref(info), C.byref(tr), _ptr(laser_targets), _ptr(sensor_targets),
            _ptr(out), _stream()))
        return out

    def scratch_bytes(self) -> int:
        """Device bytes the context's scratch arenas hold (fbp_context_scratch_bytes)."""
        return int(self.lib.fbp_context_scratch_bytes(self.handle))

    def backproject_depth(self, info, intensity, laser_targets, depths, sensor_targets=None,
                          out=None):
        """backprojectConfocal / backprojectExhaustive (depth planes)."""
        _need(depths, "depths", torch.float64)
        D = int(depths.numel())
        L = info.num_laser_targets
        exhaustive = info.capture_system != CONFOCAL
        _need_transient(info, intensity, STORAGE_F32, None, exhaustive=exhaustive)
        _need_points(laser_targets, "laser_targets", L)
        if exhaustive:
            _need_points(sensor_targets, "sensor_targets", info.num_sensor_targets)
        if out is None:
            out = torch.empty((L, D), dtype=torch.float32, device=self._dev())
        else:
            _need(out, "out", torch.float32, shape=(L, D))
        tr = self._transient(intensity, STORAGE_F32, LAYOUT_TIME_FASTEST, None)
        if info.capture_system == CONFOCAL:
            st = self.lib.fbp_backproject_depth_confocal(
                self.handle, C.byref(info), C.byref(tr), _ptr(laser_targets), _ptr(depths), D,
                _ptr(out), _stream())
        else:
            st = self.lib.fbp_backproject_depth_exhaustive(
                self.handle, C.byref(info), C.byref(tr), _ptr(laser_targets),
                _ptr(sensor_targets), _ptr(depths), D, _ptr(out), _stream())
        self.check(st)
        return out

    def depth_slices_rgba(self, activation, width, height, out_width=None, out_height=None,
                          out=None):
        """Per-depth RGBA slice images of TransientImage::save
        (FB/TransientImage.cpp:27-45, FB/Image.h:45-177): activation
        [width*height][D] -> uint8 [D][out_height][out_width][4]."""
        _need(activation, "activation", torch.float32)
        if activation.dim() != 2 or activation.shape[0] != width * height:
            raise ValueError("activation must be [width*height][num_depths]")
        D = int(activation.shape[1])
        ow = int(out_width or width)
        oh = int(out_height or height)
        if out is None:
            out = torch.empty((D, oh, ow, 4), dtype=torch.uint8, device=self._dev())
        else:
            _need(out, "out", torch.uint8, shape=(D, oh, ow, 4))
        self.check(self.lib.fbp_depth_slices_rgba(
            self.handle, _ptr(activation.contiguous()), int(width), int(height), D, ow, oh,
            _ptr(out), _stream()))
        return out

    def backproject_confocal_mis(self, info, intensity, laser_targets, noise, num_samples,
                                 mode=MIS_REFERENCE, out=None, seed=None, noise_size=None):
        """Laser::reconstructAABBConfocalMIS (FB/Laser.cu:526-611).  With
        noise=None (MIS_TILED only) the uniforms come from the seeded counter
        hash over a virtual noise array of noise_size entries (default: the
        number of voxels)."""
        _need_transient(info, intensity, STORAGE_F32, None, exhaustive=False)
        _need_points(laser_targets, "laser_targets", info.num_laser_targets)
        if noise is None:
            if mode != MIS_TILED or seed is None:
                raise ValueError("noise=None needs mode=MIS_TILED and a seed")
            r = info.voxel_resolution
            n = int(noise_size) if noise_size is not None else r.x * r.y * r.z
            if not 0 < n < 2 ** 32:
                raise ValueError("noise_size out of range")
            p = MisParams(int(num_samples), int(mode), None, n, int(seed) & (2 ** 64 - 1))
        else:
            _need(noise, "noise", torch.float32, min_numel=1)
            p = MisParams(int(num_samples), int(mode), _ptr(noise), int(noise.numel()), 0)
        vol = self._volume(info, out)
        tr = self._transient(intensity, STORAGE_F32, LAYOUT_TIME_FASTEST, None)
        self.check(self.lib.fbp_backproject_confocal_mis(
            self.handle, C.byref(info), C.byref(tr), _ptr(laser_targets), C.byref(p), _ptr(vol),
            _stream()))
        return vol

    def backproject_points(self, info, intensity, laser_targets, voxels, out=None):
        """Exact confocal voxel backprojection at the listed voxel indices
        (fbp_backproject_confocal_points); float64 values."""
        _need_transient(info, intensity, STORAGE_F32, None, exhaustive=False)
        _need_points(laser_targets, "laser_targets", info.num_laser_targets)
        _need(voxels, "voxels", (torch.int32, torch.uint32))
        n = int(voxels.numel())
        if out is None:
            out = torch.empty((n,), dtype=torch.float64, device=intensity.device)
        else:
            _need(out, "out", torch.float64, numel=n)
        tr = self._transient(intensity, STORAGE_F32, LAYOUT_TIME_FASTEST, None)
        self.check(self.lib.fbp_backproject_confocal_points(
            self.handle, C.byref(info), C.byref(tr), _ptr(laser_targets), _ptr(voxels), n,
            _ptr(out), _stream()))
        return out

    def learned_calibrate(self, info, intensity, laser_targets, volume, reference, extremes=16,
                          kernel_size=5, sigma=1.0):
        """Global calibration of a refined, normalised LoG volume
        (fbp_learned_calibrate), in place: reference is an independent
        unbiased estimate of the LoG in physical units."""
        _need_transient(info, intensity, STORAGE_F32, None, exhaustive=False)
        _need_points(laser_targets, "laser_targets", info.num_laser_targets)
        r = info.voxel_resolution
        _need(volume, "volume", torch.float32, shape=(r.z, r.y, r.x))
        _need(reference, "reference", torch.float32, shape=(r.z, r.y, r.x))
        tr = self._transient(intensity, STORAGE_F32, LAYOUT_TIME_FASTEST, None)
        self.check(self.lib.fbp_learned_calibrate(
            self.handle, C.byref(info), C.byref(tr), _ptr(laser_targets), _ptr(volume),
            _ptr(reference), int(kernel_size), float(sigma), int(extremes), _stream()))
        return volume

    def quantize(self, intensity: torch.Tensor, storage: int):
        """Per-row scaled F16/BF16 copy of a [rows][T] float32 transient."""
        _need(intensity, "intensity", torch.float32, min_numel=1)
        rows = intensity.numel() // intensity.shape[-1]
        cols = intensity.shape[-1]
        q = torch.empty(intensity.shape, dtype=torch.int16, device=intensity.device)
        sc = torch.empty((rows,), dtype=torch.float32, device=intensity.device)
        self.check(self.lib.fbp_quantize_transient(self.handle, _ptr(intensity), rows, cols,
                                                   storage, _ptr(q), _ptr(sc), _stream()))
        return q, sc

    # --------------------------------------------------------------- filters
    @staticmethod
    def _filter_io(volume, out):
        _need(volume, "volume", torch.float32)
        if volume.dim() != 3:
            raise ValueError("volume must be [Z][Y][X]")
        if out is None:
            return torch.empty_like(volume)
        _need(out, "out", torch.float32, shape=tuple(volume.shape))
        return out

    def laplacian(self, volume, kernel_size=5, out=None):
        out = self._filter_io(volume, out)
        Z, Y, X = volume.shape
        self.check(self.lib.fbp_filter_laplacian(self.handle, _ptr(volume), _ptr(out),
                                                 UVec3(X, Y, Z), kernel_size, _stream()))
        return out

    def log(self, volume, kernel_size=5, sigma=1.0, out=None):
        out = self._filter_io(volume, out)
        Z, Y, X = volume.shape
        self.check(self.lib.fbp_filter_log(self.handle, _ptr(volume), _ptr(out), UVec3(X, Y, Z),
                                           kernel_size, float(sigma), _stream()))
        return out

    def log_normalize(self, volume, kernel_size=5, sigma=1.0, out=None):
        """LoG then min-max normalisation (fbp_filter_log_normalize): the
        results of log() followed by normalize() in two launches."""
        out = self._filter_io(volume, out)
        Z, Y, X = volume.shape
        self.check(self.lib.fbp_filter_log_normalize(self.handle, _ptr(volume), _ptr(out),
                                                     UVec3(X, Y, Z), kernel_size, float(sigma),
                                                     _stream()))
        return out

    def log_fft(self, volume, sigma=1.0):
        self._filter_io(volume, volume)
        Z, Y, X = volume.shape
        self.check(self.lib.fbp_filter_log_fft(self.handle, _ptr(volume), UVec3(X, Y, Z),
                                               float(sigma), _stream()))
        return volume

    def normalize(self, v):
        _need(v, "v", torch.float32, min_numel=1)
        self.check(self.lib.fbp_normalize_minmax(self.handle, _ptr(v), v.numel(), _stream()))
        return v

    def minmax(self, v):
        _need(v, "v", torch.float32, min_numel=1)
        mn, mx = C.c_float(), C.c_float()
        self.check(self.lib.fbp_minmax(self.handle, _ptr(v), v.numel(), C.byref(mn), C.byref(mx),
                                       _stream()))
        return mn.value, mx.value

    def filter_h(self, data, delta_t, wl_mean, wl_sigma=0.0, border="zero"):
        """Laser::filter_H_cuda (FB/Laser.cu:128-240), in place on data[..., T]."""
        _need(data, "data", torch.float32, min_numel=1)
        nt = data.shape[-1]
        batch = data.numel() // nt
        b = {"zero": 0, "constant": 0, "edge": 1}[border]
        self.check(self.lib.fbp_filter_h(self.handle, _ptr(data), batch, nt, float(delta_t),
                                         float(wl_mean), float(wl_sigma), b, _stream()))
        return data

    def lct(self, data, width, range_max, snr=0.8, is_diffuse=False, is_backprop=False,
            z_trim=0, out=None, downsample=0):
        """python/lct.py cnlos_reconstruction (:95-211) on data[N][N][m]:
        `downsample` (K) rounds of pairwise time sums, then steps 1-4; returns
        [m >> K][N][N].  range_max and z_trim are in the undownsampled units,
        as in lct.py."""
        _need(data, "data", torch.float32)
        if data.dim() != 3 or data.shape[0] != data.shape[1]:
            raise ValueError("data must be (N, N, m)")
        N, _, m = data.shape
        K = int(downsample)
        if not 0 <= K <= 16 or (m >> K) < 2:
            raise ValueError("downsample must be in [0, 16] with m >> downsample >= 2")
        M = m >> K
        p = LctParams(N, m, float(width), float(range_max), float(snr), int(is_diffuse),
                      int(is_backprop), int(z_trim), K)
        if out is None:
            out = torch.empty((M, N, N), dtype=torch.float32, device=data.device)
        else:
            _need(out, "out", torch.float32, shape=(M, N, N))
        self.check(self.lib.fbp_lct(self.handle, _ptr(data), C.byref(p), _ptr(out), _stream()))
        return out

    def render_exhaustive(self, lasers, sensors, points, albedo, num_time_bins, delta, out=None):
        L, S = lasers.shape[0], sensors.shape[0]
        _need_points(lasers, "lasers", L)
        _need_points(sensors, "sensors", S)
        _need_points(points, "points", points.shape[0])
        _need(albedo, "albedo", torch.float32, numel=points.shape[0])
        if out is None:
            out = torch.empty((L, S, num_time_bins), dtype=torch.float32, device=lasers.device)
        else:
            _need(out, "out", torch.float32, numel=L * S * num_time_bins)
        self.check(self.lib.fbp_render_exhaustive(self.handle, _ptr(lasers), L, _ptr(sensors), S,
                                                  _ptr(points), _ptr(albedo), points.shape[0],
                                                  num_time_bins, float(delta), _ptr(out),
                                                  _stream()))
        return out

    def volume_residual(self, info, volume, intensity, laser_targets, rows,
                        threshold_fraction=0.6, blur=4, polarity=1) -> float:
        """1 - Pearson correlation of the measured rows and the rows rendered
        from the volume's voxels above (polarity 1) or below (polarity -1)
        the threshold (fbp_volume_residual)."""
        res = info.voxel_resolution
        _need(volume, "volume", torch.float32, shape=(res.z, res.y, res.x))
        _need_transient(info, intensity, STORAGE_F32, None, exhaustive=False)
        _need_points(laser_targets, "laser_targets", info.num_laser_targets)
        _need(rows, "rows", (torch.int32, torch.uint32))
        r = C.c_double()
        self.check(self.lib.fbp_volume_residual(
            self.handle, C.byref(info), _ptr(volume), _ptr(intensity), _ptr(laser_targets),
            _ptr(rows), int(rows.numel()), float(threshold_fraction), int(polarity), int(blur),
            C.byref(r), _stream()))
        return r.value

    def render_confocal(self, scan_points, points, albedo, num_time_bins, delta, out=None):
        L = scan_points.shape[0]
        _need_points(scan_points, "scan_points", L)
        _need_points(points, "points", points.shape[0])
        _need(albedo, "albedo", torch.float32, numel=points.shape[0])
        if out is None:
            out = torch.empty((L, num_time_bins), dtype=torch.float32, device=scan_points.device)
        else:
            _need(out, "out", torch.float32, numel=L * num_time_bins)
        self.check(self.lib.fbp_render_confocal(self.handle, _ptr(scan_points), L, _ptr(points),
                                                _ptr(albedo), points.shape[0], num_time_bins,
                                                float(delta), _ptr(out), _stream()))
        return out


class Refiner:
    """The learned refiner (fbp_refiner_*): a 3-level 3D U-Net on S^3 volumes."""

    def __init__(self, ctx: Context, size: int, seed: int = 0):
        self.ctx = ctx
        self.lib = ctx.lib
        self.size = int(size)
        h = C.c_void_p()
        ctx.check(self.lib.fbp_refiner_create(ctx.handle, self.size, int(seed), C.byref(h)))
        self.handle = h
        self.n_params = int(self.lib.fbp_refiner_param_count())

    def close(self):
        if getattr(self, "handle", None):
            self.lib.fbp_refiner_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32)
        self.ctx.check(self.lib.fbp_refiner_load(self.handle, p.ctypes.data_as(C.c_void_p), p.size))

    def save(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        self.ctx.check(self.lib.fbp_refiner_save(self.handle, p.ctypes.data_as(C.c_void_p), p.size))
        return p

    def _need_volume(self, t, name, batch=None):
        S = self.size
        shape = (S, S, S) if batch is None else None
        _need(t, name, torch.float32, shape=shape, numel=None if batch is None else batch * S ** 3)

    def forward(self, volume: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        self._need_volume(volume, "volume")
        if out is not None:
            self._need_volume(out, "out")
        out = torch.empty_like(volume) if out is None else out
        self.ctx.check(self.lib.fbp_refiner_forward(self.handle, _ptr(volume), _ptr(out), _stream()))
        return out

    def reconstruct(self, info, intensity, laser_targets, num_samples, seed, out=None,
                    kernel_size=5, sigma=1.0, extremes=16):
        """Learned form in one call (fbp_learned_reconstruct): seeded
        tile-shared MIS, LoG, refiner on its min-max normalisation,
        calibration against that LoG."""
        _need_transient(info, intensity, STORAGE_F32, None, exhaustive=False)
        _need_points(laser_targets, "laser_targets", info.num_laser_targets)
        r = info.voxel_resolution
        if (r.x, r.y, r.z) != (self.size,) * 3:
            raise ValueError(f"voxel grid must be {self.size}^3")
        if out is not None:
            self._need_volume(out, "out")
        out = torch.empty((self.size,) * 3, dtype=torch.float32,
                          device=intensity.device) if out is None else out
        tr = self.ctx._transient(intensity, STORAGE_F32, LAYOUT_TIME_FASTEST, None)
        self.ctx.check(self.lib.fbp_learned_reconstruct(
            self.handle, C.byref(info), C.byref(tr), _ptr(laser_targets), int(num_samples),
            int(seed) & (2 ** 64 - 1), int(kernel_size), float(sigma), int(extremes), _ptr(out),
            _stream()))
        return out

    def train_step(self, inputs: torch.Tensor, targets: torch.Tensor, lr: float,
                   ssim_weight: float = 1.0, want_loss: bool = True) -> float | None:
        batch = inputs.shape[0]
        self._need_volume(inputs, "inputs", batch)
        self._need_volume(targets, "targets", batch)
        loss = C.c_float()
        self.ctx.check(self.lib.fbp_refiner_train_step(
            self.handle, _ptr(inputs), _ptr(targets), batch, float(lr), float(ssim_weight),
            C.byref(loss) if want_loss else None, _stream()))
        return loss.value if want_loss else None

    def gradient(self, volume: torch.Tensor, target: torch.Tensor, ssim_weight: float = 1.0):
        self._need_volume(volume, "volume")
        self._need_volume(target, "target")
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        torch.cuda.synchronize()
        self.ctx.check(self.lib.fbp_refiner_gradient(self.handle, _ptr(volume), _ptr(target),
                                                     float(ssim_weight),
                                                     g.ctypes.data_as(C.c_void_p), C.byref(loss)))
        return g, loss.value
