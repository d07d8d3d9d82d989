"""ctypes binding of libfbp.so (the C ABI declared in include/fbp.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libfbp.so"
HEADER = PKG.parent / "include" / "fbp.h"

FBP_OK = 0
STATUS_NAMES = {
    0: "FBP_OK", 1: "FBP_ERR_INVALID_ARGUMENT", 2: "FBP_ERR_DEVICE", 3: "FBP_ERR_OUT_OF_MEMORY",
    4: "FBP_ERR_FFT", 5: "FBP_ERR_UNSUPPORTED", 6: "FBP_ERR_PROBABILITIES", 7: "FBP_ERR_IO",
}

CONFOCAL, EXHAUSTIVE = 0, 1
STORAGE_F32, STORAGE_F16, STORAGE_BF16 = 0, 1, 2
LAYOUT_TIME_FASTEST, LAYOUT_TIME_MAJOR = 0, 1
MIS_REFERENCE, MIS_UNBIASED, MIS_TILED = 0, 1, 2
OPTION_CONFOCAL_ALIGNED = 0
OPTION_VOXEL_SCATTERING = 1
OPTION_EXHAUSTIVE_LASER_GROUPS = 2


class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class UVec3(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("z", C.c_uint32)]


class ReconstructionInfo(C.Structure):
    """Mirror of fbp_reconstruction_info / ReconstructionInfo (FB/GpuStructs.cuh:8-38)."""
    _fields_ = [
        ("sensor_position", Vec3), ("capture_system", C.c_uint32),
        ("laser_position", Vec3), ("time_offset", C.c_float),
        ("num_laser_targets", C.c_uint32), ("num_sensor_targets", C.c_uint32),
        ("num_time_bins", C.c_uint32), ("time_step", C.c_float),
        ("hidden_volume_min", Vec3), ("discard_first_last_bounces", C.c_uint32),
        ("hidden_volume_max", Vec3), ("hidden_volume_size", Vec3),
        ("hidden_volume_voxel_size", Vec3),
        ("relay_wall_min_position", Vec3), ("relay_wall_size", Vec3),
        ("relay_wall_normal", Vec3),
        ("voxel_resolution", UVec3),
    ]


assert C.sizeof(ReconstructionInfo) == 148


class Transient(C.Structure):
    _fields_ = [("data", C.c_void_p), ("storage", C.c_int32), ("layout", C.c_int32),
                ("scales", C.c_void_p)]


class MisParams(C.Structure):
    _fields_ = [("num_samples", C.c_uint32), ("mode", C.c_uint32), ("noise", C.c_void_p),
                ("noise_size", C.c_uint32), ("seed", C.c_uint64)]


class LctParams(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint32), ("width", C.c_float),
                ("range_max", C.c_float), ("snr", C.c_float), ("is_diffuse", C.c_int32),
                ("is_backprop", C.c_int32), ("z_trim", C.c_uint32), ("downsample", C.c_uint32)]


_P = C.c_void_p
_S = C.c_int  # fbp_status
_CTX = C.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "fbp_status_string": (C.c_char_p, [C.c_int]),
    "fbp_context_create": (_S, [C.c_int, C.POINTER(_CTX)]),
    "fbp_context_destroy": (_S, [_CTX]),
    "fbp_context_last_error": (C.c_char_p, [_CTX]),
    "fbp_context_device": (C.c_int, [_CTX]),
    "fbp_context_scratch_bytes": (C.c_size_t, [_CTX]),
    "fbp_context_set_option": (_S, [_CTX, C.c_int32, C.c_int32]),
    "fbp_context_get_option": (_S, [_CTX, C.c_int32, C.POINTER(C.c_int32)]),
    "fbp_context_last_confocal_path": (_S, [_CTX, C.POINTER(C.c_int32)]),
    "fbp_context_last_exhaustive_kernel": (_S, [_CTX, C.POINTER(C.c_int32)]),
    "fbp_backproject_confocal_voxel": (_S, [_CTX, C.POINTER(ReconstructionInfo), C.POINTER(Transient), _P, _P, _P]),
    "fbp_backproject_confocal_splat": (_S, [_CTX, C.POINTER(ReconstructionInfo), C.POINTER(Transient), _P, _P, _P]),
    "fbp_backproject_exhaustive_voxel": (_S, [_CTX, C.POINTER(ReconstructionInfo), C.POINTER(Transient), _P, _P, _P, _P]),
    "fbp_backproject_exhaustive_voxel_noalloc": (_S, [_CTX, C.POINTER(ReconstructionInfo), C.POINTER(Transient), _P, _P, _P, _P]),
    "fbp_backproject_depth_confocal": (_S, [_CTX, C.POINTER(ReconstructionInfo), C.POINTER(Transient), _P, _P, C.c_uint32, _P, _P]),
    "fbp_backproject_depth_exhaustive": (_S, [_CTX, C.POINTER(ReconstructionInfo), C.POINTER(Transient), _P, _P, _P, C.c_uint32, _P, _P]),
    "fbp_depth_slices_rgba": (_S, [_CTX, _P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P, _P]),
    "fbp_write_png_rgba": (_S, [C.c_char_p, _P, C.c_uint32, C.c_uint32]),
    "fbp_backproject_confocal_mis": (_S, [_CTX, C.POINTER(ReconstructionInfo), C.POINTER(Transient), _P, C.POINTER(MisParams), _P, _P]),
    "fbp_build_alias_tables_host": (_S, [_P, C.c_uint32, _P, _P]),
    "fbp_quantize_transient": (_S, [_CTX, _P, C.c_uint64, C.c_uint32, C.c_int32, _P, _P, _P]),
    "fbp_filter_laplacian": (_S, [_CTX, _P, _P, UVec3, C.c_int32, _P]),
    "fbp_filter_log": (_S, [_CTX, _P, _P, UVec3, C.c_int32, C.c_float, _P]),
    "fbp_log_kernel_host": (_S, [C.c_int32, C.c_float, _P]),
    "fbp_filter_log_fft": (_S, [_CTX, _P, UVec3, C.c_float, _P]),
    "fbp_normalize_minmax": (_S, [_CTX, _P, C.c_uint64, _P]),
    "fbp_filter_log_normalize": (_S, [_CTX, _P, _P, UVec3, C.c_int32, C.c_float, _P]),
    "fbp_minmax": (_S, [_CTX, _P, C.c_uint64, C.POINTER(C.c_float), C.POINTER(C.c_float), _P]),
    "fbp_filter_h": (_S, [_CTX, _P, C.c_uint64, C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_int32, _P]),
    "fbp_filter_h_padded_length": (C.c_uint32, [C.c_uint32, C.c_float, C.c_float, C.c_float]),
    "fbp_lct": (_S, [_CTX, _P, C.POINTER(LctParams), _P, _P]),
    "fbp_render_confocal": (_S, [_CTX, _P, C.c_uint32, _P, _P, C.c_uint32, C.c_uint32, C.c_float, _P, _P]),
    "fbp_render_exhaustive": (_S, [_CTX, _P, C.c_uint32, _P, C.c_uint32, _P, _P, C.c_uint32, C.c_uint32, C.c_float, _P, _P]),
    "fbp_backproject_confocal_points": (_S, [_CTX, C.POINTER(ReconstructionInfo), C.POINTER(Transient), _P, _P, C.c_uint32, _P, _P]),
    "fbp_learned_calibrate": (_S, [_CTX, C.POINTER(ReconstructionInfo), C.POINTER(Transient), _P, _P, _P, C.c_int32, C.c_float, C.c_uint32, _P]),
    "fbp_volume_residual": (_S, [_CTX, C.POINTER(ReconstructionInfo), _P, _P, _P, _P, C.c_uint32, C.c_float, C.c_int32, C.c_int32, C.POINTER(C.c_double), _P]),
    "fbp_refiner_param_count": (C.c_size_t, []),
    "fbp_refiner_create": (_S, [_CTX, C.c_uint32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "fbp_refiner_destroy": (_S, [C.c_void_p]),
    "fbp_refiner_load": (_S, [C.c_void_p, _P, C.c_uint64]),
    "fbp_refiner_save": (_S, [C.c_void_p, _P, C.c_uint64]),
    "fbp_refiner_forward": (_S, [C.c_void_p, _P, _P, _P]),
    "fbp_learned_reconstruct": (_S, [C.c_void_p, C.POINTER(ReconstructionInfo), C.POINTER(Transient), _P, C.c_uint32, C.c_uint64, C.c_int32, C.c_float, C.c_uint32, _P, _P]),
    "fbp_refiner_train_step": (_S, [C.c_void_p, _P, _P, C.c_uint32, C.c_float, C.c_float, C.POINTER(C.c_float), _P]),
    "fbp_refiner_gradient": (_S, [C.c_void_p, _P, _P, C.c_float, _P, C.POINTER(C.c_float)]),
}


def header_symbols() -> list[str]:
    """Every function name declared in include/fbp.h."""
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fbp_[a-z0-9_]+)\s*\(", text)))


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is not built: run `python -m fasterbackprojection_amd.build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class FbpError(RuntimeError):
    """Non-zero fbp_status; mirrors the reference's std::runtime_error."""

    def __init__(self, status: int, detail: str):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {detail}")
