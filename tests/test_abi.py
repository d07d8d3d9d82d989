"""C ABI checks that need no GPU: the library loads, exports every symbol
include/fbp.h declares, and host-only entry points behave."""
import ctypes as C

import numpy as np
import pytest

import oracle
from fasterbackprojection_amd import _lib


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    names = _lib.header_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the binding knows every one of them
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_struct_layout():
    assert C.sizeof(_lib.ReconstructionInfo) == 148
    offs = {f[0]: getattr(_lib.ReconstructionInfo, f[0]).offset for f in _lib.ReconstructionInfo._fields_}
    # FB/GpuStructs.cuh:8-38 field order, glm packing
    assert offs["capture_system"] == 12 and offs["time_offset"] == 28
    assert offs["time_step"] == 44 and offs["discard_first_last_bounces"] == 60
    assert offs["voxel_resolution"] == 136


def test_status_strings():
    lib = _lib.load()
    assert lib.fbp_status_string(0) == b"ok"
    assert b"unsupported" in lib.fbp_status_string(5)


def test_null_arguments_rejected_without_gpu():
    lib = _lib.load()
    assert lib.fbp_context_create(0, None) == 1
    assert lib.fbp_context_destroy(None) == 1
    assert lib.fbp_backproject_confocal_voxel(None, None, None, None, None, None) == 1


def test_log_kernel_host_matches_fp64_closed_form():
    """The reference's float kernel builder vs the oracle's FP64 closed form."""
    import oracle
    for k, s in [(5, 1.0), (3, 0.7), (7, 1.5), (5, 0.2)]:
        w = np.empty(k ** 3, np.float32)
        assert _lib.load().fbp_log_kernel_host(k, s, w.ctypes.data_as(C.c_void_p)) == 0
        ref = oracle.log_kernel(k, s).ravel()
        np.testing.assert_allclose(w, ref, rtol=1e-5, atol=1e-6 * np.abs(ref).max())
        assert abs(float(w.astype(np.float64).sum())) < 1e-5 * np.abs(ref).max()


def test_alias_tables_match_oracle_and_reproduce_distribution():
    lib = _lib.load()
    import oracle
    rng = np.random.default_rng(0)
    for n in (1, 2, 7, 64, 1000):
        p = rng.random(n)
        p[rng.random(n) < 0.3] = 0
        if p.sum() == 0:
            p[0] = 1
        cdf = np.cumsum(p / p.sum()).astype(np.float32)
        cdf[-1] = 1.0
        a = np.empty(n, np.uint32)
        pr = np.empty(n, np.float32)
        assert lib.fbp_build_alias_tables_host(cdf.ctypes.data_as(C.c_void_p), n,
                                               a.ctypes.data_as(C.c_void_p),
                                               pr.ctypes.data_as(C.c_void_p)) == 0
        oa, op = oracle.alias_tables(cdf)
        np.testing.assert_array_equal(a, oa)
        np.testing.assert_array_equal(pr, op)
        # Vose invariant: P(k) = (prob_k + sum_{j: alias_j = k, prob_j < 1} (1 - prob_j)) / n
        probs = np.diff(np.concatenate([[0.0], cdf.astype(np.float64)]))
        got = pr.astype(np.float64).copy()
        for j in range(n):
            if pr[j] < 1.0:
                got[a[j]] += 1.0 - pr[j]
        np.testing.assert_allclose(got / n, probs, atol=5e-6)


def test_alias_tables_reject_invalid():
    lib = _lib.load()
    bad = np.array([0.5, 0.4, 0.8], np.float32)  # negative step
    a = np.empty(3, np.uint32)
    pr = np.empty(3, np.float32)
    assert lib.fbp_build_alias_tables_host(bad.ctypes.data_as(C.c_void_p), 3,
                                           a.ctypes.data_as(C.c_void_p),
                                           pr.ctypes.data_as(C.c_void_p)) == 6
    short = np.array([0.2, 0.5], np.float32)  # sums to 0.5
    assert lib.fbp_build_alias_tables_host(short.ctypes.data_as(C.c_void_p), 2,
                                           a.ctypes.data_as(C.c_void_p),
                                           pr.ctypes.data_as(C.c_void_p)) == 6


def test_entry_points_reject_null_context_without_gpu():
    """Every device entry point checks its arguments before touching the
    device: a null context (or null buffers) is FBP_ERR_INVALID_ARGUMENT (1)."""
    lib = _lib.load()
    n = None
    assert lib.fbp_backproject_exhaustive_voxel(n, n, n, n, n, n, n) == 1
    assert lib.fbp_backproject_depth_confocal(n, n, n, n, n, 1, n, n) == 1
    assert lib.fbp_backproject_depth_exhaustive(n, n, n, n, n, n, 1, n, n) == 1
    assert lib.fbp_backproject_confocal_mis(n, n, n, n, n, n, n) == 1
    assert lib.fbp_depth_slices_rgba(n, n, 4, 4, 1, 4, 4, n, n) == 1
    assert lib.fbp_filter_laplacian(n, n, n, _lib.UVec3(4, 4, 4), 5, n) == 1
    assert lib.fbp_normalize_minmax(n, n, 16, n) == 1
    assert lib.fbp_lct(n, n, n, n, n) == 1
    assert lib.fbp_backproject_confocal_points(n, n, n, n, n, 1, n, n) == 1
    assert lib.fbp_backproject_confocal_splat(n, n, n, n, n, n) == 1
    assert lib.fbp_render_confocal(n, n, 1, n, n, 1, 16, 0.01, n, n) == 1


def test_filter_h_padded_length_rule_and_rejections():
    """nt_pad = nt + 2 (t6 - 1) with t6 = round(6 sigma / dt) made even
    (42 -> 82 extra bins for wl = 10 bins; 1 -> 2 gives nt + 2); six sigmas
    that round to 0 bins give no padding and are rejected (0)."""
    lib = _lib.load()
    dt = 0.0048
    assert lib.fbp_filter_h_padded_length(1024, dt, 10 * dt, 0.0) == 1106
    for ratio in (0.6, 1.0, 1.4):   # round(6 sigma / dt) = 1
        sigma = ratio * dt / 6
        assert lib.fbp_filter_h_padded_length(1024, dt, 10 * dt, sigma) == 1026, ratio
        assert oracle.filter_h_nt_pad(1024, dt, 10 * dt, sigma) == 1026, ratio
    assert lib.fbp_filter_h_padded_length(1024, dt, 10 * dt, 0.4 * dt / 6) == 0
    assert lib.fbp_filter_h_padded_length(1024, dt, 0.1 * dt, 0.0) == 0
    assert lib.fbp_filter_h_padded_length(1024, dt, 10 * dt, -1.0) == 0
