"""Scatter form of the confocal backprojection (fbp_backproject_confocal_splat).

Every lit sample (l, t) is splatted into the voxels of its shell, the
mathematical form of DrawEllipsoids::solveBackprojection
(FB/DrawEllipsoids.cpp:332-440).  Its volume is backprojectConfocalVoxel's
(FB/transient_reconstruction.cuh:93-136) summed in another order, so the
oracle is oracle.confocal_voxel: within rel-L2 1e-5 (FP32 sums), with bins
checked pair by pair through the ramp transient of test_gpu_bins, the EPS and
safeRCP edges, volumes that straddle the wall plane (two monotone runs per
column), the time-major layout, and the full 128^2 x 1024 -> 128^3 capture
against the gather kernel.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from fasterbackprojection_amd.api import (LAYOUT_TIME_MAJOR, STORAGE_F16, FbpError, _ptr,
                                          _stream, make_info)
from fasterbackprojection_amd.scene import small_confocal
from test_gpu_bins import EPS32, _edge_transient, cuda, geom, ramp_rows, recovered_bins

pytestmark = pytest.mark.gpu

UNSUPPORTED = 5


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def assert_matches(out, ref, tol=1e-5):
    out = out.detach().cpu().numpy() if torch.is_tensor(out) else out
    assert np.all(np.isfinite(out))
    assert rel_l2(out, ref) <= tol, rel_l2(out, ref)
    assert np.max(np.abs(out - ref)) <= 1e-4 * np.abs(ref).max()


def sparse_transient(rng, L, T, density):
    I = rng.uniform(0.1, 1.0, (L, T)).astype(np.float32)
    I[rng.random((L, T)) > density] = 0.0
    return I


# (name, geometry kwargs, density): pure chain, general chain, volumes that
# cross the wall plane (depth centres on both sides of the scan points), tiles
# cut by the volume edge on every axis, shells thinner and thicker than voxels.
CASES = {
    "pure_thin_shells": (dict(n_scan=8, T=512, res=(20, 18, 37)), 0.3),
    "general_t0": (dict(n_scan=8, T=512, res=(17, 9, 12), discard=0, t0=0.37), 0.8),
    "pure_crosses_wall": (dict(n_scan=6, T=256, res=(9, 7, 40), hmin=(-0.4, -0.3, -0.45),
                               hmax=(0.45, 0.9, 0.4)), 1.0),
    "general_crosses_wall_sparse": (dict(n_scan=6, T=400, res=(9, 7, 10), discard=0, t0=0.21,
                                         hmin=(-0.4, -0.3, -0.45), hmax=(0.45, 0.9, 0.4)), 0.05),
    "thick_shells": (dict(n_scan=8, T=256, res=(33, 17, 70), hmin=(-0.1, 0.4, -0.1),
                          hmax=(0.1, 0.6, 0.1)), 0.5),
    # under 1/64 of a window's bins lit: the jump walk
    "pure_jump": (dict(n_scan=8, T=1024, res=(20, 18, 37)), 0.01),
    "general_jump": (dict(n_scan=8, T=1024, res=(17, 9, 40), discard=0, t0=0.37), 0.008),
    "pure_jump_crosses_wall": (dict(n_scan=6, T=1024, res=(9, 7, 40), hmin=(-0.4, -0.3, -0.45),
                                    hmax=(0.45, 0.9, 0.4)), 0.01),
    "thick_shells_jump": (dict(n_scan=8, T=1024, res=(33, 17, 70), hmin=(-0.1, 0.4, -0.1),
                               hmax=(0.1, 0.6, 0.1)), 0.01),
}


@pytest.mark.parametrize("name", list(CASES))
def test_splat_matches_oracle(ctx, name):
    kw, density = CASES[name]
    g = geom(**kw)
    info = make_info(g)
    rng = np.random.default_rng(len(name))
    I = sparse_transient(rng, len(g.laser_targets), g.num_time_bins, density)
    ref = oracle.confocal_voxel(info, I, g.laser_targets)
    out = ctx.backproject_confocal_splat(info, cuda(I), cuda(g.laser_targets))
    assert_matches(out, ref)


def test_splat_time_major_layout(ctx):
    g = geom(n_scan=8, T=300, res=(12, 10, 9), discard=0, t0=0.1)
    info = make_info(g)
    I = sparse_transient(np.random.default_rng(7), 64, 300, 0.6)
    ref = oracle.confocal_voxel(info, I, g.laser_targets)
    out = ctx.backproject_confocal_splat(info, cuda(I.T.copy()), cuda(g.laser_targets),
                                         layout=LAYOUT_TIME_MAJOR)
    assert_matches(out, ref)


@pytest.mark.parametrize("discard,t0", [(1, 0.0), (0, 0.37)])
def test_splat_bins_exact(ctx, discard, t0):
    """Ramp transient I[l][t] = t + 1 on one scan point at a time: every
    (voxel, scan point) pair's bin, recovered from vol * dist^2, equals the
    oracle's (int)(dist / Delta)."""
    g = geom(T=1024, discard=discard, t0=t0)
    info = make_info(g)
    lasers = cuda(g.laser_targets)
    for l in range(0, 64, 3):
        vol = ctx.backproject_confocal_splat(info, cuda(ramp_rows(64, 1024, [l])), lasers)
        bins, dist = oracle.confocal_bins(info, g.laser_targets, l)
        rec = recovered_bins(vol.cpu().numpy(), dist)
        bad = np.argwhere(rec != bins)
        assert bad.size == 0, (l, bad[:5].tolist())


@pytest.mark.parametrize("discard,t0", [(1, 0.0), (0, 0.37)])
def test_splat_bins_exact_sparse(ctx, discard, t0):
    """Ramp on two bins per scan point only (about 1 % of the window lit: the
    jump walk), chosen among the bins the voxels reach: voxels whose bin is
    lit recover it exactly, the others hold 0."""
    g = geom(T=1024, discard=discard, t0=t0)
    info = make_info(g)
    lasers = cuda(g.laser_targets)
    rng = np.random.default_rng(11)
    hits = 0
    for l in range(1, 64, 5):
        bins, dist = oracle.confocal_bins(info, g.laser_targets, l)
        reach = np.unique(bins[bins >= 0])
        lit = rng.choice(reach, size=2, replace=False)
        I = np.zeros((64, 1024), np.float32)
        I[l, lit] = lit.astype(np.float32) + 1.0
        vol = ctx.backproject_confocal_splat(info, cuda(I), lasers)
        want = np.where(np.isin(bins, lit), bins, -1)
        rec = recovered_bins(vol.cpu().numpy(), dist)
        bad = np.argwhere(rec != want)
        assert bad.size == 0, (l, bad[:5].tolist())
        hits += int((want >= 0).sum())
    assert hits >= 26


def test_splat_eps_threshold(ctx):
    """Samples of exactly 1e-5f are dropped, the next float above is kept."""
    g = geom(T=512)
    info = make_info(g)
    I = _edge_transient(np.random.default_rng(2), 64, 512)
    I[:, ::7] = EPS32
    ref = oracle.confocal_voxel(info, I, g.laser_targets)
    out = ctx.backproject_confocal_splat(info, cuda(I), cuda(g.laser_targets))
    assert_matches(out, ref)


def test_splat_safe_rcp_clamp(ctx):
    """t0 = one pair's exact 2r: that pair's weight is safeRCP(0) = FLT_MAX."""
    base = geom(n_scan=8, T=512, discard=1)
    l0, (vx0, vy0, z0) = 27, (3, 2, 1)
    d0 = oracle.confocal_distf(make_info(base), base.laser_targets, l0, vx0, vy0, z0)
    g = geom(n_scan=8, T=512, discard=1, t0=float(d0))
    info = make_info(g)
    rng = np.random.default_rng(5)
    I = rng.uniform(0.1, 1.0, (64, 512)).astype(np.float32)
    I[:, 0] = rng.uniform(1e-4, 1e-3, 64).astype(np.float32)
    ref = oracle.confocal_voxel(info, I, g.laser_targets)
    clamp = ref > 1e30
    assert clamp[z0, vy0, vx0]
    out = ctx.backproject_confocal_splat(info, cuda(I), cuda(g.laser_targets)).cpu().numpy()
    assert np.all(np.isfinite(out))
    np.testing.assert_allclose(out[clamp], ref[clamp], rtol=1e-6)
    assert rel_l2(out[~clamp], ref[~clamp]) <= 1e-5


def test_splat_small_capture_and_determinism(ctx):
    """The 32^2 x 256 capture of 200 points with Poisson noise -> 32^3:
    against the oracle, and bit-identical on a second call."""
    cap = small_confocal(seed=3)
    info = make_info(cap.geometry)
    I, lasers = cuda(cap.transient), cuda(cap.geometry.laser_targets)
    ref = oracle.confocal_voxel(info, cap.transient, cap.geometry.laser_targets)
    a = ctx.backproject_confocal_splat(info, I, lasers)
    b = ctx.backproject_confocal_splat(info, I, lasers)
    assert torch.equal(a, b)
    assert_matches(a, ref)


def test_splat_full_capture_matches_gather(ctx):
    """128^2 x 1024 -> 128^3 (4 random quads, Poisson at peak 1000): the
    splat volume equals the gather kernel's, and its first two depth slices
    the oracle's."""
    from test_gpu_full_size import confocal_capture, first_slices
    g, laser, I = confocal_capture(ctx, 128, 4, 1000.0, 11)
    info = make_info(g)
    ctx.set_aligned_path(False)
    try:
        gather = ctx.backproject_confocal(info, I, laser)
    finally:
        ctx.set_aligned_path(True)
    splat = ctx.backproject_confocal_splat(info, I, laser)
    assert rel_l2(splat.cpu().numpy(), gather.cpu().numpy()) <= 1e-5
    ref = oracle.confocal_voxel(first_slices(info, 2), I.cpu().numpy(), g.laser_targets)
    assert rel_l2(splat[:2].cpu().numpy(), ref) <= 1e-5


def test_splat_rejects_unsupported(ctx):
    """F16 storage and the voxel-scattering option have no scatter form: the
    call fails with FBP_ERR_UNSUPPORTED and the context stays usable."""
    g = geom(T=256)
    info = make_info(g)
    I = cuda(np.ones((64, 256), np.float32))
    lasers = cuda(g.laser_targets)
    q, sc = ctx.quantize(I, STORAGE_F16)
    vol = torch.empty((info.voxel_resolution.z, info.voxel_resolution.y,
                       info.voxel_resolution.x), dtype=torch.float32, device=I.device)
    tr = ctx._transient(q, STORAGE_F16, 0, sc)
    st = ctx.lib.fbp_backproject_confocal_splat(ctx.handle, C.byref(info), C.byref(tr),
                                                _ptr(lasers), _ptr(vol), _stream())
    assert st == UNSUPPORTED
    ctx.set_voxel_scattering(True)
    try:
        with pytest.raises(FbpError, match="scatter"):
            ctx.backproject_confocal_splat(info, I, lasers)
    finally:
        ctx.set_voxel_scattering(False)
    out = ctx.backproject_confocal_splat(info, I, lasers)
    assert torch.isfinite(out).all()
