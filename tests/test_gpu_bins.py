"""Bit-exact bins on the GPU, pair by pair, and the EPS / safeRCP edges.

Bin recovery: light one scan point l (one (laser, sensor) pair for the
exhaustive form) with I[l][t] = t + 1.  Every voxel then holds
(bin + 1) * safeRCP(dist^2), so round(vol * dist^2) - 1 is the bin the
kernel used for that (voxel, l) pair (0 when the bin is outside [0, T)),
which must equal the oracle's (int)(dist / Delta) of the reference's FP32
expression (FB/transient_reconstruction.cuh:102-130, 150-161) exactly.  The
depth-plane mode is unweighted, so its activation is bin + 1 directly.  A
wrong bin in a single pair fails these tests; volume tolerances cannot hide
it.

Edges: samples exactly at 1e-5f must be dropped and the next float above
kept (`I > EPS`, :133); pairs with dist^2 <= 1e-5 take safeRCP's FLT_MAX
(FB/math.cuh:23-29), and dist in (-Delta, 0) truncates to bin 0.
"""
import numpy as np
import pytest
import torch

import oracle
from fasterbackprojection_amd.api import STORAGE_BF16, STORAGE_F16, Geometry, make_info
from fasterbackprojection_amd.scene import wall_grid

pytestmark = pytest.mark.gpu

EPS32 = np.float32(1e-5)
FLT_MAX = float(np.finfo(np.float32).max)


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def geom(n_scan=8, T=512, res=(6, 5, 7), discard=1, t0=0.0, wall=1.0,
         hmin=(-0.4, 0.35, -0.45), hmax=(0.45, 1.25, 0.4)):
    scan = wall_grid(n_scan, wall)
    return Geometry(laser_targets=scan, sensor_targets=scan, num_time_bins=T, time_step=0.0096,
                    hidden_min=hmin, hidden_max=hmax, voxel_resolution=res,
                    discard_first_last_bounces=discard, time_offset=t0,
                    laser_position=(0.1, -0.8, 0.2), sensor_position=(-0.2, -0.7, 0.1),
                    wall_size=(wall, wall))


def aligned_geom(nx, wall, X, VY, ZD, T, depth0, dpitch, R=1):
    """Scan pitch wall / nx, voxel pitch R * wall / nx along world x and z:
    the aligned-grid (ring convolution) path applies."""
    pitch = R * wall / nx
    scan = wall_grid(nx, wall)
    return Geometry(laser_targets=scan, sensor_targets=scan, num_time_bins=T, time_step=0.0096,
                    hidden_min=(-wall / 2, depth0, -wall / 2),
                    hidden_max=(-wall / 2 + X * pitch, depth0 + VY * dpitch, -wall / 2 + ZD * pitch),
                    voxel_resolution=(X, VY, ZD), wall_size=(wall, wall))


def recovered_bins(vol, dist):
    """round(vol * dist^2) - 1, in FP64 (vol holds (bin + 1) / dist^2)."""
    v = vol.astype(np.float64) * dist.astype(np.float64) ** 2
    return np.rint(v).astype(np.int64) - 1


def ramp_rows(L, T, rows):
    I = np.zeros((L, T), np.float32)
    I[rows] = np.arange(1, T + 1, dtype=np.float32)
    return I


def check_confocal_bins(ctx, g, scan_points, aligned, storage=None):
    info = make_info(g)
    L, T = len(g.laser_targets), g.num_time_bins
    lasers = cuda(g.laser_targets)
    ctx.set_aligned_path(aligned)
    vols = []
    try:
        for l in scan_points:
            I = ramp_rows(L, T, [l])
            if storage is None:
                vols.append(ctx.backproject_confocal(info, cuda(I), lasers))
            else:
                q, sc = ctx.quantize(cuda(I), storage)
                vols.append(ctx.backproject_confocal(info, q, lasers, storage=storage, scales=sc))
            assert ctx.last_confocal_path_aligned() == aligned, l
        got = torch.stack(vols).cpu().numpy()
    finally:
        ctx.set_aligned_path(True)
    checked = 0
    for i, l in enumerate(scan_points):
        bins, dist = oracle.confocal_bins(info, g.laser_targets, l)
        rec = recovered_bins(got[i], dist)
        bad = np.argwhere(rec != bins)
        assert bad.size == 0, (l, bad[:5].tolist(), rec[tuple(bad[0])], bins[tuple(bad[0])])
        checked += bins.size
    return checked


@pytest.mark.parametrize("discard,t0", [(1, 0.0), (0, 0.0), (0, 0.37), (1, 0.21)])
def test_confocal_general_kernel_bins_exact(ctx, discard, t0):
    """Every (voxel, scan point) pair of a 64-point scan on the general
    kernel, pure chain (t0 = 0, bounces discarded) and general chain."""
    g = geom(T=1024, discard=discard, t0=t0)
    n = check_confocal_bins(ctx, g, list(range(64)), aligned=False)
    assert n == 64 * 6 * 5 * 7


@pytest.mark.parametrize("storage", [STORAGE_F16, STORAGE_BF16])
def test_confocal_general_kernel_bins_exact_quantized(ctx, storage):
    """F16 / BF16 storage: (t + 1) / max is a multiple of 1/256 for T = 256,
    exact in both formats, so the recovered bins are exact too."""
    g = geom(T=256)
    check_confocal_bins(ctx, g, list(range(0, 64, 3)), aligned=False, storage=storage)


ALIGNED = {
    # one scan point per voxel pitch
    "r1_64": dict(nx=64, wall=1.0, X=64, VY=16, ZD=8, T=256, depth0=0.25, dpitch=0.02),
    # two scan points per voxel pitch (four phase planes)
    "r2_64": dict(nx=128, wall=1.0, R=2, X=64, VY=16, ZD=6, T=256, depth0=0.25, dpitch=0.02),
}


def _aligned_scan_points(nx, seed):
    """Corners and centre of the scan grid (every ring offset in [-(n-1), n-1]
    on both axes is then met), both phases of each corner for n = 2X, and
    random points."""
    corners = [(0, 0), (0, nx - 1), (nx - 1, 0), (nx - 1, nx - 1), (1, 1), (1, nx - 2),
               (nx - 2, 1), (nx - 2, nx - 2), (nx // 2, nx // 2), (nx // 2 + 1, nx // 2 - 1)]
    pts = [iz * nx + ix for iz, ix in corners]
    rng = np.random.default_rng(seed)
    pts += rng.choice(nx * nx, 10, replace=False).tolist()
    return sorted(set(pts))


@pytest.mark.parametrize("name", sorted(ALIGNED))
def test_confocal_aligned_path_bins_exact(ctx, name):
    cfg = ALIGNED[name]
    g = aligned_geom(**cfg)
    pts = _aligned_scan_points(cfg["nx"], 3)
    check_confocal_bins(ctx, g, pts, aligned=True)


@pytest.mark.parametrize("storage", [STORAGE_F16, STORAGE_BF16])
def test_confocal_aligned_path_bins_exact_quantized(ctx, storage):
    cfg = ALIGNED["r1_64"]
    g = aligned_geom(**cfg)
    check_confocal_bins(ctx, g, _aligned_scan_points(cfg["nx"], 5)[:8], aligned=True,
                        storage=storage)


@pytest.mark.parametrize("discard,t0,dt,depth0", [(1, 0.0, 0.0096, 0.35), (0, 0.29, 0.0096, 0.35),
                                                  # bins >= 160: linear-weight kernel
                                                  (1, 0.0, 0.0048, 0.5), (0, 0.1, 0.0048, 0.5)])
def test_exhaustive_bins_exact(ctx, discard, t0, dt, depth0):
    """Every (voxel, laser, sensor) triple with 4 lasers x 9 sensors on
    different grids, one lit pair per call (index l*S*T + s*T + t).  With
    linear weights vol * dist^2 = (bin + 1)(1 - e), e <= 3e-5, which still
    rounds to bin + 1 for bins below 10^4."""
    lasers = wall_grid(2, 0.5)
    sensors = wall_grid(3, 1.0)
    T = 1024 if dt > 0.005 else 2048
    g = Geometry(laser_targets=lasers, sensor_targets=sensors, num_time_bins=T, time_step=dt,
                 hidden_min=(-0.4, depth0, -0.45), hidden_max=(0.45, depth0 + 0.9, 0.4),
                 voxel_resolution=(8, 6, 5), confocal=False, discard_first_last_bounces=discard,
                 time_offset=t0, laser_position=(0.1, -0.8, 0.2),
                 sensor_position=(-0.2, -0.7, 0.1), wall_size=(1.0, 1.0))
    info = make_info(g)
    cl, cs = cuda(lasers), cuda(sensors)
    vols = []
    for l in range(4):
        for s in range(9):
            I = np.zeros((4, 9, T), np.float32)
            I[l, s] = np.arange(1, T + 1, dtype=np.float32)
            vols.append(ctx.backproject_exhaustive(info, cuda(I), cl, cs))
    got = torch.stack(vols).cpu().numpy()
    for i in range(36):
        l, s = divmod(i, 9)
        bins, dist = oracle.exhaustive_bins(info, lasers, sensors, l, s)
        rec = recovered_bins(got[i], dist)
        bad = np.argwhere(rec != bins)
        assert bad.size == 0, (l, s, bad[:5].tolist())
        assert (bins >= 0).mean() > 0.9


def _depth_check(ctx, info, lasers, depths, lit, sensors=None, exhaustive=False, aligned=True):
    L, T = info.num_laser_targets, info.num_time_bins
    ctx.set_aligned_path(aligned)
    acts = []
    try:
        for item in lit:
            if exhaustive:
                l, s = item
                I = np.zeros((L, len(sensors), T), np.float32)
                I[l, s] = np.arange(1, T + 1, dtype=np.float32)
                acts.append(ctx.backproject_depth(info, cuda(I), cuda(lasers), cuda(depths),
                                                  cuda(sensors)))
            else:
                acts.append(ctx.backproject_depth(info, cuda(ramp_rows(L, T, [item])),
                                                  cuda(lasers), cuda(depths)))
                assert ctx.last_confocal_path_aligned() == aligned, item
        got = torch.stack(acts).cpu().numpy()
    finally:
        ctx.set_aligned_path(True)
    for i, item in enumerate(lit):
        l, s = item if exhaustive else (item, 0)
        bins = oracle.depth_bins(info, lasers, depths, l, sensors, s, exhaustive)
        # activation = bin + 1 exactly (integers below 2^24 in FP32); 0 when out of range
        rec = np.rint(got[i].astype(np.float64)).astype(np.int64) - 1
        assert np.array_equal(got[i], np.rint(got[i])), item
        bad = np.argwhere(rec != bins)
        assert bad.size == 0, (item, bad[:5].tolist())


@pytest.mark.parametrize("discard,t0", [(1, 0.0), (0, 0.13)])
def test_depth_confocal_gather_bins_exact(ctx, discard, t0):
    g = geom(n_scan=6, T=1024, discard=discard, t0=t0)
    info = make_info(g)
    depths = np.linspace(0.2, 1.1, 9) + 0.0013
    _depth_check(ctx, info, g.laser_targets, depths, list(range(36)), aligned=False)


def test_depth_confocal_aligned_bins_exact(ctx):
    """64^2 scan grid with the wall normal along y: the ring-convolution
    depth path, scan points at the corners, centre and random places."""
    scan = wall_grid(64, 1.0)
    g = Geometry(laser_targets=scan, sensor_targets=scan, num_time_bins=256, time_step=0.0096,
                 hidden_min=(-0.5, 0.2, -0.5), hidden_max=(0.5, 1.0, 0.5),
                 voxel_resolution=(16, 16, 16), wall_size=(1.0, 1.0))
    info = make_info(g)
    depths = np.linspace(0.3, 0.5, 12, endpoint=False)
    pts = _aligned_scan_points(64, 9)
    _depth_check(ctx, info, scan, depths, pts, aligned=True)


@pytest.mark.parametrize("discard,t0", [(1, 0.0), (0, 0.11)])
def test_depth_exhaustive_bins_exact(ctx, discard, t0):
    lasers = wall_grid(3, 0.9)
    sensors = wall_grid(2, 0.5)
    g = Geometry(laser_targets=lasers, sensor_targets=sensors, num_time_bins=1024,
                 time_step=0.0096, hidden_min=(-0.4, 0.35, -0.45), hidden_max=(0.45, 1.25, 0.4),
                 voxel_resolution=(4, 4, 4), confocal=False, discard_first_last_bounces=discard,
                 time_offset=t0, laser_position=(0.1, -0.8, 0.2),
                 sensor_position=(-0.2, -0.7, 0.1), wall_size=(1.0, 1.0))
    info = make_info(g)
    depths = np.linspace(0.2, 1.0, 11) + 0.0013
    lit = [(l, s) for l in range(9) for s in range(4)]
    _depth_check(ctx, info, lasers, depths, lit, sensors, exhaustive=True)


# ------------------------------------------------------------------- edges
def _edge_transient(rng, L, T):
    """Samples at exactly 1e-5f (dropped), the next float above (kept),
    5e-6 (dropped), 0 and 2e-5: every kept value is of the threshold's
    magnitude, so the dropped ones would move the volume visibly."""
    above = np.nextafter(EPS32, np.float32(1))
    choices = np.array([EPS32, above, np.float32(5e-6), 0.0, 2e-5], np.float32)
    return choices[rng.integers(0, len(choices), (L, T))]


def _assert_close(gpu, ref, tol=1e-5):
    gpu = gpu.detach().cpu().numpy().astype(np.float64)
    scale = np.abs(ref).max()
    assert scale > 0
    err = np.linalg.norm(gpu - ref) / np.linalg.norm(ref)
    assert err <= tol, err
    assert np.max(np.abs(gpu - ref)) <= 1e-4 * scale


def test_threshold_edge_general_and_aligned(ctx):
    """I > EPS: a sample of exactly 1e-5f must be dropped and the next float
    above it kept, on both confocal paths.  An inclusive test (>=) would
    add every 1e-5 sample and move the volume by a large fraction."""
    rng = np.random.default_rng(2)
    for g, aligned in [(geom(T=512), False), (aligned_geom(**ALIGNED["r1_64"]), True),
                       (aligned_geom(**ALIGNED["r2_64"]), True)]:
        info = make_info(g)
        I = _edge_transient(rng, len(g.laser_targets), g.num_time_bins)
        I[:, ::7] = EPS32      # a dense set of exactly-at-threshold samples
        ref = oracle.confocal_voxel(info, I, g.laser_targets)
        incl = oracle.confocal_voxel(info, np.where(I == EPS32, np.nextafter(EPS32, 1), I),
                                     g.laser_targets)
        assert np.linalg.norm(incl - ref) > 0.05 * np.linalg.norm(ref)
        ctx.set_aligned_path(aligned)
        try:
            out = ctx.backproject_confocal(info, cuda(I), cuda(g.laser_targets))
            assert ctx.last_confocal_path_aligned() == aligned
        finally:
            ctx.set_aligned_path(True)
        _assert_close(out, ref)


def test_threshold_edge_exhaustive_and_depth(ctx):
    rng = np.random.default_rng(3)
    lasers = wall_grid(2, 0.5)
    sensors = wall_grid(3, 1.0)
    g = Geometry(laser_targets=lasers, sensor_targets=sensors, num_time_bins=300,
                 time_step=0.0096, hidden_min=(-0.4, 0.35, -0.45), hidden_max=(0.45, 1.25, 0.4),
                 voxel_resolution=(8, 6, 5), confocal=False, wall_size=(1.0, 1.0))
    info = make_info(g)
    I = _edge_transient(rng, 36, 300).reshape(4, 9, 300)
    I[:, :, ::5] = EPS32
    ref = oracle.exhaustive_voxel(info, I, lasers, sensors)
    out = ctx.backproject_exhaustive(info, cuda(I), cuda(lasers), cuda(sensors))
    _assert_close(out, ref)
    depths = np.linspace(0.3, 1.0, 8) + 0.0013
    refd = oracle.depth(info, I, lasers, depths, sensors, exhaustive=True)
    outd = ctx.backproject_depth(info, cuda(I), cuda(lasers), cuda(depths), cuda(sensors))
    _assert_close(outd, refd)
    gc = geom(n_scan=6, T=300)
    infoc = make_info(gc)
    Ic = _edge_transient(rng, 36, 300)
    refc = oracle.depth(infoc, Ic, gc.laser_targets, depths)
    ctx.set_aligned_path(False)
    try:
        outc = ctx.backproject_depth(infoc, cuda(Ic), cuda(gc.laser_targets), cuda(depths))
    finally:
        ctx.set_aligned_path(True)
    _assert_close(outc, refc)


def test_safe_rcp_clamp_confocal(ctx):
    """t0 set to one pair's exact 2r: that pair's dist is 0, so its weight is
    safeRCP(0) = FLT_MAX; pairs with dist in (-Delta, 0) read bin 0.  The
    transient's bin-0 samples are small (<= 1e-3) so I * FLT_MAX stays finite."""
    base = geom(n_scan=8, T=512, discard=1)
    info0 = make_info(base)
    l0, (vx0, vy0, z0) = 27, (3, 2, 1)
    d0 = oracle.confocal_distf(info0, base.laser_targets, l0, vx0, vy0, z0)
    g = geom(n_scan=8, T=512, discard=1, t0=float(d0))
    info = make_info(g)
    bins, dist = oracle.confocal_bins(info, g.laser_targets, l0)
    assert dist[z0, vy0, vx0] == 0.0
    rng = np.random.default_rng(5)
    I = rng.uniform(0.1, 1.0, (64, 512)).astype(np.float32)
    I[:, 0] = rng.uniform(1e-4, 1e-3, 64).astype(np.float32)
    ref = oracle.confocal_voxel(info, I, g.laser_targets)
    clamp = ref > 1e30
    assert clamp[z0, vy0, vx0]
    # voxels that read bin 0 from a negative dist exist too
    neg = [(oracle.confocal_bins(info, g.laser_targets, l)[1] < 0).sum() for l in range(64)]
    assert sum(neg) > 0
    out = ctx.backproject_confocal(info, cuda(I), cuda(g.laser_targets)).cpu().numpy()
    assert np.all(np.isfinite(out))
    np.testing.assert_allclose(out[clamp], ref[clamp], rtol=1e-6)
    rest = ~clamp
    err = np.linalg.norm(out[rest] - ref[rest]) / np.linalg.norm(ref[rest])
    assert err <= 1e-5, err


def test_safe_rcp_clamp_exhaustive(ctx):
    """Exhaustive form: t0 set to |l0 - p| + |s0 - p| of one triple."""
    lasers = wall_grid(2, 0.5)
    sensors = wall_grid(3, 1.0)

    def g_of(t0):
        return Geometry(laser_targets=lasers, sensor_targets=sensors, num_time_bins=512,
                        time_step=0.0096, hidden_min=(-0.4, 0.35, -0.45),
                        hidden_max=(0.45, 1.25, 0.4), voxel_resolution=(8, 6, 5), confocal=False,
                        time_offset=t0, wall_size=(1.0, 1.0))
    _, d = oracle.exhaustive_bins(make_info(g_of(0.0)), lasers, sensors, 2, 4)
    z0, vy0, vx0 = 2, 3, 4
    info = make_info(g_of(float(d[z0, vy0, vx0])))
    _, d1 = oracle.exhaustive_bins(info, lasers, sensors, 2, 4)
    assert d1[z0, vy0, vx0] == 0.0
    rng = np.random.default_rng(6)
    I = rng.uniform(0.1, 1.0, (4, 9, 512)).astype(np.float32)
    I[:, :, 0] = rng.uniform(1e-4, 1e-3, (4, 9)).astype(np.float32)
    ref = oracle.exhaustive_voxel(info, I, lasers, sensors)
    clamp = ref > 1e30
    assert clamp[z0, vy0, vx0]
    out = ctx.backproject_exhaustive(info, cuda(I), cuda(lasers), cuda(sensors)).cpu().numpy()
    assert np.all(np.isfinite(out))
    np.testing.assert_allclose(out[clamp], ref[clamp], rtol=1e-6)
    rest = ~clamp
    err = np.linalg.norm(out[rest] - ref[rest]) / np.linalg.norm(ref[rest])
    assert err <= 1e-5, err
