"""GPU parity of the backprojection kernels against the CPU oracle.

Bins are bit-exact requirements: the transients used here have values that
vary over orders of magnitude from bin to bin, so one wrong bin in one
(voxel, scan point) pair moves that voxel far outside the FP32 tolerance.
Tolerances: rel-L2 <= 1e-4 against the FP64 oracle for FP32 storage,
per-voxel relative error <= 5e-4 on voxels above 1% of the maximum."""
import numpy as np
import pytest
import torch

import oracle
from fasterbackprojection_amd.api import (LAYOUT_TIME_MAJOR, MIS_TILED, STORAGE_BF16, STORAGE_F16,
                                          Geometry, make_info)
from fasterbackprojection_amd.scene import small_confocal, wall_grid

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def assert_volume_close(gpu, ref, tol=1e-4, vox_tol=5e-4):
    gpu = gpu.detach().cpu().numpy().astype(np.float64)
    assert gpu.shape == ref.shape
    e = rel_l2(gpu, ref)
    assert e <= tol, f"rel-L2 {e:.3e}"
    big = np.abs(ref) > 0.01 * np.abs(ref).max()
    worst = np.max(np.abs(gpu[big] - ref[big]) / np.abs(ref[big]))
    assert worst <= vox_tol, f"worst voxel rel err {worst:.3e}"


def spiky(rng, shape):
    """Transient whose adjacent bins differ by orders of magnitude."""
    v = (10.0 ** rng.uniform(-3, 3, size=shape)).astype(np.float32)
    v[rng.random(shape) < 0.2] = 0.0
    return v


def geom(n_scan=16, T=256, res=(24, 20, 28), delta=0.0096, discard=1, t0=0.0,
         hmin=(-0.4, 0.35, -0.45), hmax=(0.45, 1.25, 0.4), wall=1.0):
    scan = wall_grid(n_scan, wall)
    return Geometry(laser_targets=scan, sensor_targets=scan, num_time_bins=T, time_step=delta,
                    hidden_min=hmin, hidden_max=hmax, voxel_resolution=res,
                    discard_first_last_bounces=discard, time_offset=t0,
                    laser_position=(0.1, -0.8, 0.2), sensor_position=(-0.2, -0.7, 0.1),
                    wall_size=(wall, wall))


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("discard,t0", [(1, 0.0), (0, 0.0), (0, 0.45), (1, 0.3)])
def test_confocal_spiky_parity(ctx, discard, t0):
    """General geometry (non-multiple-of-16 volume, non-aligned voxels),
    every bin path: pure (t0 = 0, discard) and general."""
    g = geom(discard=discard, t0=t0)
    info = make_info(g)
    rng = np.random.default_rng(11 + discard)
    I = spiky(rng, (len(g.laser_targets), g.num_time_bins))
    ref = oracle.confocal_voxel(info, I, g.laser_targets)
    out = ctx.backproject_confocal(info, cuda(I), cuda(g.laser_targets))
    torch.cuda.synchronize()
    assert_volume_close(out, ref)


def test_confocal_small_config_parity(ctx):
    """32x32 scan, 256 bins, 200 points, Poisson at peak 1000, 32^3 volume."""
    cap = small_confocal(seed=1)
    info = make_info(cap.geometry)
    ref = oracle.confocal_voxel(info, cap.transient, cap.geometry.laser_targets)
    out = ctx.backproject_confocal(info, cuda(cap.transient), cuda(cap.geometry.laser_targets))
    torch.cuda.synchronize()
    assert_volume_close(out, ref)


def test_confocal_time_major_layout(ctx):
    g = geom(res=(16, 16, 16))
    info = make_info(g)
    rng = np.random.default_rng(5)
    I = spiky(rng, (len(g.laser_targets), g.num_time_bins))
    ref = oracle.confocal_voxel(info, I, g.laser_targets)
    out = ctx.backproject_confocal(info, cuda(np.ascontiguousarray(I.T)), cuda(g.laser_targets),
                                   layout=LAYOUT_TIME_MAJOR)
    torch.cuda.synchronize()
    assert_volume_close(out, ref)


def test_confocal_near_bin_edges(ctx):
    """Scan pitch and voxel pitch chosen so many distances land within a few
    ulps of bin edges (dt = voxel pitch): the exact fallback must fire."""
    scan = wall_grid(16, 1.0)
    g = Geometry(laser_targets=scan, sensor_targets=scan, num_time_bins=512, time_step=0.0625,
                 hidden_min=(-0.5, 0.5, -0.5), hidden_max=(0.5, 1.5, 0.5), voxel_resolution=(16, 16, 16),
                 wall_size=(1.0, 1.0))
    info = make_info(g)
    rng = np.random.default_rng(9)
    I = spiky(rng, (256, 512))
    ref = oracle.confocal_voxel(info, I, scan)
    out = ctx.backproject_confocal(info, cuda(I), cuda(scan))
    torch.cuda.synchronize()
    assert_volume_close(out, ref)


def test_confocal_deterministic(ctx):
    cap = small_confocal(seed=2)
    info = make_info(cap.geometry)
    a = ctx.backproject_confocal(info, cuda(cap.transient), cuda(cap.geometry.laser_targets))
    b = ctx.backproject_confocal(info, cuda(cap.transient), cuda(cap.geometry.laser_targets))
    torch.cuda.synchronize()
    assert torch.equal(a, b)


# ------------------------------------------------------------ aligned grids
def aligned_geom(nx, wall, X, VY, ZD, T, depth0, dpitch, t0=0.0, discard=1, R=1):
    """Scan grid of pitch wall / nx; voxel lattice of pitch R * wall / nx
    along world x and z, both starting at -wall / 2, so the aligned-grid
    path applies (R = 2: every voxel centre lies between two scan points,
    and odd scan columns mirror even ones).  Voxel sizes follow the
    reference's size / resolution per component: world x <- X, world y
    (depth) <- size.y / VY, world z <- size.z / ZD."""
    pitch = R * wall / nx
    scan = wall_grid(nx, wall)
    hmin = (-wall / 2, depth0, -wall / 2)
    hmax = (-wall / 2 + X * pitch, depth0 + VY * dpitch, -wall / 2 + ZD * pitch)
    return Geometry(laser_targets=scan, sensor_targets=scan, num_time_bins=T, time_step=0.0096,
                    hidden_min=hmin, hidden_max=hmax, voxel_resolution=(X, VY, ZD),
                    discard_first_last_bounces=discard, time_offset=t0,
                    laser_position=(0.1, -0.8, 0.2), sensor_position=(-0.2, -0.7, 0.1),
                    wall_size=(wall, wall))


ALIGNED = {
    "64": dict(nx=64, wall=1.0, X=64, VY=16, ZD=8, T=256, depth0=0.25, dpitch=0.02),
    "64_ragged_depth": dict(nx=64, wall=1.0, X=64, VY=32, ZD=5, T=256, depth0=0.05, dpitch=0.02),
    "128": dict(nx=128, wall=2.0, X=128, VY=16, ZD=4, T=512, depth0=0.5, dpitch=0.03),
    # Two scan points per voxel pitch: four phase planes per bin.
    "r2_64": dict(nx=128, wall=1.0, R=2, X=64, VY=16, ZD=6, T=256, depth0=0.25, dpitch=0.02),
    "r2_64_full_rows": dict(nx=128, wall=1.0, R=2, X=64, VY=64, ZD=2, T=256, depth0=0.05,
                            dpitch=0.01),
    "r2_128": dict(nx=256, wall=2.0, R=2, X=128, VY=16, ZD=3, T=512, depth0=0.5, dpitch=0.03),
}


@pytest.mark.parametrize("name", sorted(ALIGNED))
def test_confocal_aligned_path_parity(ctx, name):
    g = aligned_geom(**ALIGNED[name])
    info = make_info(g)
    rng = np.random.default_rng(31)
    I = spiky(rng, (len(g.laser_targets), g.num_time_bins))
    ref = oracle.confocal_voxel(info, I, g.laser_targets)
    out = ctx.backproject_confocal(info, cuda(I), cuda(g.laser_targets))
    assert ctx.last_confocal_path_aligned()
    assert_volume_close(out, ref)
    again = ctx.backproject_confocal(info, cuda(I), cuda(g.laser_targets))
    torch.cuda.synchronize()
    assert torch.equal(out, again)                   # deterministic
    ctx.set_aligned_path(False)
    try:
        gen = ctx.backproject_confocal(info, cuda(I), cuda(g.laser_targets))
        assert not ctx.last_confocal_path_aligned()
    finally:
        ctx.set_aligned_path(True)
    assert_volume_close(gen, ref)


def test_confocal_aligned_table_cache(ctx):
    """The offset tables are kept between calls: switching between
    geometries (different depth range; same parameters but the wall moved
    in depth; a perturbed grid that takes the general kernel) and back must
    rebuild or reuse them correctly."""
    ga = aligned_geom(**ALIGNED["64"])
    gb = aligned_geom(**dict(ALIGNED["64"], depth0=0.4))
    gc = aligned_geom(**ALIGNED["64"])
    gc.laser_targets = gc.laser_targets.copy()
    gc.laser_targets[:, 1] = np.float32(-0.05)           # wall plane moved
    gd = aligned_geom(**ALIGNED["64"])
    gd.laser_targets = gd.laser_targets.copy()
    gd.laser_targets[5, 2] = np.nextafter(gd.laser_targets[5, 2], np.float32(2.0))
    rng = np.random.default_rng(35)
    I = spiky(rng, (64 * 64, 256))
    for g, aligned in [(ga, True), (ga, True), (gb, True), (ga, True), (gc, True), (gd, False),
                       (ga, True)]:
        info = make_info(g)
        out = ctx.backproject_confocal(info, cuda(I), cuda(g.laser_targets))
        assert ctx.last_confocal_path_aligned() == aligned
        assert_volume_close(out, oracle.confocal_voxel(info, I, g.laser_targets))


def test_confocal_aligned_time_major(ctx):
    g = aligned_geom(**ALIGNED["64"])
    info = make_info(g)
    rng = np.random.default_rng(32)
    I = spiky(rng, (len(g.laser_targets), g.num_time_bins))
    ref = oracle.confocal_voxel(info, I, g.laser_targets)
    out = ctx.backproject_confocal(info, cuda(I.T), cuda(g.laser_targets),
                                   layout=LAYOUT_TIME_MAJOR)
    assert ctx.last_confocal_path_aligned()
    assert_volume_close(out, ref)


def test_confocal_aligned_check_rejects_perturbed_grid(ctx):
    """One scan point moved by one ulp breaks translation invariance: the
    device-side check must send the call to the general kernel."""
    g = aligned_geom(**ALIGNED["64"])
    scan = g.laser_targets.copy()
    scan[777, 0] = np.nextafter(scan[777, 0], np.float32(1.0))
    g.laser_targets = scan
    info = make_info(g)
    rng = np.random.default_rng(33)
    I = spiky(rng, (len(scan), g.num_time_bins))
    ref = oracle.confocal_voxel(info, I, scan)
    out = ctx.backproject_confocal(info, cuda(I), cuda(scan))
    assert not ctx.last_confocal_path_aligned()
    assert_volume_close(out, ref)


def test_confocal_aligned_geometry_general_chain(ctx):
    """t0 != 0 makes bins depend on more than the squared distance: the
    aligned path does not apply (host-side decision)."""
    g = aligned_geom(**ALIGNED["64"], t0=0.05)
    info = make_info(g)
    rng = np.random.default_rng(34)
    I = spiky(rng, (len(g.laser_targets), g.num_time_bins))
    ref = oracle.confocal_voxel(info, I, g.laser_targets)
    out = ctx.backproject_confocal(info, cuda(I), cuda(g.laser_targets))
    assert not ctx.last_confocal_path_aligned()
    assert_volume_close(out, ref)


def dequantize(q, sc, storage):
    """The FP32 samples the kernels read from F16/BF16 storage: the stored
    value widened to FP32 times the row's FP32 scale (one rounding)."""
    q = q.cpu().numpy()
    if storage == STORAGE_F16:
        v = q.view(np.float16).astype(np.float32)
    else:
        v = (q.view(np.uint16).astype(np.uint32) << 16).view(np.float32)
    return (v * sc.cpu().numpy()[:, None]).astype(np.float32)


@pytest.mark.parametrize("name", ["64", "r2_64"])
@pytest.mark.parametrize("storage", [STORAGE_F16, STORAGE_BF16])
def test_confocal_aligned_quantized(ctx, name, storage):
    """F16/BF16 storage on the aligned path (dequantised into phase planes,
    Kahan-compensated across planes): within FP32 tolerance of the FP64
    oracle run on the dequantised samples, and of the general kernel on
    the same storage."""
    g = aligned_geom(**ALIGNED[name])
    info = make_info(g)
    rng = np.random.default_rng(36)
    I = spiky(rng, (len(g.laser_targets), g.num_time_bins))
    I[I <= 1e-5] = 0.0            # quantised storage holds thresholded samples
    q, sc = ctx.quantize(cuda(I), storage)
    ref = oracle.confocal_voxel(info, dequantize(q, sc, storage), g.laser_targets)
    out = ctx.backproject_confocal(info, q, cuda(g.laser_targets), storage=storage, scales=sc)
    assert ctx.last_confocal_path_aligned()
    assert_volume_close(out, ref)
    ctx.set_aligned_path(False)
    try:
        gen = ctx.backproject_confocal(info, q, cuda(g.laser_targets), storage=storage, scales=sc)
        assert not ctx.last_confocal_path_aligned()
    finally:
        ctx.set_aligned_path(True)
    assert_volume_close(gen, ref)


def test_confocal_aligned_r2_time_major(ctx):
    g = aligned_geom(**ALIGNED["r2_64"])
    info = make_info(g)
    rng = np.random.default_rng(37)
    I = spiky(rng, (len(g.laser_targets), g.num_time_bins))
    ref = oracle.confocal_voxel(info, I, g.laser_targets)
    out = ctx.backproject_confocal(info, cuda(I.T), cuda(g.laser_targets),
                                   layout=LAYOUT_TIME_MAJOR)
    assert ctx.last_confocal_path_aligned()
    assert_volume_close(out, ref)


def test_confocal_aligned_r2_rejects_unmirrored_grid(ctx):
    """Shifting the voxel lattice by a quarter of its pitch puts voxel
    centres on even scan columns: odd columns no longer mirror even ones,
    and the device-side check must send the call to the general kernel."""
    g = aligned_geom(**ALIGNED["r2_64"])
    vs = (g.hidden_max[0] - g.hidden_min[0]) / 64
    g.hidden_min = (g.hidden_min[0] - vs / 4, g.hidden_min[1], g.hidden_min[2])
    g.hidden_max = (g.hidden_max[0] - vs / 4, g.hidden_max[1], g.hidden_max[2])
    info = make_info(g)
    rng = np.random.default_rng(38)
    I = spiky(rng, (len(g.laser_targets), g.num_time_bins))
    ref = oracle.confocal_voxel(info, I, g.laser_targets)
    out = ctx.backproject_confocal(info, cuda(I), cuda(g.laser_targets))
    assert not ctx.last_confocal_path_aligned()
    assert_volume_close(out, ref)


@pytest.mark.parametrize("storage,tol", [(STORAGE_F16, 2e-3), (STORAGE_BF16, 8e-3)])
def test_confocal_quantized_storage(ctx, storage, tol):
    """Precision form: FP16/BF16 with per-scan-point max scale."""
    cap = small_confocal(seed=3)
    info = make_info(cap.geometry)
    ref = oracle.confocal_voxel(info, cap.transient, cap.geometry.laser_targets)
    I = cuda(cap.transient)
    q, sc = ctx.quantize(I, storage)
    out = ctx.backproject_confocal(info, q, cuda(cap.geometry.laser_targets), storage=storage,
                                   scales=sc)
    torch.cuda.synchronize()
    e = rel_l2(out.cpu().numpy(), ref)
    assert e <= tol, e


def test_exhaustive_spiky_parity(ctx):
    lasers = wall_grid(4, 1.0)
    sensors = wall_grid(6, 1.0)
    g = Geometry(laser_targets=lasers, sensor_targets=sensors, num_time_bins=200,
                 time_step=0.0096, hidden_min=(-0.4, 0.35, -0.45), hidden_max=(0.45, 1.25, 0.4),
                 voxel_resolution=(20, 18, 12), confocal=False, wall_size=(1.0, 1.0))
    info = make_info(g)
    rng = np.random.default_rng(4)
    I = spiky(rng, (16, 36, 200))
    ref = oracle.exhaustive_voxel(info, I, lasers, sensors)
    out = ctx.backproject_exhaustive(info, cuda(I), cuda(lasers), cuda(sensors))
    torch.cuda.synchronize()
    assert_volume_close(out, ref)


@pytest.mark.parametrize("discard,t0", [(1, 0.0), (0, 0.1)])
def test_exhaustive_linear_weight_kernel(ctx, discard, t0):
    """Hidden box >= 0.5 m from the wall at 4.8 mm bins: every pair's bins
    are >= 160, so the linear-weight kernel runs (window entries I/m^2 and
    -2I/m^3, boundary voxels replaced by the exact chain at the end of each
    batch).  Spiky data: a wrong bin or weight moves a voxel far out of the
    tolerance."""
    lasers = wall_grid(4, 1.0)
    sensors = wall_grid(6, 1.0)
    g = Geometry(laser_targets=lasers, sensor_targets=sensors, num_time_bins=1024,
                 time_step=0.0048, hidden_min=(-0.4, 0.5, -0.45), hidden_max=(0.45, 1.4, 0.4),
                 voxel_resolution=(20, 18, 40), confocal=False,
                 discard_first_last_bounces=discard, time_offset=t0,
                 laser_position=(0.1, -0.8, 0.2), sensor_position=(-0.2, -0.7, 0.1),
                 wall_size=(1.0, 1.0))
    info = make_info(g)
    rng = np.random.default_rng(14)
    I = spiky(rng, (16, 36, 1024))
    ref = oracle.exhaustive_voxel(info, I, lasers, sensors)
    out = ctx.backproject_exhaustive(info, cuda(I), cuda(lasers), cuda(sensors))
    torch.cuda.synchronize()
    assert_volume_close(out, ref)


@pytest.mark.parametrize("discard,t0", [(0, 0.2)])
def test_exhaustive_general_chain(ctx, discard, t0):
    lasers = wall_grid(3, 1.0)
    sensors = wall_grid(5, 1.0)
    g = Geometry(laser_targets=lasers, sensor_targets=sensors, num_time_bins=300,
                 time_step=0.0096, hidden_min=(-0.4, 0.35, -0.45), hidden_max=(0.45, 1.25, 0.4),
                 voxel_resolution=(16, 16, 8), confocal=False, discard_first_last_bounces=discard,
                 time_offset=t0, laser_position=(0.1, -0.8, 0.2), sensor_position=(-0.2, -0.7, 0.1),
                 wall_size=(1.0, 1.0))
    info = make_info(g)
    rng = np.random.default_rng(8)
    I = spiky(rng, (9, 25, 300))
    ref = oracle.exhaustive_voxel(info, I, lasers, sensors)
    out = ctx.backproject_exhaustive(info, cuda(I), cuda(lasers), cuda(sensors))
    torch.cuda.synchronize()
    assert_volume_close(out, ref)


def test_depth_planes(ctx):
    g = geom(n_scan=8, res=(4, 4, 4))
    info = make_info(g)
    rng = np.random.default_rng(2)
    I = spiky(rng, (64, 256))
    depths = np.linspace(0.6, 0.8, 20, endpoint=False)
    ref = oracle.depth(info, I, g.laser_targets, depths)
    out = ctx.backproject_depth(info, cuda(I), cuda(g.laser_targets), cuda(depths))
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("L,S,D", [(16, 64, 5), (10, 73, 3), (4, 9, 40)])
def test_depth_planes_exhaustive(ctx, L, S, D):
    """Exhaustive depth planes (backprojectExhaustive): pair chunks split
    across threads (ragged last chunk, chunks crossing sensor rows) and
    reduced in FP64, against the oracle."""
    g = geom(n_scan=8, T=256)
    rng = np.random.default_rng(L * 100 + S)
    lasers = wall_grid(8, 1.0)[rng.choice(64, L, replace=False)].astype(np.float32)
    sensors = (wall_grid(16, 1.0)[rng.choice(256, S, replace=False)]).astype(np.float32)
    g = Geometry(laser_targets=lasers, sensor_targets=sensors, num_time_bins=256,
                 time_step=0.0096, hidden_min=g.hidden_min, hidden_max=g.hidden_max,
                 voxel_resolution=g.voxel_resolution, confocal=False, discard_first_last_bounces=0,
                 laser_position=(0.1, -0.8, 0.2), sensor_position=(-0.2, -0.7, 0.1),
                 wall_size=(1.0, 1.0))
    info = make_info(g)
    I = spiky(rng, (L, S, 256))
    depths = np.linspace(0.5, 0.9, D, endpoint=False)
    ref = oracle.depth(info, I, lasers, depths, sensor=sensors, exhaustive=True)
    out = ctx.backproject_depth(info, cuda(I), cuda(lasers), cuda(depths), sensor_targets=cuda(sensors))
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), ref) <= 1e-5


# Exhaustive depth-plane cases by kernel path: the staged-window fast path
# with and without the time/path offsets, pairs whose points coincide with a
# laser target (depth 0: the IEEE chain for every point), many workgroups
# (64 pixels x 50 depths), and windows too wide for the staging buffers
# (0.5 mm bins: the per-pair gather).
DEPTH_EXHAUSTIVE = {
    "pure": dict(L=16, S=64, D=5, discard=1, t0=0.0, dt=0.0096, T=256, lo=0.5, hi=0.9),
    "t0": dict(L=16, S=64, D=5, discard=1, t0=0.05, dt=0.0096, T=256, lo=0.5, hi=0.9),
    "offsets_t0": dict(L=12, S=50, D=7, discard=0, t0=0.3, dt=0.0096, T=512, lo=0.4, hi=0.8),
    "near_wall": dict(L=16, S=64, D=6, discard=1, t0=0.0, dt=0.0096, T=256, lo=0.0, hi=0.3),
    "many_blocks": dict(L=64, S=64, D=50, discard=1, t0=0.0, dt=0.0096, T=256, lo=0.5, hi=0.9),
    "wide_window": dict(L=16, S=64, D=5, discard=1, t0=0.0, dt=0.0005, T=6000, lo=0.5, hi=0.9),
    # 12 pairs: a single staged batch; 24 pairs: two (the second ragged).
    "one_batch": dict(L=3, S=4, D=7, discard=1, t0=0.0, dt=0.0096, T=256, lo=0.5, hi=0.9),
    "two_batches": dict(L=4, S=6, D=9, discard=0, t0=0.1, dt=0.0096, T=512, lo=0.5, hi=0.9),
}


@pytest.mark.parametrize("name", sorted(DEPTH_EXHAUSTIVE))
def test_depth_planes_exhaustive_paths(ctx, name):
    """Every output against the oracle: a wrong bin for one (pixel, depth,
    pair) moves that output by far more than the tolerance (spiky data)."""
    p = DEPTH_EXHAUSTIVE[name]
    L, S, D, T = p["L"], p["S"], p["D"], p["T"]
    rng = np.random.default_rng(L * 1000 + S + D)
    lasers = wall_grid(8, 1.0)[rng.choice(64, L, replace=False)].astype(np.float32)
    sensors = wall_grid(16, 1.0)[rng.choice(256, S, replace=False)].astype(np.float32)
    g = Geometry(laser_targets=lasers, sensor_targets=sensors, num_time_bins=T,
                 time_step=p["dt"], hidden_min=(-0.5, 0.2, -0.5), hidden_max=(0.5, 1.0, 0.5),
                 voxel_resolution=(8, 8, 8), confocal=False,
                 discard_first_last_bounces=p["discard"], time_offset=p["t0"],
                 laser_position=(0.1, -0.8, 0.2), sensor_position=(-0.2, -0.7, 0.1),
                 wall_size=(1.0, 1.0))
    info = make_info(g)
    I = spiky(rng, (L, S, T))
    depths = np.linspace(p["lo"], p["hi"], D, endpoint=False)
    ref = oracle.depth(info, I, lasers, depths, sensor=sensors, exhaustive=True)
    out = ctx.backproject_depth(info, cuda(I), cuda(lasers), cuda(depths),
                                sensor_targets=cuda(sensors)).cpu().numpy().astype(np.float64)
    assert out.shape == ref.shape
    assert np.abs(ref).max() > 0, "every bin out of range: the case checks nothing"
    # FP32 partial sums: ~1e-6 relative per output.
    err = np.abs(out - ref) / np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())
    assert err.max() <= 2e-5, (name, np.unravel_index(np.argmax(err), err.shape))


DEPTH_ALIGNED = {
    "64": dict(nx=64, wall=1.0, T=256, lo=0.3, hi=0.5, D=20),
    "128": dict(nx=128, wall=2.0, T=512, lo=0.6, hi=0.9, D=4),
    "64_one_depth": dict(nx=64, wall=1.0, T=256, lo=0.45, hi=0.46, D=1),
}


def depth_geom(nx, wall, T, **_):
    scan = wall_grid(nx, wall)
    return Geometry(laser_targets=scan, sensor_targets=scan, num_time_bins=T, time_step=0.0096,
                    hidden_min=(-0.5, 0.2, -0.5), hidden_max=(0.5, 1.0, 0.5),
                    voxel_resolution=(16, 16, 16), discard_first_last_bounces=1,
                    laser_position=(0.1, -0.8, 0.2), sensor_position=(-0.2, -0.7, 0.1),
                    wall_size=(wall, wall))


@pytest.mark.parametrize("name", sorted(DEPTH_ALIGNED))
def test_depth_planes_aligned_path(ctx, name):
    """Depth-plane mode on a 64^2 / 128^2 scan grid with the wall normal
    along y: the ring-convolution path (weight 1, pixels = scan points)
    against the oracle's gather, and the general kernel on the same input."""
    cfg = DEPTH_ALIGNED[name]
    g = depth_geom(**cfg)
    info = make_info(g)
    rng = np.random.default_rng(41)
    I = spiky(rng, (len(g.laser_targets), g.num_time_bins))
    depths = np.linspace(cfg["lo"], cfg["hi"], cfg["D"], endpoint=False)
    ref = oracle.depth(info, I, g.laser_targets, depths)
    out = ctx.backproject_depth(info, cuda(I), cuda(g.laser_targets), cuda(depths))
    assert ctx.last_confocal_path_aligned()
    assert_volume_close(out, ref)
    ctx.set_aligned_path(False)
    try:
        gen = ctx.backproject_depth(info, cuda(I), cuda(g.laser_targets), cuda(depths))
        assert not ctx.last_confocal_path_aligned()
    finally:
        ctx.set_aligned_path(True)
    assert_volume_close(gen, ref)


def test_depth_planes_aligned_check_and_host_conditions(ctx):
    """A scan point moved by one ulp (device check) or a wall normal with an
    x component (host check) sends the depth planes to the gather kernel;
    a voxel backprojection between depth calls rebuilds its own tables."""
    cfg = DEPTH_ALIGNED["64"]
    depths = np.linspace(cfg["lo"], cfg["hi"], 8, endpoint=False)
    rng = np.random.default_rng(42)
    g = depth_geom(**cfg)
    I = spiky(rng, (len(g.laser_targets), g.num_time_bins))
    moved = depth_geom(**cfg)
    moved.laser_targets = moved.laser_targets.copy()
    moved.laser_targets[300, 0] = np.nextafter(moved.laser_targets[300, 0], np.float32(1.0))
    tilted = depth_geom(**cfg)
    tilted.wall_normal = (0.0, 0.8, 0.6)
    vg = aligned_geom(**ALIGNED["64"])
    for gg, aligned in [(g, True), (moved, False), (tilted, False), (g, True)]:
        info = make_info(gg)
        out = ctx.backproject_depth(info, cuda(I), cuda(gg.laser_targets), cuda(depths))
        assert ctx.last_confocal_path_aligned() == aligned
        assert_volume_close(out, oracle.depth(info, I, gg.laser_targets, depths))
        vinfo = make_info(vg)
        vol = ctx.backproject_confocal(vinfo, cuda(I), cuda(vg.laser_targets))
        assert ctx.last_confocal_path_aligned()
        assert_volume_close(vol, oracle.confocal_voxel(vinfo, I, vg.laser_targets))


def test_depth_planes_aligned_table_cache(ctx):
    """Depth-plane tables are reused for the same depths and rebuilt when a
    depth value changes (by one ulp), when the number of depths changes, and
    after a voxel backprojection used the tables."""
    cfg = DEPTH_ALIGNED["64"]
    g = depth_geom(**cfg)
    info = make_info(g)
    rng = np.random.default_rng(43)
    I = spiky(rng, (len(g.laser_targets), g.num_time_bins))
    d1 = np.linspace(cfg["lo"], cfg["hi"], 6, endpoint=False)
    d2 = d1.copy()
    d2[3] = float(np.nextafter(np.float32(d1[3]), np.float32(1.0)))  # next float depth
    d3 = d1[:5].copy()
    vg = aligned_geom(**ALIGNED["64"])
    vinfo = make_info(vg)
    for depths in (d1, d1, d2, d1, d3, "voxel", d3):
        if isinstance(depths, str):
            vol = ctx.backproject_confocal(vinfo, cuda(I), cuda(vg.laser_targets))
            assert_volume_close(vol, oracle.confocal_voxel(vinfo, I, vg.laser_targets))
            continue
        out = ctx.backproject_depth(info, cuda(I), cuda(g.laser_targets), cuda(depths))
        assert ctx.last_confocal_path_aligned()
        assert_volume_close(out, oracle.depth(info, I, g.laser_targets, depths))


def test_mis_reference_mode_matches_oracle(ctx):
    cap = small_confocal(seed=4, res=16)
    info = make_info(cap.geometry)
    I = cap.transient
    L = len(cap.geometry.laser_targets)
    cdf = oracle.mis_cdf(I)
    alias, prob = oracle.alias_tables(cdf)
    rng = np.random.default_rng(0)
    noise = rng.random(16 ** 3).astype(np.float32)
    ref = oracle.mis(info, I, cap.geometry.laser_targets, cdf, alias, prob, noise, L // 4, 0)
    out = ctx.backproject_confocal_mis(info, cuda(I), cuda(cap.geometry.laser_targets),
                                       cuda(noise), L // 4, mode=0)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), ref) < 1e-4


@pytest.mark.parametrize("mode,discard,t0", [(0, 0, 0.3), (1, 1, 0.0), (1, 0, 0.3)])
def test_mis_per_voxel_modes_match_oracle(ctx, mode, discard, t0):
    """REFERENCE and UNBIASED per-voxel sampling on a volume whose voxel count
    is not a multiple of 64 (the last wave is partial), with the general
    chain's offsets: sample s of voxel v draws from noise state v + s."""
    g = geom(n_scan=12, T=256, res=(13, 11, 9), discard=discard, t0=t0)
    info = make_info(g)
    rng = np.random.default_rng(40 + mode)
    I = rng.uniform(0.0, 1.0, (144, 256)).astype(np.float32)
    I[rng.random(I.shape) < 0.5] = 0.0
    cdf = oracle.mis_cdf(I)
    alias, prob = oracle.alias_tables(cdf)
    noise = rng.random(5000).astype(np.float32)
    ref = oracle.mis(info, I, g.laser_targets, cdf, alias, prob, noise, 100, mode)
    out = ctx.backproject_confocal_mis(info, cuda(I), cuda(g.laser_targets), cuda(noise), 100,
                                       mode=mode)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("mode", [0, 1])
def test_mis_per_voxel_at_scan_points(ctx, mode):
    """Voxel centres exactly on scan points (depth slice 0 on the wall): those
    pairs have dist = 0 and safeRCP's FLT_MAX weight, which the pure-chain
    walk takes from the reference's IEEE chain."""
    g = geom(n_scan=8, T=256, res=(8, 8, 8), hmin=(-0.5, -0.025, -0.5),
             hmax=(0.5, 0.375, 0.5))
    info = make_info(g)
    rng = np.random.default_rng(3 + mode)
    I = rng.uniform(0.1, 1.0, (64, 256)).astype(np.float32)
    I[:, 0] = rng.uniform(1e-4, 1e-3, 64).astype(np.float32)
    cdf = oracle.mis_cdf(I)
    alias, prob = oracle.alias_tables(cdf)
    noise = rng.random(3000).astype(np.float32)
    ref = oracle.mis(info, I, g.laser_targets, cdf, alias, prob, noise, 48, mode)
    clamp = ref > 1e30
    assert clamp.any()
    out = ctx.backproject_confocal_mis(info, cuda(I), cuda(g.laser_targets), cuda(noise), 48,
                                       mode=mode).cpu().numpy()
    assert np.all(np.isfinite(out))
    np.testing.assert_allclose(out[clamp], ref[clamp], rtol=1e-5)
    assert rel_l2(out[~clamp], ref[~clamp]) < 1e-5


@pytest.mark.parametrize("discard,t0", [(1, 0.0), (0, 0.3)])
def test_mis_tiled_matches_oracle(ctx, discard, t0):
    """Tile-shared MIS: same samples and weights as the restatement, exact
    bins; values either 0 or >= 1e-2 so that weighting cannot cross the
    1e-5 threshold.  (1, 0.0): the pure chain's single-pass sub-tile kernel;
    (0, 0.3): the tile kernel over chunks of the lists."""
    g = geom(n_scan=16, T=256, res=(32, 32, 20), discard=discard, t0=t0)
    info = make_info(g)
    rng = np.random.default_rng(41)
    I = (10.0 ** rng.uniform(-2, 3, size=(256, 256))).astype(np.float32)
    I[rng.random(I.shape) < 0.3] = 0.0
    noise = rng.random(4096).astype(np.float32)
    ref = oracle.mis_tiled(info, I, g.laser_targets, noise, 48)
    out = ctx.backproject_confocal_mis(info, cuda(I), cuda(g.laser_targets), cuda(noise), 48,
                                       mode=MIS_TILED)
    torch.cuda.synchronize()
    assert_volume_close(out, ref)


def test_mis_tiled_near_wall(ctx):
    """Tile-shared MIS with voxel centres 2 mm in front of scan points (the
    first depth slice, x and z on the scan grid; dist 4 mm keeps safeRCP of
    dist^2 finite): those samples take the IEEE fallback of the single-pass
    kernel."""
    g = geom(n_scan=16, T=128, res=(16, 8, 16), hmin=(-0.5, -0.008, -0.5),
             hmax=(0.5, 0.152, 0.5))
    info = make_info(g)
    rng = np.random.default_rng(47)
    I = (10.0 ** rng.uniform(-2, 3, size=(256, 128))).astype(np.float32)
    noise = rng.random(2048).astype(np.float32)
    ref = oracle.mis_tiled(info, I, g.laser_targets, noise, 64)
    out = ctx.backproject_confocal_mis(info, cuda(I), cuda(g.laser_targets), cuda(noise), 64,
                                       mode=MIS_TILED)
    torch.cuda.synchronize()
    assert_volume_close(out, ref)


def _seeded_noise(seed, n):
    """The seeded virtual noise array of fbp_mis_params (its header formula)."""
    M = 0xFFFFFFFF

    def mix(x):
        x = np.asarray(x, np.uint64) & M
        x ^= x >> 16
        x = (x * 0x7FEB352D) & M
        x ^= x >> 15
        x = (x * 0x846CA68B) & M
        x ^= x >> 16
        return x

    key = mix(np.uint64(seed & M) ^ mix((np.uint64(seed >> 32) + 0x9E3779B9) & M))
    h = mix(np.arange(n, dtype=np.uint64) ^ key)
    return ((h >> 8).astype(np.float64) / 2.0 ** 24).astype(np.float32)


@pytest.mark.parametrize("seed", [0, 12345, 2 ** 40 + 77])
def test_mis_tiled_seeded_matches_oracle(ctx, seed):
    """noise=None: the same samples as the restatement fed the seeded noise
    array (the hash restated above); exact bins."""
    g = geom(n_scan=16, T=256, res=(32, 32, 20))
    info = make_info(g)
    rng = np.random.default_rng(43)
    I = (10.0 ** rng.uniform(-2, 3, size=(256, 256))).astype(np.float32)
    I[rng.random(I.shape) < 0.3] = 0.0
    n = 32 * 32 * 20
    ref = oracle.mis_tiled(info, I, g.laser_targets, _seeded_noise(seed, n), 48)
    out = ctx.backproject_confocal_mis(info, cuda(I), cuda(g.laser_targets), None, 48,
                                       mode=MIS_TILED, seed=seed)
    torch.cuda.synchronize()
    assert_volume_close(out, ref)


def test_mis_tiled_is_unbiased(ctx):
    """The mean over independent noise tables converges to the plain
    backprojection: its error must be within 4 standard errors estimated
    from the draws themselves, and shrink with the number of draws."""
    cap = small_confocal(seed=5, n_scan=16, T=256, res=16)
    info = make_info(cap.geometry)
    I = cuda(cap.transient)
    laser = cuda(cap.geometry.laser_targets)
    plain = ctx.backproject_confocal(info, I, laser).double()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(3)
    draws = []
    for _ in range(256):
        noise = torch.rand(4096, generator=gen, device="cuda")
        draws.append(ctx.backproject_confocal_mis(info, I, laser, noise, 32, mode=MIS_TILED).double())
    D = torch.stack(draws)
    mean = D.mean(0)
    se = D.std(0) / np.sqrt(len(draws))
    err = float((mean - plain).norm() / plain.norm())
    bound = float(4 * se.norm() / plain.norm())
    assert err < bound, (err, bound)
    err16 = float((D[:16].mean(0) - plain).norm() / plain.norm())
    assert err < err16


def test_rejects_bad_arguments(ctx):
    from fasterbackprojection_amd.api import FbpError
    g = geom()
    info = make_info(g)
    info.time_step = 0.0
    I = torch.zeros((len(g.laser_targets), g.num_time_bins), device="cuda")
    with pytest.raises(FbpError):
        ctx.backproject_confocal(info, I, cuda(g.laser_targets))


# ------------------------------------------------ voxel-scattering variant
@pytest.fixture
def scatter_ctx(ctx):
    ctx.set_voxel_scattering(True)
    assert ctx.option(1) == 1
    try:
        yield ctx
    finally:
        ctx.set_voxel_scattering(False)


@pytest.mark.parametrize("discard,t0", [(1, 0.0), (0, 0.3)])
def test_confocal_scatter_matches_oracle(scatter_ctx, discard, t0):
    """Every bin of each voxel's walk (walks that leave [0, T) included:
    distances here exceed T * Delta), spiky values so that one wrong bin
    fails the tolerance."""
    g = geom(discard=discard, t0=t0)
    info = make_info(g)
    rng = np.random.default_rng(51)
    I = spiky(rng, (len(g.laser_targets), g.num_time_bins))
    ref = oracle.confocal_voxel_scatter(info, I, g.laser_targets)
    out = scatter_ctx.backproject_confocal(info, cuda(I), cuda(g.laser_targets))
    assert not scatter_ctx.last_confocal_path_aligned()
    assert_volume_close(out, ref)


def test_confocal_scatter_time_major_and_aligned_grid(scatter_ctx):
    """An aligned-grid geometry still takes the scattering kernel; the
    time-major layout reads the same samples."""
    g = aligned_geom(**ALIGNED["64"])
    info = make_info(g)
    rng = np.random.default_rng(52)
    I = spiky(rng, (len(g.laser_targets), g.num_time_bins))
    ref = oracle.confocal_voxel_scatter(info, I, g.laser_targets)
    out = scatter_ctx.backproject_confocal(info, cuda(I), cuda(g.laser_targets))
    assert not scatter_ctx.last_confocal_path_aligned()
    assert_volume_close(out, ref)
    out_tm = scatter_ctx.backproject_confocal(info, cuda(I.T), cuda(g.laser_targets),
                                              layout=LAYOUT_TIME_MAJOR)
    assert_volume_close(out_tm, ref)


def test_confocal_scatter_quantized(scatter_ctx):
    g = geom()
    info = make_info(g)
    rng = np.random.default_rng(53)
    I = spiky(rng, (len(g.laser_targets), g.num_time_bins))
    I[I <= 1e-5] = 0.0
    q, sc = scatter_ctx.quantize(cuda(I), STORAGE_F16)
    ref = oracle.confocal_voxel_scatter(info, dequantize(q, sc, STORAGE_F16), g.laser_targets)
    out = scatter_ctx.backproject_confocal(info, q, cuda(g.laser_targets), storage=STORAGE_F16,
                                           scales=sc)
    assert_volume_close(out, ref)


def test_mis_reference_scatter_matches_oracle(scatter_ctx):
    cap = small_confocal(seed=6, res=16)
    info = make_info(cap.geometry)
    I = cap.transient
    L = len(cap.geometry.laser_targets)
    cdf = oracle.mis_cdf(I)
    alias, prob = oracle.alias_tables(cdf)
    noise = np.random.default_rng(1).random(16 ** 3).astype(np.float32)
    ref = oracle.mis(info, I, cap.geometry.laser_targets, cdf, alias, prob, noise, L // 4, 0,
                     scatter=True)
    plain = oracle.mis(info, I, cap.geometry.laser_targets, cdf, alias, prob, noise, L // 4, 0)
    out = scatter_ctx.backproject_confocal_mis(info, cuda(I), cuda(cap.geometry.laser_targets),
                                               cuda(noise), L // 4, mode=0)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), ref) < 1e-4
    assert rel_l2(plain, ref) > 0.1       # the walk adds several bins per sample


def test_mis_tiled_scatter_matches_oracle(scatter_ctx):
    g = geom(n_scan=16, T=256, res=(32, 32, 20))
    info = make_info(g)
    rng = np.random.default_rng(54)
    I = (10.0 ** rng.uniform(-2, 3, size=(256, 256))).astype(np.float32)
    I[rng.random(I.shape) < 0.3] = 0.0
    noise = rng.random(4096).astype(np.float32)
    ref = oracle.mis_tiled(info, I, g.laser_targets, noise, 48, scatter=True)
    out = scatter_ctx.backproject_confocal_mis(info, cuda(I), cuda(g.laser_targets), cuda(noise),
                                               48, mode=MIS_TILED)
    torch.cuda.synchronize()
    assert_volume_close(out, ref)


# --------------------------------------------------- exact backprojection at points
@pytest.mark.parametrize("discard,t0", [(1, 0.0), (0, 0.3)])
def test_backproject_points_matches_oracle(ctx, discard, t0):
    g = geom(discard=discard, t0=t0)
    info = make_info(g)
    rng = np.random.default_rng(61)
    I = spiky(rng, (len(g.laser_targets), g.num_time_bins))
    ref = oracle.confocal_voxel(info, I, g.laser_targets).reshape(-1)
    idx = rng.choice(ref.size, 300, replace=False).astype(np.int32)
    out = ctx.backproject_points(info, cuda(I), cuda(g.laser_targets), cuda(idx))
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.cpu().numpy(), ref[idx], rtol=2e-6, atol=0)
