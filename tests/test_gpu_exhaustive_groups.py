"""Laser-group kernel of the exhaustive voxel backprojection
(backprojectExhaustiveVoxel, FB/transient_reconstruction.cuh:138-166): taken
for linear-weight captures (every bin >= 160) with at least 8 lasers and 64
sensors.  Sensor distances are shared by groups of eight lasers and the bin
estimate's error bound is folded into the distance terms; voxels whose
estimate lies within the bound of a bin boundary take the reference's IEEE
chain after each batch.

The captures here put the hidden box 0.5 m from the wall at 4.8 mm bins and
use 9 or 144 lasers (a partial last group) and 81 or 144 sensors (one or two
sensor parts per chunk).  Voxel pitches keep a 16 x 16 x 32 tile's bin
window within the kernel's 512 bins."""
import numpy as np
import pytest
import torch

import oracle
from fasterbackprojection_amd.api import Geometry, make_info
from fasterbackprojection_amd.scene import wall_grid

pytestmark = pytest.mark.gpu


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def spiky(rng, shape):
    """Adjacent bins differ by orders of magnitude: a wrong bin moves a voxel
    far outside the tolerance."""
    v = (10.0 ** rng.uniform(-3, 3, size=shape)).astype(np.float32)
    v[rng.random(shape) < 0.2] = 0.0
    return v


def geometry(n_las, n_sen, T=1024, res=(32, 40, 32), discard=1, t0=0.0, depth0=0.5):
    return Geometry(laser_targets=wall_grid(n_las, 1.0), sensor_targets=wall_grid(n_sen, 1.0),
                    num_time_bins=T, time_step=0.0048,
                    hidden_min=(-0.4, depth0, -0.45), hidden_max=(0.45, depth0 + 0.9, 0.4),
                    voxel_resolution=res, confocal=False, discard_first_last_bounces=discard,
                    time_offset=t0, laser_position=(0.1, -0.8, 0.2),
                    sensor_position=(-0.2, -0.7, 0.1), wall_size=(1.0, 1.0))


def run(ctx, info, I, g, groups=True, **kw):
    ctx.set_exhaustive_laser_groups(groups)
    try:
        out = ctx.backproject_exhaustive(info, I, cuda(g.laser_targets), cuda(g.sensor_targets),
                                         **kw)
        used = ctx.last_exhaustive_laser_groups()
    finally:
        ctx.set_exhaustive_laser_groups(True)
    return out, used


@pytest.mark.parametrize("n_las,discard,t0", [(3, 1, 0.0), (3, 0, 0.1), (5, 1, 0.0), (5, 0, 0.1)])
def test_group_kernel_matches_oracle(ctx, n_las, discard, t0):
    g = geometry(n_las, 9, discard=discard, t0=t0)
    info = make_info(g)
    I = spiky(np.random.default_rng(21), (n_las * n_las, 81, 1024))
    ref = oracle.exhaustive_voxel(info, I, g.laser_targets, g.sensor_targets)
    out, used = run(ctx, info, cuda(I), g)
    assert used
    got = out.cpu().numpy().astype(np.float64)
    assert rel_l2(got, ref) <= 1e-4
    big = np.abs(ref) > 0.01 * np.abs(ref).max()
    assert np.max(np.abs(got[big] - ref[big]) / np.abs(ref[big])) <= 5e-4
    # The per-pair kernel on the same input: same bins, weights to FP32 rounding.
    other, used2 = run(ctx, info, cuda(I), g, groups=False)
    assert not used2
    assert rel_l2(got, other.cpu().numpy()) <= 1e-5


def test_group_kernel_needs_the_linear_range(ctx):
    """Bins below 160 (box 0.3 m from the wall): the per-pair kernel runs."""
    g = geometry(3, 9, depth0=0.3)
    info = make_info(g)
    I = spiky(np.random.default_rng(22), (9, 81, 1024))
    out, used = run(ctx, info, cuda(I), g)
    assert not used
    ref = oracle.exhaustive_voxel(info, I, g.laser_targets, g.sensor_targets)
    assert rel_l2(out.cpu().numpy(), ref) <= 1e-4


@pytest.mark.parametrize("n_las,n_sen,discard,t0", [(3, 9, 1, 0.0), (3, 12, 0, 0.1), (3, 12, 1, 0.0),
                                                   (5, 9, 1, 0.0), (5, 8, 0, 0.1)])
def test_group_kernel_bins_exact(ctx, n_las, n_sen, discard, t0):
    """One lit (laser, sensor) pair per call, I[l][s][t] = t + 1: with the
    linear weights vol * dist^2 = (bin + 1)(1 + e), |e| <= 3e-5, which rounds
    to bin + 1 below 10^4.  9 lasers (one full group of 8 and a partial one)
    or 25 (three full groups); pairs from full and partial laser groups and
    from both sensor parts."""
    T = 2048
    g = geometry(n_las, n_sen, T=T, res=(24, 40, 24), discard=discard, t0=t0)
    info = make_info(g)
    L, S = len(g.laser_targets), len(g.sensor_targets)
    rng = np.random.default_rng(5)
    pairs = [(0, 0), (7, S - 1), (8, S // 2), (3, 1), (5, S - 2), (L - 1, 0)]
    if L > 16:
        pairs += [(15, 3), (16, S - 1), (L - 2, 7)]
    pairs += [(int(l), int(s)) for l, s in zip(rng.integers(0, L, 10), rng.integers(0, S, 10))]
    cl, cs = cuda(g.laser_targets), cuda(g.sensor_targets)
    vols = []
    for l, s in pairs:
        I = np.zeros((L, S, T), np.float32)
        I[l, s] = np.arange(1, T + 1, dtype=np.float32)
        vols.append(ctx.backproject_exhaustive(info, cuda(I), cl, cs))
        assert ctx.last_exhaustive_laser_groups()
    got = torch.stack(vols).cpu().numpy()
    for i, (l, s) in enumerate(pairs):
        bins, dist = oracle.exhaustive_bins(info, g.laser_targets, g.sensor_targets, l, s)
        rec = np.rint(got[i].astype(np.float64) * dist.astype(np.float64) ** 2).astype(np.int64) - 1
        bad = np.argwhere(rec != bins)
        assert bad.size == 0, (l, s, bad[:5].tolist())
        assert (bins >= 160).all()


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_group_kernel_quantized_storage(ctx, storage):
    """F16/BF16 storage with per-series scales: the group kernel against the
    per-pair kernel on the same quantised samples, and against the oracle on
    the dequantised samples."""
    from fasterbackprojection_amd._lib import STORAGE_BF16, STORAGE_F16
    st = STORAGE_F16 if storage == "f16" else STORAGE_BF16
    g = geometry(3, 9)
    info = make_info(g)
    I = spiky(np.random.default_rng(23), (9, 81, 1024))
    q, sc = ctx.quantize(cuda(I), st)
    out, used = run(ctx, info, q, g, storage=st, scales=sc)
    assert used
    other, _ = run(ctx, info, q, g, groups=False, storage=st, scales=sc)
    got = out.cpu().numpy()
    assert rel_l2(got, other.cpu().numpy()) <= 1e-5
    qh = q.cpu().numpy().reshape(-1, 1024)
    if st == STORAGE_F16:
        v = qh.view(np.float16).astype(np.float32)
    else:
        v = (qh.view(np.uint16).astype(np.uint32) << 16).view(np.float32)
    deq = (v * sc.cpu().numpy()[:, None]).astype(np.float32).reshape(I.shape)
    ref = oracle.exhaustive_voxel(info, deq, g.laser_targets, g.sensor_targets)
    assert rel_l2(got, ref) <= 1e-4


# ------------------------------------------------------------ depth planes
def depth_geometry(n_las, n_sen, T=1024, discard=1, t0=0.0):
    return Geometry(laser_targets=wall_grid(n_las, 1.0), sensor_targets=wall_grid(n_sen, 1.0),
                    num_time_bins=T, time_step=0.0048, hidden_min=(-0.5, 0.2, -0.5),
                    hidden_max=(0.5, 1.0, 0.5), voxel_resolution=(8, 8, 8), confocal=False,
                    discard_first_last_bounces=discard, time_offset=t0,
                    laser_position=(0.1, -0.8, 0.2), sensor_position=(-0.2, -0.7, 0.1),
                    wall_size=(1.0, 1.0))


def run_depth(ctx, info, I, g, depths, groups=True):
    ctx.set_exhaustive_laser_groups(groups)
    try:
        out = ctx.backproject_depth(info, I, cuda(g.laser_targets), cuda(depths),
                                    sensor_targets=cuda(g.sensor_targets))
        used = ctx.last_exhaustive_laser_groups()
    finally:
        ctx.set_exhaustive_laser_groups(True)
    return out.cpu().numpy().astype(np.float64), used


@pytest.mark.parametrize("discard,t0", [(1, 0.0), (0, 0.1)])
def test_depth_group_matches_oracle_and_pair_kernel(ctx, discard, t0):
    """Depth planes (backprojectExhaustive, FB/transient_reconstruction.cuh:66-91)
    through the laser-group form: 36 pixels (a full tile of 32 and a partial
    one; 5 laser groups, the last of 4), 81 sensors, 300 depths (two depth
    tiles), against the oracle and the staged-window kernel."""
    g = depth_geometry(6, 9, discard=discard, t0=t0)
    info = make_info(g)
    rng = np.random.default_rng(31)
    I = spiky(rng, (36, 81, 1024))
    depths = np.linspace(0.3, 0.9, 300, endpoint=False)
    ref = oracle.depth(info, I, g.laser_targets, depths, sensor=g.sensor_targets, exhaustive=True)
    out, used = run_depth(ctx, info, cuda(I), g, depths)
    assert used
    assert np.abs(ref).max() > 0
    err = np.abs(out - ref) / np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())
    assert err.max() <= 2e-5, np.unravel_index(np.argmax(err), err.shape)
    other, used2 = run_depth(ctx, info, cuda(I), g, depths, groups=False)
    assert not used2
    assert rel_l2(out, other) <= 1e-6


@pytest.mark.parametrize("discard,t0", [(1, 0.0), (0, 0.11)])
def test_depth_group_bins_exact(ctx, discard, t0):
    """One lit (laser, sensor) pair per call with I[l][s][t] = t + 1: each
    activation is the single sample bin + 1 (exact in FP32), so every
    (pixel, depth) bin is compared with the oracle's restatement of the
    reference's IEEE chain; pairs from the full and the partial laser group."""
    T = 1024
    g = depth_geometry(3, 8, T=T, discard=discard, t0=t0)
    info = make_info(g)
    L, S = len(g.laser_targets), len(g.sensor_targets)
    depths = np.linspace(0.2, 1.0, 37) + 0.0013
    cl, cs, cd = cuda(g.laser_targets), cuda(g.sensor_targets), cuda(depths)
    pairs = [(0, 0), (7, S - 1), (8, 5), (4, 33), (8, S - 1), (2, 17)]
    acts = []
    for l, s in pairs:
        I = np.zeros((L, S, T), np.float32)
        I[l, s] = np.arange(1, T + 1, dtype=np.float32)
        acts.append(ctx.backproject_depth(info, cuda(I), cl, cd, sensor_targets=cs))
        assert ctx.last_exhaustive_laser_groups()
    got = torch.stack(acts).cpu().numpy()
    for i, (l, s) in enumerate(pairs):
        bins = oracle.depth_bins(info, g.laser_targets, depths, l, g.sensor_targets, s, True)
        assert np.array_equal(got[i], np.rint(got[i])), (l, s)
        rec = np.rint(got[i].astype(np.float64)).astype(np.int64) - 1
        bad = np.argwhere(rec != bins)
        assert bad.size == 0, (l, s, bad[:5].tolist())
        assert (bins >= 0).mean() > 0.9
