"""Parity at the full workload sizes: 128^2 x 1024 -> 128^3 confocal (aligned
path), 256^2 x 1024 F16/BF16 -> 128^3, 16^2 x 64^2 x 1024 exhaustive -> 64^3
and 200 depth planes of the 128^2 capture.  The FP64 oracle runs on the
first depth slices (a bounded sample of the same volume); the rest of each
volume is checked through properties that do not depend on size: agreement
of the two independent confocal implementations (ring convolution and the
general gather kernel), linearity in the transient, and the peak of a point
scatterer."""
import numpy as np
import pytest
import torch

import oracle
from oracle import lct as olct
from fasterbackprojection_amd.api import STORAGE_BF16, STORAGE_F16, make_info
from fasterbackprojection_amd.scene import (exhaustive_geometry, full_confocal_geometry,
                                            random_quads, random_spheres)

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def confocal_capture(ctx, n_scan, n_quads, peak, seed):
    g = full_confocal_geometry(n_scan, 1024, 128)
    dev = torch.device("cuda", 0)
    sc = random_quads(np.random.default_rng(seed), n_quads, 20000)
    laser = torch.from_numpy(g.laser_targets).to(dev)
    clean = ctx.render_confocal(laser, torch.from_numpy(sc.points.astype(np.float32)).to(dev),
                                torch.from_numpy(sc.albedo.astype(np.float32)).to(dev), 1024,
                                g.time_step)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    I = torch.poisson(clean * (peak / clean.max()), generator=gen).contiguous()
    return g, laser, I


def first_slices(info, nz):
    sub = type(info).from_buffer_copy(info)
    sub.voxel_resolution.z = nz
    return sub


@pytest.fixture(scope="module")
def c2(ctx):
    g, laser, I = confocal_capture(ctx, 128, 4, 1000.0, 11)
    return g, make_info(g), laser, I


def test_c2_full_matches_oracle_slices(ctx, c2):
    g, info, laser, I = c2
    vol = ctx.backproject_confocal(info, I, laser)
    assert ctx.last_confocal_path_aligned()
    ref = oracle.confocal_voxel(first_slices(info, 2), I.cpu().numpy(), g.laser_targets)
    got = vol[:2].cpu().numpy()
    assert rel_l2(got, ref) <= 1e-4
    big = np.abs(ref) > 0.01 * np.abs(ref).max()
    assert np.max(np.abs(got[big] - ref[big]) / np.abs(ref[big])) <= 5e-4


def test_c2_full_aligned_equals_general(ctx, c2):
    g, info, laser, I = c2
    a = ctx.backproject_confocal(info, I, laser).cpu().numpy()
    ctx.set_aligned_path(False)
    try:
        b = ctx.backproject_confocal(info, I, laser).cpu().numpy()
        assert not ctx.last_confocal_path_aligned()
    finally:
        ctx.set_aligned_path(True)
    assert rel_l2(a, b) <= 1e-5


def test_c2_full_linearity(ctx, c2):
    g, info, laser, I = c2
    J = torch.roll(I, shifts=37, dims=1).contiguous()  # a second, different transient
    both = ctx.backproject_confocal(info, (2.0 * I + 3.0 * J).contiguous(), laser)
    a = ctx.backproject_confocal(info, I, laser)
    b = ctx.backproject_confocal(info, J, laser)
    assert rel_l2(both.cpu().numpy(), (2.0 * a + 3.0 * b).cpu().numpy()) <= 1e-6


def test_c2_full_point_scatterer_peak(ctx):
    """One point at a voxel centre: the backprojection peaks at that voxel
    (or a face neighbour: the ellipsoids are quantised to 4.8 mm bins)."""
    g = full_confocal_geometry(128, 1024, 128)
    info = make_info(g)
    vs = np.asarray([info.hidden_volume_voxel_size.x, info.hidden_volume_voxel_size.y,
                     info.hidden_volume_voxel_size.z], np.float32)
    hmin = np.asarray([info.hidden_volume_min.x, info.hidden_volume_min.y,
                       info.hidden_volume_min.z], np.float32)
    vx, vy, z = 70, 45, 60            # world x <- vx, world y (depth) <- z, world z <- vy
    p = hmin + (np.array([vx, z, vy], np.float32) + np.float32(0.5)) * vs
    dev = torch.device("cuda", 0)
    laser = torch.from_numpy(g.laser_targets).to(dev)
    I = ctx.render_confocal(laser, torch.from_numpy(p.reshape(1, 3)).to(dev),
                            torch.ones(1, device=dev), 1024, g.time_step)
    vol = ctx.backproject_confocal(info, I.contiguous(), laser).cpu().numpy()
    peak = np.unravel_index(np.argmax(vol), vol.shape)
    assert np.abs(np.array(peak) - np.array([z, vy, vx])).sum() <= 1, peak


def test_c4_full_quantized_matches_oracle_slices(ctx):
    g, laser, I = confocal_capture(ctx, 256, 8, 500.0, 12)
    info = make_info(g)
    host = I.cpu().numpy()
    ref = oracle.confocal_voxel(first_slices(info, 2), host, g.laser_targets)
    for storage in (STORAGE_F16, STORAGE_BF16):
        q, sc = ctx.quantize(I, storage)
        vol = ctx.backproject_confocal(info, q, laser, storage=storage, scales=sc)
        assert ctx.last_confocal_path_aligned()
        # Storage bar against the FP64 oracle of the FP32 transient.
        assert rel_l2(vol[:2].cpu().numpy(), ref) <= 2e-3


def test_depth_planes_full_matches_oracle(ctx, c2):
    g, info, laser, I = c2
    focus = abs((np.float32(g.hidden_min[1]) + np.float32(g.hidden_max[1])) * np.float32(0.5))
    lo = float(np.float32(focus - np.float32(0.1)))
    hi = float(np.float32(focus + np.float32(0.1)))
    depths = lo + (hi - lo) / 200 * np.arange(200)
    act = ctx.backproject_depth(info, I, laser, torch.from_numpy(depths).cuda())
    assert ctx.last_confocal_path_aligned()
    got = act.cpu().numpy()
    for cols in ([0, 1], [117, 199]):
        ref = oracle.depth(info, I.cpu().numpy(), g.laser_targets, depths[cols])
        assert rel_l2(got[:, cols], ref) <= 1e-6


def test_c3_full_exhaustive_matches_oracle_slice(ctx):
    g = exhaustive_geometry()
    info = make_info(g)
    dev = torch.device("cuda", 0)
    lasers = torch.from_numpy(g.laser_targets).to(dev)
    sensors = torch.from_numpy(g.sensor_targets).to(dev)
    sc = random_spheres(np.random.default_rng(13), 8, 8000)
    I = ctx.render_exhaustive(lasers, sensors, torch.from_numpy(sc.points.astype(np.float32)).to(dev),
                              torch.from_numpy(sc.albedo.astype(np.float32)).to(dev),
                              g.num_time_bins, g.time_step)
    I.mul_(1000.0 / I.max())
    gen = torch.Generator(device=dev)
    gen.manual_seed(13)
    I = torch.poisson(I, generator=gen)
    vol = ctx.backproject_exhaustive(info, I, lasers, sensors)
    ref = oracle.exhaustive_voxel(first_slices(info, 1), I.cpu().numpy(), g.laser_targets,
                                  g.sensor_targets)
    assert rel_l2(vol[:1].cpu().numpy(), ref) <= 1e-4
    assert bool(torch.isfinite(vol).all())


def test_c3_full_depth_planes_match_oracle(ctx):
    """200 depth planes of the 16^2 x 64^2 x 1024 capture (staged-window
    kernel, row-wrapping workgroups included) against the oracle on two of
    the planes, output by output: one wrong bin moves an output far beyond
    the tolerance."""
    g = exhaustive_geometry()
    info = make_info(g)
    dev = torch.device("cuda", 0)
    lasers = torch.from_numpy(g.laser_targets).to(dev)
    sensors = torch.from_numpy(g.sensor_targets).to(dev)
    sc = random_spheres(np.random.default_rng(14), 8, 8000)
    I = ctx.render_exhaustive(lasers, sensors, torch.from_numpy(sc.points.astype(np.float32)).to(dev),
                              torch.from_numpy(sc.albedo.astype(np.float32)).to(dev),
                              g.num_time_bins, g.time_step)
    I.mul_(1000.0 / I.max())
    gen = torch.Generator(device=dev)
    gen.manual_seed(14)
    I = torch.poisson(I, generator=gen)
    focus = abs((np.float32(g.hidden_min[1]) + np.float32(g.hidden_max[1])) * np.float32(0.5))
    lo = float(np.float32(focus - np.float32(0.1)))
    hi = float(np.float32(focus + np.float32(0.1)))
    depths = lo + (hi - lo) / 200 * np.arange(200)
    act = ctx.backproject_depth(info, I, lasers, torch.from_numpy(depths).to(dev),
                                sensor_targets=sensors).cpu().numpy().astype(np.float64)
    cols = [3, 151]
    ref = oracle.depth(info, I.cpu().numpy(), g.laser_targets, depths[cols],
                       sensor=g.sensor_targets, exhaustive=True)
    got = act[:, cols]
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())
    assert err.max() <= 1e-5
    assert np.isfinite(act).all()


def test_lct_full_c2_matches_restatement(ctx, c2):
    """LCT of the full capture (128 x 128 x 1024 padded to 256 x 256 x 2048,
    snr 8) against the NumPy/SciPy restatement of python/lct.py."""
    g, info, laser, I = c2
    data = I.reshape(128, 128, 1024).contiguous()
    width, rmax = 1.0, 1024 * float(g.time_step)
    out = ctx.lct(data, width, rmax, 8.0).cpu().numpy()
    ref = olct.lct(data.cpu().numpy(), width, rmax, 8.0)
    assert rel_l2(out, ref) <= 1e-4
