"""fbp_backproject_exhaustive_voxel_noalloc (partial sums in the library's
fixed per-GPU array, no allocation) against fbp_backproject_exhaustive_voxel
(partial sums in the context's arena), through the C ABI.

Two calls in a row, the second with a smaller input: the array keeps the
first call's partial sums, which the second must neither read nor need
cleared.  The first call's volume is large enough that the array caps its
pair chunks below the allocating entry point's count, so the two paths
reduce differently split partial sums (results equal to FP32 rounding)."""
import numpy as np
import pytest
import torch

import oracle
from fasterbackprojection_amd.api import Context, FbpError, Geometry, make_info
from fasterbackprojection_amd.scene import wall_grid

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _case(n_las, n_sen, T, res, seed):
    lasers, sensors = wall_grid(n_las, 1.0), wall_grid(n_sen, 1.0)
    g = Geometry(laser_targets=lasers, sensor_targets=sensors, num_time_bins=T, time_step=0.0096,
                 hidden_min=(-0.5, 0.3, -0.5), hidden_max=(0.5, 1.3, 0.5), voxel_resolution=res,
                 confocal=False, wall_size=(1.0, 1.0))
    rng = np.random.default_rng(seed)
    I = (10.0 ** rng.uniform(-3, 2, (len(lasers), len(sensors), T))).astype(np.float32)
    I[rng.random(I.shape) < 0.3] = 0.0
    return g, I


def _cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_noalloc_matches_the_allocating_entry_point():
    noalloc_ctx = Context(0)     # its arena must stay empty
    alloc_ctx = Context(0)
    try:
        # 128 x 128 x 64 voxels = 128 tiles (4 MiB of partial sums per pair
        # chunk): the 32 MiB array fits 8 chunks, the allocating path uses 16.
        cases = [_case(4, 8, 256, (128, 128, 64), 1), _case(3, 4, 300, (24, 20, 40), 2)]
        for i, (g, I) in enumerate(cases):
            info = make_info(g)
            r = info.voxel_resolution
            It, lt, st = _cuda(I), _cuda(g.laser_targets), _cuda(g.sensor_targets)
            out = torch.full((r.z, r.y, r.x), float("nan"), dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            mem = torch.cuda.memory_allocated()
            noalloc_ctx.backproject_exhaustive_noalloc(info, It, lt, st, out)
            torch.cuda.synchronize()
            assert torch.cuda.memory_allocated() == mem
            assert noalloc_ctx.scratch_bytes() == 0
            ref = alloc_ctx.backproject_exhaustive(info, It, lt, st).cpu().numpy()
            got = out.cpu().numpy()
            assert np.all(np.isfinite(got)), i
            e = rel_l2(got, ref)
            assert e <= 1e-5, (i, e)
            big = np.abs(ref) > 0.01 * np.abs(ref).max()
            assert np.max(np.abs(got[big] - ref[big]) / np.abs(ref[big])) <= 1e-4, i
            if i == 1:   # small enough for the FP64 oracle
                exact = oracle.exhaustive_voxel(info, I, g.laser_targets, g.sensor_targets)
                assert rel_l2(got, exact) <= 1e-4
        assert alloc_ctx.scratch_bytes() > 0
    finally:
        noalloc_ctx.close()
        alloc_ctx.close()


def test_noalloc_refuses_volumes_beyond_its_array():
    """256^3 voxels need 64 MiB of partial sums per chunk: refused with
    FBP_ERR_UNSUPPORTED before any launch (the output stays untouched)."""
    ctx = Context(0)
    try:
        g, I = _case(1, 1, 16, (256, 256, 256), 3)
        info = make_info(g)
        out = torch.full((256, 256, 256), 7.0, dtype=torch.float32, device="cuda")
        with pytest.raises(FbpError) as e:
            ctx.backproject_exhaustive_noalloc(info, _cuda(I), _cuda(g.laser_targets),
                                               _cuda(g.sensor_targets), out)
        assert e.value.status == 5
        torch.cuda.synchronize()
        assert bool((out == 7.0).all())
        assert ctx.scratch_bytes() == 0
    finally:
        ctx.close()


def _linear_case(n_las, n_sen, res, seed):
    """Box 0.5 m from the wall at 4.8 mm bins (every bin >= 160): the
    laser-group kernel of the exhaustive backprojection runs."""
    lasers, sensors = wall_grid(n_las, 1.0), wall_grid(n_sen, 1.0)
    g = Geometry(laser_targets=lasers, sensor_targets=sensors, num_time_bins=1024,
                 time_step=0.0048, hidden_min=(-0.45, 0.5, -0.45), hidden_max=(0.45, 1.4, 0.45),
                 voxel_resolution=res, confocal=False, wall_size=(1.0, 1.0))
    rng = np.random.default_rng(seed)
    I = (10.0 ** rng.uniform(-3, 2, (len(lasers), len(sensors), 1024))).astype(np.float32)
    I[rng.random(I.shape) < 0.3] = 0.0
    return g, I


def test_noalloc_laser_group_kernel():
    """Both entry points on the laser-group kernel: 144 lasers (18 groups of
    8) x 64 sensors -> 128 x 128 x 64 voxels, whose 4 MiB of partial sums per
    chunk leave the 32 MiB array 8 chunks (three laser groups each); then a
    smaller capture (9 lasers, a partial second group) checked against the
    FP64 oracle."""
    noalloc_ctx = Context(0)
    alloc_ctx = Context(0)
    try:
        cases = [_linear_case(12, 8, (128, 128, 64), 4), _linear_case(3, 8, (24, 40, 24), 5)]
        for i, (g, I) in enumerate(cases):
            info = make_info(g)
            r = info.voxel_resolution
            It, lt, st = _cuda(I), _cuda(g.laser_targets), _cuda(g.sensor_targets)
            out = torch.full((r.z, r.y, r.x), float("nan"), dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            mem = torch.cuda.memory_allocated()
            noalloc_ctx.backproject_exhaustive_noalloc(info, It, lt, st, out)
            torch.cuda.synchronize()
            assert torch.cuda.memory_allocated() == mem
            assert noalloc_ctx.scratch_bytes() == 0
            assert noalloc_ctx.last_exhaustive_laser_groups(), i
            ref = alloc_ctx.backproject_exhaustive(info, It, lt, st).cpu().numpy()
            assert alloc_ctx.last_exhaustive_laser_groups(), i
            got = out.cpu().numpy()
            assert np.all(np.isfinite(got)), i
            assert rel_l2(got, ref) <= 1e-6, i
            if i == 1:
                exact = oracle.exhaustive_voxel(info, I, g.laser_targets, g.sensor_targets)
                assert rel_l2(got, exact) <= 1e-4
    finally:
        noalloc_ctx.close()
        alloc_ctx.close()
