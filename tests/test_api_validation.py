"""Host-side argument checks of the Python wrappers (CPU only).

Each wrapper must refuse a mistyped or short buffer with ValueError before
any pointer reaches the C ABI: the context used here has no library behind
it (lib = None), so a call that got past validation would fail with
AttributeError instead.
"""
import numpy as np
import pytest
import torch

from fasterbackprojection_amd.api import (STORAGE_F16, Context, Geometry, Refiner, make_info)
from fasterbackprojection_amd.scene import wall_grid


@pytest.fixture
def ctx():
    c = Context.__new__(Context)
    c.lib = None
    c.handle = None
    c.device = 0
    return c


def _info(confocal=True, L=16, S=16, T=32, res=(4, 5, 6)):
    lasers = wall_grid(int(np.sqrt(L)), 1.0)
    sensors = wall_grid(int(np.sqrt(S)), 1.0)
    g = Geometry(laser_targets=lasers, sensor_targets=sensors, num_time_bins=T, time_step=0.01,
                 hidden_min=(-0.4, 0.4, -0.4), hidden_max=(0.4, 1.2, 0.4), voxel_resolution=res,
                 confocal=confocal)
    return make_info(g), torch.from_numpy(lasers), torch.from_numpy(sensors)


def _vol(info, dtype=torch.float32):
    r = info.voxel_resolution
    return torch.zeros((r.z, r.y, r.x), dtype=dtype)


def test_confocal_backprojection_rejects_bad_buffers(ctx):
    info, lasers, _ = _info()
    good = torch.zeros((16, 32))
    out = _vol(info)
    with pytest.raises(ValueError, match="intensity must hold 512"):
        ctx.backproject_confocal(info, torch.zeros((16, 31)), lasers, out=out)
    with pytest.raises(ValueError, match="intensity must be torch.float32"):
        ctx.backproject_confocal(info, good.double(), lasers, out=out)
    with pytest.raises(ValueError, match="laser_targets"):
        ctx.backproject_confocal(info, good, lasers[:15], out=out)
    with pytest.raises(ValueError, match="laser_targets"):
        ctx.backproject_confocal(info, good, lasers.double(), out=out)
    with pytest.raises(ValueError, match="out must have shape"):
        ctx.backproject_confocal(info, good, lasers, out=torch.zeros((6, 5, 5)))
    with pytest.raises(ValueError, match="out must be torch.float32"):
        ctx.backproject_confocal(info, good, lasers, out=_vol(info, torch.float64))
    # F16 storage: two-byte samples and one float32 scale per scan point
    with pytest.raises(ValueError, match="intensity must be"):
        ctx.backproject_confocal(info, good, lasers, storage=STORAGE_F16,
                                 scales=torch.ones(16), out=out)
    with pytest.raises(ValueError, match="scales must hold 16"):
        ctx.backproject_confocal(info, good.half(), lasers, storage=STORAGE_F16,
                                 scales=torch.ones(15), out=out)
    with pytest.raises(ValueError, match="scales must be"):
        ctx.backproject_confocal(info, good.half(), lasers, storage=STORAGE_F16, scales=None,
                                 out=out)
    # valid arguments get through validation (and stop at the missing library)
    with pytest.raises((AttributeError, ValueError), match="GPU|NoneType|attribute"):
        ctx.backproject_confocal(info, good, lasers, out=out)


def test_exhaustive_backprojection_rejects_bad_buffers(ctx):
    info, lasers, sensors = _info(confocal=False, L=4, S=9)
    out = _vol(info)
    with pytest.raises(ValueError, match="intensity must hold 1152"):
        ctx.backproject_exhaustive(info, torch.zeros((4, 32)), lasers, sensors, out=out)
    with pytest.raises(ValueError, match="sensor_targets"):
        ctx.backproject_exhaustive(info, torch.zeros((4, 9, 32)), lasers, sensors[:8], out=out)


def test_depth_planes_reject_bad_buffers(ctx):
    info, lasers, sensors = _info()
    I = torch.zeros((16, 32))
    with pytest.raises(ValueError, match="depths must be torch.float64"):
        ctx.backproject_depth(info, I, lasers, torch.zeros(5, dtype=torch.float32))
    with pytest.raises(ValueError, match="out must have shape"):
        ctx.backproject_depth(info, I, lasers, torch.zeros(5, dtype=torch.float64),
                              out=torch.zeros((16, 4)))
    einfo, el, es = _info(confocal=False, L=4, S=9)
    with pytest.raises(ValueError, match="intensity must hold"):
        ctx.backproject_depth(einfo, torch.zeros((4, 32)), el, torch.zeros(5, dtype=torch.float64),
                              es, out=torch.zeros((4, 5)))


def test_mis_and_points_reject_bad_buffers(ctx):
    info, lasers, _ = _info()
    I = torch.zeros((16, 32))
    with pytest.raises(ValueError, match="noise must be torch.float32"):
        ctx.backproject_confocal_mis(info, I, lasers, torch.zeros(8, dtype=torch.float64), 4,
                                     out=_vol(info))
    with pytest.raises(ValueError, match="noise must hold at least 1"):
        ctx.backproject_confocal_mis(info, I, lasers, torch.zeros(0), 4, out=_vol(info))
    with pytest.raises(ValueError, match="voxels must be"):
        ctx.backproject_points(info, I, lasers, torch.zeros(3, dtype=torch.int64),
                               out=torch.zeros(3, dtype=torch.float64))
    with pytest.raises(ValueError, match="out must hold 3"):
        ctx.backproject_points(info, I, lasers, torch.zeros(3, dtype=torch.int32),
                               out=torch.zeros(2, dtype=torch.float64))


def test_filters_reject_bad_buffers(ctx):
    v = torch.zeros((6, 5, 4))
    with pytest.raises(ValueError, match="out must have shape"):
        ctx.laplacian(v, 5, out=torch.zeros((6, 5, 3)))
    with pytest.raises(ValueError, match="out must have shape"):
        ctx.log(v, 5, 1.0, out=torch.zeros(120))
    with pytest.raises(ValueError, match="volume must be torch.float32"):
        ctx.log(v.double(), 5, 1.0, out=torch.zeros((6, 5, 4)))
    with pytest.raises(ValueError, match=r"volume must be \[Z\]\[Y\]\[X\]"):
        ctx.laplacian(torch.zeros(120), 5, out=torch.zeros(120))
    with pytest.raises(ValueError, match="volume must be torch.float32"):
        ctx.log_fft(v.half())
    with pytest.raises(ValueError, match="v must be torch.float32"):
        ctx.normalize(v.double())
    with pytest.raises(ValueError, match="data must be torch.float32"):
        ctx.filter_h(torch.zeros((3, 8), dtype=torch.float64), 0.01, 0.1)


def test_lct_rejects_bad_buffers(ctx):
    d = torch.zeros((8, 8, 64))
    with pytest.raises(ValueError, match="out must have shape"):
        ctx.lct(d, 0.5, 1.0, out=torch.zeros((64, 8, 8)), downsample=1)
    with pytest.raises(ValueError, match="downsample"):
        ctx.lct(d, 0.5, 1.0, downsample=6)
    with pytest.raises(ValueError, match=r"data must be \(N, N, m\)"):
        ctx.lct(torch.zeros((8, 4, 64)), 0.5, 1.0)
    with pytest.raises(ValueError, match="data must be torch.float32"):
        ctx.lct(d.double(), 0.5, 1.0)


def test_renderers_and_residual_reject_bad_buffers(ctx):
    pts = torch.zeros((5, 3))
    with pytest.raises(ValueError, match="albedo must hold 5"):
        ctx.render_confocal(torch.zeros((4, 3)), pts, torch.zeros(4), 16, 0.01,
                            out=torch.zeros((4, 16)))
    with pytest.raises(ValueError, match="out must hold 64"):
        ctx.render_confocal(torch.zeros((4, 3)), pts, torch.zeros(5), 16, 0.01,
                            out=torch.zeros((4, 15)))
    with pytest.raises(ValueError, match="out must hold"):
        ctx.render_exhaustive(torch.zeros((2, 3)), torch.zeros((3, 3)), pts, torch.zeros(5), 16,
                              0.01, out=torch.zeros((2, 3, 15)))
    info, lasers, _ = _info()
    with pytest.raises(ValueError, match="volume must have shape"):
        ctx.volume_residual(info, torch.zeros((6, 5, 5)), torch.zeros((16, 32)), lasers,
                            torch.zeros(4, dtype=torch.int32))


def test_refiner_rejects_bad_volumes(ctx):
    r = Refiner.__new__(Refiner)
    r.ctx, r.lib, r.size, r.handle = ctx, None, 16, None
    with pytest.raises(ValueError, match="volume must have shape"):
        r.forward(torch.zeros((16, 16, 15)))
    with pytest.raises(ValueError, match="out must have shape"):
        r.forward(torch.zeros((16, 16, 16)), out=torch.zeros((16, 16, 8)))
    with pytest.raises(ValueError, match="targets must hold 8192"):
        r.train_step(torch.zeros((2, 16, 16, 16)), torch.zeros((1, 16, 16, 16)), 1e-3)
