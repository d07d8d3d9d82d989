"""fbp_depth_slices_rgba against the oracle's TransientImage::save
restatement (FB/TransientImage.cpp:27-45, FB/Image.h:45-177): bit-exact
bytes for the reference's output sizes, plain copies and non-square
images."""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


def activation(P, D, seed):
    rng = np.random.default_rng(seed)
    a = rng.random((P, D), dtype=np.float32)
    a.flat[::97] = 0.0
    a.flat[5::101] = 1.0
    a.flat[7::211] = -0.25      # negative -> 0
    a.flat[11::307] = 3.0       # saturates
    a.flat[13::401] = np.nan    # -> 0
    return a


@pytest.mark.parametrize("W,H,D,OW,OH", [
    (16, 16, 5, 256, 256),      # confocal 16^2: x16
    (32, 32, 3, 256, 256),      # confocal 32^2: x8
    (128, 128, 200, 256, 256),  # confocal 128^2, 200 depths
    (256, 256, 2, 256, 256),    # no resize
    (8, 8, 4, 32, 32),          # exhaustive: x4
    (12, 20, 3, 48, 80),        # non-square
    (100, 100, 1, 200, 200),    # 256 / 100 = 2
    (3, 3, 65, 5, 5),           # ratio 0.5, more than one depth tile
])
def test_slices_match_oracle(ctx, W, H, D, OW, OH):
    a = activation(W * H, D, W * 7 + D)
    got = ctx.depth_slices_rgba(torch.from_numpy(a).cuda(), W, H, OW, OH).cpu().numpy()
    ref = oracle.depth_slices_rgba(a, W, H, OW, OH)
    assert got.shape == ref.shape
    bad = np.argwhere(got != ref)
    assert bad.size == 0, (bad[:5], got[tuple(bad[0])], ref[tuple(bad[0])])


def test_slices_reject_degenerate_resize(ctx):
    a = torch.zeros((4, 1), device="cuda")
    with pytest.raises(RuntimeError):
        ctx.depth_slices_rgba(a, 2, 2, 1, 1)
    with pytest.raises(ValueError):
        ctx.depth_slices_rgba(a, 3, 2)
