"""GPU refiner against the FP64 oracle: forward (with the GPU's BF16 rounding
points emulated, which pins the MFMA operand layout), gradients (central
finite differences of the FP64 loss) and a training step that lowers the
loss."""
import numpy as np
import pytest
import torch

from fasterbackprojection_amd.api import Refiner
from oracle import refiner as orf

pytestmark = pytest.mark.gpu


def _params(seed, scale=0.15):
    rng = np.random.default_rng(seed)
    p = np.zeros(orf.param_count(), np.float32)
    o = 0
    for ci, co, k in orf.LAYERS:
        n = k ** 3 * ci * co
        p[o:o + n] = rng.normal(size=n) * np.sqrt(2.0 / (k ** 3 * ci))
        p[o + n:o + n + co] = rng.normal(size=co) * scale
        o += n + co
    return p


@pytest.mark.parametrize("S", [16, 32, 64])
def test_forward_matches_oracle(ctx, S):
    """S = 64 is the learned form's volume (the size the benchmark times)."""
    p = _params(S)
    ref = Refiner(ctx, S)
    ref.load(p)
    rng = np.random.default_rng(1)
    vol = rng.random((S, S, S)).astype(np.float32) * 3 - 1
    out = ref.forward(torch.from_numpy(vol).cuda()).cpu().numpy()
    exp = orf.forward(p, vol, emulate_bf16=True)
    e = np.linalg.norm(out - exp) / np.linalg.norm(exp)
    assert e < 5e-3, e
    exp64 = orf.forward(p, vol)
    assert np.linalg.norm(out - exp64) / np.linalg.norm(exp64) < 5e-2
    ref.close()


def test_gradient_matches_finite_differences(ctx):
    S = 16
    p = _params(3)
    ref = Refiner(ctx, S)
    ref.load(p)
    rng = np.random.default_rng(2)
    vol = rng.random((S, S, S)).astype(np.float32)
    tgt = rng.random((S, S, S)).astype(np.float32)
    g, loss = ref.gradient(torch.from_numpy(vol).cuda(), torch.from_numpy(tgt).cuda())
    assert abs(loss - orf.loss(p, vol, tgt)) < 2e-2 * abs(orf.loss(p, vol, tgt))
    # parameters spread over all layers, weights and biases
    offs, idx, o = [], [], 0
    for ci, co, k in orf.LAYERS:
        n = k ** 3 * ci * co
        offs.append((o, n, co))
        o += n + co
    r = np.random.default_rng(5)
    for (o, n, co) in offs:
        idx += list(o + r.choice(n, 4, replace=False)) + [o + n + int(r.integers(co))]
    fd = orf.fd_grad(p, vol, tgt, idx, h=2e-3)
    gi = g[idx]
    big = np.abs(fd) > 1e-3 * np.abs(fd).max()
    rel = np.abs(gi[big] - fd[big]) / np.abs(fd[big])
    cos = gi @ fd / (np.linalg.norm(gi) * np.linalg.norm(fd))
    assert cos > 0.99, (cos, gi, fd)
    assert np.median(rel) < 0.05, rel
    ref.close()


def test_training_reduces_loss(ctx):
    S = 32
    ref = Refiner(ctx, S, seed=7)
    rng = np.random.default_rng(3)
    x = torch.from_numpy(rng.random((2, S, S, S)).astype(np.float32)).cuda()
    t = torch.from_numpy(np.clip(x.cpu().numpy() * 0.5 + 0.2, 0, 1).astype(np.float32)).cuda()
    first = ref.train_step(x, t, lr=1e-3)
    for _ in range(40):
        last = ref.train_step(x, t, lr=1e-3)
    assert last < 0.7 * first, (first, last)
    ref.close()
