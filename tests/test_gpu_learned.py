"""Learned form with the committed weights: it must beat the same method
without the refiner (MIS + LoG + normalisation) against the plain target,
and its forward-model residual check must fall back to the plain path when
forced (threshold 0) and on an out-of-distribution capture."""
import numpy as np
import pytest
import torch

from fasterbackprojection_amd.learned import WEIGHTS, LearnedReconstructor, rel_l2

pytestmark = pytest.mark.gpu

need_weights = pytest.mark.skipif(not (WEIGHTS / "refiner_c5.f32").exists(),
                                  reason="no trained refiner weights")


@need_weights
def test_refiner_beats_unrefined_input(ctx):
    lrn = LearnedReconstructor(ctx)
    p = lrn.p
    el, eb = [], []
    for s in range(3_000_000, 3_000_008):
        I = p.render(s)
        y = p.plain(I).clone()
        out, fb, _ = lrn(I, s, check=False)
        el.append(rel_l2(out, y))
        eb.append(rel_l2(p.mis_input(I, s), y))
    assert np.mean(el) < np.mean(eb), (np.mean(el), np.mean(eb))


@need_weights
def test_forced_fallback_returns_plain_path(ctx):
    lrn = LearnedReconstructor(ctx, threshold=0.0)
    I = lrn.p.render(3_000_100)
    y = lrn.p.plain(I).clone()
    out, fell_back, r = lrn(I, 3_000_100)
    assert fell_back and r > 0.0
    assert torch.equal(out, y)


@need_weights
def test_out_of_distribution_falls_back(ctx):
    lrn = LearnedReconstructor(ctx)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    I = torch.poisson(torch.full((64 * 64, 512), 3.0, device="cuda"), generator=gen).contiguous()
    _, fell_back, r = lrn(I, 0)
    assert fell_back, r


@need_weights
@pytest.mark.parametrize("kind", ["shuffled_scan_points", "sparse_noise"])
def test_corrupted_captures_fall_back(ctx, kind):
    """A capture whose scan points were permuted, or one of sparse
    background counts only (mean 0.3 per bin), is inconsistent with any
    volume the refiner produces for it: the check rejects it.  (A capture
    delayed in time is not such a case: it is consistent with the scene
    moved farther away.)"""
    lrn = LearnedReconstructor(ctx)
    g = torch.Generator(device="cuda")
    g.manual_seed(8)
    if kind == "shuffled_scan_points":
        I = lrn.p.render(3_000_300)
        I = I[torch.randperm(I.shape[0], generator=g, device="cuda")].contiguous()
    else:
        I = torch.poisson(torch.full((64 * 64, 512), 0.3, device="cuda"), generator=g)
    _, fell_back, r = lrn(I.contiguous(), 0)
    assert fell_back, r


@need_weights
def test_in_distribution_mostly_accepted(ctx):
    lrn = LearnedReconstructor(ctx)
    acc = 0
    for s in range(3_000_200, 3_000_216):
        I = lrn.p.render(s)
        _, fell_back, _ = lrn(I, s)
        acc += not fell_back
    assert acc >= 13, acc
