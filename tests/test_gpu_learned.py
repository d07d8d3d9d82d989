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


def test_calibration_recovers_an_affine_distortion(ctx):
    """With the exact LoG as the reference, the calibration maps any
    a*v + b (a > 0) of the plain path's normalised volume v back onto v:
    the fit is exact and the extremes keep their voxels."""
    from fasterbackprojection_amd.learned import Pipeline, geometry
    p = Pipeline(ctx, geometry())
    I = p.render(3_000_400)
    v = p.plain(I).clone()
    ctx.backproject_confocal(p.info, I, p.laser, out=p.vol)
    L = torch.empty_like(p.vol)
    ctx.log(p.vol, 5, 1.0, out=L)
    for a, b in ((0.73, 0.11), (1.0, 0.0), (2.5, -0.4)):
        d = (a * v + b).contiguous()
        ctx.learned_calibrate(p.info, I, p.laser, d, L, 16)
        assert rel_l2(d, v) < 1e-5, (a, b, rel_l2(d, v))


@need_weights
def test_learned_path_within_two_percent(ctx):
    """Calibrated learned output against the plain path on 16 held-out seeds:
    median rel-L2 below 2 % (the mean over 256 seeds is recorded with the
    weights; a scene whose true LoG extreme lies outside the probed voxels
    can miss the global scale, so the mean of 16 gets a looser bound)."""
    lrn = LearnedReconstructor(ctx)
    errs = []
    for s in range(3_000_500, 3_000_516):
        I = lrn.p.render(s)
        y = lrn.p.plain(I).clone()
        out, _, _ = lrn(I, s, check=False)
        errs.append(rel_l2(out, y))
    assert np.median(errs) < 0.02 and np.mean(errs) < 0.03, errs


@pytest.mark.parametrize("res,ties,pure", [((16, 16, 16), False, True), ((128, 128, 40), False, True),
                                           ((128, 128, 40), True, True), ((16, 16, 16), False, False)])
def test_calibration_matches_cpu_selection(ctx, res, ties, pure):
    """fbp_learned_calibrate against a CPU restatement: least squares of the
    reference on the volume (FP64), the k lowest and k highest voxels by
    (value, index) order (np.lexsort), the exact LoG there from the FP64
    oracle backprojection and LoG filter, and the affine map.  The volume is
    noise independent of the scene, so the range depends on exactly which
    voxels are selected; 128 x 128 x 40 = 655360 voxels makes every block of
    the GPU's selection pass sort two 512-voxel chunks, and the tied case
    (16 distinct values) checks the index order of ties; pure=False (first
    and last bounces kept, t0 = 0.05 m) takes the IEEE chain for the exact
    neighbourhoods instead of the threshold row."""
    import dataclasses
    import oracle
    from fasterbackprojection_amd.api import make_info
    from fasterbackprojection_amd.scene import C_LIGHT, confocal_geometry

    k = 16
    g = confocal_geometry(4, 1.0, 96, C_LIGHT * 128e-12, (-0.4, 0.3, -0.4), (0.4, 0.9, 0.4), res)
    if not pure:
        g = dataclasses.replace(g, discard_first_last_bounces=0, time_offset=0.05,
                                laser_position=(0.1, -0.8, 0.2), sensor_position=(-0.2, -0.7, 0.1))
    info = make_info(g)
    rng = np.random.default_rng(71)
    I = (10.0 ** rng.uniform(-2, 1, size=(len(g.laser_targets), 96))).astype(np.float32)
    bp = oracle.confocal_voxel(info, I, g.laser_targets)
    w = np.empty(125, np.float32)
    ctx.lib.fbp_log_kernel_host(5, 1.0, w.ctypes.data)
    L = oracle.log_filter(bp, 5, 1.0, weights=w.astype(np.float64).reshape(5, 5, 5)).ravel()
    if ties:
        v = (np.floor(rng.random(L.size) * 16) / 16).astype(np.float32)
    else:
        v = rng.random(L.size).astype(np.float32)
    r = (L + 0.1 * np.std(L) * rng.standard_normal(L.size)).astype(np.float32)
    vd, rd = v.astype(np.float64), r.astype(np.float64)
    n = float(v.size)
    S = [n, vd.sum(), (vd * vd).sum(), rd.sum(), (vd * rd).sum()]
    det = S[0] * S[2] - S[1] ** 2
    a = (S[0] * S[4] - S[1] * S[3]) / det
    b = (S[3] - a * S[1]) / S[0]
    idx = np.arange(v.size)
    lows = np.lexsort((idx, v))[:k]
    highs = np.lexsort((idx, -v))[:k]
    mn, mx = L[lows].min(), L[highs].max()
    c0, c1 = np.float32(a / (mx - mn)), np.float32((b - mn) / (mx - mn))
    want = (np.float64(c0) * vd + np.float64(c1)).astype(np.float32)
    shape = (res[2], res[1], res[0])
    d = torch.from_numpy(v.reshape(shape)).cuda()
    ctx.learned_calibrate(info, torch.from_numpy(I).cuda(), torch.from_numpy(g.laser_targets).cuda(),
                          d, torch.from_numpy(r.reshape(shape)).cuda(), k)
    got = d.cpu().numpy().ravel()
    scale = np.abs(want).max()
    assert np.max(np.abs(got - want)) <= 2e-5 * scale, np.max(np.abs(got - want)) / scale


@need_weights
def test_one_call_equals_the_steps(ctx):
    """fbp_learned_reconstruct gives what its steps give through the
    separate entry points: the seeded MIS draw, LoG, min-max normalisation,
    fbp_refiner_forward, fbp_learned_calibrate against the draw's LoG."""
    from fasterbackprojection_amd.api import MIS_TILED
    from fasterbackprojection_amd.learned import CALIB_EXTREMES, SAMPLE_FRACTION
    lrn = LearnedReconstructor(ctx)
    p = lrn.p
    for s in (3_000_600, 3_000_601):
        I = p.render(s)
        one = p.learned(lrn.net, I, s).clone()
        mis = ctx.backproject_confocal_mis(p.info, I, p.laser, None, p.L // SAMPLE_FRACTION,
                                           mode=MIS_TILED, seed=10_000_019 + s)
        ref = ctx.log(mis, 5, 1.0)
        x = ref.clone()
        ctx.normalize(x)
        y = lrn.net.forward(x)
        ctx.learned_calibrate(p.info, I, p.laser, y, ref, CALIB_EXTREMES, 5, 1.0)
        torch.cuda.synchronize()
        assert torch.equal(one, y), float((one - y).abs().max())
