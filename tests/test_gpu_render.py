"""GPU scene renderers (synthetic transients for tests, benchmarks and the
learned form's training data) against NumPy restatements.

The GPU computes distances in FP32 (correctly rounded sqrt and division,
no contraction).  Two checks per renderer: against a NumPy emulation of
that FP32 arithmetic (same bins, weights in FP64: rel-L2 1e-6), and against
the FP64 renderer of fasterbackprojection_amd.scene, where samples within an
FP32 rounding of a bin boundary may land one bin over (energy per histogram
to 1e-5, histograms after a +-1 bin box blur to rel-L2 1e-2)."""
import numpy as np
import pytest
import torch

from fasterbackprojection_amd.scene import (C_LIGHT, random_points, random_spheres,
                                            render_confocal, render_exhaustive, wall_grid)

pytestmark = pytest.mark.gpu
F = np.float32


def blur1(a):
    p = np.pad(a.astype(np.float64), [(0, 0)] * (a.ndim - 1) + [(1, 1)])
    return p[..., :-2] + p[..., 1:-1] + p[..., 2:]


def check(gpu, ref64, emu):
    gpu = gpu.detach().cpu().numpy().astype(np.float64)
    assert np.linalg.norm(gpu - emu) / np.linalg.norm(emu) < 1e-6
    ref = ref64.astype(np.float64)
    np.testing.assert_allclose(gpu.sum(-1), ref.sum(-1), rtol=1e-5, atol=1e-30)
    bg, br = blur1(gpu), blur1(ref)
    assert np.linalg.norm(bg - br) / np.linalg.norm(br) < 1e-2


def _r2(p, q):
    d = (p - q).astype(F)
    return ((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]).astype(F) + d[:, 2] * d[:, 2]).astype(F)


def emulate_confocal(P, A, scan, T, delta):
    out = np.zeros((len(scan), T))
    for l in range(len(scan)):
        r2 = _r2(P, scan[l])
        b = np.rint(((F(2) * np.sqrt(r2)).astype(F) / F(delta)).astype(F)).astype(np.int64)
        ok = (b >= 0) & (b < T) & (r2 > 0)
        np.add.at(out[l], b[ok], A[ok].astype(np.float64) / r2[ok].astype(np.float64) ** 2)
    return out


def emulate_exhaustive(P, A, lasers, sensors, T, delta):
    out = np.zeros((len(lasers), len(sensors), T))
    for l in range(len(lasers)):
        ra2 = _r2(P, lasers[l])
        for s in range(len(sensors)):
            rb2 = _r2(P, sensors[s])
            b = np.rint(((np.sqrt(ra2) + np.sqrt(rb2)).astype(F) / F(delta)).astype(F))
            b = b.astype(np.int64)
            ok = (b >= 0) & (b < T) & (ra2 > 0) & (rb2 > 0)
            w = A[ok].astype(np.float64) / (ra2[ok].astype(np.float64) * rb2[ok])
            np.add.at(out[l, s], b[ok], w)
    return out


def test_render_confocal(ctx):
    rng = np.random.default_rng(11)
    sc = random_points(rng, 300)
    scan = wall_grid(12, 1.0)
    delta = C_LIGHT * 32e-12
    ref = render_confocal(sc, scan, 256, delta)
    emu = emulate_confocal(sc.points.astype(F), sc.albedo.astype(F), scan, 256, delta)
    out = ctx.render_confocal(torch.from_numpy(scan).cuda(),
                              torch.from_numpy(sc.points.astype(np.float32)).cuda(),
                              torch.from_numpy(sc.albedo.astype(np.float32)).cuda(), 256, delta)
    torch.cuda.synchronize()
    assert out.shape == (144, 256)
    check(out, ref, emu)


def test_render_exhaustive(ctx):
    rng = np.random.default_rng(12)
    sc = random_spheres(rng, 3, 600)
    lasers = wall_grid(4, 2.0)
    sensors = wall_grid(8, 2.0)
    delta = C_LIGHT * 16e-12
    ref = render_exhaustive(sc, lasers, sensors, 1024, delta)
    emu = emulate_exhaustive(sc.points.astype(F), sc.albedo.astype(F), lasers, sensors, 1024,
                             delta)
    out = ctx.render_exhaustive(torch.from_numpy(lasers).cuda(), torch.from_numpy(sensors).cuda(),
                                torch.from_numpy(sc.points.astype(np.float32)).cuda(),
                                torch.from_numpy(sc.albedo.astype(np.float32)).cuda(), 1024, delta)
    torch.cuda.synchronize()
    assert out.shape == (16, 64, 1024)
    check(out, ref, emu)


# ------------------------------------------------------------ residual check
def _centres(info):
    """float32 world coordinates of voxel centres, volume index order
    [z (depth, world y)][vy (world z)][vx (world x)]."""
    f = np.float32
    res = (info.voxel_resolution.x, info.voxel_resolution.y, info.voxel_resolution.z)
    hmin = (info.hidden_volume_min.x, info.hidden_volume_min.y, info.hidden_volume_min.z)
    vs = (info.hidden_volume_voxel_size.x, info.hidden_volume_voxel_size.y,
          info.hidden_volume_voxel_size.z)
    c = [(f(hmin[a]) + np.arange(n, dtype=f) * f(vs[a])) + f(0.5) * f(vs[a])
         for a, n in zip(range(3), (res[0], res[2], res[1]))]
    cx, cy, cz = c      # world x <- vx, world y <- z, world z <- vy
    Z, Y, X = res[2], res[1], res[0]
    wx = np.broadcast_to(cx[None, None, :], (Z, Y, X))
    wy = np.broadcast_to(cy[:, None, None], (Z, Y, X))
    wz = np.broadcast_to(cz[None, :, None], (Z, Y, X))
    return wx, wy, wz


def _residual_numpy(info, vol, I, scan, rows, frac, blur):
    f = np.float32
    T = info.num_time_bins
    delta = f(info.time_step)
    thr = f(vol.min() + f(frac) * f(vol.max() - vol.min()))
    wx, wy, wz = _centres(info)
    a = (vol - thr).astype(f)
    sel = a > 0
    sums = np.zeros(5)
    for l in rows:
        dx = (wx[sel] - f(scan[l, 0])).astype(f)
        dy = (wy[sel] - f(scan[l, 1])).astype(f)
        dz = (wz[sel] - f(scan[l, 2])).astype(f)
        r2 = ((dx * dx + dy * dy).astype(f) + dz * dz).astype(f)
        b = np.rint((f(2) * np.sqrt(r2)).astype(f) / delta).astype(np.int64)
        ok = (b >= 0) & (b < T) & (r2 > 0)
        w = a[sel][ok].astype(np.float64) / (r2[ok].astype(np.float64) ** 2)
        fix = np.floor(w * 2.0 ** 32 + 0.5)
        h = np.zeros(T)
        np.add.at(h, b[ok], fix)
        h /= 2.0 ** 32
        k = np.ones(2 * blur + 1)
        p = np.convolve(h, k, mode="same").astype(np.float32).astype(np.float64)
        m = np.convolve(I[l].astype(np.float64), k, mode="same").astype(np.float32).astype(np.float64)
        sums += [p.sum(), m.sum(), (p * m).sum(), (p * p).sum(), (m * m).sum()]
    return sums


def _residual_rows(info, vol, I, scan, rows, frac, blur):
    T = info.num_time_bins
    cov = vp = vm = 0.0
    for l in rows:
        s = _residual_numpy(info, vol, I, scan, [l], frac, blur)
        cov += s[2] - s[0] * s[1] / T
        vp += s[3] - s[0] ** 2 / T
        vm += s[4] - s[1] ** 2 / T
    return 1.0 - cov / np.sqrt(vp * vm)


def _small_geometry():
    from fasterbackprojection_amd.scene import confocal_geometry
    g = confocal_geometry(8, 1.0, 160, C_LIGHT * 32e-12, (-0.4, 0.3, -0.4), (0.4, 1.1, 0.4),
                          (12, 10, 14))
    from fasterbackprojection_amd.api import make_info
    return g, make_info(g)


def test_residual_matches_numpy(ctx):
    g, info = _small_geometry()
    rng = np.random.default_rng(21)
    vol = rng.random((14, 10, 12)).astype(np.float32)
    I = rng.poisson(3.0, (64, 160)).astype(np.float32)
    rows = np.array([0, 5, 17, 40, 63], np.uint32)
    r = ctx.volume_residual(info, torch.from_numpy(vol).cuda(), torch.from_numpy(I).cuda(),
                            torch.from_numpy(g.laser_targets).cuda(),
                            torch.from_numpy(rows.astype(np.int32)).cuda(), 0.6, 3)
    exp = _residual_rows(info, vol, I, g.laser_targets, rows, 0.6, 3)
    assert abs(r - exp) < 1e-6, (r, exp)


def test_residual_of_self_rendered_transient_is_zero(ctx):
    """Known answer: a transient rendered (render_confocal) from exactly the
    voxels the check renders, with albedo = excess over the threshold,
    correlates perfectly with the check's own rendering."""
    g, info = _small_geometry()
    rng = np.random.default_rng(22)
    vol = np.zeros((14, 10, 12), np.float32)
    idx = rng.choice(vol.size, 12, replace=False)
    vol.flat[idx] = rng.uniform(0.5, 1.0, 12).astype(np.float32)
    vol.flat[idx[0]] = 1.0
    thr = np.float32(0.0 + np.float32(0.3) * np.float32(1.0))
    wx, wy, wz = _centres(info)
    sel = vol > thr
    pts = np.stack([wx[sel], wy[sel], wz[sel]], 1).astype(np.float32)
    alb = (vol[sel] - thr).astype(np.float32)
    scan = torch.from_numpy(g.laser_targets).cuda()
    I = ctx.render_confocal(scan, torch.from_numpy(pts).cuda(), torch.from_numpy(alb).cuda(),
                            info.num_time_bins, info.time_step)
    rows = torch.arange(0, 64, 3, dtype=torch.int32, device="cuda")
    r = ctx.volume_residual(info, torch.from_numpy(vol).cuda(), I, scan, rows, 0.3, 2)
    assert abs(r) < 1e-6, r


def test_residual_of_constant_rows_is_one(ctx):
    g, info = _small_geometry()
    vol = np.random.default_rng(23).random((14, 10, 12)).astype(np.float32)
    I = torch.full((64, 160), 4.0, device="cuda")
    rows = torch.arange(64, dtype=torch.int32, device="cuda")
    r = ctx.volume_residual(info, torch.from_numpy(vol).cuda(), I,
                            torch.from_numpy(g.laser_targets).cuda(), rows, 0.6, 0)
    assert r == 1.0
