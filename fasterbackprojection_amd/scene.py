"""Seeded synthetic NLOS captures (the reference ships no data).

Coordinate frame of the reference kernels: the relay wall is the x-z plane at
y = 0 with normal +y; depth is +y.  Scan grids are cell-centred and their
coordinates are computed with the same float expression as voxel centres,
(min + float(i) * step) + 0.5f * step, so a voxel grid with the wall's
lateral pitch lines up with the scan grid bit for bit.

Transients are rendered as in the benchmark descriptions: every surface
sample at distance r from a scan point adds albedo / r^4 at bin
round(2 r / delta) (confocal) or round((|l-p| + |p-s|) / delta) (exhaustive),
then Poisson noise at a given peak count.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .api import Geometry

C_LIGHT = 3.0e8  # m/s, as in python/lct.py:26


def grid_coord(lo: float, step: float, n: int) -> np.ndarray:
    """(lo + float(i) * step) + 0.5f * step in float32, i = 0..n-1."""
    lo = np.float32(lo)
    step = np.float32(step)
    i = np.arange(n, dtype=np.float32)
    return ((lo + i * step) + np.float32(0.5) * step).astype(np.float32)


def wall_grid(n: int, size: float, nz: int | None = None) -> np.ndarray:
    """Scan points of an n x nz cell-centred grid on a size x size wall,
    index l = iz * n + ix, position (x, 0, z)."""
    nz = nz or n
    xs = grid_coord(-size / 2, size / n, n)
    zs = grid_coord(-size / 2, size / nz, nz)
    pts = np.zeros((nz, n, 3), np.float32)
    pts[:, :, 0] = xs[None, :]
    pts[:, :, 2] = zs[:, None]
    return pts.reshape(-1, 3)


# ------------------------------------------------------------------ scenes
@dataclass
class Scene:
    points: np.ndarray   # (P, 3) float64 surface samples
    albedo: np.ndarray   # (P,) float64
    kind: str = ""
    params: dict = field(default_factory=dict)


def random_points(rng: np.random.Generator, n: int, xz=(-0.4, 0.4), depth=(0.4, 1.2),
                  albedo=(0.2, 1.0)) -> Scene:
    p = np.empty((n, 3))
    p[:, 0] = rng.uniform(*xz, n)
    p[:, 1] = rng.uniform(*depth, n)
    p[:, 2] = rng.uniform(*xz, n)
    return Scene(p, rng.uniform(*albedo, n), "points")


def _random_rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def random_quads(rng: np.random.Generator, n_quads: int, samples: int, side=(0.2, 0.6),
                 depth=(0.5, 1.5), xz=(-0.5, 0.5), albedo=(0.5, 1.0)) -> Scene:
    """Planar square quads with random orientation, sampled uniformly."""
    per = np.full(n_quads, samples // n_quads)
    per[: samples - per.sum()] += 1
    pts, alb = [], []
    for q in range(n_quads):
        s = rng.uniform(*side)
        c = np.array([rng.uniform(*xz), rng.uniform(*depth), rng.uniform(*xz)])
        R = _random_rotation(rng)
        uv = rng.uniform(-s / 2, s / 2, size=(per[q], 2))
        local = np.stack([uv[:, 0], np.zeros(per[q]), uv[:, 1]], 1)
        pts.append(c + local @ R.T)
        alb.append(np.full(per[q], rng.uniform(*albedo)))
    return Scene(np.concatenate(pts), np.concatenate(alb), "quads")


def random_spheres(rng: np.random.Generator, n_spheres: int, samples: int, radius=(0.05, 0.2),
                   depth=(0.6, 1.4), xz=(-0.5, 0.5), albedo=(0.5, 1.0)) -> Scene:
    per = np.full(n_spheres, samples // n_spheres)
    per[: samples - per.sum()] += 1
    pts, alb = [], []
    for k in range(n_spheres):
        r = rng.uniform(*radius)
        c = np.array([rng.uniform(*xz), rng.uniform(*depth), rng.uniform(*xz)])
        d = rng.normal(size=(per[k], 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        pts.append(c + r * d)
        alb.append(np.full(per[k], rng.uniform(*albedo)))
    return Scene(np.concatenate(pts), np.concatenate(alb), "spheres")


def mixed_scene(rng: np.random.Generator, samples: int = 4000) -> Scene:
    """1-6 random quads and spheres (training scenes of the learned form)."""
    n = int(rng.integers(1, 7))
    parts = []
    for _ in range(n):
        if rng.random() < 0.5:
            parts.append(random_quads(rng, 1, samples // n, side=(0.1, 0.4), depth=(0.3, 0.9),
                                      xz=(-0.35, 0.35)))
        else:
            parts.append(random_spheres(rng, 1, samples // n, radius=(0.05, 0.15),
                                        depth=(0.3, 0.9), xz=(-0.35, 0.35)))
    return Scene(np.concatenate([p.points for p in parts]),
                 np.concatenate([p.albedo for p in parts]), "mixed")


# --------------------------------------------------------------- rendering
def render_confocal(scene: Scene, scan: np.ndarray, T: int, delta: float,
                    chunk: int = 2048) -> np.ndarray:
    """float32 [L][T]: sum of albedo / r^4 at bin round(2 r / delta), FP64."""
    L = len(scan)
    out = np.zeros((L, T), np.float64)
    P = scene.points
    for a in range(0, L, chunk):
        s = scan[a:a + chunk].astype(np.float64)
        d = P[None, :, :] - s[:, None, :]
        r2 = np.einsum("lpk,lpk->lp", d, d)
        r = np.sqrt(r2)
        b = np.rint(2.0 * r / delta).astype(np.int64)
        w = scene.albedo[None, :] / (r2 * r2)
        ok = (b >= 0) & (b < T)
        rows = np.broadcast_to(np.arange(len(s))[:, None], b.shape)
        np.add.at(out[a:a + chunk], (rows[ok], b[ok]), w[ok])
    return out.astype(np.float32)


def render_exhaustive(scene: Scene, lasers: np.ndarray, sensors: np.ndarray, T: int,
                      delta: float) -> np.ndarray:
    """float32 [L][S][T]: albedo / (|l-p|^2 |p-s|^2) at bin round((|l-p| + |p-s|) / delta)."""
    L, S = len(lasers), len(sensors)
    out = np.zeros((L, S, T), np.float64)
    P = scene.points
    ds = np.linalg.norm(P[None, :, :] - sensors[:, None, :].astype(np.float64), axis=2)  # S,P
    for li in range(L):
        dl = np.linalg.norm(P - lasers[li].astype(np.float64), axis=1)  # P
        tot = dl[None, :] + ds
        b = np.rint(tot / delta).astype(np.int64)
        w = scene.albedo[None, :] / (dl[None, :] ** 2 * ds ** 2)
        ok = (b >= 0) & (b < T)
        rows = np.broadcast_to(np.arange(S)[:, None], b.shape)
        np.add.at(out[li], (rows[ok], b[ok]), w[ok])
    return out.astype(np.float32)


def poisson(transient: np.ndarray, peak: float, rng: np.random.Generator) -> np.ndarray:
    """Scales to the given peak count and draws Poisson counts (float32)."""
    m = float(transient.max())
    if m <= 0:
        return transient.copy()
    lam = transient.astype(np.float64) * (peak / m)
    return rng.poisson(lam).astype(np.float32)


# ----------------------------------------------------------------- configs
@dataclass
class Capture:
    geometry: Geometry
    transient: np.ndarray      # float32, [L][T] or [L][S][T]
    scene: Scene | None = None
    name: str = ""


def confocal_geometry(n_scan: int, wall: float, T: int, delta: float, hidden_min, hidden_max,
                      res) -> Geometry:
    scan = wall_grid(n_scan, wall)
    return Geometry(laser_targets=scan, sensor_targets=scan, num_time_bins=T, time_step=delta,
                    hidden_min=tuple(hidden_min), hidden_max=tuple(hidden_max),
                    voxel_resolution=tuple(res), confocal=True, wall_size=(wall, wall))


def small_confocal(seed: int = 0, n_scan: int = 32, T: int = 256, res: int = 32,
                   n_points: int = 200, peak: float = 1000.0, noise: bool = True) -> Capture:
    """Plain confocal method, small: 32x32 scan on a 1.0 m wall, 256 bins of
    32 ps (c*dt = 9.6 mm), 200 Lambertian points, Poisson at peak 1000,
    32^3 volume over [-0.4,0.4] x [0.4,1.2] x [-0.4,0.4]."""
    rng = np.random.default_rng(seed)
    delta = C_LIGHT * 32e-12
    g = confocal_geometry(n_scan, 1.0, T, delta, (-0.4, 0.4, -0.4), (0.4, 1.2, 0.4), (res,) * 3)
    sc = random_points(rng, n_points)
    tr = render_confocal(sc, g.laser_targets, T, delta)
    if noise:
        tr = poisson(tr, peak, rng)
    return Capture(g, tr, sc, "confocal-small")


def full_confocal_geometry(n_scan: int = 128, T: int = 1024, res: int = 128) -> Geometry:
    """128x128 scan on a 2.0 m wall, 1024 bins of 16 ps (4.8 mm); hidden volume
    [-1,1] x [0.25,1.75] x [-1,1] so the voxel pitch in x and z equals the
    scan pitch."""
    delta = C_LIGHT * 16e-12
    return confocal_geometry(n_scan, 2.0, T, delta, (-1.0, 0.25, -1.0), (1.0, 1.75, 1.0),
                             (res,) * 3)


def exhaustive_geometry(n_lasers: int = 16, n_sensors: int = 64, T: int = 1024,
                        res: int = 64) -> Geometry:
    """Exhaustive (non-confocal) capture: n_lasers^2 laser targets and
    n_sensors^2 sensor targets on a 2.0 m wall, 1024 bins of 16 ps
    (4.8 mm); hidden volume [-0.8,0.8] x [0.4,1.6] x [-0.8,0.8] holds the
    random spheres of random_spheres (depth 0.6-1.4, x/z within 0.7 m)."""
    delta = C_LIGHT * 16e-12
    return Geometry(laser_targets=wall_grid(n_lasers, 2.0), sensor_targets=wall_grid(n_sensors, 2.0),
                    num_time_bins=T, time_step=delta, hidden_min=(-0.8, 0.4, -0.8),
                    hidden_max=(0.8, 1.6, 0.8), voxel_resolution=(res,) * 3, confocal=False,
                    wall_size=(2.0, 2.0))
