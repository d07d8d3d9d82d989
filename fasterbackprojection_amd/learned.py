"""Learned form: MIS-sampled backprojection -> 3D U-Net refiner -> the plain
(all scan points) backprojection + LoG volume, with a forward-model
residual check that falls back to the plain path.

Workload: 64 x 64 confocal scan on a 1.0 m wall, 512 bins of 16 ps; 64^3
volume over [-0.5, 0.5] x [0.2, 1.0] x [-0.5, 0.5].  Scenes: 1-6 random
quads and spheres (scene.mixed_scene), 1/r^4 falloff, Poisson noise at a
peak of 1000 counts.  Input: unbiased MIS estimate with L/16 samples per
voxel, shared by the voxels of each 16^3 tile (FBP_MIS_TILED) (1/16 of the scan-point x bin pairs).  Target: plain
backprojection, LoG (sigma 1, k 5), min-max normalised.  The input goes
through the same LoG + normalisation, and the refiner predicts a correction
of it (its last layer adds the input), so the un-refined MIS path is the
starting point the training improves on.

The refined volume matches the target up to a global affine map (the
normalisation of the noisy input is set by its extremes, which the
convolutions cannot see); fbp_learned_calibrate fixes that map: the draw's
own LoG in physical units regressed on the refined volume over all voxels
gives the LoG's scale and offset (its noise is about 5 % of the signal, so
the bias it puts into the fit is of order 0.3 %), and exact LoG values at
the refined volume's CALIB_EXTREMES lowest and highest voxels (exact
backprojections of their 5^3 neighbourhoods) give the true range.

The learned call is one C-ABI call, fbp_learned_reconstruct (Refiner.
reconstruct): seeded MIS draw, LoG with its min / max, the refiner's five
fused kernels, calibration.  Training data use torch's RNG for the noise
buffer (the same sampler fed from a buffer); every step runs in libfbp.so
kernels, torch holds device memory and draws the Poisson counts.

    python -m fasterbackprojection_amd.learned train --out DIR [--scenes N --steps K]
    python -m fasterbackprojection_amd.learned calibrate [--weights FILE]
"""
from __future__ import annotations

import argparse
import json
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from .api import MIS_TILED, Context, Geometry, Refiner, make_info
from .scene import C_LIGHT, confocal_geometry, mixed_scene

WEIGHTS = Path(__file__).resolve().parent / "weights"
SIZE = 64
SCAN = 64
BINS = 512
SAMPLE_FRACTION = 16          # MIS samples per voxel = L / SAMPLE_FRACTION
CALIB_EXTREMES = 16           # exact LoG at the k lowest and k highest voxels


def geometry() -> Geometry:
    return confocal_geometry(SCAN, 1.0, BINS, C_LIGHT * 16e-12, (-0.5, 0.2, -0.5),
                             (0.5, 1.0, 0.5), (SIZE,) * 3)


@dataclass
class Pipeline:
    """Device-resident pieces shared by the plain and learned paths."""
    ctx: Context
    g: Geometry

    def __post_init__(self):
        self.info = make_info(self.g)
        self.dev = torch.device("cuda", self.ctx.device)
        self.laser = torch.from_numpy(self.g.laser_targets).to(self.dev)
        self.L = len(self.g.laser_targets)
        self.vol = torch.empty((SIZE,) * 3, dtype=torch.float32, device=self.dev)
        self.filt = torch.empty_like(self.vol)
        self.mis = torch.empty_like(self.vol)
        self.ref = torch.empty_like(self.vol)
        # residual check: every 4th scan point in x and z
        n = SCAN
        idx = np.array([iz * n + ix for iz in range(0, n, 4) for ix in range(0, n, 4)])
        self.sub_idx = torch.from_numpy(idx.astype(np.int32)).to(self.dev)

    def render(self, seed: int, peak: float = 1000.0) -> torch.Tensor:
        rng = np.random.default_rng(seed)
        sc = mixed_scene(rng, 4000)
        pts = torch.from_numpy(sc.points.astype(np.float32)).to(self.dev)
        alb = torch.from_numpy(sc.albedo.astype(np.float32)).to(self.dev)
        clean = self.ctx.render_confocal(self.laser, pts, alb, BINS, self.g.time_step)
        gen = torch.Generator(device=self.dev)
        gen.manual_seed(seed)
        return torch.poisson(clean * (peak / clean.max()), generator=gen).contiguous()

    def plain(self, I: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """Backprojection over all scan points + LoG + min-max normalisation."""
        self.ctx.backproject_confocal(self.info, I, self.laser, out=self.vol)
        out = self.filt if out is None else out
        return self.ctx.log_normalize(self.vol, 5, 1.0, out=out)

    def mis_input(self, I: torch.Tensor, seed: int, out: torch.Tensor | None = None) -> torch.Tensor:
        """Refiner input: unbiased MIS backprojection + LoG + min-max
        normalisation (the same filter chain as the plain path).  self.ref
        keeps the draw's LoG in physical units (calibrate's reference)."""
        gen = torch.Generator(device=self.dev)
        gen.manual_seed(10_000_019 + seed)
        noise = torch.rand(SIZE ** 3, generator=gen, device=self.dev)
        self.ctx.backproject_confocal_mis(self.info, I, self.laser, noise,
                                          self.L // SAMPLE_FRACTION, mode=MIS_TILED,
                                          out=self.mis)
        self.ctx.log(self.mis, 5, 1.0, out=self.ref)
        out = torch.empty_like(self.vol) if out is None else out
        out.copy_(self.ref)
        self.ctx.normalize(out)
        return out

    def calibrate(self, volume: torch.Tensor, I: torch.Tensor) -> torch.Tensor:
        """Global calibration of a refined volume (fbp_learned_calibrate), in
        place, against the LoG of the draw that made the refiner's input
        (self.ref, set by mis_input)."""
        return self.ctx.learned_calibrate(self.info, I, self.laser, volume, self.ref,
                                          CALIB_EXTREMES, 5, 1.0)

    def learned(self, net: Refiner, I: torch.Tensor, seed: int,
                out: torch.Tensor | None = None) -> torch.Tensor:
        """The learned form in one call (fbp_learned_reconstruct), MIS
        uniforms from seed."""
        return net.reconstruct(self.info, I, self.laser, self.L // SAMPLE_FRACTION,
                               10_000_019 + seed, out=out, extremes=CALIB_EXTREMES)

    def residual(self, volume: torch.Tensor, I: torch.Tensor) -> float:
        """Forward-model residual on every 16th scan point (fbp_volume_residual):
        the voxels below 30% of the volume's range (surfaces are the minima
        of the LoG-filtered, normalised volume) are rendered and correlated
        with the measured rows, both blurred over +-4 bins."""
        return self.ctx.volume_residual(self.info, volume, I, self.laser, self.sub_idx,
                                        threshold_fraction=0.3, blur=4, polarity=-1)


def rel_l2(a: torch.Tensor, b: torch.Tensor) -> float:
    return float((a - b).norm() / b.norm())


def build_dataset(p: Pipeline, seeds, log=print):
    X = torch.empty((len(seeds), SIZE, SIZE, SIZE), dtype=torch.float32, device=p.dev)
    Y = torch.empty_like(X)
    t = time.time()
    for i, s in enumerate(seeds):
        I = p.render(int(s))
        p.plain(I, out=Y[i])
        p.mis_input(I, int(s), out=X[i])
        if (i + 1) % 500 == 0:
            torch.cuda.synchronize()
            log(f"dataset {i + 1}/{len(seeds)} ({time.time() - t:.1f} s)")
    torch.cuda.synchronize()
    return X, Y


def train(out_dir: Path, scenes: int = 4000, held_out: int = 256, steps: int = 3000,
          batch: int = 4, lr: float = 1e-3, log=print) -> dict:
    ctx = Context(0)
    p = Pipeline(ctx, geometry())
    torch.manual_seed(0)
    Xtr, Ytr = build_dataset(p, range(scenes), log)
    Xte, Yte = build_dataset(p, range(1_000_000, 1_000_000 + held_out), log)
    net = Refiner(ctx, SIZE, seed=1)
    rng = np.random.default_rng(0)
    t0 = time.time()
    hist = []
    for step in range(steps):
        idx = torch.from_numpy(rng.choice(scenes, batch, replace=False)).to(p.dev)
        lr_t = lr * (0.5 * (1 + np.cos(np.pi * step / steps)))
        want = (step % 100 == 0) or step == steps - 1
        loss = net.train_step(Xtr[idx].contiguous(), Ytr[idx].contiguous(), lr_t, want_loss=want)
        if want:
            hist.append((step, loss))
            log(f"step {step} loss {loss:.4f} ({time.time() - t0:.0f} s)")
    torch.cuda.synchronize()
    params = net.save()
    # held-out evaluation and residual calibration
    train_seconds = time.time() - t0
    stats = evaluate(p, net, range(1_000_000, 1_000_000 + held_out), Xte, Yte)
    stats["residual_flat_noise"] = flat_noise_residual(p, net)
    out_dir.mkdir(parents=True, exist_ok=True)
    params.tofile(out_dir / "refiner_c5.f32")
    meta = {"size": SIZE, "scan": SCAN, "bins": BINS, "sample_fraction": SAMPLE_FRACTION,
            "scenes": scenes, "held_out": held_out, "steps": steps, "batch": batch, "lr": lr,
            "train_seconds": train_seconds, "loss_history": hist, **stats}
    (out_dir / "refiner_c5.json").write_text(json.dumps(meta, indent=1))
    net.close()
    ctx.close()
    return meta


def flat_noise_residual(p: Pipeline, net: Refiner) -> float:
    """Residual of an out-of-distribution capture (flat Poisson noise, mean 3
    counts); it should lie above the calibrated threshold."""
    gen = torch.Generator(device=p.dev)
    gen.manual_seed(5)
    I = torch.poisson(torch.full((SCAN * SCAN, BINS), 3.0, device=p.dev), generator=gen)
    I = I.contiguous()
    return p.residual(p.learned(net, I, 0), I)


def calibrate(weights: Path, held_out: int = 256, log=print) -> dict:
    """Re-evaluates committed weights on the held-out seeds and rewrites the
    residual statistics and threshold in their metadata file."""
    ctx = Context(0)
    p = Pipeline(ctx, geometry())
    net = Refiner(ctx, SIZE)
    net.load(np.fromfile(weights, np.float32))
    stats = evaluate(p, net, range(1_000_000, 1_000_000 + held_out))
    stats["residual_flat_noise"] = flat_noise_residual(p, net)
    meta_path = weights.with_suffix(".json")
    meta = json.loads(meta_path.read_text()) if meta_path.exists() else {}
    meta.update(stats)
    meta_path.write_text(json.dumps(meta, indent=1))
    log(f"calibrated: {stats}")
    net.close()
    ctx.close()
    return meta


def evaluate(p: Pipeline, net: Refiner, seeds, X=None, Y=None) -> dict:
    """Held-out rel-L2 of the learned output and of the un-learned method
    (MIS input + LoG + normalisation) against the plain target, and the
    residual distribution that calibrates the fallback threshold."""
    learned, baseline, res, res_plain = [], [], [], []
    out = torch.empty((SIZE,) * 3, dtype=torch.float32, device=p.dev)
    for i, s in enumerate(seeds):
        if X is None:
            I = p.render(int(s))
            y = p.plain(I).clone()
            x = p.mis_input(I, int(s))
        else:
            I = p.render(int(s))
            x, y = X[i], Y[i]
        p.learned(net, I, int(s), out=out)
        learned.append(rel_l2(out, y))
        baseline.append(rel_l2(x, y))
        res.append(p.residual(out, I))
        res_plain.append(p.residual(y, I))
    res = np.array(res)
    return {"rel_l2_learned_mean": float(np.mean(learned)),
            "rel_l2_learned_median": float(np.median(learned)),
            "rel_l2_without_refiner_mean": float(np.mean(baseline)),
            "residual_median": float(np.median(res)),
            "residual_p95": float(np.percentile(res, 95)),
            "residual_plain_median": float(np.median(res_plain)),
            "residual_threshold": float(min(0.999, np.percentile(res, 99) * 1.05))}


class LearnedReconstructor:
    """Learned path with fallback: MIS -> refiner; when the forward-model
    residual of the result exceeds the calibrated threshold, the plain path
    (all scan points + LoG) runs instead."""

    def __init__(self, ctx: Context, weights: Path | None = None, threshold: float | None = None):
        weights = Path(weights or WEIGHTS / "refiner_c5.f32")
        meta_path = weights.with_suffix(".json")
        meta = json.loads(meta_path.read_text()) if meta_path.exists() else {}
        self.threshold = threshold if threshold is not None else meta.get("residual_threshold", 0.5)
        self.p = Pipeline(ctx, geometry())
        self.net = Refiner(ctx, SIZE)
        self.net.load(np.fromfile(weights, np.float32))
        self.out = torch.empty((SIZE,) * 3, dtype=torch.float32, device=self.p.dev)
        self.fallbacks = 0

    def __call__(self, I: torch.Tensor, seed: int = 0, check: bool = True):
        self.p.learned(self.net, I, seed, out=self.out)
        if check:
            r = self.p.residual(self.out, I)
            if r > self.threshold:
                self.fallbacks += 1
                return self.p.plain(I), True, r
            return self.out, False, r
        return self.out, False, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "calibrate"])
    ap.add_argument("--weights", type=Path, default=WEIGHTS / "refiner_c5.f32")
    ap.add_argument("--out", type=Path, default=WEIGHTS)
    ap.add_argument("--scenes", type=int, default=4000)
    ap.add_argument("--held-out", type=int, default=256)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    a = ap.parse_args()
    if a.cmd == "calibrate":
        meta = calibrate(a.weights, a.held_out, log=lambda m: print(f"[learned] {m}", flush=True))
        print(json.dumps({k: v for k, v in meta.items() if k != "loss_history"}))
        return
    meta = train(a.out, a.scenes, a.held_out, a.steps, a.batch, a.lr,
                 log=lambda m: print(f"[learned] {m}", flush=True))
    print(json.dumps({k: v for k, v in meta.items() if k != "loss_history"}))


if __name__ == "__main__":
    main()
