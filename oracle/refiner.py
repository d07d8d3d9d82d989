"""FP64 NumPy restatement of the learned refiner (csrc/refiner.hip).

TEST INFRASTRUCTURE ONLY.  The network has no reference counterpart; this
file is the checker for the GPU forward pass (optionally emulating the BF16
rounding of weights and activations at the points where the GPU stores
BF16) and, through central finite differences of its FP64 loss, for the GPU
gradients.  Layer table, parameter layout and loss follow the comments at
the top of csrc/refiner.hip.
"""
from __future__ import annotations

import numpy as np

LAYERS = [(1, 16, 3), (16, 32, 3), (32, 64, 3), (64, 32, 3), (32, 16, 1), (16, 1, 3)]


def param_count() -> int:
    return sum(k ** 3 * ci * co + co for ci, co, k in LAYERS)


def unpack(p):
    out, o = [], 0
    for ci, co, k in LAYERS:
        n = k ** 3 * ci * co
        w = np.asarray(p[o:o + n], np.float64).reshape(k ** 3, co, ci)
        o += n
        b = np.asarray(p[o:o + co], np.float64)
        o += co
        out.append((w, b, k))
    return out


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def conv(x, w, b, k):
    """x: (S, S, S, Cin); w: (k^3, Cout, Cin); zero padding k // 2."""
    r = k // 2
    D, H, W, _ = x.shape
    xp = np.pad(x, ((r, r), (r, r), (r, r), (0, 0)))
    out = np.broadcast_to(b, (D, H, W, len(b))).copy()
    t = 0
    for dz in range(k):
        for dy in range(k):
            for dx in range(k):
                out += np.einsum("zyxc,oc->zyxo", xp[dz:dz + D, dy:dy + H, dx:dx + W], w[t])
                t += 1
    return out


def pool(x):
    D, H, W, C = x.shape
    return x.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(1, 3, 5))


def up(x):
    return x.repeat(2, 0).repeat(2, 1).repeat(2, 2)


def forward(p, vol, emulate_bf16=False):
    """vol: (S, S, S) raw input; returns (S, S, S).  Input normalised by its
    max |value| as on the GPU."""
    q = (lambda a: bf16(a)) if emulate_bf16 else (lambda a: a)
    L = unpack(p)
    x = np.asarray(vol, np.float64)
    m = np.abs(x).max()
    x = (x / m if m > 0 else x * 0)[..., None]
    if emulate_bf16:
        x = np.asarray(x, np.float32).astype(np.float64)
    relu = lambda a: np.maximum(a, 0)
    mw = lambda i: (q(L[i][0]), L[i][1], L[i][2])   # MFMA layers use BF16 weights
    e1 = q(relu(conv(x, *L[0])))
    p1 = q(pool(e1))
    e2 = q(relu(conv(p1, *mw(1))))
    p2 = q(pool(e2))
    e3 = q(relu(conv(p2, *mw(2))))
    u2 = up(e3)
    d2 = q(relu(conv(u2, *mw(3))) + e2)
    u1 = up(d2)
    d1 = q(relu(conv(u1, *mw(4))) + e1)
    return conv(d1, *L[5])[..., 0] + x[..., 0]


def box(a, r=2):
    """Zero-padded (2r+1)^3 box mean (normalised by the full window)."""
    for ax in range(3):
        pad = [(0, 0)] * a.ndim
        pad[ax] = (r, r)
        ap = np.pad(a, pad)
        s = sum(np.take(ap, range(i, i + a.shape[ax]), axis=ax) for i in range(2 * r + 1))
        a = s / (2 * r + 1)
    return a


def ssim_mean(x, t):
    C1, C2 = 1e-4, 9e-4
    mx, my = box(x), box(t)
    vx = box(x * x) - mx * mx
    vy = box(t * t) - my * my
    cxy = box(x * t) - mx * my
    s = ((2 * mx * my + C1) * (2 * cxy + C2)) / ((mx * mx + my * my + C1) * (vx + vy + C2))
    return s.mean()


def loss(p, vol, target, ssim_weight=1.0):
    y = forward(p, vol)
    t = np.asarray(target, np.float64)
    return np.abs(y - t).mean() + ssim_weight * (1 - ssim_mean(y, t))


def fd_grad(p, vol, target, idx, h=1e-3, ssim_weight=1.0):
    """Central finite differences of the FP64 loss at parameter indices idx."""
    p = np.asarray(p, np.float64)
    g = []
    for i in idx:
        a = p.copy()
        a[i] += h
        b = p.copy()
        b[i] -= h
        g.append((loss(a, vol, target, ssim_weight) - loss(b, vol, target, ssim_weight)) / (2 * h))
    return np.array(g)
