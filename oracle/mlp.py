"""ORACLE -- test infrastructure only: the surrogate network restated in numpy float64.

The learned surrogate has no counterpart in the reference; its definition is the one the library documents
(include/agwn.h, antipodal_amd/csrc/agwn_mlp.hip): 192 Fourier features of the box-normalised position, 4 hidden
layers of 256 hard-SiLU units (z clamp(z/4 + 1/2, 0, 1)), a linear output, mean squared error, Adam. This module evaluates that definition in float64:

  forward / loss_grad(..., emulate=False)   exact float64 math (pinned by central finite differences and by closed
                                            forms in tests/test_oracle_mlp.py)
  forward / loss_grad(..., emulate=True)    the same with the library's BF16 rounding points applied (features,
                                            weights, activations, backward activations), for comparing GPU results
  adam_step                                 one Adam update with bias correction

Only tests/ import this module, as the checker.
"""
from __future__ import annotations

import numpy as np

FREQ = 32
IN = 3 * FREQ * 2
H = 256
LAYERS = 4
IN_DIM = (IN, H, H, H)


def layout() -> dict:
    """Offsets of W_l [fan_in x 256] (row-major), b_l, wo, bo in the flat parameter vector, and its length."""
    o = 0
    lay = {"w": [], "b": []}
    for l in range(LAYERS):
        lay["w"].append(o)
        o += IN_DIM[l] * H
        lay["b"].append(o)
        o += H
    lay["wo"] = o
    o += H
    lay["bo"] = o
    o += 1
    lay["total"] = o
    return lay


LAYOUT = layout()


def unpack(p: np.ndarray):
    p = np.asarray(p, dtype=np.float64).reshape(-1)
    Ws = [p[LAYOUT["w"][l]: LAYOUT["w"][l] + IN_DIM[l] * H].reshape(IN_DIM[l], H) for l in range(LAYERS)]
    bs = [p[LAYOUT["b"][l]: LAYOUT["b"][l] + H] for l in range(LAYERS)]
    return Ws, bs, p[LAYOUT["wo"]: LAYOUT["wo"] + H], float(p[LAYOUT["bo"]])


def pack(Ws, bs, wo, bo) -> np.ndarray:
    p = np.zeros(LAYOUT["total"])
    for l in range(LAYERS):
        p[LAYOUT["w"][l]: LAYOUT["w"][l] + IN_DIM[l] * H] = np.asarray(Ws[l]).reshape(-1)
        p[LAYOUT["b"][l]: LAYOUT["b"][l] + H] = bs[l]
    p[LAYOUT["wo"]: LAYOUT["wo"] + H] = wo
    p[LAYOUT["bo"]] = bo
    return p


def bf16(x) -> np.ndarray:
    """Round to bfloat16 (round to nearest, ties to even), returned as float64. Via float32: the float64 -> float32
    step can only double-round when the float64 value lies within 2^-24 relative of a bf16 tie, which the callers'
    values (sums of products) reach with negligible probability."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> np.uint64(16)) & np.uint64(1)
    r = ((u + np.uint64(0x7FFF) + lsb) >> np.uint64(16)) << np.uint64(16)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def frequencies(max_freq: float = 128.0) -> np.ndarray:
    """w_k = max_freq^(k/31), k = 0..31, rounded to float32."""
    return np.exp2(np.log2(float(max_freq)) * np.arange(FREQ) / (FREQ - 1)).astype(np.float32)


def features(pts, lo, hi, max_freq: float = 128.0) -> np.ndarray:
    """(n, 192): feature pair j = 32 a + k is (sin, cos)(pi w_k u_a), u_a = (p_a - c_a) / h_a. The argument w_k u_a is
    formed in float32 as the kernels do (c = (lo + hi) / 2, 1/h = 1 / ((hi - lo) / 2), all float32); sin and cos are
    evaluated in float64."""
    p = np.asarray(pts, dtype=np.float32).reshape(-1, 3)
    lo = np.asarray(lo, dtype=np.float32)
    hi = np.asarray(hi, dtype=np.float32)
    c = np.float32(0.5) * (lo + hi)
    inv_h = np.float32(1.0) / (np.float32(0.5) * (hi - lo))
    u = (p - c) * inv_h  # float32
    w = frequencies(max_freq)
    arg = (w[None, None, :] * u[:, :, None]).astype(np.float64)  # (n, 3, 32), exact float32 products
    out = np.empty((len(p), 3, FREQ, 2))
    out[..., 0] = np.sin(np.pi * arg)
    out[..., 1] = np.cos(np.pi * arg)
    return out.reshape(len(p), IN)


def hsilu(z):
    """Hard SiLU: z clamp(z/4 + 1/2, 0, 1)."""
    z = np.asarray(z, dtype=np.float64)
    return z * np.clip(0.25 * z + 0.5, 0.0, 1.0)


def hsilu_grad(z):
    """z/2 + 1/2 for |z| < 2, 1 for z >= 2, 0 for z <= -2 (the library's one-sided choices at the kinks)."""
    z = np.asarray(z, dtype=np.float64)
    return np.where(np.abs(z) < 2.0, 0.5 * z + 0.5, np.where(z >= 2.0, 1.0, 0.0))


def forward(p, pts, lo, hi, emulate: bool = False, max_freq: float = 128.0):
    """Predictions and the cache of the backward pass."""
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    Ws, bs, wo, bo = unpack(p)
    x = r(features(pts, lo, hi, max_freq))
    xs, zs = [x], []
    for l in range(LAYERS):
        z = x @ r(Ws[l]) + bs[l]
        x = r(hsilu(z))
        zs.append(z)
        xs.append(x)
    pred = x @ wo + bo
    return pred, (xs, zs)


def loss_grad(p, pts, y, lo, hi, emulate: bool = False, max_freq: float = 128.0):
    """(mean squared error, gradient in the flat layout, predictions)."""
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    Ws, bs, wo, bo = unpack(p)
    y = np.asarray(y, dtype=np.float64).reshape(-1)
    n = len(y)
    pred, (xs, zs) = forward(p, pts, lo, hi, emulate, max_freq)
    err = pred - y
    loss = float(np.mean(err * err))
    g = 2.0 * err / n
    gWs, gbs = [None] * LAYERS, [None] * LAYERS
    gwo = xs[LAYERS].T @ g
    gbo = float(g.sum())
    d = r(g[:, None] * wo[None, :] * hsilu_grad(zs[LAYERS - 1]))
    for l in range(LAYERS - 1, -1, -1):
        gWs[l] = xs[l].T @ d
        gbs[l] = d.sum(axis=0)
        if l > 0:
            d = r((d @ r(Ws[l]).T) * hsilu_grad(zs[l - 1]))
    return loss, pack(gWs, gbs, gwo, gbo), pred


def adam_step(p, m, v, g, t: int, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    """One Adam update (step number t >= 1) -> (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
