"""Pins the numpy restatement of the surrogate network (oracle/mlp.py), on the CPU.

* bf16 rounding: against torch's float32 -> bfloat16 conversion (round to nearest even), ties included;
* features: closed forms at the box centre and corners, the frequency ladder 1 .. 128;
* backward pass: central finite differences of the float64 loss, for coordinates in every parameter block;
* Adam: the first step moves every coordinate by lr * sign(g) (bias correction), a second step by the closed form;
* layout: the parameter count the library reports.
A dropped term, a transposed weight or a wrong activation derivative fails the finite-difference check.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp as om

LO = np.array([-1.5, -1.2, -0.4], dtype=np.float32)
HI = np.array([1.4, 1.3, 0.45], dtype=np.float32)


def init_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    Ws = [rng.uniform(-1, 1, (om.IN_DIM[l], om.H)) * np.sqrt(6.0 / om.IN_DIM[l]) * scale for l in range(om.LAYERS)]
    bs = [rng.normal(0, 0.1, om.H) for _ in range(om.LAYERS)]
    wo = rng.uniform(-1, 1, om.H) * 0.1
    return om.pack(Ws, bs, wo, 0.5)


def test_bf16_matches_torch_rne():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    x = np.concatenate([
        rng.normal(0, 1, 20000).astype(np.float32),
        rng.normal(0, 1e-3, 5000).astype(np.float32),
        # exact ties: bf16 value + half an ulp, both parities
        (np.arange(1, 2000, dtype=np.float32) * np.float32(2.0 ** -7) + np.float32(1.0) + np.float32(2.0 ** -8)),
        np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8), 0.0, -0.0], dtype=np.float32),
    ])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    np.testing.assert_array_equal(om.bf16(x), want)
    assert om.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0  # tie to even (down)
    assert om.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6  # tie to even (up)


def test_features_closed_forms():
    w = om.frequencies()
    assert w[0] == 1.0 and w[-1] == 128.0
    np.testing.assert_allclose(w[1:] / w[:-1], 2.0 ** (7 / 31), rtol=1e-6)
    c = 0.5 * (LO + HI)
    f = om.features(c[None, :], LO, HI).reshape(3, om.FREQ, 2)
    np.testing.assert_allclose(f[..., 0], 0.0, atol=1e-6)
    np.testing.assert_allclose(f[..., 1], 1.0, atol=1e-12)
    # at the low corner u = -1: sin(-pi w) and cos(pi w); w_0 = 1 gives (0, -1), w_31 = 128 gives (0, 1)
    f = om.features(LO[None, :], LO, HI).reshape(3, om.FREQ, 2)
    np.testing.assert_allclose(f[:, 0, :], [[0.0, -1.0]] * 3, atol=1e-6)
    np.testing.assert_allclose(f[:, -1, :], [[0.0, 1.0]] * 3, atol=1e-4)
    np.testing.assert_allclose(f[..., 0] ** 2 + f[..., 1] ** 2, 1.0, atol=1e-12)


def fd_mismatches():
    """Parameters (a seeded sample from every block) whose backward-pass gradient misses the central difference."""
    rng = np.random.default_rng(5)
    p = init_params(1)
    pts = LO + (HI - LO) * rng.random((12, 3)).astype(np.float32)
    y = rng.random(12)
    loss, g, _ = om.loss_grad(p, pts, y, LO, HI, emulate=False)
    lay = om.LAYOUT
    idx = [lay["bo"]]
    for l in range(om.LAYERS):
        idx += list(lay["w"][l] + rng.integers(0, om.IN_DIM[l] * om.H, 6))
        idx += list(lay["b"][l] + rng.integers(0, om.H, 3))
    idx += list(lay["wo"] + rng.integers(0, om.H, 4))
    bad = []
    for i in idx:
        h = 1e-5 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (om.loss_grad(pp, pts, y, LO, HI)[0] - om.loss_grad(pm, pts, y, LO, HI)[0]) / (2 * h)
        if not abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g[i])) + 1e-9:
            bad.append((int(i), fd, g[i]))
    return bad, len(idx)


def test_backward_matches_finite_differences():
    bad, n = fd_mismatches()
    assert n > 40 and not bad, bad


def test_planted_derivative_error_is_caught(monkeypatch):
    # the clamped slope alone, without the product rule's z/4 term: a plausible slip in the derivative
    monkeypatch.setattr(om, "hsilu_grad", lambda z: np.clip(0.25 * np.asarray(z, dtype=np.float64) + 0.5, 0.0, 1.0))
    bad, n = fd_mismatches()
    assert len(bad) > n // 4


def test_emulated_pass_close_to_exact():
    """The BF16 rounding points change the loss and gradient by about the BF16 precision (2^-8), not more."""
    rng = np.random.default_rng(9)
    p = init_params(2)
    pts = LO + (HI - LO) * rng.random((256, 3)).astype(np.float32)
    y = rng.random(256)
    l0, g0, p0 = om.loss_grad(p, pts, y, LO, HI, emulate=False)
    l1, g1, p1 = om.loss_grad(p, pts, y, LO, HI, emulate=True)
    assert np.max(np.abs(p1 - p0)) < 0.05 * np.max(np.abs(p0 - 0.5)) + 1e-3
    assert np.linalg.norm(g1 - g0) < 0.05 * np.linalg.norm(g0)


def test_adam_closed_forms():
    rng = np.random.default_rng(1)
    p = rng.normal(size=1000)
    g = rng.normal(size=1000)
    lr = 1e-3
    p1, m, v = om.adam_step(p, np.zeros_like(p), np.zeros_like(p), g, 1, lr)
    np.testing.assert_allclose(p1 - p, -lr * g / (np.abs(g) + 1e-8), rtol=1e-9)  # = -lr sign(g) for |g| >> eps
    g2 = rng.normal(size=1000)
    p2, _, _ = om.adam_step(p1, m, v, g2, 2, lr)
    mh = (0.9 * 0.1 * g + 0.1 * g2) / (1 - 0.81)
    vh = (0.999 * 0.001 * g * g + 0.001 * g2 * g2) / (1 - 0.999 ** 2)
    np.testing.assert_allclose(p2 - p1, -lr * mh / (np.sqrt(vh) + 1e-8), rtol=1e-9)


def test_layout_matches_library(agwn):
    from antipodal_amd.surrogate import num_params

    assert om.LAYOUT["total"] == 192 * 256 + 256 + 3 * (256 * 256 + 256) + 256 + 1
    assert num_params() == om.LAYOUT["total"]


def test_activation_closed_forms():
    z = np.array([-5.0, -2.0, -1.0, 0.0, 0.5, 2.0, 3.0])
    np.testing.assert_allclose(om.hsilu(z), [0.0, 0.0, -0.25, 0.0, 0.3125, 2.0, 3.0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(om.hsilu_grad(z), [0.0, 0.0, 0.0, 0.5, 0.75, 1.0, 1.0], rtol=0, atol=1e-15)
    # away from the kinks the derivative is the central difference of the function
    zz = np.linspace(-3.0, 3.0, 601)
    zz = zz[np.abs(np.abs(zz) - 2.0) > 1e-3]
    h = 1e-6
    np.testing.assert_allclose(om.hsilu_grad(zz), (om.hsilu(zz + h) - om.hsilu(zz - h)) / (2 * h), atol=1e-8)
