"""Fits the odd minimax polynomial of the FP64 arctangent in csrc/agwn_math.hpp (atan2_f64) and checks both device
arctangent polynomials' error bounds. Run from the repository root:

    python antipodal_amd/csrc/atan_fit.py            # prints the FP64 coefficients and the achieved error

Method: the Remez exchange algorithm (Remez 1934; e.g. Muller, "Elementary Functions", ch. 3) in 40-digit arithmetic
(mpmath) for the relative error of atan(t) ~ t P(t^2) on |t| <= T, with T = 0.2, a little above tan(pi/16) =
0.19891..., the largest |t| that atan2_f64's three-interval reduction produces. With u = t^2 the problem is the
weighted minimax approximation of f(u) = atan(sqrt(u)) / sqrt(u) by a polynomial P of degree n on [0, T^2] with weight
1 / f (relative error of atan). The exchange starts from Chebyshev extrema, solves for the n + 1 coefficients and the
levelled error E on n + 2 reference points, moves the references to the local extrema of the error (a dense grid,
refined by golden-section search), and stops when the extrema are equal to 1e-6 of E.

The coefficients are rounded to double; the rounded polynomial's relative error over 2e5 points of [0, T] is
reported (Horner in double with fused multiply-adds emulated exactly, as the device evaluates it). The FP32
polynomial of agwn_frac.hip (degree 15 on [0, 1], Lawson's iteratively reweighted least squares) is checked the same
way against its stated bound, 1.1e-7 absolute in FP32 evaluation.
"""
from __future__ import annotations

import mpmath as mp
import numpy as np

mp.mp.dps = 40
T = mp.mpf("0.2")
U = T * T


def f(u):
    """atan(sqrt(u)) / sqrt(u), with its limit 1 at u = 0."""
    if u == 0:
        return mp.mpf(1)
    s = mp.sqrt(u)
    return mp.atan(s) / s


def horner(c, u):
    r = mp.mpf(0)
    for a in reversed(c):
        r = r * u + a
    return r


def rel_err(c, u):
    return (horner(c, u) - f(u)) / f(u)


def remez(n: int, iters: int = 40):
    """Coefficients c[0..n] of P and the levelled relative error E."""
    m = n + 2
    ref = [U * (1 - mp.cos(mp.pi * k / (m - 1))) / 2 for k in range(m)]  # Chebyshev extrema on [0, U]
    E = mp.mpf(0)
    c = [mp.mpf(0)] * (n + 1)
    for _ in range(iters):
        A = mp.matrix(m, m)
        b = mp.matrix(m, 1)
        for i, u in enumerate(ref):
            fu = f(u)
            for j in range(n + 1):
                A[i, j] = u ** j
            A[i, n + 1] = (-1) ** i * fu
            b[i] = fu
        x = mp.lu_solve(A, b)
        c = [x[j] for j in range(n + 1)]
        E = x[n + 1]
        # local extrema of the error between sign changes, on a dense grid refined by golden-section search
        grid = [U * k / 4000 for k in range(4001)]
        e = [rel_err(c, u) for u in grid]
        ext = []
        k = 0
        while k < len(grid):
            s = mp.sign(e[k]) or 1
            best = k
            while k < len(grid) and (mp.sign(e[k]) or s) == s:
                if abs(e[k]) > abs(e[best]):
                    best = k
                k += 1
            lo, hi = grid[max(best - 1, 0)], grid[min(best + 1, len(grid) - 1)]
            g = (mp.sqrt(5) - 1) / 2
            for _ in range(60):
                a, bb = hi - g * (hi - lo), lo + g * (hi - lo)
                if abs(rel_err(c, a)) > abs(rel_err(c, bb)):
                    hi = bb
                else:
                    lo = a
            ext.append((lo + hi) / 2)
        if len(ext) < m:
            raise RuntimeError(f"degree {n}: {len(ext)} alternating extrema, {m} needed")
        # keep m alternating extrema with the largest errors (endpoints included)
        while len(ext) > m:
            ea = [abs(rel_err(c, u)) for u in ext]
            drop = 0 if ea[0] < ea[-1] else len(ext) - 1
            ext.pop(drop)
        ref = ext
        errs = [abs(rel_err(c, u)) for u in ref]
        if max(errs) - min(errs) <= mp.mpf("1e-6") * abs(E):
            break
    return c, abs(E)


def fma_horner_double(c, t: np.ndarray) -> np.ndarray:
    """t P(t^2) evaluated as the device does: u = t*t (rounded), Horner with fused multiply-adds (each step rounded
    once; emulated exactly in 80-bit arithmetic, whose 64-bit significand holds a double product and sum except in the
    last bits, below the 2^-53 rounding being emulated), then t * P (rounded)."""
    u = (t * t).astype(np.float64)
    acc = np.full_like(u, c[-1])
    for a in reversed(c[:-1]):
        acc = (acc.astype(np.longdouble) * u + np.longdouble(a)).astype(np.float64)
    return (t * acc).astype(np.float64)


FP32_COEFFS = [0.9999993443489075, -0.33329862356185913, 0.19946569204330444, -0.13908647000789642,
               0.09642244875431061, -0.055913012474775314, 0.021863453090190887, -0.004054709803313017]


def check_fp32() -> float:
    """Max absolute error of the FP32 polynomial r P(r^2) on [0, 1], evaluated in float32 with fused multiply-adds."""
    r = np.linspace(0.0, 1.0, 200001, dtype=np.float32)
    t = (r * r).astype(np.float32)
    p = np.full_like(t, np.float32(FP32_COEFFS[-1]))
    for a in reversed(FP32_COEFFS[:-1]):
        p = (p.astype(np.float64) * t + np.float64(np.float32(a))).astype(np.float32)
    v = (r * p).astype(np.float32).astype(np.longdouble)
    return float(np.max(np.abs(v - np.arctan(r.astype(np.longdouble)))))


def main() -> None:
    for n in (7, 8, 9):
        c, E = remez(n)
        cd = [float(x) for x in c]
        t = np.linspace(0.0, float(T), 200001)
        v = fma_horner_double(cd, t).astype(np.longdouble)
        ref = np.arctan(t.astype(np.longdouble))
        rel = np.abs(v[1:] - ref[1:]) / ref[1:]
        print(f"degree {2 * n + 1}: levelled relative error {mp.nstr(E, 6)}; rounded, evaluated in double: "
              f"max relative error {float(rel.max()):.3e} = {float(rel.max()) / 2.0**-53:.2f} ulp")
        for k, x in enumerate(cd):
            print(f"    c{k} = {x!r}")
    print(f"FP32 polynomial (agwn_frac.hip atan2_f32): max |error| on [0, 1] = {check_fp32():.3e}")


if __name__ == "__main__":
    main()
