"""Far-field bound of one closed boundary loop's fractional contribution -- CPU restatement (test infrastructure only:
imported by tests/, never by the library).

No reference counterpart: this is the mathematics behind the library's far-field decision (the labelling stage in
csrc/agwn_frac.hip), restated in float64 numpy so that the CPU tests can check it against the oracle's own fractional
sum (Oracle.frac, which follows math/common.hh:181-189 and gwn_mesh.hh:34-46 and is pinned against the reference).

A closed loop L (points P_0..P_m, P_m = P_0) spanned by the fan D from its centroid c lies in the ball (c, rho). For a
query q outside the ball such that the line through q along x0 misses the ball, the loop's fractional term is
    frac_L(q) = -Omega_D(q) / 4 pi,   -Omega_D(q) = a.u / d^2 + (tr M - 3 u.M.u) / d^3 + R,
    |R| <= 3 I / (d - rho)^4 <= 3 rho^2 A / (d - rho)^4,
with x = c - q, d = |x|, u = x / d, a = the loop's vector area (1/2) sum u_i x u_i+1 (u_i = P_i - c), A the fan's area,
M = sum over fan triangles of (centroid - c) (x) (n dA) and I = the integral of |y - c|^2 dA over the fan (the
second-order Taylor term of the kernel at y is at most 3 |y - c|^2 / (d - rho)^4; over a triangle (0, u, w) the
integral of |y|^2 dA is area (|u|^2 + |w|^2 + u.w) / 6). The expansion and the bound hold for any surface inside the
ball, so a group of loops (cluster_table: the union of their fans, moments about a common centre) is estimated the same
way.
"""
from __future__ import annotations

import numpy as np


def unit_loops(seg7: np.ndarray):
    """Greedy walks over the unit-weight expansion of a boundary (negative weights reversed), in input order; returns
    the walks as point arrays and whether every walk is closed."""
    seg7 = np.asarray(seg7, dtype=np.float64).reshape(-1, 7)
    units = []
    for s in seg7:
        w = int(round(s[6]))
        a, b = tuple(s[0:3]), tuple(s[3:6])
        if w < 0:
            a, b, w = b, a, -w
        units += [(a, b)] * w
    out: dict = {}
    for i, (a, _) in enumerate(units):
        out.setdefault(a, []).append(i)
    used = np.zeros(len(units), dtype=bool)
    walks = []
    for i in range(len(units)):
        if used[i]:
            continue
        pts = [units[i][0]]
        cur = i
        while True:
            used[cur] = True
            e = units[cur][1]
            pts.append(e)
            nxt = [j for j in out.get(e, []) if not used[j]]
            if not nxt:
                break
            cur = nxt[0]
        walks.append(np.array(pts, dtype=np.float64))
    closed = all(np.array_equal(w[0], w[-1]) for w in walks)
    return walks, closed


def split_loop(points: np.ndarray, arc: int = 16):
    """The library's split of a long closed loop (more than `arc` segments): arcs of `arc` segments, each closed by a
    chord back to its first point, plus the skeleton loop through the split points (its chords cancel the arcs')."""
    p = np.asarray(points, dtype=np.float64)
    m = len(p) - 1
    if m <= arc:
        return [p]
    ks = list(range(0, m, arc)) + [m]
    pieces = [np.concatenate([p[ks[j]: ks[j + 1] + 1], p[ks[j]: ks[j] + 1]]) for j in range(len(ks) - 1)]
    pieces.append(p[ks])
    return pieces


def loop_table(points: np.ndarray) -> dict:
    """Centre, radius, vector area, fan area and first moment of one closed loop (last point = first)."""
    p = np.asarray(points, dtype=np.float64)
    c = p[:-1].mean(axis=0)
    u = p[:-1] - c
    w = p[1:] - c
    t = 0.5 * np.cross(u, w)
    g = (u + w) / 3.0
    at = np.sqrt((t * t).sum(axis=1))
    return {
        "c": c,
        "rho": float(np.sqrt((u * u).sum(axis=1)).max()),
        "a": t.sum(axis=0),
        "area": float(at.sum()),
        "m": (g[:, :, None] * t[:, None, :]).sum(axis=0),
        "inertia": float((at * ((u * u).sum(axis=1) + (w * w).sum(axis=1) + (u * w).sum(axis=1)) / 6.0).sum()),
        "s": (at[:, None] * g).sum(axis=0),
    }


def shifted(t: dict, c: np.ndarray) -> tuple:
    """(second moment, first moment) of a record's fans about another centre c: I + 2 dc.s + |dc|^2 A, s + A dc with
    dc = c_t - c (the library's rule for its loop-tree nodes)."""
    dc = t["c"] - c
    return t["inertia"] + 2.0 * float(dc @ t["s"]) + float(dc @ dc) * t["area"], t["s"] + t["area"] * dc


def cluster_table(tabs: list, shift: bool = True) -> dict:
    """One record for a group of loops about their mean centre: radius = max |c_L - c| + rho_L (a ball holding every
    member's ball, hence every fan), vector area and fan area summed, first moment about the new centre
    sum_L M_L + (c_L - c) (x) a_L -- the combination rule of the library's loop-tree nodes (merge_tables, which picks
    the smallest enclosing ball of two children instead of the mean centre). `shift=False` drops the shift term, only
    so that tests can plant the mistake."""
    c = np.mean([t["c"] for t in tabs], axis=0)
    m = np.zeros((3, 3))
    inertia, s = 0.0, np.zeros(3)
    for t in tabs:
        m += t["m"] + (np.outer(t["c"] - c, t["a"]) if shift else 0.0)
        i_t, s_t = shifted(t, c)
        inertia += i_t
        s += s_t
    return {
        "c": c,
        "rho": float(max(np.linalg.norm(t["c"] - c) + t["rho"] for t in tabs)),
        "a": np.sum([t["a"] for t in tabs], axis=0),
        "area": float(sum(t["area"] for t in tabs)),
        "m": m,
        "inertia": inertia,
        "s": s,
    }


def merge_tables(ta: dict, tb: dict) -> dict:
    """The library's loop-tree node over two children: the smallest ball holding both children's balls, vector areas
    and fan areas summed, first moments shifted to the new centre (M_child + (c_child - c) (x) a_child)."""
    dv = tb["c"] - ta["c"]
    dist = float(np.linalg.norm(dv))
    if dist + tb["rho"] <= ta["rho"]:
        c, rho = ta["c"].copy(), ta["rho"]
    elif dist + ta["rho"] <= tb["rho"]:
        c, rho = tb["c"].copy(), tb["rho"]
    else:
        c = ta["c"] + (dist + tb["rho"] - ta["rho"]) / (2.0 * dist) * dv
        rho = 0.5 * (dist + ta["rho"] + tb["rho"])
    m = np.zeros((3, 3))
    inertia, s = 0.0, np.zeros(3)
    for t in (ta, tb):
        rho = max(rho, float(np.linalg.norm(t["c"] - c)) + t["rho"])
        m += t["m"] + np.outer(t["c"] - c, t["a"])
        i_t, s_t = shifted(t, c)
        inertia += i_t
        s += s_t
    return {"c": c, "rho": rho, "a": ta["a"] + tb["a"], "area": ta["area"] + tb["area"], "m": m, "inertia": inertia,
            "s": s}


def loop_segments(points: np.ndarray) -> np.ndarray:
    """The loop as unit-weight weighted_segment3 rows (p0, p1, w = 1)."""
    p = np.asarray(points, dtype=np.float64)
    return np.concatenate([p[:-1], p[1:], np.ones((len(p) - 1, 1))], axis=1)


def applicable(tab: dict, q: np.ndarray, x0: np.ndarray) -> np.ndarray:
    """Queries (n, 3) for which the expansion holds: outside the ball, and the line along x0 misses it."""
    e = np.asarray(x0, dtype=np.float64) / np.linalg.norm(x0)
    x = tab["c"][None, :] - np.asarray(q, dtype=np.float64)
    d = np.linalg.norm(x, axis=1)
    line = np.linalg.norm(np.cross(x, e[None, :]), axis=1)
    return (d > tab["rho"]) & (line > tab["rho"])


def estimate(tab: dict, q: np.ndarray, sign: float = 1.0):
    """(estimate of frac_L, remainder bound) at queries q (n, 3), both in units of the fractional term. `sign` exists
    only so that tests can plant a sign error and see the bound check catch it."""
    x = tab["c"][None, :] - np.asarray(q, dtype=np.float64)
    d = np.linalg.norm(x, axis=1)
    u = x / d[:, None]
    m = tab["m"]
    mq = np.einsum("ni,ij,nj->n", u, m, u)
    om = sign * (u @ tab["a"]) / d ** 2 + (np.trace(m) - 3.0 * mq) / d ** 3
    bound = 3.0 * min(tab["inertia"], tab["rho"] ** 2 * tab["area"]) / (d - tab["rho"]) ** 4
    return om / (4.0 * np.pi), bound / (4.0 * np.pi)
