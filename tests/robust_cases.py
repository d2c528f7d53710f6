"""The geometries, probes and thresholds of the reference's test/test_gwn_mesh_robust.cc, as data.

Each case returns (verts, idx, boundary, points) in float64; thresholds are the reference's.
"""
from __future__ import annotations

import numpy as np

CONFIGS = [  # test_gwn_mesh_robust.cc:40-45
    ("float32", "int64_grid"),
    ("float64", "int64_grid"),
    ("float32", "int128_grid"),
    ("float64", "int128_grid"),
    ("float32", "exact_float"),
    ("float64", "exact_float"),
]
DOUBLE_CONFIGS = ["int64_grid", "int128_grid", "exact_float"]


def eps_for(dtype) -> float:  # :24-31
    return 1e-4 if np.dtype(dtype) == np.float32 else 1e-5


def near_tol(policy: str) -> float:  # :48-52
    return 1e-6 if policy == "int64_grid" else 1e-10


CUBE_V = np.array([[-.5, -.5, -.5], [.5, -.5, -.5], [-.5, .5, -.5], [.5, .5, -.5], [-.5, -.5, .5], [.5, -.5, .5],
                   [-.5, .5, .5], [.5, .5, .5]])
CUBE_CLOSED = np.array([0, 2, 1, 1, 2, 3, 4, 5, 6, 6, 5, 7, 0, 4, 2, 2, 4, 6, 1, 3, 5, 5, 3, 7, 0, 1, 4, 4, 1, 5,
                        2, 6, 3, 3, 6, 7], dtype=np.int32)
CUBE_OPEN_Y = CUBE_CLOSED[:30].copy()
CUBE_OPEN_Y_BOUNDARY = np.array([np.r_[CUBE_V[6], CUBE_V[2], 1], np.r_[CUBE_V[7], CUBE_V[6], 1],
                                 np.r_[CUBE_V[3], CUBE_V[7], 1], np.r_[CUBE_V[2], CUBE_V[3], 1]])
# probes on and off the y+z=0 diagonal (:114-128)
CUBE_POINTS = np.array([[0, 0, 0], [.1, -.2, .3], [-.4, .4, -.4], [.49, 0, 0], [2, 0, 0], [-2, -2, -2], [0, 1.5, 0],
                        [1, 1, 1]], dtype=np.float64)

TRI_V = np.array([[0, 0, 0], [0, 1, 0], [0, 0, 1.0]])  # single open triangle in x = 0, hypotenuse y + z = 1
TRI_I = np.array([0, 1, 2], dtype=np.int32)


def grazing_sweep():
    """:258-270 -- 1001 queries sweeping +-5e-7 across y = 0.7 at x = 100, z = 0.3."""
    n, center, hw = 1001, 0.7, 5e-7
    s = center - hw + 2.0 * hw * (np.arange(n) / (n - 1))
    return np.stack([np.full(n, 100.0), s, np.full(n, 0.3)], axis=1)


def near_open_edge_points():
    """:307-316 -- within 1e-9 of the hypotenuse."""
    return np.array([[x, 0.3, 0.7 + dz] for x in (1e-8, -1e-8, 1e-10, -1e-10) for dz in (1e-9, -1e-9)])


EDGE_RAY_V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 1.0]])  # :330


def edge_ray_points():
    """:336-346 -- an edge along x seen from just off its line."""
    pts = []
    for x in (1.5, 3.0, 10.0):
        for off in (1e-4, 1e-6, 1e-8, 1e-10, 1e-12):
            for sy, sz in ((1.0, 0.0), (0.0, 1.0), (1.0, -1.0), (-1.0, 0.3)):
                pts.append([x, off * sy, off * sz])
    return np.array(pts)


VERTEX_RAY_V = np.array([[0, 0, 0], [0, 1, 0], [0, 0, 1], [0.5, 1, 1.0]])  # :355
VERTEX_RAY_I = np.array([0, 1, 2, 1, 3, 2], dtype=np.int32)


def vertex_ray_points():
    """:360-368 -- the -x ray passes exactly through each vertex."""
    return np.array([[v[0] + d, v[1], v[2]] for v in VERTEX_RAY_V for d in (0.25, 1.0, 3.0, 17.0)])


def far_points():
    """:387-394 -- (d, point) pairs up to 1e9 away."""
    out = []
    for d in (10.0, 100.0, 300.0, 1e3, 5e3, 1e5, 1e9):
        for p in ([0.5, d, 0.2], [d, 0.2, 0.2], [-d, 0.2, 0.2], [d, -d, d]):
            out.append((d, p))
    return np.array([d for d, _ in out]), np.array([p for _, p in out])


def far_tol(policy: str, d: np.ndarray) -> np.ndarray:
    near = 1e-6 if policy == "int64_grid" else 1e-12
    clamped = 1e-6 if policy != "exact_float" else 1e-12
    return np.where(d <= 300.0, near, clamped)


# :405-417 -- closed tetrahedron whose face abc contains the x direction (exact coordinates)
TETRA_V = np.array([
    [0, 0.4303914952519497, 0.64558724287792457],
    [1, 0.68425246596985911, 1.0263786989547887],
    [0.29999999999999999, 0.19952205828266512, 0.29928308742399767],
    [0.5, 0.83805533983482472, 0.25708300975223697],
])
TETRA_I = np.array([0, 1, 2, 0, 3, 1, 1, 3, 2, 2, 3, 0], dtype=np.int32)


def tetra_points():
    """:417-428 -- the query that used to count 1, then the wall's line z = 1.5 y swept on and just off it."""
    b, c = TETRA_V[1], TETRA_V[2]
    pts = [[5, 0.20194571032110109, 0.30291856548165164]]
    for i in range(201):
        y = c[1] + (b[1] - c[1]) * (i / 200.0)
        for off in (0.0, 3e-7, -3e-7):
            pts.append([5, y + off, 1.5 * y - off])
    return np.array(pts)
