"""Seeded, bit-reproducible synthetic inputs for the batch-GWN workloads.

Every random number comes from a counter-based SplitMix64 stream (Steele, Lea & Flood 2014: state_i = seed' + i *
0x9E3779B97F4A7C15, output = the SplitMix64 finaliser of state_i), converted to doubles as (z >> 11) * 2^-53. Gaussians
use Box-Muller on two such streams. The same (seed, stream) gives the same bits on every machine with this numpy, so
CPU and GPU runs see identical inputs; FP32 inputs are float32(FP64 inputs).

Workloads (sizes as named in the benchmark plan):
  icosphere_cap_mesh()  icosphere level 3 (1280 triangles, radius 1) minus every triangle whose centroid has z > 0.6
  holed_torus_mesh()    torus R=1, r=0.35, nu x nv quads (2 triangles each), vertices perturbed by isotropic Gaussian
                        noise sigma = 0.01 r, then every triangle whose centroid lies within `hole_radius` of one of
                        `holes` seeded surface points removed. Defaults (1353 x 473, 400 holes of radius 0.05, seed 1)
                        leave 1,004,810 triangles and 31,376 unit-weight boundary segments: the 1M-triangle open mesh.
                        (400 disks on this torus cannot give much more boundary at 1M triangles: the total rim length
                        peaks near hole radius 0.074, at about 31k segments.)
  grid_points()         a dense n^3 lattice over a bounding box expanded by 10 % (5 % each side)
  uniform_points()      uniform points in a box
  surface_points()      area-uniform points on a mesh with their unit face normals
  precision_stress()    the precision-stress configuration: the torus family at 3040 x 1063 quads with 5000 holes of
                        radius 0.02 (4,038,497 triangles, 429,869 boundary segments), queries half uniform in the
                        10 %-expanded box and half at |N(0, 1e-3)| along the outward normal of surface points
  batch_workload()      the uneven many-mesh configuration: 4096 open spheres/tori (100..2e5 triangles log-uniform,
                        1..200 holes), 1e3..1e6 queries per mesh log-uniform (592,815,454 in all)
"""
from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def splitmix64(seed: int, stream: int, n: int) -> np.ndarray:
    """n uint64 outputs of the counter-based SplitMix64 stream (seed, stream)."""
    with np.errstate(over="ignore"):
        base = _mix(np.array([(seed * 0x9E3779B97F4A7C15 + stream * 0xD1B54A32D192ED03) & (2**64 - 1)], dtype=np.uint64))[0]
        ctr = np.arange(1, n + 1, dtype=np.uint64)
        return _mix(base + ctr * _GOLDEN)


def uniform01(seed: int, stream: int, n: int) -> np.ndarray:
    """n doubles in [0, 1): (z >> 11) * 2^-53."""
    return (splitmix64(seed, stream, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """n standard normal doubles (Box-Muller, cosine branch)."""
    u1 = 1.0 - uniform01(seed, 2 * stream + 1, n)  # (0, 1]
    u2 = uniform01(seed, 2 * stream + 2, n)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def uniform_points(seed: int, n: int, lo, hi, stream: int = 0) -> np.ndarray:
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    u = uniform01(seed, stream, 3 * n).reshape(n, 3)
    return lo + u * (hi - lo)


# ---------------------------------------------------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------------------------------------------------
def icosphere(level: int = 3, radius: float = 1.0):
    """Icosahedron subdivided `level` times, vertices projected to the sphere; outward (CCW-from-outside) triangles."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache: dict[tuple[int, int], int] = {}

        def mid(a: int, b: int) -> int:
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    v = np.array(verts, dtype=np.float64) * radius
    f = np.array(faces, dtype=np.int32)
    # orient outward: the icosahedron table above is CCW from outside; verify and fix per face
    c = v[f].mean(axis=1)
    n = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    flip = np.einsum("ij,ij->i", n, c) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return v, f.reshape(-1)


def _compact(verts: np.ndarray, idx: np.ndarray):
    used = np.unique(idx)
    remap = np.full(len(verts), -1, dtype=np.int64)
    remap[used] = np.arange(len(used))
    return np.ascontiguousarray(verts[used]), remap[idx].astype(np.int32)


def icosphere_cap_mesh(level: int = 3, radius: float = 1.0, zcut: float = 0.6):
    """The small open mesh: icosphere minus the triangles whose centroid z exceeds `zcut` (one boundary loop)."""
    v, idx = icosphere(level, radius)
    tri = idx.reshape(-1, 3)
    keep = v[tri].mean(axis=1)[:, 2] <= zcut
    return _compact(v, tri[keep].reshape(-1))


def torus_mesh(nu: int, nv: int, R: float = 1.0, r: float = 0.35):
    i = np.arange(nu)
    j = np.arange(nv)
    u = 2.0 * math.pi * i / nu
    w = 2.0 * math.pi * j / nv
    uu, ww = np.meshgrid(u, w, indexing="ij")
    rad = R + r * np.cos(ww)
    verts = np.stack([rad * np.cos(uu), rad * np.sin(uu), r * np.sin(ww)], axis=-1).reshape(-1, 3)
    ii, jj = np.meshgrid(i, j, indexing="ij")
    a = ii * nv + jj
    b = ((ii + 1) % nu) * nv + jj
    c = ((ii + 1) % nu) * nv + (jj + 1) % nv
    d = ii * nv + (jj + 1) % nv
    tris = np.stack([np.stack([a, b, c], -1), np.stack([a, c, d], -1)], axis=2).reshape(-1, 3)
    return verts, tris.astype(np.int32).reshape(-1)


def holed_torus_mesh(seed: int = 1, nu: int = 1353, nv: int = 473, R: float = 1.0, r: float = 0.35,
                     noise: float = 0.01, holes: int = 400, hole_radius: float = 0.05):
    """Noisy torus with `holes` disk-shaped holes (triangles whose centroid is within hole_radius of a hole centre).

    Hole centres are the centroids of seeded uniformly chosen triangles of the noisy torus.
    """
    verts, idx = torus_mesh(nu, nv, R, r)
    nvtx = len(verts)
    verts = verts + (noise * r) * normal(seed, 1, 3 * nvtx).reshape(nvtx, 3)
    tri = idx.reshape(-1, 3)
    cent = verts[tri].mean(axis=1)
    if holes > 0:
        pick = (uniform01(seed, 2, holes) * len(tri)).astype(np.int64)
        centres = cent[pick]
        from scipy.spatial import cKDTree

        d, _ = cKDTree(centres).query(cent, k=1)
        tri = tri[d > hole_radius]
    return _compact(verts, tri.reshape(-1))


def surface_points(verts, idx, n: int, seed: int, stream: int = 10):
    """n area-uniform points on the triangles (seeded) and their unit face normals (right-hand winding)."""
    v = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
    tri = np.asarray(idx, dtype=np.int64).reshape(-1, 3)
    a, b, c = v[tri[:, 0]], v[tri[:, 1]], v[tri[:, 2]]
    nrm = np.cross(b - a, c - a)
    area = np.linalg.norm(nrm, axis=1)
    ok = area > 0
    nrm[ok] /= area[ok, None]
    cdf = np.cumsum(area)
    k = np.minimum(np.searchsorted(cdf, uniform01(seed, stream, n) * cdf[-1], side="right"), len(tri) - 1)
    r1 = np.sqrt(uniform01(seed, stream + 1, n))
    r2 = uniform01(seed, stream + 2, n)
    p = (1 - r1)[:, None] * a[k] + (r1 * (1 - r2))[:, None] * b[k] + (r1 * r2)[:, None] * c[k]
    return p, nrm[k]


def offset_surface_points(verts, idx, n: int, seed: int, stream: int, sigma: float = 1e-3) -> np.ndarray:
    """n points at |N(0, sigma)| along the outward normal of seeded area-uniform surface points (float64); the
    surface points use streams stream..stream+2, the offsets stream + 10."""
    p, nrm = surface_points(verts, idx, n, seed, stream=stream)
    off = np.abs(normal(seed, stream + 10, n)) * sigma
    return p + off[:, None] * nrm


def precision_stress(n_queries: int = 10_000_000, seed: int = 1):
    """(verts, idx, queries float64 (n, 3)) of the precision-stress configuration."""
    verts, idx = holed_torus_mesh(seed=seed, nu=3040, nv=1063, holes=5000, hole_radius=0.02)
    lo, hi = expanded_bbox(verts)
    nu = n_queries // 2
    q_uniform = uniform_points(seed, nu, lo, hi, stream=20)
    p, nrm = surface_points(verts, idx, n_queries - nu, seed, stream=30)
    off = np.abs(normal(seed, 40, n_queries - nu)) * 1e-3
    q_near = p + off[:, None] * nrm
    return verts, idx, np.concatenate([q_uniform, q_near])


def uv_sphere_mesh(nu: int, nv: int, radius: float = 1.0):
    """Latitude-longitude sphere: nu longitudes x nv latitude bands, poles as single vertices (2 nu (nv - 1) tris)."""
    nu, nv = max(nu, 3), max(nv, 2)
    th = np.pi * np.arange(1, nv) / nv
    ph = 2.0 * np.pi * np.arange(nu) / nu
    tt, pp = np.meshgrid(th, ph, indexing="ij")  # (nv - 1, nu)
    ring = np.stack([np.sin(tt) * np.cos(pp), np.sin(tt) * np.sin(pp), np.cos(tt)], -1).reshape(-1, 3)
    verts = np.concatenate([[[0.0, 0.0, 1.0]], ring, [[0.0, 0.0, -1.0]]]) * radius
    ring_id = lambda i, j: 1 + i * nu + (j % nu)
    tris = []
    j = np.arange(nu)
    tris.append(np.stack([np.zeros(nu, np.int64), ring_id(0, j), ring_id(0, j + 1)], -1))
    for i in range(nv - 2):
        a, b, c, d = ring_id(i, j), ring_id(i + 1, j), ring_id(i + 1, j + 1), ring_id(i, j + 1)
        tris.append(np.stack([a, b, c], -1))
        tris.append(np.stack([a, c, d], -1))
    last = len(verts) - 1
    tris.append(np.stack([np.full(nu, last), ring_id(nv - 2, j + 1), ring_id(nv - 2, j)], -1))
    return verts, np.concatenate(tris).astype(np.int32).reshape(-1)


def open_mesh(kind: str, ntri: int, holes: int, seed: int, noise: float = 0.01):
    """A noisy sphere or torus of about `ntri` triangles with `holes` disk holes (radius scaled so that the holes cover
    about a quarter of the surface at most), seeded."""
    if kind == "sphere":
        nv = max(2, int(round(math.sqrt(ntri / 4.0))))
        nu = max(3, int(round(ntri / (2.0 * max(nv - 1, 1)))))
        verts, idx = uv_sphere_mesh(nu, nv)
        scale, area = 1.0, 4.0 * math.pi
    else:
        nvv = max(3, int(round(math.sqrt(ntri / 2.0 * 0.35))))
        nuu = max(3, int(round(ntri / (2.0 * nvv))))
        verts, idx = torus_mesh(nuu, nvv)
        scale, area = 0.35, 4.0 * math.pi ** 2 * 0.35
    nvtx = len(verts)
    verts = verts + (noise * scale) * normal(seed, 1, 3 * nvtx).reshape(nvtx, 3)
    tri = idx.reshape(-1, 3)
    cent = verts[tri].mean(axis=1)
    if holes > 0:
        rad = math.sqrt(0.25 * area / (math.pi * holes))
        rad = min(rad, 0.5 * scale)
        pick = (uniform01(seed, 2, holes) * len(tri)).astype(np.int64)
        from scipy.spatial import cKDTree

        d, _ = cKDTree(cent[pick]).query(cent, k=1)
        tri = tri[d > rad]
    return _compact(verts, tri.reshape(-1))


def batch_workload(k: int = 4096, seed: int = 7, workers: int = 8):
    """The uneven many-mesh workload: k seeded open meshes, noisy spheres (even m) and tori (odd m) alternating, with
    triangle counts log-uniform in [100, 2e5] and 1..200 holes, and a query count per mesh log-uniform in [1e3, 1e6].
    The draws come from numpy's default_rng(seed), three per mesh in mesh order (triangles, holes, queries), so the
    workload does not depend on `workers` (threads that build the meshes; numpy and scipy release the GIL).
    Returns (meshes: [(float32 vertices (nv, 3), int32 indices)], query counts, [(box lo, box hi)] of each mesh's
    10 %-expanded bounding box)."""
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(seed)
    specs, qn = [], []
    for m in range(k):
        ntri = int(math.exp(rng.uniform(math.log(100), math.log(2e5))))
        holes = int(rng.integers(1, 201))
        qn.append(int(math.exp(rng.uniform(math.log(1e3), math.log(1e6)))))
        specs.append(("sphere" if m % 2 == 0 else "torus", ntri, holes, 100003 + m))

    def one(s):
        v, idx = open_mesh(*s)
        v = v.astype(np.float32)
        return (v, idx), expanded_bbox(v)

    with ThreadPoolExecutor(max(1, workers)) as ex:
        res = list(ex.map(one, specs))
    return [r[0] for r in res], qn, [r[1] for r in res]


def expanded_bbox(verts: np.ndarray, frac: float = 0.10):
    lo = verts.min(axis=0)
    hi = verts.max(axis=0)
    pad = 0.5 * frac * (hi - lo)
    return lo - pad, hi + pad


def grid_axes(lo, hi, n: int):
    """start, dx, dy, dz of an n^3 lattice p(i,j,k) = start + i dx + j dy + k dz spanning [lo, hi]."""
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    step = (hi - lo) / (n - 1)
    return lo, np.array([step[0], 0, 0]), np.array([0, step[1], 0]), np.array([0, 0, step[2]])


def grid_points(lo, hi, n: int, dtype=np.float64) -> np.ndarray:
    """n^3 lattice points in ((k*n)+j)*n+i order (x fastest), computed in `dtype` as start + i dx + j dy + k dz."""
    start, dx, dy, dz = (np.asarray(a, dtype=dtype) for a in grid_axes(lo, hi, n))
    r = np.arange(n, dtype=dtype)
    pts = np.empty((n, n, n, 3), dtype=dtype)
    # left-to-right evaluation order of start + dx*i + dy*j + dz*k, per component
    pts[...] = start
    pts += (dx[None, None, None, :] * r[None, None, :, None])
    pts += (dy[None, None, None, :] * r[None, :, None, None])
    pts += (dz[None, None, None, :] * r[:, None, None, None])
    return pts.reshape(-1, 3)


def grid_points_at(lo, hi, n: int, flat, dtype=np.float64) -> np.ndarray:
    """grid_points(lo, hi, n, dtype)[flat] without the whole lattice: the same per-component operations
    ((start + dx i) + dy j) + dz k in `dtype` (i fastest in the flat index)."""
    start, dx, dy, dz = (np.asarray(a, dtype=dtype) for a in grid_axes(lo, hi, n))
    flat = np.asarray(flat, dtype=np.int64)
    i, j, k = (flat % n).astype(dtype), ((flat // n) % n).astype(dtype), (flat // (n * n)).astype(dtype)
    pts = np.empty((len(flat), 3), dtype=dtype)
    pts[...] = start
    pts += dx[None, :] * i[:, None]
    pts += dy[None, :] * j[:, None]
    pts += dz[None, :] * k[:, None]
    return pts


EXAMPLE_X0 = (0.31, 0.57, 0.76)  # examples/main.cc:205 (normalised by the caller)


def example_x0(dtype=np.float64) -> np.ndarray:
    """normalize(0.31, 0.57, 0.76) evaluated in `dtype` the way the reference's normalize does (a / sqrt(dot(a,a)))."""
    a = np.array(EXAMPLE_X0, dtype=dtype)
    d = dtype(dtype(a[0] * a[0]) + dtype(a[1] * a[1]))
    d = dtype(d + dtype(a[2] * a[2]))
    l = dtype(np.sqrt(d))
    return np.array([a[0] / l, a[1] / l, a[2] / l], dtype=dtype)
