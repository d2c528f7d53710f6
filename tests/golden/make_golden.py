"""Writes the golden fixtures in tests/golden/ from the REFERENCE's own code (oracle/_ref/libref.so, built by
`make -C oracle` from the reference headers where they lie). Run in the container that has the reference sources:

    python tests/golden/make_golden.py

Every fixture is data: seeded inputs (or the generator parameters that reproduce them bit for bit, plus a SHA-256 of
the generated arrays) and the reference's outputs on them.

  pair_kernel.npz        signed_spherical_tri_area_half_unorm (math/common.hh:181-189) on random and special triples
  c1_icosphere_cap.npz   config 1: open icosphere, boundary (reference order), 100k seed-0 queries in [-1.5,1.5]^3,
                         x0 = normalize(0.31,0.57,0.76); counts for all queries, fractional/full on a 4096 subsample,
                         SHA-256 of the full f64/f32 outputs; f32 and f64
  cube.npz               the unit-cube fixtures of test/test_gwn_mesh.cc and test/test_intersector.cc with the
                         reference's NaiveIntersector counts, full GWN and test/gwn_baseline.hh values
  predicates.npz         robust_predicates.hh edge_sign / atan2_num sign / project_into_tri on the reference tests'
                         random schemes (20-bit int grid, 50-bit grid, 2^-20 double lattice incl. collinear cases)
  torus_sub.npz          the 1M-triangle holed torus (generator parameters + SHA-256), 384 queries of its 256^3 grid:
                         reference fractional (f32, f64) and NaiveIntersector counts (f32, f64)
  torus_ref.npz          the same torus: 16,384 seeded lattice points of its 256^3 grid + 4,096 points 1e-3 off its
                         surface (regenerated from the seeds; SHA-256 of the positions kept), the reference's counts,
                         fractional terms, full GWN and labels in FP32 and FP64, and its boundary multiset's SHA-256

  c5_cap.npz             4096 seeded points of the 512^3 lattice over the config-1 cap: the reference's full GWN and
                         labels (x0 = +x), FP32 and FP64

    python tests/golden/make_golden.py [fixture ...]   (only the named fixtures, e.g. torus_ref)
"""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from antipodal_amd import meshgen as mg  # noqa: E402
from oracle import Reference  # noqa: E402

OUT = Path(__file__).resolve().parent


def sha(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def pair_kernel(ref: Reference) -> None:
    rng = np.random.default_rng(20251203)
    n = 2000
    x1 = rng.normal(size=(n, 3))
    x1 /= np.linalg.norm(x1, axis=1, keepdims=True)
    v0 = rng.normal(size=(n, 3)) * 10.0 ** rng.uniform(-3, 2, size=(n, 1))
    v1 = rng.normal(size=(n, 3)) * 10.0 ** rng.uniform(-3, 2, size=(n, 1))
    # special cases from test/test_common.cc: degenerate, octant, flips, equator loop, scaling
    sp = [
        ((0, 0, 1), (1, 0, 0), (1, 0, 0)),
        ((0, 0, 1), (0, 0, 1), (1, 0, 0)),
        ((0, 0, 1), (1, 0, 0), (0, 0, 1)),
        ((0, 0, 1), (1, 0, 0), (0, 1, 0)),
        ((0, 0, 1), (0, 1, 0), (1, 0, 0)),
        ((0, 0, -1), (1, 0, 0), (0, 1, 0)),
        ((0, 0, 1), (0, 1, 0), (-1, 0, 0)),
        ((0, 0, 1), (-1, 0, 0), (0, -1, 0)),
        ((0, 0, 1), (0, -1, 0), (1, 0, 0)),
        ((0, 0, 1), (1.3, -0.4, 0.2), (-0.7, 1.1, 0.5)),
        ((0, 0, 1), (9.1, -2.8, 1.4), (-0.007, 0.011, 0.005)),
    ]
    x1 = np.concatenate([x1, np.array([s[0] for s in sp], float)])
    v0 = np.concatenate([v0, np.array([s[1] for s in sp], float)])
    v1 = np.concatenate([v1, np.array([s[2] for s in sp], float)])
    out64 = np.array([ref.pair_half_area(a, b, c, np.float64) for a, b, c in zip(x1, v0, v1)])
    f = [a.astype(np.float32) for a in (x1, v0, v1)]
    out32 = np.array([ref.pair_half_area(a, b, c, np.float32) for a, b, c in zip(*f)], dtype=np.float32)
    np.savez_compressed(OUT / "pair_kernel.npz", x1=x1, v0=v0, v1=v1, out_f64=out64, out_f32=out32,
                        n_special=np.int64(len(sp)))


def c1(ref: Reference) -> None:
    verts, idx = mg.icosphere_cap_mesh()
    pos = mg.uniform_points(0, 100_000, [-1.5] * 3, [1.5] * 3)
    sub = np.arange(0, 100_000, 100_000 // 4096)[:4096]
    d = {"verts": verts, "idx": idx, "sub": sub}
    for dt, s in ((np.float64, "f64"), (np.float32, "f32")):
        v = verts.astype(dt)
        p = pos.astype(dt)
        x0 = mg.example_x0(dt)
        seg = ref.boundary(v, idx, dt)
        frac = ref.frac(p, seg, x0, dt, mode=1)
        cnt = ref.naive_count(v, idx, p, x0, dt, mode=1)
        full = ref.full(v, idx, p, x0, seg, dt, mode=1)
        base = ref.gwn_baseline(v, idx, p[sub], dt)
        d.update({f"seg_{s}": seg, f"x0_{s}": x0, f"count_{s}": cnt.astype(np.int8), f"frac_sub_{s}": frac[sub],
                  f"full_sub_{s}": full[sub], f"baseline_sub_{s}": base,
                  f"sha_frac_{s}": np.bytes_(sha(frac)), f"sha_full_{s}": np.bytes_(sha(full)),
                  f"sha_pos_{s}": np.bytes_(sha(p))})
    np.savez_compressed(OUT / "c1_icosphere_cap.npz", **d)


def cube(ref: Reference) -> None:
    v = np.array([[-.5, -.5, -.5], [.5, -.5, -.5], [-.5, .5, -.5], [.5, .5, -.5], [-.5, -.5, .5], [.5, -.5, .5],
                  [-.5, .5, .5], [.5, .5, .5]])
    closed = np.array([0, 2, 1, 1, 2, 3, 4, 5, 6, 6, 5, 7, 0, 4, 2, 2, 4, 6, 1, 3, 5, 5, 3, 7, 0, 1, 4, 4, 1, 5,
                       2, 6, 3, 3, 6, 7], dtype=np.int32)
    open_y = closed[:30].copy()
    # hand-written open-Y boundary of test/test_gwn_mesh.cc:62-78: v6->v2, v7->v6, v3->v7, v2->v3, weight 1
    bnd = np.array([np.r_[v[6], v[2], 1], np.r_[v[7], v[6], 1], np.r_[v[3], v[7], 1], np.r_[v[2], v[3], 1]])
    pts = np.array([[0, 0, 0], [.1, -.2, .3], [-.4, .4, -.4], [.49, 0, 0], [2, 0, 0], [-2, -2, -2], [0, 1.5, 0],
                    [1, 1, 1]])
    dirs = np.array([[.31, .57, .76], [-.6, .4, -.69], [.83, -.21, .51], [1.0, 0.0, 0.0]])
    d = {"verts": v, "closed": closed, "open_y": open_y, "boundary": bnd, "points": pts, "dirs": dirs}
    for dt, s in ((np.float64, "f64"), (np.float32, "f32")):
        vv, pp, bb = v.astype(dt), pts.astype(dt), bnd.astype(dt)
        x0 = mg.example_x0(dt)
        d[f"x0_{s}"] = x0
        for name, idx in (("closed", closed), ("open_y", open_y)):
            d[f"count_{name}_{s}"] = np.stack([ref.naive_count(vv, idx, pp, dd.astype(dt), dt) for dd in dirs])
            d[f"baseline_{name}_{s}"] = ref.gwn_baseline(vv, idx, pp, dt)
            seg = bb if name == "open_y" else np.zeros((0, 7), dt)
            d[f"full_{name}_{s}"] = ref.full(vv, idx, pp, x0, seg, dt)
        d[f"boundary_ref_{s}"] = ref.boundary(vv, open_y, dt)
    np.savez_compressed(OUT / "cube.npz", **d)


def predicates(ref: Reference) -> None:
    rng = np.random.default_rng(1234)
    n = 2500
    # int64_grid scheme of test/math/test_robust_predicates.cc:44-58 (coordinates in +-2^20)
    i32 = rng.integers(-(1 << 20), (1 << 20) + 1, size=(n, 12), dtype=np.int64).astype(np.int32)
    # near-collinear cases on the 20-bit grid: q on the line through a, b (plus 0/1 unit nudges)
    a = rng.integers(-(1 << 19), 1 << 19, size=(n, 3))
    dvec = rng.integers(-512, 513, size=(n, 3))
    k = rng.integers(-64, 65, size=(n, 1))
    q = a + k * dvec
    q[:, 1] += rng.integers(-1, 2, size=n)
    col = np.concatenate([q, a, a + dvec, rng.integers(-(1 << 19), 1 << 19, size=(n, 3))], axis=1)
    i32 = np.concatenate([i32, col.astype(np.int32)])
    # int128_grid scheme (50-bit coordinates), incl. collinear ones
    big = rng.integers(-(1 << 50), 1 << 50, size=(n, 12), dtype=np.int64)
    a = rng.integers(-(1 << 49), 1 << 49, size=(n, 3), dtype=np.int64)
    dvec = rng.integers(-1024, 1025, size=(n, 3), dtype=np.int64)
    k = rng.integers(-1024, 1025, size=(n, 1), dtype=np.int64)
    q = a + k * dvec
    q[:, 1] += rng.integers(-1, 2, size=n)
    i64 = np.concatenate([big, np.concatenate([q, a, a + dvec, rng.integers(-(1 << 49), 1 << 49, size=(n, 3))],
                                              axis=1)])
    # exact_float scheme of :136-169: integers times 2^-20 mixing 2^50 and 2^10 magnitudes, collinear +- one unit
    a = rng.integers(-(1 << 50), 1 << 50, size=(n, 3), dtype=np.int64)
    a[:, 0] = 0
    dvec = rng.integers(-1024, 1025, size=(n, 3), dtype=np.int64)
    dvec[:, 0] = 0
    k = rng.integers(-1024, 1025, size=(n, 1), dtype=np.int64)
    q = a + k * dvec
    q[:, 1] += rng.integers(-1, 2, size=n)
    small = rng.integers(-1024, 1025, size=(n, 3), dtype=np.int64)
    small[:, 0] = 0
    q[::4] = small[::4]
    c = rng.integers(-(1 << 50), 1 << 50, size=(n, 3), dtype=np.int64)
    f64 = np.ldexp(np.concatenate([q, a, a + dvec, c], axis=1).astype(np.float64), -20)
    # plus generic doubles
    f64 = np.concatenate([f64, rng.normal(size=(n, 12))])
    d = {"i32": i32, "i64": i64, "f64": f64}
    for key, ct in (("i32", np.int32), ("i64", np.int64), ("f64", np.float64)):
        es, ns, ins, w = ref.preds(d[key], ct)
        d.update({f"{key}_edge_sign": es, f"{key}_num_sign": ns, f"{key}_inside": ins, f"{key}_w": w})
    np.savez_compressed(OUT / "predicates.npz", **d)


def torus_sub(ref: Reference) -> None:
    verts, idx = mg.holed_torus_mesh()
    lo, hi = mg.expanded_bbox(verts)
    n = 256
    rng = np.random.default_rng(7)
    flat = np.sort(rng.choice(n ** 3, size=384, replace=False))
    d = {"mesh_sha": np.bytes_(sha(verts, idx)), "ntri": np.int64(len(idx) // 3), "flat": flat, "grid_n": np.int64(n)}
    for dt, s in ((np.float64, "f64"), (np.float32, "f32")):
        pts = mg.grid_points(lo, hi, n, dtype=dt)[flat]
        v = verts.astype(dt)
        x0 = mg.example_x0(dt)
        seg = ref.boundary(v, idx, dt)
        d[f"nseg_{s}"] = np.int64(len(seg))
        d[f"pos_{s}"] = pts
        d[f"frac_{s}"] = ref.frac(pts, seg, x0, dt, mode=1)
        d[f"count_{s}"] = ref.naive_count(v, idx, pts, x0, dt, mode=1)
    np.savez_compressed(OUT / "torus_sub.npz", **d)


def torus_ref(ref: Reference) -> None:
    """16,384 seeded points of the 256^3 lattice over the 1M-triangle torus plus 4,096 points 1e-3 off its surface:
    the reference's NaiveIntersector counts, fractional terms and full-GWN labels (GWN > 0.5), FP32 and FP64
    (OpenMPDispatcher), and the SHA-256 of its boundary multiset (segments sorted lexicographically)."""
    from oracle import sort_segments

    verts, idx = mg.holed_torus_mesh()
    lo, hi = mg.expanded_bbox(verts)
    n = 256
    flat = np.sort(np.random.default_rng(17).choice(n ** 3, size=16384, replace=False))
    near = mg.offset_surface_points(verts, idx, 4096, seed=5, stream=50)
    d = {"mesh_sha": np.bytes_(sha(verts, idx)), "flat": flat, "grid_n": np.int64(n), "n_grid": np.int64(len(flat)),
         "n_near": np.int64(len(near))}
    for dt, s in ((np.float64, "f64"), (np.float32, "f32")):
        pts = np.concatenate([mg.grid_points_at(lo, hi, n, flat, dtype=dt), near.astype(dt)])
        v = verts.astype(dt)
        x0 = mg.example_x0(dt)
        seg = ref.boundary(v, idx, dt)
        d[f"boundary_sha_{s}"] = np.bytes_(sha(sort_segments(seg)))
        d[f"nseg_{s}"] = np.int64(len(seg))
        d[f"pos_sha_{s}"] = np.bytes_(sha(pts))
        d[f"frac_{s}"] = ref.frac(pts, seg, x0, dt, mode=1)
        d[f"count_{s}"] = ref.naive_count(v, idx, pts, x0, dt, mode=1).astype(np.int8)
        full = ref.full(v, idx, pts, x0, seg, dt, mode=1)
        d[f"label_{s}"] = np.packbits(full > 0.5)
        d[f"full_{s}"] = full
        print(f"torus_ref {s}: {len(pts)} points", flush=True)
    np.savez_compressed(OUT / "torus_ref.npz", **d)


def c5_cap(ref: Reference) -> None:
    """4096 seeded points of the 512^3 lattice over the config-1 cap (its 10 %-expanded box), the lattice the
    surrogate labels: the reference's full GWN (NaiveIntersector, x0 = +x, the lattice's integration axis) and its
    labels, FP32 on the float32 lattice points and FP64 on the float64 ones."""
    verts, idx = mg.icosphere_cap_mesh()
    lo, hi = mg.expanded_bbox(verts)
    n = 512
    flat = np.sort(np.random.default_rng(29).choice(n ** 3, size=4096, replace=False))
    d = {"flat": flat, "grid_n": np.int64(n)}
    for dt, s in ((np.float64, "f64"), (np.float32, "f32")):
        pts = mg.grid_points_at(lo, hi, n, flat, dtype=dt)
        v = verts.astype(dt)
        x0 = np.array([1.0, 0.0, 0.0], dtype=dt)
        seg = ref.boundary(v, idx, dt)
        full = ref.full(v, idx, pts, x0, seg, dt, mode=1)
        d[f"pos_sha_{s}"] = np.bytes_(sha(pts))
        d[f"full_{s}"] = full
        d[f"label_{s}"] = np.packbits(full > 0.5)
    np.savez_compressed(OUT / "c5_cap.npz", **d)


def main() -> None:
    ref = Reference()
    if not ref.available():
        raise SystemExit("oracle/_ref/libref.so missing: run `make -C oracle` where the reference sources are")
    only = set(sys.argv[1:])
    for f in (pair_kernel, c1, cube, predicates, torus_sub, torus_ref, c5_cap):
        if only and f.__name__ not in only:
            continue
        f(ref)
        print("wrote", f.__name__)


if __name__ == "__main__":
    main()
