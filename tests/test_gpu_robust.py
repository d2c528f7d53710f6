"""Parity of the robust HIP path (RobustMeshGwn on the GPU) through the C ABI.

Bars: signed_intersection_count bit-exact with the robust restatement (every sign is exact, so the count is a pure
function of the inputs); fractional within 1e-12 (FP64 sqrt/atan2 last-bit differences only: the same exact-sign
branches, the same FP64 operations in the same order, summed in the same order); and the reference test file's own
thresholds against the solid-angle baseline (test/test_gwn_mesh_robust.cc).
"""
from __future__ import annotations

import numpy as np
import pytest

import robust_cases as rc
from antipodal_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def _both(gpu, oracle, v, idx, bnd, p, pol, dt):
    dt = np.dtype(dt)
    v, bnd, p = v.astype(dt), bnd.astype(dt), p.astype(dt)
    with gpu.RobustMeshGwn(v, idx, bnd, precision=pol, dtype=dt) as r:
        fr = r.fractional(p)
        cnt = r.signed_intersection_count(p)
        ev = gpu.eval_gwnr_mesh_batch_robust(r, p)
    ofr, ocnt, oev = oracle.robust(v, idx, bnd, p, pol, dt)
    return (fr, cnt, ev), (ofr, ocnt, oev)


def _check_parity(got, ref, dt):
    (fr, cnt, ev), (ofr, ocnt, oev) = got, ref
    assert np.array_equal(cnt, ocnt), f"{int((cnt != ocnt).sum())} counts differ"
    tol = 1e-12 if np.dtype(dt) == np.float64 else 2e-7  # FP32 outputs: one rounding of the same FP64 value
    assert np.max(np.abs(fr.astype(np.float64) - ofr.astype(np.float64))) <= tol
    assert np.max(np.abs(ev.astype(np.float64) - oev.astype(np.float64))) <= tol


@pytest.mark.parametrize("dt,pol", rc.CONFIGS)
def test_cube(gpu, oracle, dt, pol):
    for idx, bnd in ((rc.CUBE_CLOSED, np.zeros((0, 7))), (rc.CUBE_OPEN_Y, rc.CUBE_OPEN_Y_BOUNDARY)):
        got, ref = _both(gpu, oracle, rc.CUBE_V, idx, bnd, rc.CUBE_POINTS, pol, dt)
        _check_parity(got, ref, dt)
        base = oracle.gwn_baseline(rc.CUBE_V, idx, rc.CUBE_POINTS)
        assert np.max(np.abs(got[2].astype(np.float64) - base)) < 2 * rc.eps_for(dt)


@pytest.mark.parametrize("pol", rc.DOUBLE_CONFIGS)
def test_reference_robust_cases(gpu, oracle, pol):
    tri_b = oracle.boundary(rc.TRI_V, rc.TRI_I)
    # grazing sweep (payoff test)
    got, ref = _both(gpu, oracle, rc.TRI_V, rc.TRI_I, tri_b, rc.grazing_sweep(), pol, np.float64)
    _check_parity(got, ref, np.float64)
    assert np.max(np.abs(np.diff(got[2]))) < 0.1
    # near an open edge
    p = rc.near_open_edge_points()
    got, ref = _both(gpu, oracle, rc.TRI_V, rc.TRI_I, tri_b, p, pol, np.float64)
    _check_parity(got, ref, np.float64)
    err = np.max(np.abs(got[2] - oracle.gwn_baseline(rc.TRI_V, rc.TRI_I, p)))
    assert (err > 1e-2) if pol == "int64_grid" else (err < 1e-6)
    # edge along the ray
    idx = np.array([0, 1, 2], dtype=np.int32)
    p = rc.edge_ray_points()
    got, ref = _both(gpu, oracle, rc.EDGE_RAY_V, idx, oracle.boundary(rc.EDGE_RAY_V, idx), p, pol, np.float64)
    _check_parity(got, ref, np.float64)
    assert np.max(np.abs(got[2] - oracle.gwn_baseline(rc.EDGE_RAY_V, idx, p))) < rc.near_tol(pol)
    # ray through a boundary vertex
    p = rc.vertex_ray_points()
    got, ref = _both(gpu, oracle, rc.VERTEX_RAY_V, rc.VERTEX_RAY_I, oracle.boundary(rc.VERTEX_RAY_V, rc.VERTEX_RAY_I),
                     p, pol, np.float64)
    _check_parity(got, ref, np.float64)
    assert np.max(np.abs(got[2] - oracle.gwn_baseline(rc.VERTEX_RAY_V, rc.VERTEX_RAY_I, p))) < 1e-6
    # far-away queries (incl. the grid policies' clamp)
    d, p = rc.far_points()
    got, ref = _both(gpu, oracle, rc.TRI_V, rc.TRI_I, tri_b, p, pol, np.float64)
    _check_parity(got, ref, np.float64)
    assert np.all(np.abs(got[2] - oracle.gwn_baseline(rc.TRI_V, rc.TRI_I, p)) < rc.far_tol(pol, d))
    # face parallel to the ray: count 0 everywhere
    got, ref = _both(gpu, oracle, rc.TETRA_V, rc.TETRA_I, np.zeros((0, 7)), rc.tetra_points(), pol, np.float64)
    _check_parity(got, ref, np.float64)
    assert np.all(got[1] == 0)


@pytest.mark.parametrize("dt,pol", rc.CONFIGS)
def test_c1_mesh_random(gpu, oracle, dt, pol):
    v, idx = mg.icosphere_cap_mesh()
    bnd = oracle.boundary(v, idx)
    p = mg.uniform_points(21, 20000, [-1.5] * 3, [1.5] * 3)
    print(f"case robust c1 {dt} {pol}", flush=True)
    got, ref = _both(gpu, oracle, v, idx, bnd, p, pol, dt)
    _check_parity(got, ref, dt)


@pytest.mark.parametrize("dt,pol", [("float32", "int64_grid"), ("float64", "exact_float"), ("float64", "int128_grid")])
def test_torus_subsample(gpu, oracle, dt, pol):
    """The 1M-triangle holed torus: 96 grid queries + 32 near-surface ones, counts bit-exact, fractional to 1e-12."""
    verts, idx = mg.holed_torus_mesh()
    lo, hi = mg.expanded_bbox(verts)
    flat = (mg.uniform01(3, 7, 96) * 256 ** 3).astype(np.int64)
    grid = mg.grid_points(lo, hi, 256)[flat]
    tri = idx.reshape(-1, 3)
    pick = (mg.uniform01(3, 8, 32) * len(tri)).astype(np.int64)
    near = verts[tri[pick]].mean(axis=1) + 1e-4 * mg.normal(3, 9, 96).reshape(32, 3)
    p = np.concatenate([grid, near])
    bnd = oracle.boundary(verts, idx)
    print(f"case robust torus {dt} {pol}", flush=True)
    got, ref = _both(gpu, oracle, verts, idx, bnd, p, pol, dt)
    _check_parity(got, ref, dt)


@pytest.mark.parametrize("dt", ["float32", "float64"])
def test_int64_grid_fp64_form_at_its_limit(gpu, oracle, dt):
    """int64_grid computes a wave's determinants, dot products and squared lengths in FP64 when every quantized query
    of the wave lies within 2^25 in each coordinate (mesh coordinates lie within 2^20, so every such value is an
    integer below 2^53); other waves take the 64-bit integer form. 64 queries just inside the limit (31.9 mesh
    half-extents from the centre along the max norm, the largest exact products) and 64 just outside (32.1) fill one
    wave each; both equal the robust restatement: counts bit-exact, fractional to the policy's bar."""
    v, idx = mg.icosphere_cap_mesh()
    bnd = oracle.boundary(v, idx)
    lo, hi = v.min(axis=0), v.max(axis=0)
    centre, half = 0.5 * (lo + hi), 0.5 * float(np.max(hi - lo))
    dirs = mg.normal(3, 31, 3 * 128).reshape(128, 3)
    dirs /= np.abs(dirs).max(axis=1, keepdims=True)
    radius = np.concatenate([np.full(64, 31.9), np.full(64, 32.1)]) * half
    p = centre + dirs * radius[:, None]
    print(f"case robust int64_grid limit {dt}", flush=True)
    got, ref = _both(gpu, oracle, v, idx, bnd, p, "int64_grid", dt)
    _check_parity(got, ref, dt)


@pytest.mark.parametrize("dt,pol", rc.CONFIGS)
def test_stacked_squares(gpu, oracle, dt, pol):
    """80 parallel unit squares x = const (two triangles each, sharing the diagonal y = z): rays along the robust axis
    cross up to 80 of them, and the lattice queries with y = z meet the shared diagonals exactly, where only the exact
    signs and the symbolic perturbation decide which triangle counts. Counts bit-exact, fractional to the bar."""
    n = 80
    xs = np.linspace(-0.9, 0.9, n)
    corners = np.array([[0, -0.5, -0.5], [0, 0.5, -0.5], [0, 0.5, 0.5], [0, -0.5, 0.5]])
    v = np.concatenate([corners + [x, 0, 0] for x in xs])
    q = 4 * np.arange(n)[:, None]
    idx = np.concatenate([q + [0, 1, 2], q + [0, 2, 3]], axis=1).reshape(-1).astype(np.int32)
    bnd = oracle.boundary(v, idx)
    t = np.linspace(-0.8, 0.8, 9)
    yy, zz = np.meshgrid(t, t, indexing="ij")
    xq = np.array([-1.2, -0.45, 0.0, 0.37, 1.2])
    p = np.array([[x, y, z] for x in xq for y, z in zip(yy.ravel(), zz.ravel())])
    print(f"case robust stacked squares {dt} {pol}", flush=True)
    got, ref = _both(gpu, oracle, v, idx, bnd, p, pol, dt)
    _check_parity(got, ref, dt)
    assert np.abs(ref[1]).max() >= 40


def test_empty_inputs(gpu):
    with gpu.RobustMeshGwn(np.zeros((0, 3)), np.zeros(0, dtype=np.int32), np.zeros((0, 7))) as r:
        p = mg.uniform_points(1, 100, [-1] * 3, [1] * 3)
        assert np.all(r.eval(p) == 0.0)
        assert len(r.eval(np.zeros((0, 3)))) == 0
