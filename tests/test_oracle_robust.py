"""Pins the robust restatement (orc_robust_eval, gwn_mesh_robust.hh) on the reference's own robust test cases
(test/test_gwn_mesh_robust.cc), against the per-triangle solid-angle baseline (test/gwn_baseline.hh, FP64) with the
reference's thresholds. The reference's RobustMeshGwn itself needs Embree (absent), so these geometries and thresholds
are what pin it; its predicates are pinned bit-for-bit against the reference's robust_predicates.hh in test_oracle.py.
"""
from __future__ import annotations

import numpy as np
import pytest

import robust_cases as rc


def _baseline(oracle, v, i, p):
    return oracle.gwn_baseline(v, i, p, np.float64)


@pytest.mark.parametrize("dt,pol", rc.CONFIGS)
def test_cube_closed_and_open(oracle, dt, pol):
    dt = np.dtype(dt)
    for idx, bnd in ((rc.CUBE_CLOSED, np.zeros((0, 7))), (rc.CUBE_OPEN_Y, rc.CUBE_OPEN_Y_BOUNDARY)):
        v = rc.CUBE_V.astype(dt)
        p = rc.CUBE_POINTS.astype(dt)
        _, _, ev = oracle.robust(v, idx, bnd.astype(dt), p, pol, dt)
        ref = oracle.gwn_baseline(v, idx, p, dt)
        assert np.max(np.abs(ev.astype(np.float64) - ref)) < rc.eps_for(dt) * (1 + np.abs(ref)).max()


@pytest.mark.parametrize("pol", rc.DOUBLE_CONFIGS)
def test_grazing_sweep_is_smooth(oracle, pol):
    """:224-287 the payoff test: no spurious ~1 jump across a grazing boundary edge."""
    bnd = oracle.boundary(rc.TRI_V, rc.TRI_I)
    p = rc.grazing_sweep()
    _, _, ev = oracle.robust(rc.TRI_V, rc.TRI_I, bnd, p, pol)
    assert np.max(np.abs(np.diff(ev))) < 0.1


def test_grazing_sweep_nonrobust_jumps(oracle):
    """The same sweep evaluated the non-robust way (FP64 fractional along -x + FP32 ray count) does jump by ~1 --
    the reason the robust path exists (the reference uses Embree's FP32 count; Moeller-Trumbore in FP32 stands in)."""
    bnd = oracle.boundary(rc.TRI_V, rc.TRI_I)
    p = rc.grazing_sweep()
    x0 = np.array([-1.0, 0.0, 0.0])
    frac = oracle.frac(p, bnd, x0)
    cnt = oracle.naive_count(rc.TRI_V.astype(np.float32), rc.TRI_I, p.astype(np.float32), x0.astype(np.float32),
                             np.float32)
    nonrobust = frac + cnt
    assert np.max(np.abs(np.diff(nonrobust))) > 0.5


def test_near_open_edge_accuracy(oracle):
    """:291-322: int128_grid and exact_float accurate within an int64_grid cell of an open edge; int64_grid is not."""
    bnd = oracle.boundary(rc.TRI_V, rc.TRI_I)
    p = rc.near_open_edge_points()
    ref = _baseline(oracle, rc.TRI_V, rc.TRI_I, p)
    err = {pol: np.max(np.abs(oracle.robust(rc.TRI_V, rc.TRI_I, bnd, p, pol)[2] - ref)) for pol in rc.DOUBLE_CONFIGS}
    assert err["int128_grid"] < 1e-6 and err["exact_float"] < 1e-6
    assert err["int64_grid"] > 1e-2


@pytest.mark.parametrize("pol", rc.DOUBLE_CONFIGS)
def test_edge_seen_along_the_ray(oracle, pol):
    """:326-348"""
    idx = np.array([0, 1, 2], dtype=np.int32)
    bnd = oracle.boundary(rc.EDGE_RAY_V, idx)
    p = rc.edge_ray_points()
    _, _, ev = oracle.robust(rc.EDGE_RAY_V, idx, bnd, p, pol)
    assert np.max(np.abs(ev - _baseline(oracle, rc.EDGE_RAY_V, idx, p))) < rc.near_tol(pol)


@pytest.mark.parametrize("pol", rc.DOUBLE_CONFIGS)
def test_ray_through_boundary_vertex(oracle, pol):
    """:351-369"""
    bnd = oracle.boundary(rc.VERTEX_RAY_V, rc.VERTEX_RAY_I)
    p = rc.vertex_ray_points()
    _, _, ev = oracle.robust(rc.VERTEX_RAY_V, rc.VERTEX_RAY_I, bnd, p, pol)
    assert np.max(np.abs(ev - _baseline(oracle, rc.VERTEX_RAY_V, rc.VERTEX_RAY_I, p))) < 1e-6


@pytest.mark.parametrize("pol", rc.DOUBLE_CONFIGS)
def test_far_away_queries(oracle, pol):
    """:374-395"""
    bnd = oracle.boundary(rc.TRI_V, rc.TRI_I)
    d, p = rc.far_points()
    _, _, ev = oracle.robust(rc.TRI_V, rc.TRI_I, bnd, p, pol)
    err = np.abs(ev - _baseline(oracle, rc.TRI_V, rc.TRI_I, p))
    assert np.all(err < rc.far_tol(pol, d))


@pytest.mark.parametrize("pol", rc.DOUBLE_CONFIGS)
def test_face_parallel_to_ray(oracle, pol):
    """:401-430 -- every query in front of the closed tetrahedron: count 0."""
    n = np.cross(rc.TETRA_V[1] - rc.TETRA_V[0], rc.TETRA_V[2] - rc.TETRA_V[0])
    assert n[0] == 0.0  # the precondition the reference test is about
    p = rc.tetra_points()
    _, cnt, _ = oracle.robust(rc.TETRA_V, rc.TETRA_I, np.zeros((0, 7)), p, pol)
    assert np.all(cnt == 0)


def test_exact_float_fractional_equals_plain_along_minus_x(oracle):
    """The robust fractional term (swapped boundary, gwn_mesh_robust.hh:198-206) equals the plain FP64 fractional term
    with x0 = axis() = (-1, 0, 0) away from grazing configurations; the counts add up to the baseline."""
    from antipodal_amd import meshgen as mg

    v, idx = mg.icosphere_cap_mesh()
    bnd = oracle.boundary(v, idx)
    p = mg.uniform_points(5, 3000, [-1.5] * 3, [1.5] * 3)
    fr, cnt, ev = oracle.robust(v, idx, bnd, p, "exact_float")
    plain = oracle.frac(p, bnd, np.array([-1.0, 0.0, 0.0]))
    assert np.max(np.abs(fr - plain)) < 1e-12
    assert np.max(np.abs(ev - _baseline(oracle, v, idx, p))) < 1e-9
    # the integer part is genuinely exercised
    assert (cnt != 0).sum() > 100
