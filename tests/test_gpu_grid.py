"""The dense-grid evaluator (HipGridGwn; OpenGLMeshGwn's interface, gwn_mesh_opengl.hh:426-643) and the C++
drop-in (include/antipodal_hip/gwn_mesh_hip.hh) on the GPU.

Bars: every cell equals eval_gwnr_mesh_batch of the same lattice point with the same x0 (the GL evaluator's choice:
normalize(step of the smallest axis)) bit for bit; integer term of the cells bit-exact with NaiveIntersector<float>;
against the reference's CPU float path with its own x0 within the GL test's 3e-3 for cells more than two steps from
the surface (test/test_gwn_mesh_opengl.cc:190-200).
"""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

import robust_cases as rc
from antipodal_amd import meshgen as mg

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]


def rotated_basis():
    """test_gwn_mesh_opengl.cc:106-124, in float32."""
    f = np.float32

    def nrm(a):
        a = np.asarray(a, dtype=f)
        return a / f(np.sqrt(f(a[0] * a[0] + a[1] * a[1]) + f(a[2] * a[2])))

    ex = nrm([0.60, 0.70, 0.39])
    ez = nrm(np.cross(ex, np.array([-0.40, 0.20, 0.90], dtype=f)).astype(f))
    ey = np.cross(ez, ex).astype(f)
    return ex, ey, ez


def lattice(start, dx, dy, dz, nx, ny, nz):
    """start + dx*ix + dy*iy + dz*iz in float32, left to right, in ((iz*ny)+iy)*nx+ix order."""
    f = np.float32
    iz, iy, ix = np.meshgrid(np.arange(nz, dtype=f), np.arange(ny, dtype=f), np.arange(nx, dtype=f), indexing="ij")
    pts = ((start[None, :] + dx[None, :] * ix.reshape(-1, 1)) + dy[None, :] * iy.reshape(-1, 1)) + dz[None, :] * iz.reshape(-1, 1)
    return pts.astype(f)


def point_tri_dist(p, a, b, c):
    """Distance from points p (n,3) to triangle (a,b,c) by projection + edge clamping (vectorised)."""
    ab, ac = b - a, c - a
    n = np.cross(ab, ac)
    n /= np.linalg.norm(n)
    d = (p - a) @ n
    q = p - d[:, None] * n
    # barycentrics of q
    v0, v1, v2 = ab, ac, q - a
    d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
    d20, d21 = v2 @ v0, v2 @ v1
    den = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den
    w = (d00 * d21 - d01 * d20) / den
    inside = (v >= 0) & (w >= 0) & (v + w <= 1)

    def seg(p0, p1):
        t = np.clip(((p - p0) @ (p1 - p0)) / ((p1 - p0) @ (p1 - p0)), 0, 1)
        return np.linalg.norm(p - (p0 + t[:, None] * (p1 - p0)), axis=1)

    edge = np.minimum(np.minimum(seg(a, b), seg(b, c)), seg(c, a))
    return np.where(inside, np.abs(d), edge)


@pytest.mark.parametrize("name,e,n", [("closed", 0.15, (20, 20, 20)), ("open_y", 0.15, (20, 20, 20)),
                                      ("closed_slice", 0.13, (24, 1, 24)), ("closed_zsmall", 0.15, (20, 20, 6))])
def test_cube_grids(gpu, oracle, name, e, n):
    """The four subcases of test_gwn_mesh_opengl.cc:215-245."""
    idx = rc.CUBE_OPEN_Y if name == "open_y" else rc.CUBE_CLOSED
    v = rc.CUBE_V.astype(np.float32)
    bnd = gpu.build_boundary_segments(v, idx)
    ex, ey, ez = rotated_basis()
    f = np.float32
    nx, ny, nz = n
    dx, dy, dz = ex * f(e), ey * f(e), ez * f(e)
    start = (dx * f(-0.5 * (nx - 1)) + dy * f(-0.5 * (ny - 1))) + dz * f(-0.5 * (nz - 1))
    with gpu.HipGridGwn() as g:
        g.set_scene(v, idx, bnd)
        got = g.eval_grid(start, dx, dy, dz, nx, ny, nz)
    pts = lattice(start, dx, dy, dz, nx, ny, nz)
    # same cells through the batch path with the evaluator's x0
    counts = [nx, ny, nz]
    ia = int(np.argmin(counts))
    a = [dx, dy, dz][ia]
    x0 = a / f(np.sqrt(f(a[0] * a[0] + a[1] * a[1]) + f(a[2] * a[2])))
    with gpu.HipMesh(v, idx) as mesh:
        batch = gpu.eval_gwnr_mesh_batch(mesh, pts, x0, bnd)
        cnt = gpu.eval_gwnr_mesh_batch_integer(mesh, pts, x0)
    assert np.array_equal(got, batch)
    assert np.array_equal(cnt, oracle.naive_count(v, idx, pts, x0, np.float32))
    # the GL test's comparison: reference float path with its own x0, cells away from the surface
    ref = oracle.full(v, idx, pts, mg.example_x0(np.float32), bnd, np.float32)
    tri = v[idx.reshape(-1, 3)].astype(np.float64)
    dist = np.min(np.stack([point_tri_dist(pts.astype(np.float64), *t) for t in tri]), axis=0)
    far = dist >= 2 * e
    assert far.sum() > 0
    assert np.max(np.abs(got[far] - ref[far])) < 3e-3


def test_torus_grid_matches_batch(gpu, oracle):
    """A 48^3 grid on the 1M-triangle holed torus: cell values == batch path on the same points (bit-exact), and the
    extreme cells plus a few seeded ones against the per-triangle solid-angle sum (test/gwn_baseline.hh, FP64)."""
    verts, idx = mg.holed_torus_mesh()
    v = verts.astype(np.float32)
    bnd = gpu.build_boundary_segments(v, idx)
    lo, hi = mg.expanded_bbox(verts)
    n = 48
    start, dx, dy, dz = (np.asarray(a, dtype=np.float32) for a in mg.grid_axes(lo, hi, n))
    with gpu.HipGridGwn() as g:
        g.set_scene(v, idx, bnd)
        got = g.eval_grid(start, dx, dy, dz, n, n, n)
    pts = lattice(start, dx, dy, dz, n, n, n)
    x0 = dx / np.float32(np.sqrt(np.float32(dx[0] * dx[0] + dx[1] * dx[1]) + np.float32(dx[2] * dx[2])))
    with gpu.HipMesh(v, idx) as mesh:
        batch = gpu.eval_gwnr_mesh_batch(mesh, pts, x0, bnd)
    assert np.array_equal(got, batch)
    # the holed torus's GWN leaves [0, 1] near its holes (below -1 where several hole rims are seen at once)
    cells = np.concatenate([[np.argmin(got), np.argmax(got)], (mg.uniform01(5, 1, 6) * len(got)).astype(np.int64)])
    base = oracle.gwn_baseline(verts, idx, pts[cells].astype(np.float64), np.float64, compensated=True)
    assert np.max(np.abs(got[cells] - base)) < 5e-4


def _sheets(n: int):
    """n parallel squares x = const, y and z in [-0.5, 0.5] (two triangles each, split along y = z): an open surface
    that a lattice row along x through the squares crosses n times."""
    f = np.float32
    xs = np.linspace(-0.9, 0.9, n)
    corners = np.array([[0, -0.5, -0.5], [0, 0.5, -0.5], [0, 0.5, 0.5], [0, -0.5, 0.5]])
    v = np.concatenate([corners + [x, 0, 0] for x in xs]).astype(f)
    q = 4 * np.arange(n)[:, None]
    idx = np.concatenate([q + [0, 1, 2], q + [0, 2, 3]], axis=1).reshape(-1).astype(np.int32)
    return v, idx


@pytest.mark.parametrize("case", ["plus_x", "minus_x", "rows_along_y", "rows_along_minus_z", "one_point_rows",
                                  "descending_y_z", "second_axis", "third_axis", "sheets"])
def test_grid_rows_along_the_ray(gpu, oracle, case):
    """Axis-aligned lattices, whose rows along the integration axis (the lattice axis with the fewest cells, x0 = +-e_k)
    take the row-list integer term: each triangle is binned into the rows that its leaf box spans and where it passes
    the u and v tests, and each point tests its row's triangles. Rows along +x, -x, +y and -z, rows of one point, rows
    whose other two coordinates descend with their indices, rows along the lattice's second and third axes, and a
    stack of 80 squares that every row through it crosses 80 times (more triangles than a row's list holds: those
    rows' points take the traversal). Every cell equals the batch path on the same points with the same x0 bit for
    bit, and the integer terms equal NaiveIntersector<float>'s."""
    f = np.float32
    if case == "sheets":
        v, idx = _sheets(80)
    else:
        verts, idx = mg.holed_torus_mesh(nu=96, nv=40, holes=20, hole_radius=0.1)
        v = verts.astype(f)
    bnd = gpu.build_boundary_segments(v, idx)
    h, hz = f(3.0 / 15), f(1.0 / 17)
    ex, ey, ez = np.eye(3, dtype=f)
    nx, ny, nz = 16, 20, 18
    start = np.array([-1.5, -1.5, -0.5], dtype=f)
    dx, dy, dz = ex * h, ey * f(3.0 / 19), ez * hz
    if case == "minus_x":
        start, dx = np.array([1.5, -1.5, -0.5], dtype=f), ex * -h
    elif case == "rows_along_y":
        dx, dy = ey * h, ex * f(3.0 / 19)
    elif case == "rows_along_minus_z":
        start = np.array([-1.5, -1.5, 0.5], dtype=f)
        dx, dz = ez * -hz, ex * f(3.0 / 19)
    elif case == "one_point_rows":
        nx = 1
    elif case == "descending_y_z":  # the rows' other two coordinates fall with their indices
        start = np.array([-1.5, 1.5, 0.5], dtype=f)
        dy, dz = ey * f(-3.0 / 19), ez * -hz
    elif case == "second_axis":  # the lattice's y axis has the fewest cells and steps along world x
        nx, ny, nz = 20, 12, 18
        dx, dy = ey * f(3.0 / 19), ex * f(3.0 / 11)
    elif case == "third_axis":  # the lattice's z axis has the fewest cells and steps along world -z
        nx, ny, nz = 20, 18, 10
        start = np.array([-1.5, -1.5, 0.5], dtype=f)
        dx, dy, dz = ex * f(3.0 / 19), ey * f(3.0 / 17), ez * f(-1.0 / 9)
    elif case == "sheets":
        start = np.array([-1.2, -0.8, -0.8], dtype=f)
        dx, dy, dz = ex * f(2.4 / 15), ey * f(1.6 / 19), ez * f(1.6 / 19)
        nz = 20
    with gpu.HipGridGwn() as g:
        g.set_scene(v, idx, bnd)
        got = g.eval_grid(start, dx, dy, dz, nx, ny, nz)
    pts = lattice(start, dx, dy, dz, nx, ny, nz)
    a = [dx, dy, dz][int(np.argmin([nx, ny, nz]))]  # the evaluator's integration axis: fewest cells, ties to the first
    x0 = a / f(np.sqrt(f(a[0] * a[0] + a[1] * a[1]) + f(a[2] * a[2])))
    assert np.count_nonzero(x0) == 1 and np.abs(x0).max() == 1.0
    with gpu.HipMesh(v, idx) as mesh:
        batch = gpu.eval_gwnr_mesh_batch(mesh, pts, x0, bnd)
        cnt = gpu.eval_gwnr_mesh_batch_integer(mesh, pts, x0)
    ref_cnt = oracle.naive_count(v, idx, pts, x0, np.float32)
    assert np.array_equal(cnt, ref_cnt)
    assert np.array_equal(got, batch)
    assert np.count_nonzero(ref_cnt) > 0
    if case == "sheets":
        assert np.abs(ref_cnt).max() >= 65  # rows whose triangles overflow a row's list (64)


def test_grid_rejects_bad_arguments(gpu):
    with gpu.HipGridGwn() as g:
        with pytest.raises(gpu.AgwnError):
            g.eval_grid([0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], 2, 2, 2)  # no scene yet
        g.set_scene(rc.CUBE_V, rc.CUBE_CLOSED, None)
        with pytest.raises(gpu.AgwnError):
            g.eval_grid([0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], 0, 2, 2)


def test_cpp_plugin_drop_in(gpu):
    """tests/cpp/test_plugin: the reference's own test cases compiled against its API with the HIP backend swapped
    in (HipDispatcher, HipMeshIntersector, RobustMeshGwn alias, HipGridGwn)."""
    exe = ROOT / "tests" / "cpp" / "test_plugin"
    if not exe.exists():
        pytest.skip("tests/cpp/test_plugin not built (needs the reference headers at build time)")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout
