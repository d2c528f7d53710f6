"""Range of the FP32 fractional kernels: scale invariance and far queries, on the GPU through the C ABI.

The GWN is invariant under a common scaling of the mesh and the queries, and scaling float data by a power of two is
exact, so every FP32 kernel must give bit-identical results on the config-1 cap and its queries scaled by 2^-34 and
2^+34 (the mathematics fixes this; a kernel whose products overflow or reach subnormals at either scale fails it):
the chain kernel (unit weights), the per-segment kernel (general weights), the grid evaluator and the batched
launch. Far queries (10 to 1e13 away from a unit-size mesh) must stay finite and within 1e-5 of the FP64
evaluation of the same inputs (the reference's own far-query range test, test_gwn_mesh_robust.cc:374-395, runs to
1e9 with the robust evaluator; the plain FP32 path of common.hh:181-189 is evaluated there too).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden
from antipodal_amd import meshgen as mg

pytestmark = pytest.mark.gpu

SCALES = [2.0**-34, 2.0**34]


def _cap32():
    g = load_golden("c1_icosphere_cap")
    return g["verts"].astype(np.float32), g["idx"], g["x0_f32"].astype(np.float32)


@pytest.mark.parametrize("weights", ["unit", "general"])
def test_fractional_and_full_bit_identical_under_power_of_two_scaling(gpu, weights):
    verts, idx, x0 = _cap32()
    seg = gpu.build_boundary_segments(verts, idx)
    if weights == "general":
        seg = seg.copy()
        seg[:, 6] = seg[:, 6] * np.float32(0.75) + np.float32(0.125) * (np.arange(len(seg)) % 3)
    pos = mg.uniform_points(0, 100_000, [-1.5] * 3, [1.5] * 3).astype(np.float32)
    frac0 = gpu.eval_gwnr_mesh_batch_fractional(pos, x0, seg)
    with gpu.HipMesh(verts, idx) as mesh:
        full0 = gpu.eval_gwnr_mesh_batch(mesh, pos, x0, seg)
    assert np.all(np.isfinite(frac0))
    for sc in SCALES:
        print(f"case scale {sc:g} weights={weights}", flush=True)
        f = np.float32(sc)
        s_seg = seg.copy()
        s_seg[:, :6] *= f
        frac = gpu.eval_gwnr_mesh_batch_fractional(pos * f, x0, s_seg)
        assert np.array_equal(frac, frac0), f"fractional: {int((frac != frac0).sum())} values differ at {sc:g}"
        with gpu.HipMesh(verts * f, idx) as mesh:
            full = gpu.eval_gwnr_mesh_batch(mesh, pos * f, x0, s_seg)
        assert np.array_equal(full, full0), f"full: {int((full != full0).sum())} values differ at {sc:g}"


def test_grid_bit_identical_under_power_of_two_scaling(gpu):
    verts, idx, _ = _cap32()
    seg = gpu.build_boundary_segments(verts, idx)
    start = np.array([-1.25, -1.3, -1.2], np.float32)
    dx, dy, dz = (np.eye(3, dtype=np.float32) * np.float32(0.05))
    with gpu.HipGridGwn() as g:
        ref = g.set_scene(verts, idx, seg).eval_grid(start, dx, dy, dz, 50, 52, 48)
    for sc in SCALES:
        f = np.float32(sc)
        print(f"case grid scale {sc:g}", flush=True)
        s_seg = seg.copy()
        s_seg[:, :6] *= f
        with gpu.HipGridGwn() as g:
            got = g.set_scene(verts * f, idx, s_seg).eval_grid(start * f, dx * f, dy * f, dz * f, 50, 52, 48)
        assert np.array_equal(got, ref), f"{int((got != ref).sum())} cells differ at {sc:g}"


def test_batch_bit_identical_under_power_of_two_scaling(gpu):
    import torch

    verts, idx, x0 = _cap32()
    seg = gpu.build_boundary_segments(verts, idx)
    tv, ti = mg.open_mesh("torus", 4000, 12, seed=11)
    tv = tv.astype(np.float32)
    tseg = gpu.build_boundary_segments(tv, ti)
    base = [(verts, idx, seg), (tv, ti, tseg)]
    qs = [mg.uniform_points(1, 20_000, [-1.5] * 3, [1.5] * 3).astype(np.float32),
          mg.uniform_points(2, 30_000, *mg.expanded_bbox(tv)).astype(np.float32)]
    # each mesh three times: unscaled, 2^-34, 2^34 (with its queries scaled alike)
    meshes, queries = [], []
    for sc in [1.0] + SCALES:
        f = np.float32(sc)
        for (v, i, b), q in zip(base, qs):
            bb = b.copy()
            bb[:, :6] *= f
            meshes.append((v * f, i, bb))
            queries.append(q * f)
    q_off = np.concatenate([[0], np.cumsum([len(q) for q in queries])]).astype(np.int64)
    pos = torch.from_numpy(np.concatenate(queries)).cuda()
    print(f"case batch scaled: {len(meshes)} meshes, {int(q_off[-1])} queries", flush=True)
    with gpu.HipMeshBatch(meshes) as batch:
        got = batch.eval(pos, q_off, x0).cpu().numpy()
    parts = [got[q_off[m]:q_off[m + 1]] for m in range(len(meshes))]
    for m in range(2, len(meshes)):
        assert np.array_equal(parts[m], parts[m % 2]), f"mesh {m}: {int((parts[m] != parts[m % 2]).sum())} differ"
    # and the unscaled meshes equal the single-mesh path
    for m in range(2):
        v, i, b = meshes[m]
        with gpu.HipMesh(v, i) as mesh:
            one = gpu.eval_gwnr_mesh_batch(mesh, queries[m], x0, b)
        assert np.array_equal(parts[m], one)


def test_far_queries_finite_and_accurate(gpu, oracle):
    verts, idx, x0 = _cap32()
    seg = gpu.build_boundary_segments(verts, idx)
    rng = np.random.default_rng(7)
    dists = 10.0 ** np.arange(1, 14)
    d = rng.normal(size=(len(dists) * 64, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = (d * np.repeat(dists, 64)[:, None]).astype(np.float32)
    print(f"case far queries: {len(pos)} at 1e1..1e13", flush=True)
    got = gpu.eval_gwnr_mesh_batch_fractional(pos, x0, seg)
    assert np.all(np.isfinite(got))
    # FP64 evaluation of the same (float) inputs, in extended precision
    ext = oracle.frac_ext(pos.astype(np.float64), seg.astype(np.float64), x0.astype(np.float64))
    err = np.abs(got.astype(np.float64) - ext)
    worst = [float(err[k * 64:(k + 1) * 64].max()) for k in range(len(dists))]
    print("worst error per distance:", dict(zip([f"{x:.0e}" for x in dists], [f"{w:.1e}" for w in worst])))
    assert err.max() < 1e-5
    # a unit-size boundary seen from far subtends almost nothing: the fractional term is ~0 unless the pole ray
    # through the query meets the cap (then ~ -count); either way it is within 1e-5 of an integer there
    far = np.repeat(dists, 64) >= 1e6
    assert np.max(np.abs(got[far] - np.round(got[far]))) < 1e-5
