"""The far-field decision of surrogate labelling (frac_decide), on the GPU.

Every query is routed (delta = +inf), so each label comes from the exact stage: integer term, then the far-field
decision where the boundary consists of closed loops, then the full FP32 boundary sum for the points it leaves open.
Bar: labels identical to those of the full evaluation (HipMesh.full > 0.5) on every query, for queries next to the
surface, around the boundary's rims (where the GWN is near 0.5 and the decision must defer) and uniform in the box;
and for a boundary whose walks are not closed (no decision stage: the full sum for every point).
"""
from __future__ import annotations

import numpy as np
import pytest

from antipodal_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def queries(v, idx, seg, n, seed):
    rng = np.random.default_rng(seed)
    k = n // 3
    sp, nrm = mg.surface_points(v, idx, k, seed=seed)
    near = sp + nrm * rng.normal(0.0, 0.004, k)[:, None]
    mid = 0.5 * (seg[:, 0:3] + seg[:, 3:6])
    rim = mid[rng.integers(0, len(mid), k)] + rng.normal(0.0, 0.03, (k, 3))
    lo, hi = mg.expanded_bbox(v)
    uni = lo + (hi - lo) * rng.random((n - 2 * k, 3))
    return np.concatenate([near, rim, uni]).astype(np.float32)


@pytest.fixture(scope="module")
def torus(gpu):
    v, idx = mg.holed_torus_mesh()
    v32 = v.astype(np.float32)
    seg = gpu.build_boundary_segments(v32, idx)
    return v32, idx, seg


def exact_and_routed(gpu, v32, idx, seg, q, x0):
    import torch

    p = torch.from_numpy(q).cuda()
    segt = torch.from_numpy(np.ascontiguousarray(seg)).cuda()
    lo, hi = mg.expanded_bbox(v32)
    with gpu.HipMesh(v32, idx, dtype=np.float32) as mesh:
        full = torch.empty(len(q), dtype=torch.float32, device="cuda")
        mesh.full(p, x0, segt, out=full)
        exact = (full > 0.5).to(torch.int8)
        net = gpu.GwnSurrogate(lo.astype(np.float32), hi.astype(np.float32), seed=3)
        lab, st = net.label(mesh, p, x0, segt, float("inf"))
        torch.cuda.synchronize()
    return exact, lab, st, full


def test_decision_labels_equal_full_evaluation(gpu, torus):
    import torch

    v32, idx, seg = torus
    q = queries(v32, idx, seg, 300_000, seed=4)
    x0 = mg.example_x0(np.float32)
    exact, lab, st, full = exact_and_routed(gpu, v32, idx, seg, q, x0)
    assert st["routed"] == len(q)
    bad = int((lab != exact).sum())
    near_half = int(((full - 0.5).abs() < 0.05).sum())
    print(f"decision: {len(q)} queries, {near_half} with |GWN - 0.5| < 0.05, mislabels {bad}")
    assert bad == 0
    assert torch.equal(lab, exact)


def test_open_walks_take_the_full_sum(gpu, torus):
    v32, idx, seg = torus
    part = np.ascontiguousarray(seg[:-7])  # some loops lose segments: their walks no longer close
    q = queries(v32, idx, part, 30_000, seed=5)
    x0 = mg.example_x0(np.float32)
    exact, lab, st, _ = exact_and_routed(gpu, v32, idx, part, q, x0)
    assert int((lab != exact).sum()) == 0


@pytest.mark.parametrize("kind,ntri,holes,seed", [("cap", 0, 0, 0), ("sphere", 3000, 5, 11), ("torus", 20000, 40, 12)])
def test_decision_on_other_boundaries(gpu, kind, ntri, holes, seed):
    """Other loop trees: the cap's single 36-segment walk (arcs of 16, 16 and 4 segments and a skeleton), and seeded
    open meshes whose holes give walks of every length -- among them walks of 17 to 31 segments, whose skeleton goes
    out to the split point and back (no fan area, bound 0). Labels equal the full evaluation's on every query."""
    if kind == "cap":
        v, idx = mg.icosphere_cap_mesh()
    else:
        v, idx = mg.open_mesh(kind, ntri, holes, seed=seed)
    v32 = np.asarray(v, dtype=np.float32)
    seg = gpu.build_boundary_segments(v32, idx)
    q = queries(v32, idx, seg, 60_000, seed=seed + 1)
    x0 = mg.example_x0(np.float32)
    exact, lab, st, _ = exact_and_routed(gpu, v32, idx, seg, q, x0)
    assert st["routed"] == len(q)
    assert int((lab != exact).sum()) == 0
