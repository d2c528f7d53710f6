"""Reference parity on the 1M-triangle holed torus, through the C ABI on the GPU, against the reference's own outputs
(tests/golden/torus_ref.npz, written by the reference's code: NaiveIntersector, eval_gwnr_mesh_batch_fractional,
eval_gwnr_mesh_batch, build_boundary_segments).

Queries: 16,384 seeded points of the 256^3 lattice over the mesh's 10 %-expanded box and 4,096 points at
|N(0, 1e-3)| off its surface, in FP32 and FP64, x0 = normalize(0.31, 0.57, 0.76).

Bars: integer counts bit-exact (intersector.hh:61-100); FP64 fractional and full GWN within 1e-10 of the reference's
(its sum runs in its boundary's unordered_map order); FP32 fractional within 1e-5 of the extended-precision
evaluation of the same FP32 inputs (the reference's own FP32 sum is off by up to 2e-2 on the near-surface points, so
it is not the bar); every inside/outside label (GWN > 0.5) equal to the reference's in the same precision; the GPU
boundary extraction equal to the reference's multiset (SHA-256 of the sorted segments).
"""
from __future__ import annotations

import hashlib

import numpy as np
import pytest

from conftest import load_golden
from antipodal_amd import meshgen as mg
from oracle import sort_segments

pytestmark = pytest.mark.gpu

CASES = [("f64", np.float64), ("f32", np.float32)]


def sha(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="module")
def torus():
    g = load_golden("torus_ref")
    verts, idx = mg.holed_torus_mesh()
    assert sha(verts, idx) == g["mesh_sha"].item().decode(), "mesh generator drifted"
    lo, hi = mg.expanded_bbox(verts)
    near = mg.offset_surface_points(verts, idx, int(g["n_near"]), seed=5, stream=50)
    pos = {}
    for s, dt in CASES:
        pos[s] = np.concatenate([mg.grid_points_at(lo, hi, int(g["grid_n"]), g["flat"], dtype=dt), near.astype(dt)])
        assert sha(pos[s]) == g[f"pos_sha_{s}"].item().decode(), "query generator drifted"
    return g, verts, idx, pos


def labels(g, s, n):
    return np.unpackbits(g[f"label_{s}"])[:n].astype(bool)


@pytest.mark.parametrize("s,dt", CASES)
def test_torus_counts_bit_exact(gpu, torus, s, dt):
    g, verts, idx, pos = torus
    print(f"case torus_ref counts {s}: {len(pos[s])} queries", flush=True)
    with gpu.HipMesh(verts.astype(dt), idx) as mesh:
        cnt = gpu.eval_gwnr_mesh_batch_integer(mesh, pos[s], mg.example_x0(dt))
    ref = g[f"count_{s}"].astype(np.int32)
    assert np.array_equal(cnt, ref), f"{int((cnt != ref).sum())} counts differ"


@pytest.mark.parametrize("s,dt", CASES)
def test_torus_fractional_full_and_labels(gpu, oracle, torus, s, dt):
    g, verts, idx, pos = torus
    p = pos[s]
    x0 = mg.example_x0(dt)
    seg = gpu.build_boundary_segments(verts.astype(dt), idx)
    print(f"case torus_ref fractional/full {s}: {len(p)} queries, {len(seg)} segments", flush=True)
    frac = gpu.eval_gwnr_mesh_batch_fractional(p, x0, seg)
    with gpu.HipMesh(verts.astype(dt), idx) as mesh:
        full = gpu.eval_gwnr_mesh_batch(mesh, p, x0, seg)
    if dt == np.float64:
        ef = np.max(np.abs(frac - g["frac_f64"]))
        ew = np.max(np.abs(full - g["full_f64"]))
        print(f"FP64 vs reference: fractional {ef:.2e}, full {ew:.2e}")
        assert ef < 1e-10 and ew < 1e-10
    else:
        ext = oracle.frac_ext(p.astype(np.float64), seg.astype(np.float64), x0.astype(np.float64))
        e = np.abs(frac.astype(np.float64) - ext)
        er = np.abs(frac.astype(np.float64) - g["frac_f32"])
        print(f"FP32 vs extended precision: max {e.max():.2e} (grid {e[:16384].max():.2e}, near {e[16384:].max():.2e});"
              f" vs the reference's FP32 sum: max {er.max():.2e}")
        assert e.max() < 1e-5
    lab = full > 0.5
    ref = labels(g, s, len(p))
    assert np.array_equal(lab, ref), f"{int((lab != ref).sum())} labels differ"


@pytest.mark.parametrize("s,dt", CASES)
def test_torus_boundary_multiset_on_device(gpu, torus, s, dt):
    import torch

    g, verts, idx, _ = torus
    vt = torch.from_numpy(np.ascontiguousarray(verts, dtype=dt)).cuda()
    it = torch.from_numpy(np.ascontiguousarray(idx, dtype=np.int32)).cuda()
    seg = gpu.build_boundary_segments(vt, it).cpu().numpy()
    assert len(seg) == int(g[f"nseg_{s}"])
    assert sha(sort_segments(seg)) == g[f"boundary_sha_{s}"].item().decode()
