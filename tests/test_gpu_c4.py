"""The batched uneven workload at its full size: 4096 seeded open meshes (noisy spheres and tori, 100..2e5 triangles
log-uniform, 1..200 holes) with 1e3..1e6 queries each (592,815,454 in all, uniform in each mesh's 10 %-expanded box),
FP32, x0 = normalize(0.31, 0.57, 0.76) -- one HipMeshBatch evaluation (persistent launches, autotuned shape, graph).

Bars: a seeded sample of 64 meshes is bit-identical to the single-mesh path (eval_gwnr_mesh_batch with the same mesh,
boundary, queries and x0: the per-query arithmetic is the same); the smallest meshes of the batch, on up to 2000 of
their queries, equal NaiveIntersector<float>'s counts (the oracle, bit-exact) plus the extended-precision fractional
term of the same inputs within 1e-5 (the FP32 bar); a graph replay reproduces the first evaluation.
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from antipodal_amd import meshgen as mg

pytestmark = pytest.mark.gpu

K = 4096


@pytest.mark.timeout(900)
def test_c4_batch_of_4096_meshes(gpu, oracle):
    import torch

    workers = min(16, os.cpu_count() or 1)
    print(f"case c4: generating {K} meshes", flush=True)
    geo, qn, boxes = mg.batch_workload(K, seed=7, workers=workers)
    with ThreadPoolExecutor(workers) as ex:
        segs = list(ex.map(lambda g: gpu.build_boundary_segments(*g), geo))
    meshes = [(v, idx, b) for (v, idx), b in zip(geo, segs)]
    q_off = np.concatenate([[0], np.cumsum(qn)]).astype(np.int64)
    nq = int(q_off[-1])
    ntri = sum(len(m[1]) // 3 for m in meshes)
    print(f"case c4: {K} meshes, {ntri} triangles, {sum(len(s) for s in segs)} segments, {nq} queries", flush=True)
    assert 5.5e8 < nq < 6.5e8 and 7e7 < ntri < 1e8
    # queries on the device, uniform in each mesh's box (seeded)
    dev = torch.device("cuda:0")
    mid = torch.repeat_interleave(torch.arange(K, device=dev), torch.tensor(qn, device=dev))
    lo_k = torch.from_numpy(np.array([b[0] for b in boxes], np.float32)).to(dev)
    hi_k = torch.from_numpy(np.array([b[1] for b in boxes], np.float32)).to(dev)
    gen = torch.Generator(device=dev).manual_seed(11)
    pos = torch.rand((nq, 3), device=dev, generator=gen)
    pos.mul_(hi_k[mid] - lo_k[mid]).add_(lo_k[mid])
    del mid
    x0 = mg.example_x0(np.float32)
    with gpu.HipMeshBatch(meshes) as batch:
        out = batch.eval(pos, q_off, x0)
        info = batch.info()
        print(f"case c4: tuned {info}", flush=True)
        assert len(info["frac_table"]) == gpu.HipMeshBatch.FRAC_SHAPES
        assert len(info["leaf_table"]) == gpu.HipMeshBatch.LEAF_SIZES
        first = out.clone()
        batch.eval(pos, q_off, x0, out=out)  # graph replay
        torch.cuda.synchronize()
        assert torch.equal(out, first)
    got = first.cpu().numpy()
    del first, out
    assert np.all(np.isfinite(got))
    # 64 seeded meshes against the single-mesh path, bit for bit
    sample = np.random.default_rng(123).choice(K, 64, replace=False)
    for m in sample:
        v, idx, b = meshes[m]
        q = pos[q_off[m]:q_off[m + 1]].cpu().numpy()
        with gpu.HipMesh(v, idx) as mesh:
            one = gpu.eval_gwnr_mesh_batch(mesh, q, x0, b)
        part = got[q_off[m]:q_off[m + 1]]
        assert np.array_equal(part, one), f"mesh {m}: {int((part != one).sum())} of {len(one)} differ"
    # the smallest meshes against the oracle: NaiveIntersector<float> counts + extended-precision fractional term
    small = np.argsort([len(m[1]) for m in meshes], kind="stable")[:4]
    for m in small:
        v, idx, b = meshes[m]
        q = pos[q_off[m]:q_off[m] + min(2000, qn[m])].cpu().numpy()
        cnt = oracle.naive_count(v, idx, q, x0, dtype=np.float32)
        frac = oracle.frac_ext(q.astype(np.float64), b.astype(np.float64), x0.astype(np.float64))
        part = got[q_off[m]:q_off[m] + len(q)].astype(np.float64)
        err = np.abs(part - (cnt + frac))
        print(f"case c4 oracle: mesh {m} ({len(idx) // 3} triangles, {len(b)} segments): max error {err.max():.2e}")
        assert err.max() < 1e-5
