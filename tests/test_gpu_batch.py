"""Batched evaluation of many meshes with uneven query counts (HipMeshBatch / agwn_batch_*) on the GPU.

Bars: every query's full GWN equals the single-mesh path's (HipMesh.full, FP32, same x0) bit for bit -- the batch
runs the same per-query arithmetic in persistent launches; graph replays reproduce the first result; a replay after
the queries moved farther out (same buffers) re-prepares and stays exact; the autotuner reports a time for every
fractional kernel shape and leaf-bucket size, and every shape and size gives the same bits. Meshes include a closed
one (no boundary), a mesh without triangles and a mesh without queries.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from antipodal_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def make_batch(k: int, seed: int, qmin: int = 1):
    rng = np.random.default_rng(seed)
    meshes, queries = [], []
    for m in range(k):
        kind = "sphere" if m % 2 == 0 else "torus"
        ntri = int(math.exp(rng.uniform(math.log(100), math.log(20000))))
        holes = 0 if m == 3 else int(rng.integers(1, 40))
        v, idx = mg.open_mesh(kind, ntri, holes, seed=seed * 1000 + m)
        if m == 5:
            idx = np.zeros(0, dtype=np.int32)  # no triangles
        v = v.astype(np.float32)
        b = mg_boundary(v, idx)
        nq = 0 if m == 7 else int(math.exp(rng.uniform(math.log(qmin), math.log(20000))))
        lo, hi = mg.expanded_bbox(v)
        q = (lo + (hi - lo) * rng.random((nq, 3))).astype(np.float32)
        meshes.append((v, idx, b))
        queries.append(q)
    return meshes, queries


def mg_boundary(v, idx):
    import antipodal_amd as agwn

    return agwn.build_boundary_segments(v, idx) if len(idx) else np.zeros((0, 7), np.float32)


def test_batch_matches_single_mesh_path(gpu):
    import torch

    meshes, queries = make_batch(40, seed=5)
    x0 = mg.example_x0(np.float32)
    q_off = np.concatenate([[0], np.cumsum([len(q) for q in queries])]).astype(np.int64)
    pos = torch.from_numpy(np.concatenate(queries)).cuda()
    with gpu.HipMeshBatch(meshes) as batch:
        out = batch.eval(pos, q_off, x0)
        info = batch.info()
        print("batch:", info, "queries", int(q_off[-1]))
        assert len(info["frac_table"]) == gpu.HipMeshBatch.FRAC_SHAPES and info["queries_per_lane"] in (1, 2, 4)
        assert len(info["leaf_table"]) == gpu.HipMeshBatch.LEAF_SIZES and info["leaf_size"] in (1, 2, 4, 8, 16)
        got = out.cpu().numpy()
        for m, (v, idx, b) in enumerate(meshes):
            if q_off[m + 1] == q_off[m]:
                continue
            p = pos[q_off[m]: q_off[m + 1]].contiguous()
            want = torch.empty(len(p), dtype=torch.float32, device="cuda")
            if len(idx):
                with gpu.HipMesh(v, idx, dtype=np.float32) as mesh:
                    seg = torch.from_numpy(b).cuda() if len(b) else None
                    mesh.full(p, x0, seg, out=want)
                w = want.cpu().numpy()
            else:
                w = np.zeros(len(p), dtype=np.float32)
            np.testing.assert_array_equal(got[q_off[m]: q_off[m + 1]], w, err_msg=f"mesh {m}")
        # replay of the captured graph
        out2 = torch.full_like(out, -7.0)
        batch.eval(pos, q_off, x0, out=out2)
        assert torch.equal(out, out2)
        # same buffers, queries moved 3x farther out: the prepared states are re-checked and rebuilt
        pos.mul_(3.0)
        batch.eval(pos, q_off, x0, out=out2)
        m = 1
        p = pos[q_off[m]: q_off[m + 1]].contiguous()
        v, idx, b = meshes[m]
        with gpu.HipMesh(v, idx, dtype=np.float32) as mesh:
            want = torch.empty(len(p), dtype=torch.float32, device="cuda")
            mesh.full(p, x0, torch.from_numpy(b).cuda(), out=want)
        assert torch.equal(out2[q_off[m]: q_off[m + 1]], want)


def test_batch_rejects_general_weights(gpu):
    v, idx = mg.icosphere_cap_mesh()
    b = mg_boundary(v.astype(np.float32), idx)
    b[:, 6] = 0.5
    with pytest.raises(gpu.AgwnError, match="integer boundary weights"):
        gpu.HipMeshBatch([(v, idx, b)])


def test_sorted_batch_counts_match_single_mesh_path(gpu):
    """A whole batch of 2^16 queries or more sorts its rays for the integer term (per mesh, by the projected origin)
    and counts them in packets: every query's full GWN still equals the single-mesh path's bit for bit, on replay
    too."""
    import torch

    meshes, queries = make_batch(24, seed=9, qmin=4000)
    x0 = mg.example_x0(np.float32)
    q_off = np.concatenate([[0], np.cumsum([len(q) for q in queries])]).astype(np.int64)
    assert q_off[-1] >= 1 << 16
    pos = torch.from_numpy(np.concatenate(queries)).cuda()
    with gpu.HipMeshBatch(meshes) as batch:
        out = batch.eval(pos, q_off, x0)
        got = out.cpu().numpy()
        for m, (v, idx, b) in enumerate(meshes):
            if q_off[m + 1] == q_off[m] or len(idx) == 0:
                assert np.all(got[q_off[m]: q_off[m + 1]] == 0.0) or len(idx) > 0
                continue
            p = pos[q_off[m]: q_off[m + 1]].contiguous()
            want = torch.empty(len(p), dtype=torch.float32, device="cuda")
            with gpu.HipMesh(v, idx, dtype=np.float32) as mesh:
                seg = torch.from_numpy(b).cuda() if len(b) else None
                mesh.full(p, x0, seg, out=want)
            np.testing.assert_array_equal(got[q_off[m]: q_off[m + 1]], want.cpu().numpy(), err_msg=f"mesh {m}")
        out2 = torch.full_like(out, -7.0)
        batch.eval(pos, q_off, x0, out=out2)
        assert torch.equal(out, out2)


@pytest.mark.parametrize("qmin", [1, 4000])
def test_every_tuning_choice_gives_the_same_bits(gpu, qmin):
    """Each fractional kernel shape (queries per lane x block x LDS stage) and each leaf-bucket size, set instead of
    tuned, reproduces the tuned evaluation bit for bit: the shapes change which lanes hold which queries and how many
    stream points a barrier stages, the buckets which triangles are tested without their own boxes -- not the
    per-query arithmetic. qmin = 4000: a batch of 2^16 queries or more (sorted rays, packet traversal); qmin = 1 has
    small meshes (per-lane traversal when the whole batch is small)."""
    import torch

    meshes, queries = make_batch(24 if qmin > 1 else 16, seed=13, qmin=qmin)
    x0 = mg.example_x0(np.float32)
    q_off = np.concatenate([[0], np.cumsum([len(q) for q in queries])]).astype(np.int64)
    pos = torch.from_numpy(np.concatenate(queries)).cuda()
    B = gpu.HipMeshBatch
    with B(meshes) as batch:
        ref = batch.eval(pos, q_off, x0).clone()
        info = batch.info()
        print(f"tuned: shape {info['frac_shape']} leaf {info['leaf_size']}; queries {int(q_off[-1])}", flush=True)
        out = torch.empty_like(ref)
        choices = [(v, 0) for v in range(B.FRAC_SHAPES)] + [(0, l) for l in range(1, B.LEAF_SIZES)]
        for v, l in choices:
            print(f"shape {B.frac_shape(v)} leaf {B.leaf_size(l)}", flush=True)
            batch.set_tuning(v, l)
            out.fill_(-7.0)
            batch.eval(pos, q_off, x0, out=out)
            assert torch.equal(out, ref), f"shape {v}, leaf index {l}: {int((out != ref).sum())} differ"
        with pytest.raises(gpu.AgwnError, match="out of range"):
            batch.set_tuning(B.FRAC_SHAPES, 0)
