"""The learned surrogate at the configuration's size, and what it gains, on the GPU.

C5 (config-1 mesh): the open icosphere cap, a per-mesh network trained on 2^24 exact GWN samples (70 % within 0.05
of the surface, 30 % uniform), inference over the 512^3 lattice of its 10 %-expanded box (134,217,728 points),
labels GWN > 0.5 with the band / near-surface / audit points sent to the exact path. Bars: every one of the 134 M
labels equals the plain path's (eval_grid of the same lattice; the exact path of gwn_mesh.hh:201-219), no audited
point disagrees, and 4096 seeded lattice points carry the reference's own labels (tests/golden/c5_cap.npz, written
by the reference's eval_gwnr_mesh_batch in FP32 and FP64). The fallback fraction and the time against plain are
reported: the cap's 36 boundary segments make its plain evaluation cheap, so no gain is asserted there.

Gain: the 1M-triangle holed torus (31,376 boundary segments) on its 256^3 lattice with a reduced fit (2^22 training
samples): 0 mislabels against plain on every point, and the labelling step faster than the plain evaluation.

Both use the benchmark's fit settings: 16 epochs, batch 16384, lr 3e-3, top frequency F = 64, mask voxels of a
quarter of the mask radius, x0 = +x (the lattice's integration axis). The mask radius is the benchmark's 0.1 % of the
box diagonal on the torus, whose box is flat, and the library's default 0.25 % on the cap, whose nearly cubic box
would need 1.2e10 voxels at 0.1 % (above the mask's 2^33 limit).
"""
from __future__ import annotations

import time

import numpy as np
import pytest

from conftest import load_golden
from antipodal_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def _setup(gpu, verts, idx, nres):
    import torch

    v32 = verts.astype(np.float32)
    seg32 = gpu.build_boundary_segments(v32, idx)
    lo, hi = mg.expanded_bbox(verts)
    start, dx, dy, dz = mg.grid_axes(lo, hi, nres)
    grid = gpu.HipGridGwn().set_scene(v32, idx, seg32)
    mesh = gpu.HipMesh(v32, idx, dtype=np.float32)
    nq = nres ** 3
    w = torch.empty(nq, dtype=torch.float32, device="cuda")
    grid.eval_grid(start, dx, dy, dz, nres, nres, nres, out=w)  # warm-up (per-direction BVH state)
    torch.cuda.synchronize()
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        grid.eval_grid(start, dx, dy, dz, nres, nres, nres, out=w)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    exact = (w > 0.5).to(torch.int8)
    del w
    return v32, seg32, lo, hi, (start, dx, dy, dz), grid, mesh, exact, 1e3 * float(np.median(ts))


def _fit(gpu, v32, idx, seg32, lo, hi, mesh, n_train, n_val, eps_frac):
    diag = float(np.linalg.norm(hi - lo))
    x0 = np.array([1.0, 0.0, 0.0], dtype=np.float32)
    return gpu.fit_surrogate(v32, idx, seg32, x0, lo.astype(np.float32), hi.astype(np.float32), n_train=n_train,
                             n_val=n_val, epochs=16, batch=16384, lr=3e-3, seed=0, mesh=mesh, max_freq=64.0,
                             eps=eps_frac * diag, voxel=0.25 * eps_frac * diag)


def _label_steps(net, grid, axes, nres, delta, steps):
    import torch

    lab = torch.empty(nres ** 3, dtype=torch.int8, device="cuda")
    net.label_grid(grid, *axes, nres, nres, nres, delta, audit_every=1024, audit_seed=99, labels=lab)  # warm-up
    torch.cuda.synchronize()
    ts, sts = [], []
    for k in range(steps):
        t0 = time.perf_counter()
        _, st = net.label_grid(grid, *axes, nres, nres, nres, delta, audit_every=1024, audit_seed=k, labels=lab)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
        sts.append(st)
    return lab, sts, 1e3 * float(np.median(ts))


@pytest.mark.timeout(900)
def test_c5_cap_512_grid_zero_mislabels(gpu):
    verts, idx = mg.icosphere_cap_mesh()
    nres = 512
    print(f"case c5: cap {len(idx) // 3} triangles, {nres}^3 lattice", flush=True)
    v32, seg32, lo, hi, axes, grid, mesh, exact, plain_ms = _setup(gpu, verts, idx, nres)
    print(f"case c5: plain {plain_ms:.1f} ms; fitting on 2^24 samples", flush=True)
    net, delta, fit = _fit(gpu, v32, idx, seg32, lo, hi, mesh, 1 << 24, 1 << 23, 0.0025)
    print(f"case c5: fit {fit['fit_s']:.1f} s (train {fit['train_s']:.1f} s), n_train {fit['n_train']}, "
          f"validation wrong {fit['val_wrong']}/{fit['val_points']}, delta {delta:.4f}", flush=True)
    assert fit["n_train"] == 1 << 24
    try:
        lab, sts, step_ms = _label_steps(net, grid, axes, nres, delta, 3)
        mislabels = int((lab != exact).sum())
        st = sts[-1]
        print(f"case c5: step {step_ms:.1f} ms vs plain {plain_ms:.1f} ms ({plain_ms / step_ms:.2f}x), "
              f"fallback fraction {st['fallback_fraction']:.4f}, stats {st}, mislabels {mislabels}", flush=True)
        assert mislabels == 0
        assert all(s["audit_disagree"] == 0 for s in sts)
        g = load_golden("c5_cap")
        flat = g["flat"]
        sub = lab.cpu().numpy()[flat].astype(bool)
        for s in ("f32", "f64"):
            ref = np.unpackbits(g[f"label_{s}"])[:len(flat)].astype(bool)
            assert np.array_equal(sub, ref), f"{int((sub != ref).sum())} labels differ from the reference's {s}"
    finally:
        net.close()
        grid.close()
        mesh.close()


@pytest.mark.timeout(900)
def test_surrogate_gain_on_torus_256(gpu):
    verts, idx = mg.holed_torus_mesh()
    nres = 256
    print(f"case gain: torus {len(idx) // 3} triangles, {nres}^3 lattice", flush=True)
    v32, seg32, lo, hi, axes, grid, mesh, exact, plain_ms = _setup(gpu, verts, idx, nres)
    print(f"case gain: plain {plain_ms:.1f} ms; fitting on 2^22 samples", flush=True)
    net, delta, fit = _fit(gpu, v32, idx, seg32, lo, hi, mesh, 1 << 22, 1 << 22, 0.0010)
    print(f"case gain: fit {fit['fit_s']:.1f} s, delta {delta:.4f}", flush=True)
    try:
        lab, sts, step_ms = _label_steps(net, grid, axes, nres, delta, 5)
        mislabels = int((lab != exact).sum())
        st = sts[-1]
        print(f"case gain: surrogate + fallback {step_ms:.2f} ms vs plain {plain_ms:.1f} ms = "
              f"{plain_ms / step_ms:.1f}x, fallback fraction {st['fallback_fraction']:.4f}, mislabels {mislabels}; "
              f"one-off fit {fit['fit_s']:.1f} s = {1e3 * fit['fit_s'] / (plain_ms - step_ms):.0f} evaluations' saving",
              flush=True)
        assert mislabels == 0
        assert all(s["audit_disagree"] == 0 for s in sts)
        assert step_ms < plain_ms
        # the gain comes from the network labelling most points: the band must stay narrow
        assert st["fallback_fraction"] < 0.2, (st, fit.get("val_exact_disagree"), fit.get("val_max_wrong"))
    finally:
        net.close()
        grid.close()
        mesh.close()
