"""The learned surrogate on the GPU: MFMA forward/backward against the float64 oracle (with the kernels' BF16 rounding
points emulated), Adam, the fused inference kernel, training, and surrogate + exact-fallback labelling.

Bars:
* gradient: per parameter block, |g_gpu - g_oracle| / |g_oracle| <= 1e-2 (2-norms) -- the kernels accumulate in FP32
  and round intermediate activations to BF16 exactly where the oracle does, so what remains is FP32 summation order
  and the BF16 rounding flips it causes; predictions within 1e-3 absolute, loss within 1e-3 relative;
* fused inference == the training kernels' forward pass (1e-6);
* Adam's first step: every weight moves by -lr g/(|g| + eps) of the GPU gradient (float32 rounding);
* labels: 0 mislabels against the exact path on every point labelled; routed points take the exact label bit for bit.
"""
from __future__ import annotations

import time

import numpy as np
import pytest

from antipodal_amd import meshgen as mg
from oracle import mlp as om

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    v, idx = mg.icosphere_cap_mesh()
    lo, hi = mg.expanded_bbox(v, 0.10)
    return v, idx, lo.astype(np.float32), hi.astype(np.float32)


def block_errors(g, ref):
    lay = om.LAYOUT
    out = {}
    for l in range(om.LAYERS):
        for name, off, cnt in ((f"W{l}", lay["w"][l], om.IN_DIM[l] * om.H), (f"b{l}", lay["b"][l], om.H)):
            a, b = g[off: off + cnt], ref[off: off + cnt]
            out[name] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
    a, b = g[lay["wo"]: lay["wo"] + om.H], ref[lay["wo"]: lay["wo"] + om.H]
    out["wo"] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
    out["bo"] = float(abs(g[lay["bo"]] - ref[lay["bo"]]) / max(abs(ref[lay["bo"]]), 1e-30))
    return out


def test_grad_matches_oracle(gpu, c1):
    v, idx, lo, hi = c1
    rng = np.random.default_rng(0)
    n = 1000  # not a multiple of 128: exercises the padded rows
    pts = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    y = rng.random(n).astype(np.float32)
    with gpu.GwnSurrogate(lo, hi, seed=3) as net:
        p = net.get_params()
        g, pred, loss = net.grad(pts, y)
    loss_o, g_o, pred_o = om.loss_grad(p.astype(np.float64), pts, y, lo, hi, emulate=True)
    errs = block_errors(g.astype(np.float64), g_o)
    print("gradient block errors:", {k: f"{e:.2e}" for k, e in errs.items()})
    assert max(errs.values()) <= 1e-2, errs
    assert np.max(np.abs(pred - pred_o)) <= 1e-3
    assert abs(loss - loss_o) <= 1e-3 * loss_o
    # and the emulation matters less than the bar: the exact float64 gradient is close too
    _, g_x, _ = om.loss_grad(p.astype(np.float64), pts, y, lo, hi, emulate=False)
    assert np.linalg.norm(g - g_x) <= 0.05 * np.linalg.norm(g_x)


def test_predict_equals_training_forward(gpu, c1):
    import torch

    v, idx, lo, hi = c1
    rng = np.random.default_rng(1)
    n = 3000
    pts = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    with gpu.GwnSurrogate(lo, hi, seed=5) as net:
        _, pred_train, _ = net.grad(pts, np.zeros(n, dtype=np.float32))
        pred = net.predict(torch.from_numpy(pts).cuda()).cpu().numpy()
    np.testing.assert_allclose(pred, pred_train, atol=1e-6, rtol=0)


def test_adam_first_step(gpu, c1):
    import torch

    v, idx, lo, hi = c1
    rng = np.random.default_rng(2)
    n = 512
    pts = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    y = rng.random(n).astype(np.float32)
    lr = 1e-3
    with gpu.GwnSurrogate(lo, hi, seed=7) as net:
        p0 = net.get_params()
        g, _, _ = net.grad(pts, y)
        losses = net.train(torch.from_numpy(pts).cuda(), torch.from_numpy(y).cuda(), batch=n, epochs=1, lr=lr)
        p1 = net.get_params()
    assert np.isfinite(losses).all()
    want = p0.astype(np.float64) - lr * g / (np.abs(g) + 1e-8)
    # the training step sums the batch in another order: coordinates whose gradient is within float32 noise of 0
    # may move differently; all others move by lr to float32 precision
    big = np.abs(g) > 1e-4 * np.max(np.abs(g))
    err = np.abs(p1 - want)
    assert np.max(err[big]) <= 1e-5 * lr + 4 * np.spacing(np.abs(p0[big]).astype(np.float32)).max()
    assert np.mean(err > 2e-2 * lr) < 2e-3


def test_gradient_of_a_large_batch(gpu, c1):
    """A batch of 131,072 samples (the weight-gradient GEMM's batch chunks are then 2048 long: eight LDS stages each)
    made of 32 copies of 4096 samples has the mean-squared-error gradient of those 4096 (one stage per chunk):
    within 1e-4 of the gradient's norm per parameter block (float32 sums in another grouping)."""
    v, idx, lo, hi = c1
    rng = np.random.default_rng(4)
    n = 4096
    pts = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    y = rng.random(n).astype(np.float32)
    with gpu.GwnSurrogate(lo, hi, seed=9) as net:
        g1, p1, l1 = net.grad(pts, y)
        g32, p32, l32 = net.grad(np.tile(pts, (32, 1)), np.tile(y, 32))
    assert np.array_equal(p32.reshape(32, n)[7], p1)  # the forward pass does not depend on the batch
    errs = block_errors(g32.astype(np.float64), g1.astype(np.float64))
    print("large-batch gradient block errors:", {k: f"{e:.1e}" for k, e in errs.items()})
    assert max(errs.values()) <= 1e-4, errs
    assert abs(l32 - l1) <= 1e-5 * l1


@pytest.fixture(scope="module")
def trained_c1(gpu, c1):
    """A network fitted to the small open mesh (icosphere cap): 2^18 training and 2^16 validation samples, top
    frequency 32 (the sphere is smooth at the box scale)."""
    v, idx, lo, hi = c1
    bnd = gpu.build_boundary_segments(v.astype(np.float32), idx)
    x0 = mg.example_x0(np.float32)
    t = time.perf_counter()
    net, delta, info = gpu.fit_surrogate(v.astype(np.float32), idx, bnd, x0, lo, hi, n_train=1 << 18, n_val=1 << 16,
                                         epochs=30, batch=4096, lr=2e-3, seed=11, max_freq=32.0)
    print(f"fit {time.perf_counter() - t:.2f}s", info)
    yield net, delta, info, bnd, x0
    net.close()


def test_training_reduces_loss(trained_c1):
    net, delta, info, bnd, x0 = trained_c1
    L = info["losses"]
    assert all(np.isfinite(L))
    assert L[-1] < 0.2 * L[0], L
    assert info["val_wrong"] < 0.02 * info["val_points"], info
    assert delta < 0.5, info


def test_grid_labels_zero_mislabels(gpu, c1, trained_c1):
    import torch

    v, idx, lo, hi = c1
    net, delta, info, bnd, x0 = trained_c1
    nres = 64
    start, dx, dy, dz = mg.grid_axes(lo, hi, nres)
    with gpu.HipGridGwn() as grid:
        grid.set_scene(v.astype(np.float32), idx, bnd)
        w = torch.empty(nres ** 3, dtype=torch.float32, device="cuda")
        t = time.perf_counter()
        grid.eval_grid(start, dx, dy, dz, nres, nres, nres, out=w)
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t
        exact = (w > 0.5).to(torch.int8)
        t = time.perf_counter()
        lab, st = net.label_grid(grid, start, dx, dy, dz, nres, nres, nres, delta, audit_every=64, audit_seed=1)
        torch.cuda.synchronize()
        t_sur = time.perf_counter() - t
        mislabels = int((lab != exact).sum())
        print(f"grid {nres}^3: plain {t_plain * 1e3:.1f} ms, surrogate+fallback {t_sur * 1e3:.1f} ms, stats {st}, "
              f"mislabels {mislabels}")
        assert mislabels == 0
        assert st["audit_disagree"] == 0
        assert st["fallback_fraction"] < 0.5, st
        # everything routed: the labels are the exact path's, bit for bit
        lab_all, st_all = net.label_grid(grid, start, dx, dy, dz, nres, nres, nres, float("inf"))
        assert st_all["band"] + st_all["near"] == nres ** 3 and st_all["routed"] == nres ** 3
        assert torch.equal(lab_all, exact)


def test_grid_labels_follow_the_routing_rule(gpu, c1, trained_c1):
    """On a 256^3 grid (2^24 queries) every label is what the routing rule gives -- the exact label where the point is
    in the near-surface mask or in the band |pred - 0.5| <= delta, the network's elsewhere (pred from the lattice
    prediction entry, whose first layer is the labelling's: per-axis tables) -- and the stats count those sets; with delta = +inf every point is
    routed and the labels are the exact path's bit for bit. (This network was calibrated on 2^16 points, so on a grid
    64 times the 64^3 one a few confident network errors are expected: they are counted, not failed.)"""
    import torch

    v, idx, lo, hi = c1
    net, delta, info, bnd, x0 = trained_c1
    nres = 256
    start, dx, dy, dz = mg.grid_axes(lo, hi, nres)
    f = np.float32
    r = np.arange(nres, dtype=f)
    s0, a, b, c = (np.asarray(t, dtype=f) for t in (start, dx, dy, dz))
    pts = (((s0[None, None, None, :] + a * r[None, None, :, None]) + b * r[None, :, None, None])
           + c * r[:, None, None, None]).astype(f).reshape(-1, 3)  # the kernels' float32 lattice, x fastest
    P = torch.from_numpy(pts).cuda()
    with gpu.HipGridGwn() as grid:
        grid.set_scene(v.astype(np.float32), idx, bnd)
        w = torch.empty(nres ** 3, dtype=torch.float32, device="cuda")
        grid.eval_grid(start, dx, dy, dz, nres, nres, nres, out=w)
        exact = (w > 0.5).to(torch.int8)
        near = net.near_surface(P).bool()
        pred = net.predict_grid(start, dx, dy, dz, nres, nres, nres)
        dev = (pred - net.predict(P)).abs()
        print(f"lattice vs per-point predictions: max {dev.max().item():.2e}, "
              f"99.9 % {torch.quantile(dev[::16], 0.999).item():.2e}")
        assert dev.max().item() <= 1e-2  # below the band's 0.02 slack
        band = ~near & ((pred - 0.5).abs() <= delta)
        want = torch.where(near | band, exact, (pred > 0.5).to(torch.int8))
        lab, st = net.label_grid(grid, start, dx, dy, dz, nres, nres, nres, delta)
        torch.cuda.synchronize()
        wrong = int((lab != exact).sum())
        print(f"grid {nres}^3: stats {st}, network mislabels {wrong}")
        assert torch.equal(lab, want)
        assert st["near"] == int(near.sum()) and st["band"] == int(band.sum()) and st["audit"] == 0
        assert st["routed"] == st["near"] + st["band"]
        assert wrong < 1e-5 * nres ** 3
        lab_all, st_all = net.label_grid(grid, start, dx, dy, dz, nres, nres, nres, float("inf"))
        assert st_all["routed"] == nres ** 3
        assert torch.equal(lab_all, exact)


def _lattice(start, dx, dy, dz, nx, ny, nz):
    """The kernels' float32 lattice points, x index fastest (start + dx ix + dy iy + dz iz, left to right)."""
    f = np.float32
    s0, a, b, c = (np.asarray(t, dtype=f) for t in (start, dx, dy, dz))
    return (((s0[None, None, None, :] + a * np.arange(nx, dtype=f)[None, None, :, None])
             + b * np.arange(ny, dtype=f)[None, :, None, None])
            + c * np.arange(nz, dtype=f)[:, None, None, None]).astype(f).reshape(-1, 3)


def test_lattice_inference_matches_per_point_and_oracle(gpu, c1, trained_c1):
    """The lattice inference entry: on axis-aligned lattices (the first layer summed from per-axis tables of the same
    BF16 feature x weight products) its outputs lie within 1e-2 of the per-point kernel's on the same float32 points
    and within 1e-2 of the float64 oracle with the kernels' BF16 rounding points emulated (99.9 % of points within
    1e-3 of the per-point kernel: the few larger gaps are BF16 rounding flips of activations whose FP32 pre-activation
    lies within an ulp of a rounding boundary, carried through three layers); on lattices whose
    coordinates step along several axes (no tables) it is the per-point kernel's output bit for bit. Cases: 128^3
    (whole 128-point tiles), a ragged 40 x 33 x 17, permuted axes (x index -> z coordinate ...), a coordinate that
    steps along no axis (z constant, the z index repeating the plane), and a rotated lattice."""
    import torch

    v, idx, lo, hi = c1
    net, delta, info, bnd, x0 = trained_c1
    f = np.float32
    ext = (hi - lo).astype(f)
    e = np.eye(3, dtype=f)
    q, _ = np.linalg.qr(np.random.default_rng(5).standard_normal((3, 3)))
    rot = (q * float(ext.max()) / 40).astype(f)
    centre = (0.5 * (lo + hi)).astype(f)
    cases = [
        ("128^3", lo, *(e[k] * f(ext[k] / 127) for k in range(3)), 128, 128, 128, True),
        ("40x33x17", lo, *(e[k] * f(ext[k] / (n - 1)) for k, n in enumerate((40, 33, 17))), 40, 33, 17, True),
        ("permuted", lo, e[2] * f(ext[2] / 23), e[0] * f(ext[0] / 31), e[1] * f(ext[1] / 19), 24, 32, 20, True),
        ("z constant", np.array([lo[0], lo[1], centre[2]], f), e[0] * f(ext[0] / 15), e[1] * f(ext[1] / 63),
         np.zeros(3, f), 16, 64, 64, True),
        ("rotated", (centre - 20 * (rot[0] + rot[1] + rot[2])).astype(f), rot[0], rot[1], rot[2], 40, 40, 40, False),
    ]
    params = net.get_params().astype(np.float64)
    for name, start, dx, dy, dz, nx, ny, nz, tables in cases:
        pts = _lattice(start, dx, dy, dz, nx, ny, nz)
        P = torch.from_numpy(pts).cuda()
        g = net.predict_grid(start, dx, dy, dz, nx, ny, nz)
        a = net.predict(P)
        dev = (g - a).abs()
        d = dev.max().item()
        d999 = torch.quantile(dev, 0.999).item() if len(dev) <= 1 << 24 else float("nan")
        same = float((g == a).float().mean().item())
        sub = np.random.default_rng(1).choice(len(pts), size=min(len(pts), 2048), replace=False)
        ref, _ = om.forward(params, pts[sub].astype(np.float64), lo, hi, emulate=True, max_freq=32.0)
        do = float(np.max(np.abs(g.cpu().numpy()[sub] - ref)))
        print(f"{name}: |grid - per point| max {d:.2e}, 99.9 % {d999:.2e}, equal {same:.4f}; "
              f"max |grid - oracle| = {do:.2e}")
        if tables:
            assert d <= 1e-2 and d999 <= 1e-3 and do <= 1e-2
        else:
            assert torch.equal(g, a)


def test_grid_labels_across_call_sizes(gpu, c1, trained_c1):
    """A grid handle keeps its routed-query list between labelling calls and grows it when a call needs more: calls of
    96^3, then a smaller ragged 40 x 33 x 17 lattice, then 128^3 on one handle each give the routing rule's labels
    (exact where near or in the band, the network's elsewhere), and with delta = +inf the exact labels."""
    import torch

    v, idx, lo, hi = c1
    net, delta, info, bnd, x0 = trained_c1
    f = np.float32
    with gpu.HipGridGwn() as grid:
        grid.set_scene(v.astype(np.float32), idx, bnd)
        for nx, ny, nz in ((96, 96, 96), (40, 33, 17), (128, 128, 128)):
            print(f"lattice {nx} x {ny} x {nz}", flush=True)
            start = lo.astype(f)
            dx, dy, dz = (np.eye(3, dtype=f)[k] * f((hi[k] - lo[k]) / (n - 1)) for k, n in enumerate((nx, ny, nz)))
            pts = (((start[None, None, None, :] + dx * np.arange(nx, dtype=f)[None, None, :, None])
                    + dy * np.arange(ny, dtype=f)[None, :, None, None])
                   + dz * np.arange(nz, dtype=f)[:, None, None, None]).astype(f).reshape(-1, 3)
            P = torch.from_numpy(pts).cuda()
            w = torch.empty(nx * ny * nz, dtype=torch.float32, device="cuda")
            grid.eval_grid(start, dx, dy, dz, nx, ny, nz, out=w)
            exact = (w > 0.5).to(torch.int8)
            near = net.near_surface(P).bool()
            pred = net.predict_grid(start, dx, dy, dz, nx, ny, nz)
            band = ~near & ((pred - 0.5).abs() <= delta)
            want = torch.where(near | band, exact, (pred > 0.5).to(torch.int8))
            lab, st = net.label_grid(grid, start, dx, dy, dz, nx, ny, nz, delta)
            assert torch.equal(lab, want)
            assert st["routed"] == int((near | band).sum())
            lab_all, st_all = net.label_grid(grid, start, dx, dy, dz, nx, ny, nz, float("inf"))
            assert st_all["routed"] == nx * ny * nz
            assert torch.equal(lab_all, exact)


def test_grid_labels_over_many_poles(gpu, c1, trained_c1):
    """The far-field decision keeps its tables per pole with the boundary's plan (eight poles; a ninth builds its own
    per call). Nine lattices with seeded random orientations -- nine integration axes, so nine poles -- labelled on one
    grid handle with every point routed give the exact labels of eval_grid on the same lattice; the first lattice,
    labelled again at the end, gives the same labels as the first time."""
    import torch

    v, idx, lo, hi = c1
    net, delta, info, bnd, x0 = trained_c1
    rng = np.random.default_rng(12)
    nres = 40
    centre = 0.5 * (lo + hi)
    h = float(np.max(hi - lo)) / (nres - 1)
    with gpu.HipGridGwn() as grid:
        grid.set_scene(v.astype(np.float32), idx, bnd)
        first = None
        frames = []
        for k in range(9):
            q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
            frames.append(q.astype(np.float32) * np.float32(h))
        frames.append(frames[0])
        for k, ax in enumerate(frames):
            print(f"orientation {k}", flush=True)
            dx, dy, dz = ax[0], ax[1], ax[2]
            start = (centre - 0.5 * (nres - 1) * (dx + dy + dz)).astype(np.float32)
            w = torch.empty(nres ** 3, dtype=torch.float32, device="cuda")
            grid.eval_grid(start, dx, dy, dz, nres, nres, nres, out=w)
            exact = (w > 0.5).to(torch.int8)
            lab, st = net.label_grid(grid, start, dx, dy, dz, nres, nres, nres, float("inf"))
            assert st["routed"] == nres ** 3
            assert torch.equal(lab, exact), k
            if k == 0:
                first = lab.clone()
            if k == len(frames) - 1:
                assert torch.equal(lab, first)


def test_grid_labels_on_mirrored_lattices(gpu, c1, trained_c1):
    """Lattices whose first axis (their integration axis) steps along -x, or along +y with the other axes descending,
    and one whose third axis does (along -z): x0 is then -e_x, +e_y or -e_z, and the routed points' integer terms come
    from the lattice-row lists with rays against or across the other lattices' direction. Every point routed (delta =
    +inf): the labels equal eval_grid's on the same lattice."""
    import torch

    v, idx, lo, hi = c1
    net, delta, info, bnd, x0 = trained_c1
    f = np.float32
    ex, ey, ez = np.eye(3, dtype=f)
    ext = (hi - lo).astype(f)
    with gpu.HipGridGwn() as grid:
        grid.set_scene(v.astype(np.float32), idx, bnd)
        for name, nx, ny, nz in (("minus_x", 40, 44, 48), ("plus_y", 36, 41, 45), ("minus_z", 42, 38, 30)):
            if name == "minus_x":
                start = np.array([hi[0], lo[1], lo[2]], dtype=f)
                dx, dy, dz = ex * f(-ext[0] / (nx - 1)), ey * f(ext[1] / (ny - 1)), ez * f(ext[2] / (nz - 1))
            elif name == "minus_z":
                start = np.array([lo[0], lo[1], hi[2]], dtype=f)
                dx, dy, dz = ex * f(ext[0] / (nx - 1)), ey * f(ext[1] / (ny - 1)), ez * f(-ext[2] / (nz - 1))
            else:
                start = np.array([hi[0], lo[1], hi[2]], dtype=f)
                dx, dy, dz = ey * f(ext[1] / (nx - 1)), ex * f(-ext[0] / (ny - 1)), ez * f(-ext[2] / (nz - 1))
            print(f"lattice {name} {nx} x {ny} x {nz}", flush=True)
            w = torch.empty(nx * ny * nz, dtype=torch.float32, device="cuda")
            grid.eval_grid(start, dx, dy, dz, nx, ny, nz, out=w)
            exact = (w > 0.5).to(torch.int8)
            lab, st = net.label_grid(grid, start, dx, dy, dz, nx, ny, nz, float("inf"))
            assert st["routed"] == nx * ny * nz
            assert torch.equal(lab, exact), name
            assert 0 < int(exact.sum()) < nx * ny * nz


def test_point_labels_and_audit(gpu, c1, trained_c1):
    import torch

    v, idx, lo, hi = c1
    net, delta, info, bnd, x0 = trained_c1
    rng = np.random.default_rng(4)
    n = 200_000
    pts = torch.from_numpy((lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)).cuda()
    seg = torch.from_numpy(np.ascontiguousarray(bnd, dtype=np.float32)).cuda()
    with gpu.HipMesh(v.astype(np.float32), idx, dtype=np.float32) as mesh:
        w = torch.empty(n, dtype=torch.float32, device="cuda")
        mesh.full(pts, x0, seg, out=w)
        exact = (w > 0.5).to(torch.int8)
        lab, st = net.label(mesh, pts, x0, seg, delta, audit_every=32, audit_seed=9)
        assert int((lab != exact).sum()) == 0, st
        # no band: network labels except for the near-surface points and the audit sample, which take the exact labels
        pred = net.predict(pts)
        netlab = (pred > 0.5).to(torch.int8)
        near = net.near_surface(pts).bool()
        lab0, st0 = net.label(mesh, pts, x0, seg, 0.0, audit_every=8, audit_seed=2)
        assert st0["near"] == int(near.sum())
        assert st0["audit"] > (n - st0["near"]) // 16 and st0["band"] <= 16
        assert torch.equal(lab0[near], exact[near])
        changed = (lab0 != netlab) & ~near
        assert torch.equal(lab0[changed], exact[changed])
        assert int(changed.sum()) == st0["audit_disagree"]


def test_near_surface_mask_is_conservative(gpu, c1):
    """Every point within eps of a triangle is in the mask; points outside the box are too."""
    import torch

    v, idx, lo, hi = c1
    rng = np.random.default_rng(8)
    tri = v[idx.reshape(-1, 3)]
    k = rng.integers(0, len(tri), 100_000)
    r1, r2 = np.sqrt(rng.random(len(k))), rng.random(len(k))
    p = (1 - r1)[:, None] * tri[k, 0] + (r1 * (1 - r2))[:, None] * tri[k, 1] + (r1 * r2)[:, None] * tri[k, 2]
    nrm = np.cross(tri[k, 1] - tri[k, 0], tri[k, 2] - tri[k, 0])
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    eps = 0.02
    d = rng.uniform(-0.999, 0.999, len(k)) * eps
    # plus points displaced in an arbitrary direction by less than eps
    u = rng.normal(size=(len(k), 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    pts = np.concatenate([p + d[:, None] * nrm, p + 0.999 * eps * rng.random(len(k))[:, None] * u])
    outside = np.array([[hi[0] + 0.1, 0, 0], [lo[0] - 1e-3, 0, 0], [0, 0, hi[2] + 1e-3]], dtype=np.float32)
    with gpu.GwnSurrogate(lo, hi, max_freq=32.0) as net:
        net.set_surface(v.astype(np.float32), idx, eps)
        flags = net.near_surface(torch.from_numpy(pts.astype(np.float32)).cuda()).cpu().numpy()
        assert flags.all(), int((flags == 0).sum())
        assert net.near_surface(torch.from_numpy(outside).cuda()).cpu().numpy().all()
        far = (lo + (hi - lo) * rng.random((100_000, 3))).astype(np.float32)
        frac = float(net.near_surface(torch.from_numpy(far).cuda()).float().mean())
        print(f"near-surface mask: {frac:.3f} of the box (eps {eps})")
        assert frac < 0.25


def test_bad_arguments(gpu, c1):
    import torch

    v, idx, lo, hi = c1
    with pytest.raises(gpu.AgwnError):
        gpu.GwnSurrogate(hi, lo)
    with gpu.GwnSurrogate(lo, hi) as net:
        pts = torch.zeros((256, 3), dtype=torch.float32, device="cuda")
        y = torch.zeros(256, dtype=torch.float32, device="cuda")
        with pytest.raises(gpu.AgwnError, match="multiple of 128"):
            net.train(pts, y, batch=100, epochs=1)
        with gpu.HipMesh(v.astype(np.float32), idx, dtype=np.float32) as mesh:
            with pytest.raises(gpu.AgwnError, match="delta"):
                net.label(mesh, pts, mg.example_x0(np.float32), None, -1.0)
        with gpu.HipMesh(v, idx, dtype=np.float64) as mesh64:
            with pytest.raises(gpu.AgwnError, match="f32 mesh"):
                net.label(mesh64, pts, mg.example_x0(np.float32), None, 0.1)
        # the lattice entry validates the lattice and the handle before any launch
        e = np.eye(3, dtype=np.float32)
        with pytest.raises(gpu.AgwnError, match="positive"):
            net.predict_grid(lo, e[0], e[1], e[2], 0, 4, 4)
        small = torch.empty(8, dtype=torch.float32, device="cuda")
        with pytest.raises(gpu.AgwnError, match="too large"):
            net.predict_grid(lo, e[0], e[1], e[2], 1 << 30, 1 << 30, 1 << 2, out=small)
    from antipodal_amd._lib import lib, ptr

    out = torch.empty(8, dtype=torch.float32, device="cuda")
    vs = [np.ascontiguousarray(a, dtype=np.float32) for a in (lo, e[0], e[1], e[2])]
    assert lib.agwn_mlp_predict_grid_dev(None, *[ptr(a) for a in vs], 2, 2, 2, ptr(out), None) != 0
    assert b"mlp handle is null" in lib.agwn_last_error()
