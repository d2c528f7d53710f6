"""Learned inside/outside surrogate over the C ABI (libagwn.so): a per-mesh Fourier-feature MLP whose training,
inference and routing all run in the library's HIP kernels (BF16 MFMA, FP32 accumulation).

  GwnSurrogate(lo, hi, max_freq, seed)   the network for positions in the box [lo, hi]
    .train(pts, target, ...)        Adam on the mean squared error (device tensors)
    .grad(pts, target)              one forward/backward pass without an update (host arrays; gradient checks)
    .predict(pts)                   fused inference
    .predict_grid(start, dx, ...)   fused inference over a lattice
    .label(mesh, pts, x0, boundary, delta, ...)           surrogate + exact fallback labels of points
    .label_grid(grid, start, dx, dy, dz, nx, ny, nz, delta, ...)   the same over an eval_grid lattice
  TrainingSampler(verts, idx, lo, hi)(n, ...)   the training distribution on the device: points near the surface and
                                    uniform in the box
  fit_surrogate(...)                samples -> exact GWN targets (HipMesh.full) -> training on jittered inputs ->
                                    band calibration on unjittered validation points

Labels are GWN > 0.5. The exact FP32 path (the reference's eval_gwnr_mesh_batch, gwn_mesh.hh:201-219, as evaluated by
HipMesh / HipGridGwn) decides every point whose prediction lies within `delta` of 0.5, and an audit sample of the rest.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from ._lib import check, lib, ptr
from .api import HipGridGwn, HipMesh, _is_dev, _stream_ptr, _x0


def num_params() -> int:
    return int(lib.agwn_mlp_num_params())


class GwnSurrogate:
    """The network for positions in the box [lo, hi]: Fourier frequencies max_freq^(k/31), k = 0..31, per axis;
    deterministic initialisation from `seed`."""

    def __init__(self, lo, hi, max_freq: float = 128.0, seed: int = 0):
        self.lo = np.ascontiguousarray(np.asarray(lo, dtype=np.float32).reshape(3))
        self.hi = np.ascontiguousarray(np.asarray(hi, dtype=np.float32).reshape(3))
        self.max_freq = float(max_freq)
        h = C.c_void_p()
        check(lib.agwn_mlp_create(ptr(self.lo), ptr(self.hi), C.c_float(max_freq), C.c_uint64(seed), C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def get_params(self) -> np.ndarray:
        out = np.empty(num_params(), dtype=np.float32)
        check(lib.agwn_mlp_get_params(self._h, ptr(out)))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(np.asarray(p, dtype=np.float32).reshape(-1))
        if p.size != num_params():
            raise ValueError(f"expected {num_params()} parameters, got {p.size}")
        check(lib.agwn_mlp_set_params(self._h, ptr(p)))

    def train(self, pts, target, batch: int = 16384, epochs: int = 8, lr: float = 2e-3, seed: int = 0,
              stream=None) -> np.ndarray:
        """Adam over device tensors pts (n, 3) float32 and target (n,) float32; returns each epoch's mean loss."""
        if not (_is_dev(pts) and _is_dev(target)):
            raise TypeError("train takes device tensors")
        n = target.numel()
        if pts.numel() != 3 * n:
            raise ValueError("pts.size() != 3 * target.size()")
        losses = np.zeros(max(epochs, 1), dtype=np.float32)
        check(lib.agwn_mlp_train_dev(self._h, ptr(pts), ptr(target), n, int(batch), int(epochs), C.c_float(lr),
                                     C.c_uint64(seed), ptr(losses), _stream_ptr(stream)))
        return losses[:epochs]

    def grad(self, pts, target):
        """(gradient float32[num_params], predictions float32[n], mean squared error) of host arrays."""
        p = np.ascontiguousarray(np.asarray(pts, dtype=np.float32).reshape(-1, 3))
        y = np.ascontiguousarray(np.asarray(target, dtype=np.float32).reshape(-1))
        if len(p) != len(y):
            raise ValueError("pts and target sizes differ")
        g = np.empty(num_params(), dtype=np.float32)
        pred = np.empty(len(y), dtype=np.float32)
        loss = C.c_float(0.0)
        check(lib.agwn_mlp_grad(self._h, ptr(p), ptr(y), len(y), ptr(g), ptr(pred), C.byref(loss)))
        return g, pred, float(loss.value)

    def set_surface(self, verts, idx, eps: float, voxel: float = 0.0) -> None:
        """Near-surface mask: voxels (edge `voxel`, default eps/2) within eps of a triangle are routed to the exact
        path, as are points outside the box."""
        v = np.ascontiguousarray(np.asarray(verts, dtype=np.float32).reshape(-1, 3))
        i = np.ascontiguousarray(np.asarray(idx, dtype=np.int32).reshape(-1))
        check(lib.agwn_mlp_set_surface(self._h, ptr(v), len(v), ptr(i), i.size, C.c_float(eps), C.c_float(voxel)))
        self.eps = float(eps)

    def near_surface(self, pts, out=None, stream=None):
        """int8 flags of device points: 1 = in the near-surface mask or outside the box."""
        import torch

        n = pts.numel() // 3
        if out is None:
            out = torch.empty(n, dtype=torch.int8, device=pts.device)
        check(lib.agwn_mlp_near_surface_dev(self._h, ptr(pts), n, ptr(out), _stream_ptr(stream)))
        return out

    def predict(self, pts, out=None, stream=None):
        """Network output of device points (n, 3) float32."""
        import torch

        n = pts.numel() // 3
        if out is None:
            out = torch.empty(n, dtype=torch.float32, device=pts.device)
        check(lib.agwn_mlp_predict_dev(self._h, ptr(pts), n, ptr(out), _stream_ptr(stream)))
        return out

    def predict_grid(self, start, dx, dy, dz, nx: int, ny: int, nz: int, out=None, stream=None):
        """Network output over an eval_grid lattice, in its ((iz*ny)+iy)*nx+ix layout. On axis-aligned lattices the
        first layer is summed from per-axis tables (same BF16 products, FP32 sums in another order)."""
        import torch

        n = nx * ny * nz
        if out is None:
            out = torch.empty(n, dtype=torch.float32, device="cuda")
        vs = [np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(3)) for a in (start, dx, dy, dz)]
        check(lib.agwn_mlp_predict_grid_dev(self._h, *[ptr(a) for a in vs], nx, ny, nz, ptr(out),
                                            _stream_ptr(stream)))
        return out

    def label(self, mesh: HipMesh, pts, x0, boundary, delta: float, audit_every: int = 0, audit_seed: int = 0,
              labels=None, stream=None):
        """Inside/outside labels (int8) of device points; returns (labels, stats dict)."""
        import torch

        n = pts.numel() // 3
        if labels is None:
            labels = torch.empty(n, dtype=torch.int8, device=pts.device)
        nseg = 0 if boundary is None else boundary.numel() // 7
        st = np.zeros(5, dtype=np.int64)
        check(lib.agwn_label_surrogate_dev(mesh.handle, self._h, ptr(pts), n, ptr(_x0(x0, np.float32)),
                                           ptr(boundary) if nseg else None, nseg, C.c_float(delta), int(audit_every),
                                           C.c_uint64(audit_seed), ptr(labels), ptr(st), _stream_ptr(stream)))
        return labels, _stats(st, n)

    def label_grid(self, grid: HipGridGwn, start, dx, dy, dz, nx: int, ny: int, nz: int, delta: float,
                   audit_every: int = 0, audit_seed: int = 0, labels=None, stream=None):
        """Labels of an eval_grid lattice, in its ((iz*ny)+iy)*nx+ix layout; returns (labels, stats dict)."""
        import torch

        n = nx * ny * nz
        if labels is None:
            labels = torch.empty(n, dtype=torch.int8, device="cuda")
        vs = [np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(3)) for a in (start, dx, dy, dz)]
        st = np.zeros(5, dtype=np.int64)
        check(lib.agwn_grid_label_surrogate_dev(grid._h, self._h, *[ptr(a) for a in vs], nx, ny, nz, C.c_float(delta),
                                                int(audit_every), C.c_uint64(audit_seed), ptr(labels), ptr(st),
                                                _stream_ptr(stream)))
        return labels, _stats(st, n)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.agwn_mlp_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _stats(st: np.ndarray, n: int) -> dict:
    return {"n": n, "band": int(st[0]), "near": int(st[1]), "audit": int(st[2]), "routed": int(st[3]),
            "audit_disagree": int(st[4]), "fallback_fraction": float(st[3]) / max(n, 1)}


class TrainingSampler:
    """The surrogate's training distribution on the device (`agwn_training_samples_dev`): near_frac of the points at
    signed normal offsets band * s * u^power (s = +-1, u ~ U(0, 1)) from area-uniform surface points, so all within
    `band` of the surface (power > 1 puts more of them close to it), the rest uniform in [lo, hi]. Point i of a call
    depends on (seed, i) only (counter-based SplitMix64 streams). The mesh and its cumulative triangle areas (doubled
    areas |(b - a) x (c - a)| in float64, summed in triangle order) are uploaded once."""

    def __init__(self, verts, idx, lo, hi):
        import torch

        v = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
        tri = np.asarray(idx, dtype=np.int64).reshape(-1, 3)
        a, b, c = v[tri[:, 0]], v[tri[:, 1]], v[tri[:, 2]]
        area = np.linalg.norm(np.cross(b - a, c - a), axis=1)
        self._v = torch.from_numpy(np.ascontiguousarray(v, dtype=np.float32)).cuda()
        self._i = torch.from_numpy(np.ascontiguousarray(tri, dtype=np.int32)).cuda()
        self._cdf = torch.from_numpy(np.cumsum(area)).cuda()
        self._ntri = len(tri)
        self._lo = np.ascontiguousarray(np.asarray(lo, dtype=np.float32).reshape(3))
        self._hi = np.ascontiguousarray(np.asarray(hi, dtype=np.float32).reshape(3))

    def __call__(self, n: int, near_frac: float = 0.7, band: float = 0.05, seed: int = 0, power: float = 2.0,
                 out=None, stream=None):
        import torch

        if out is None:
            out = torch.empty((n, 3), dtype=torch.float32, device="cuda")
        check(lib.agwn_training_samples_dev(ptr(self._v), ptr(self._i), ptr(self._cdf), self._ntri, ptr(self._lo),
                                            ptr(self._hi), n, C.c_double(near_frac), C.c_double(band),
                                            C.c_double(power), C.c_uint64(seed), ptr(out), _stream_ptr(stream)))
        return out


def calibrate_delta(pred: np.ndarray, exact: np.ndarray, factor: float = 2.0, margin: float = 0.02,
                    skip: np.ndarray | None = None) -> float:
    """Band half-width from validation predictions: factor x the largest |pred - 0.5| of a wrongly labelled
    validation point (those flagged in `skip` left out), plus margin."""
    wrong = (pred > 0.5) != (exact > 0.5)
    if skip is not None:
        wrong &= ~skip
    worst = float(np.max(np.abs(pred[wrong] - 0.5))) if wrong.any() else 0.0
    return factor * worst + margin


def exact_disagreements(verts, idx, boundary, x0, pts, w32):
    """Flags the points (device (m, 3) float32) whose FP32 full GWN w32 (device) differs by more than 1/2 from the
    FP64 evaluation of the same float32 geometry: the non-watertight ray test of the integer term (the reference's
    Moller-Trumbore, kept bit for bit) counted a crossing differently in the two precisions, a ray grazing an edge or
    vertex. Host bool array."""
    import torch

    if len(pts) == 0:
        return np.zeros(0, dtype=bool)
    v64 = np.asarray(verts, dtype=np.float32).astype(np.float64)
    seg64 = None
    if boundary is not None and len(boundary):
        seg64 = torch.from_numpy(np.asarray(boundary, dtype=np.float32).reshape(-1, 7).astype(np.float64)).cuda()
    w64 = torch.empty(len(pts), dtype=torch.float64, device="cuda")
    with HipMesh(v64, idx, dtype=np.float64) as m64:
        m64.full(pts.double().contiguous(), np.asarray(x0, dtype=np.float64), seg64, out=w64)
    return ((w64 - w32.double()).abs() > 0.5).cpu().numpy()


def fit_surrogate(verts, idx, boundary, x0, lo, hi, n_train: int = 1 << 24, n_val: int = 1 << 22,
                  epochs: int = 8, batch: int = 16384, lr: float = 2e-3, seed: int = 0, mesh: HipMesh | None = None,
                  max_freq: float = 64.0, eps: float | None = None, voxel: float = 0.0, jitter: float = 0.0,
                  target_clip=(-0.25, 1.25), factor: float = 2.0, margin: float = 0.02, power: float = 1.0,
                  log=None):
    """Trains a GwnSurrogate for the mesh on exact FP32 GWN values computed by the library, and picks the routing band.

    The surrogate only ever labels points farther than `eps` from the surface (default 0.25 % of the box diagonal):
    a conservative voxel mask of the surface's eps-neighbourhood sends everything closer -- where the GWN jumps and a
    noisy surface can fold -- to the exact path. Off the surface the GWN is harmonic and smooth, which is what the
    network is trained on (samples inside the mask are dropped) and calibrated on: delta = factor x the largest
    |pred - 0.5| of a wrongly labelled unmasked validation point, plus margin. `jitter` (default 0) displaces the
    training inputs by Gaussian noise of that standard deviation, keeping the targets of the undisplaced points.

    Returns (surrogate, delta, info) -- info holds the timings, per-epoch losses and validation statistics.
    """
    import torch

    info: dict = {}
    t0 = time.perf_counter()
    own_mesh = mesh is None
    if own_mesh:
        mesh = HipMesh(np.asarray(verts, dtype=np.float32), idx, dtype=np.float32)
    diag = float(np.linalg.norm(np.asarray(hi, dtype=np.float64) - np.asarray(lo, dtype=np.float64)))
    if eps is None:
        eps = 0.0025 * diag
    net = GwnSurrogate(lo, hi, max_freq=max_freq, seed=seed)
    net.set_surface(verts, idx, eps, voxel)
    seg = torch.from_numpy(np.ascontiguousarray(np.asarray(boundary, dtype=np.float32).reshape(-1, 7))).cuda()
    # candidates, then drop those the mask routes anyway; oversample so that enough remain
    want = n_train + n_val
    sampler = TrainingSampler(verts, idx, lo, hi)
    keep_parts, have, rnd = [], 0, 0
    while have < want and rnd < 4:
        cand = sampler(int(want * 1.6) + 1024, seed=seed + 7919 * rnd, power=power)
        ok = net.near_surface(cand) == 0
        keep_parts.append(cand[ok])
        have += int(ok.sum())
        rnd += 1
    pts = torch.cat(keep_parts)[:want].contiguous()
    if len(pts) < want:
        n_train = max(1, int(len(pts) * n_train / want))
        n_val = len(pts) - n_train
    info["sample_s"] = time.perf_counter() - t0
    t1 = time.perf_counter()
    w = torch.empty(len(pts), dtype=torch.float32, device="cuda")
    mesh.full(pts, x0, seg if len(seg) else None, out=w)
    torch.cuda.synchronize()
    info["label_s"] = time.perf_counter() - t1
    tr_p = pts[:n_train]
    if jitter > 0:
        gen = torch.Generator(device="cuda").manual_seed(seed + 1)
        tr_p = tr_p + jitter * torch.randn((n_train, 3), device="cuda", generator=gen)
    tr_p = tr_p.contiguous()
    va_p = pts[n_train:].contiguous()
    tr_y = w[:n_train].clamp(*target_clip).contiguous()
    va_y = w[n_train:]
    info.update(eps=eps, voxel=voxel, jitter=jitter, max_freq=max_freq, n_train=n_train)
    t2 = time.perf_counter()
    losses = net.train(tr_p, tr_y, batch=batch, epochs=epochs, lr=lr, seed=seed)
    torch.cuda.synchronize()
    info["train_s"] = time.perf_counter() - t2
    info["losses"] = [float(x) for x in losses]
    pred = net.predict(va_p).cpu().numpy()
    exact = va_y.cpu().numpy()
    wrong = (pred > 0.5) != (exact > 0.5)
    # The band calibrates the network's error, not the exact path's: a wrongly labelled validation point whose FP32
    # value is off by an integer (the integer term's ray grazed an edge; the FP64 evaluation counts it otherwise) would
    # set delta near 1 by itself, so such points are left out of the calibration (and counted).
    skip = np.zeros(len(pred), dtype=bool)
    iw = np.nonzero(wrong)[0]
    if len(iw):
        iw_t = torch.from_numpy(iw).cuda()
        skip[iw] = exact_disagreements(verts, idx, boundary, x0, va_p[iw_t], va_y[iw_t])
    delta = calibrate_delta(pred, exact, factor, margin, skip)
    counted = wrong & ~skip
    info["val_pred"], info["val_exact"] = pred, exact
    info.update(val_points=n_val, val_wrong=int(wrong.sum()), val_exact_disagree=int(skip.sum()),
                val_max_wrong=float(np.max(np.abs(pred[counted] - 0.5))) if counted.any() else 0.0, delta=delta,
                val_band_fraction=float(np.mean(np.abs(pred - 0.5) <= delta)), fit_s=time.perf_counter() - t0)
    if log:
        log("surrogate: " + str({k: v for k, v in info.items() if k not in ("val_pred", "val_exact")}))
    if own_mesh:
        mesh.close()
    return net, delta, info
