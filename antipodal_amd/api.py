"""Host-side Python mirror of the reference's batch-GWN entry points, over the C ABI.

  eval_gwnr_mesh_batch_fractional(positions, x0, boundary)   gwn_mesh.hh:67-83
  eval_gwnr_mesh_batch_integer(mesh, positions, x0)          gwn_mesh.hh:127-143
  eval_gwnr_mesh_batch(mesh, positions, x0, boundary)        gwn_mesh.hh:201-219
  build_boundary_segments(vertices, indices)                 math/common.hh:123-174
  HipMesh(vertices, indices)                                 the Intersector concept, intersector/intersector.hh:8-43
  RobustMeshGwn(vertices, indices, boundary, precision)      gwn_mesh_robust.hh:170-424
  eval_gwnr_mesh_batch_robust(robust, positions)             gwn_mesh_robust.hh:441-456
  HipGridGwn().set_scene(...).eval_grid(...)                 OpenGLMeshGwn, gwn_mesh_opengl.hh:426-643

Arrays are numpy (host; results come back as numpy) or torch CUDA tensors (device; the *_dev entry points run on the
given stream and write into `out`). Layouts are the reference's: positions (n, 3), boundary (nseg, 7) with the weight
last, vertices (nv, 3), indices (3*ntri,) int32. Size mismatches raise ValueError where the reference asserts
(gwn_mesh.hh:74,134,209); a failing library call raises AgwnError.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib, ptr, sfx


def _host(a, dtype, cols: int | None = None, name: str = "array") -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=dtype))
    if cols is not None:
        a = a.reshape(-1, cols)
    return a


def _is_dev(a) -> bool:
    return hasattr(a, "is_cuda") and bool(a.is_cuda)


def _x0(x0, dtype) -> np.ndarray:
    x = np.ascontiguousarray(np.asarray(x0, dtype=dtype).reshape(3))
    return x


def _stream_ptr(stream) -> int | None:
    if stream is None:
        try:
            import torch

            return torch.cuda.current_stream().cuda_stream
        except Exception:  # pragma: no cover - torch without a device
            return None
    return getattr(stream, "cuda_stream", stream)


def build_boundary_segments(vertices, indices, stream=None):
    """Open boundary as (nseg, 7) weighted oriented segments, sorted by (min index, max index). Host numpy in ->
    numpy out (host extraction); torch CUDA vertices (nv, 3) and int32 indices -> a device tensor (GPU extraction)."""
    if _is_dev(vertices):
        import torch

        s = sfx(vertices.dtype)
        if not _is_dev(indices) or indices.dtype != torch.int32:
            raise TypeError("device extraction takes int32 device indices")
        nidx = indices.numel()
        if nidx % 3:
            raise ValueError("indices.size() must be a multiple of 3")
        out = torch.empty((max(nidx, 1), 7), dtype=vertices.dtype, device=vertices.device)
        n = C.c_int64(0)
        fn = getattr(lib, f"agwn_boundary_{s}_dev")
        check(fn(ptr(vertices), vertices.numel() // 3, ptr(indices), nidx, ptr(out), nidx, C.byref(n),
                 _stream_ptr(stream)))
        return out[: n.value]
    v = np.ascontiguousarray(np.asarray(vertices))
    if v.dtype not in (np.float32, np.float64):
        v = v.astype(np.float64)
    v = v.reshape(-1, 3)
    idx = _host(indices, np.int32).reshape(-1)
    if idx.size % 3:
        raise ValueError("indices.size() must be a multiple of 3")
    fn = getattr(lib, f"agwn_boundary_{sfx(v.dtype)}")
    cap = idx.size
    out = np.empty((max(cap, 1), 7), dtype=v.dtype)
    n = C.c_int64(0)
    check(fn(ptr(v), len(v), ptr(idx), idx.size, ptr(out), cap, C.byref(n)))
    return out[: n.value].copy()


def eval_gwnr_mesh_batch_fractional(positions, x0, boundary, out=None, stream=None):
    """Fractional GWN of every position (gwn_mesh.hh:67-83). Host numpy in -> numpy out; torch CUDA in -> `out`."""
    if _is_dev(positions):
        s = sfx(positions.dtype)
        n = positions.numel() // 3
        if out is None or out.numel() != n:
            raise ValueError("positions.size() != out_wnrs.size()")
        nseg = 0 if boundary is None else boundary.numel() // 7
        x = _x0(x0, np.float32 if s == "f32" else np.float64)
        fn = getattr(lib, f"agwn_frac_batch_{s}_dev")
        check(fn(ptr(positions), n, ptr(boundary) if nseg else None, nseg, ptr(x), ptr(out), _stream_ptr(stream)))
        return out
    pos = np.asarray(positions)
    dt = np.float32 if pos.dtype == np.float32 else np.float64
    pos = _host(pos, dt, 3)
    seg = _host(boundary if boundary is not None else np.zeros((0, 7)), dt, 7)
    res = np.empty(len(pos), dtype=dt) if out is None else out
    if res.shape != (len(pos),) or res.dtype != dt:
        raise ValueError("positions.size() != out_wnrs.size()")
    fn = getattr(lib, f"agwn_frac_batch_{sfx(dt)}")
    check(fn(ptr(pos), len(pos), ptr(seg) if len(seg) else None, len(seg), ptr(_x0(x0, dt)), ptr(res)))
    return res


class HipMesh:
    """Prepared triangle mesh on the GPU (LBVH). Satisfies the reference's Intersector concept in batch form:
    signed_intersection_count(positions, dir) returns, per position, sum sign(N . dir) over every triangle the ray
    (p, dir) hits at t >= 0 -- NaiveIntersector's Moeller-Trumbore arithmetic, bit for bit (intersector.hh:61-100)."""

    def __init__(self, vertices, indices, dtype=None):
        v = np.asarray(vertices)
        dt = np.dtype(dtype) if dtype is not None else (np.float32 if v.dtype == np.float32 else np.float64)
        self.dtype = dt
        self._v = _host(v, dt, 3)
        self._i = _host(indices, np.int32).reshape(-1)
        if self._i.size % 3:
            raise ValueError("indices.size() must be a multiple of 3")
        h = C.c_void_p()
        fn = getattr(lib, f"agwn_mesh_create_{sfx(dt)}")
        check(fn(ptr(self._v), len(self._v), ptr(self._i), self._i.size, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def info(self) -> dict:
        ntri, nodes, sb = C.c_int64(), C.c_int64(), C.c_int()
        check(lib.agwn_mesh_info(self._h, C.byref(ntri), C.byref(nodes), C.byref(sb)))
        return {"ntri": ntri.value, "nodes": nodes.value, "scalar_bytes": sb.value}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.agwn_mesh_destroy(self._h)
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

    def signed_intersection_count(self, positions, direction, out=None, stream=None):
        s = sfx(self.dtype)
        if _is_dev(positions):
            n = positions.numel() // 3
            if out is None or out.numel() != n:
                raise ValueError("positions.size() != out_counts.size()")
            fn = getattr(lib, f"agwn_int_batch_{s}_dev")
            check(fn(self._h, ptr(positions), n, ptr(_x0(direction, self.dtype)), ptr(out), _stream_ptr(stream)))
            return out
        pos = _host(positions, self.dtype, 3)
        res = np.empty(len(pos), dtype=np.int32)
        fn = getattr(lib, f"agwn_int_batch_{s}")
        check(fn(self._h, ptr(pos), len(pos), ptr(_x0(direction, self.dtype)), ptr(res)))
        return res

    def full(self, positions, x0, boundary, out=None, stream=None):
        s = sfx(self.dtype)
        if _is_dev(positions):
            n = positions.numel() // 3
            if out is None or out.numel() != n:
                raise ValueError("positions.size() != out_wnrs.size()")
            nseg = 0 if boundary is None else boundary.numel() // 7
            fn = getattr(lib, f"agwn_full_batch_{s}_dev")
            check(fn(self._h, ptr(positions), n, ptr(_x0(x0, self.dtype)), ptr(boundary) if nseg else None, nseg,
                     ptr(out), _stream_ptr(stream)))
            return out
        pos = _host(positions, self.dtype, 3)
        seg = _host(boundary if boundary is not None else np.zeros((0, 7)), self.dtype, 7)
        res = np.empty(len(pos), dtype=self.dtype)
        fn = getattr(lib, f"agwn_full_batch_{s}")
        check(fn(self._h, ptr(pos), len(pos), ptr(_x0(x0, self.dtype)), ptr(seg) if len(seg) else None, len(seg),
                 ptr(res)))
        return res


def eval_gwnr_mesh_batch_integer(mesh: HipMesh, positions, x0, out=None, stream=None):
    """Integer GWN term of every position (gwn_mesh.hh:127-143)."""
    return mesh.signed_intersection_count(positions, x0, out=out, stream=stream)


def eval_gwnr_mesh_batch(mesh: HipMesh, positions, x0, boundary, out=None, stream=None):
    """Full GWN = fractional + integer of every position (gwn_mesh.hh:201-219)."""
    return mesh.full(positions, x0, boundary, out=out, stream=stream)


PRECISIONS = {"int64_grid": 0, "int128_grid": 1, "exact_float": 2}


def default_precision(dtype) -> str:
    """robust_predicates.hh:220-221 (default_precision): int64_grid for float, exact_float for double."""
    return "int64_grid" if np.dtype(dtype) == np.float32 else "exact_float"


class RobustMeshGwn:
    """Prepared robust GWN evaluator (gwn_mesh_robust.hh:170-424) on the GPU.

    RobustMeshGwn(vertices, indices, boundary, precision=None, dtype=None): T = dtype (float32 or float64, default the
    vertices' dtype); precision one of "int64_grid", "int128_grid", "exact_float" (default per T). Owns its device
    copies after construction. The ray is fixed to -x (axis()); eval = fractional + signed_intersection_count.
    """

    def __init__(self, vertices, indices, boundary, precision: str | None = None, dtype=None):
        v = np.asarray(vertices)
        dt = np.dtype(dtype) if dtype is not None else (np.dtype(np.float32) if v.dtype == np.float32 else np.dtype(np.float64))
        self.dtype = dt
        self.precision = precision or default_precision(dt)
        if self.precision not in PRECISIONS:
            raise ValueError(f"unknown precision {self.precision!r}")
        vv = _host(v, dt, 3)
        ii = _host(indices, np.int32).reshape(-1)
        if ii.size % 3:
            raise ValueError("indices.size() must be a multiple of 3")
        bb = _host(boundary if boundary is not None else np.zeros((0, 7)), dt, 7)
        h = C.c_void_p()
        check(lib.agwn_robust_create(dt.itemsize, PRECISIONS[self.precision], ptr(vv), len(vv), ptr(ii), ii.size,
                                     ptr(bb) if len(bb) else None, len(bb), C.byref(h)))
        self._h = h

    @staticmethod
    def axis():
        """gwn_mesh_robust.hh:226 -- the fixed ray direction."""
        return np.array([-1.0, 0.0, 0.0])

    def info(self) -> dict:
        sb, pr, nt, ns = C.c_int(), C.c_int(), C.c_int64(), C.c_int64()
        cen = np.zeros(3)
        sc = C.c_double()
        check(lib.agwn_robust_info(self._h, C.byref(sb), C.byref(pr), C.byref(nt), C.byref(ns), ptr(cen),
                                   C.byref(sc)))
        return {"scalar_bytes": sb.value, "precision": pr.value, "ntri": nt.value, "nseg": ns.value,
                "center": cen, "scale": sc.value}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.agwn_robust_destroy(self._h)
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

    def _run(self, what: str, positions, out, stream):
        is_count = what == "count"
        if _is_dev(positions):
            n = positions.numel() // 3
            if out is None or out.numel() != n:
                raise ValueError("positions.size() != out.size()")
            fn = getattr(lib, f"agwn_robust_{what}_batch_dev")
            check(fn(self._h, ptr(positions), n, ptr(out), _stream_ptr(stream)))
            return out
        pos = _host(positions, self.dtype, 3)
        res = np.empty(len(pos), dtype=np.int32 if is_count else self.dtype)
        fn = getattr(lib, f"agwn_robust_{what}_batch")
        check(fn(self._h, ptr(pos), len(pos), ptr(res)))
        return res

    def eval(self, positions, out=None, stream=None):
        """gwn_mesh_robust.hh:328 per position."""
        return self._run("eval", positions, out, stream)

    def fractional(self, positions, out=None, stream=None):
        """gwn_mesh_robust.hh:264-325 per position."""
        return self._run("fractional", positions, out, stream)

    def signed_intersection_count(self, positions, out=None, stream=None):
        """gwn_mesh_robust.hh:231-258 per position (int32)."""
        return self._run("count", positions, out, stream)


def eval_gwnr_mesh_batch_robust(robust: RobustMeshGwn, positions, out=None, stream=None):
    """gwn_mesh_robust.hh:441-456."""
    return robust.eval(positions, out=out, stream=stream)


class HipGridGwn:
    """Dense-grid evaluator (OpenGLMeshGwn, gwn_mesh_opengl.hh:426-643) on HIP. FP32.

    set_scene(vertices, indices, boundary) uploads the scene; eval_grid(start, dx, dy, dz, nx, ny, nz, out=None)
    returns (or, for a torch CUDA `out`, writes on `stream`) the full GWN in ((iz*ny)+iy)*nx+ix order.
    """

    def __init__(self):
        h = C.c_void_p()
        check(lib.agwn_grid_create(C.byref(h)))
        self._h = h

    def set_scene(self, vertices, indices, boundary):
        v = _host(vertices, np.float32, 3)
        i = _host(indices, np.int32).reshape(-1)
        if i.size % 3:
            raise ValueError("indices.size() must be a multiple of 3")
        b = _host(boundary if boundary is not None else np.zeros((0, 7)), np.float32, 7)
        check(lib.agwn_grid_set_scene(self._h, ptr(v), len(v), ptr(i), i.size, ptr(b) if len(b) else None, len(b)))
        return self

    def eval_grid(self, start, dx, dy, dz, nx: int, ny: int, nz: int, out=None, stream=None):
        vs = [np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(3)) for a in (start, dx, dy, dz)]
        if out is not None and _is_dev(out):
            if out.numel() != nx * ny * nz:
                raise ValueError("out.size() != nx*ny*nz")
            check(lib.agwn_grid_eval_dev(self._h, *[ptr(a) for a in vs], nx, ny, nz, ptr(out), _stream_ptr(stream)))
            return out
        res = np.empty(nx * ny * nz, dtype=np.float32) if out is None else out
        if res.shape != (nx * ny * nz,) or res.dtype != np.float32:
            raise ValueError("out.size() != nx*ny*nz")
        check(lib.agwn_grid_eval(self._h, *[ptr(a) for a in vs], nx, ny, nz, ptr(res)))
        return res

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.agwn_grid_destroy(self._h)
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


class HipMeshBatch:
    """Many meshes with uneven query counts, evaluated together (agwn_batch_*): eval(positions, q_off, x0) returns the
    full GWN (eval_gwnr_mesh_batch, gwn_mesh.hh:201-219) of each mesh's own queries, in two persistent launches
    captured in a HIP graph. meshes: a list of (vertices, indices, boundary) with float32 data and boundary weights
    that are small integers (as build_boundary_segments returns)."""

    def __init__(self, meshes):
        vs, ids, segs = [], [], []
        voff, ioff, soff = [0], [0], [0]
        for v, i, b in meshes:
            v = _host(v, np.float32, 3)
            i = _host(i, np.int32).reshape(-1)
            b = _host(b if b is not None else np.zeros((0, 7)), np.float32, 7)
            vs.append(v)
            ids.append(i)
            segs.append(b)
            voff.append(voff[-1] + len(v))
            ioff.append(ioff[-1] + i.size)
            soff.append(soff[-1] + len(b))
        self.k = len(meshes)
        cat = lambda parts, dt, cols: (np.ascontiguousarray(np.concatenate(parts).astype(dt))
                                       if parts else np.zeros((0, cols), dt))
        v = cat(vs, np.float32, 3)
        i = cat(ids, np.int32, 1).reshape(-1)
        b = cat(segs, np.float32, 7)
        self._offs = [np.asarray(o, dtype=np.int64) for o in (voff, ioff, soff)]
        h = C.c_void_p()
        check(lib.agwn_batch_create(self.k, ptr(v) if v.size else None, ptr(self._offs[0]), ptr(i) if i.size else None,
                                    ptr(self._offs[1]), ptr(b) if b.size else None, ptr(self._offs[2]), C.byref(h)))
        self._h = h

    def tune(self, positions, q_off, x0, out, stream=None) -> dict:
        q = np.ascontiguousarray(np.asarray(q_off, dtype=np.int64))
        check(lib.agwn_batch_tune(self._h, ptr(positions), ptr(q), ptr(_x0(x0, np.float32)), ptr(out),
                                  _stream_ptr(stream)))
        return self.info()

    def eval(self, positions, q_off, x0, out=None, stream=None):
        """positions: torch CUDA float32 (n, 3), each mesh's queries contiguous in mesh order; q_off: k + 1 offsets."""
        import torch

        q = np.ascontiguousarray(np.asarray(q_off, dtype=np.int64))
        if len(q) != self.k + 1 or q[0] != 0 or positions.numel() != 3 * q[-1]:
            raise ValueError("q_off must have k + 1 offsets from 0 to the number of positions")
        if out is None:
            out = torch.empty(int(q[-1]), dtype=torch.float32, device=positions.device)
        check(lib.agwn_batch_eval_dev(self._h, ptr(positions), ptr(q), ptr(_x0(x0, np.float32)), ptr(out),
                                      _stream_ptr(stream)))
        return out

    # the tuner's search space (agwn.h: AGWN_BATCH_FRAC_SHAPES, AGWN_BATCH_LEAF_SIZES)
    FRAC_SHAPES = 15
    LEAF_SIZES = 5

    @staticmethod
    def frac_shape(v: int) -> tuple[int, int, int]:
        """(queries per lane, workgroup size, stream points per LDS stage in workgroup sizes) of shape v."""
        q, b, s = C.c_int(), C.c_int(), C.c_int()
        check(lib.agwn_batch_frac_shape(int(v), C.byref(q), C.byref(b), C.byref(s)))
        return q.value, b.value, s.value

    @staticmethod
    def leaf_size(l: int) -> int:
        """Largest LBVH subtree (triangles) the counting traversal tests as one leaf bucket, for size index l."""
        v = C.c_int()
        check(lib.agwn_batch_leaf_size(int(l), C.byref(v)))
        return v.value

    def set_tuning(self, frac_shape: int, leaf: int) -> None:
        """Uses fractional shape `frac_shape` and leaf-bucket size index `leaf` instead of tuning."""
        check(lib.agwn_batch_set_tuning(self._h, int(frac_shape), int(leaf)))

    def info(self) -> dict:
        k, t, u = C.c_int64(), C.c_int64(), C.c_int64()
        fv, lv = C.c_int(), C.c_int()
        fms = np.zeros(self.FRAC_SHAPES, dtype=np.float64)
        lms = np.zeros(self.LEAF_SIZES, dtype=np.float64)
        check(lib.agwn_batch_info(self._h, C.byref(k), C.byref(t), C.byref(u), C.byref(fv), C.byref(lv), ptr(fms),
                                  ptr(lms)))
        q, b, s = self.frac_shape(fv.value) if fv.value >= 0 else (0, 0, 0)
        return {"meshes": k.value, "triangles": t.value, "unit_segments": u.value, "frac_shape": fv.value,
                "queries_per_lane": q, "block": b, "stage": s,
                "leaf_size": self.leaf_size(lv.value) if lv.value >= 0 else 0,
                "frac_table": [{"shape": v, "q": self.frac_shape(v)[0], "block": self.frac_shape(v)[1],
                                "stage": self.frac_shape(v)[2], "ms": float(fms[v])} for v in range(self.FRAC_SHAPES)
                               if fms[v] > 0],
                "leaf_table": [{"leaf": self.leaf_size(l), "ms": float(lms[l])} for l in range(self.LEAF_SIZES)
                               if lms[l] > 0]}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.agwn_batch_destroy(self._h)
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
