"""ORACLE -- test infrastructure only (see oracle/oracle.cc).

numpy wrappers over liboracle.so (the CPU restatement) and, where it was built, _ref/libref.so (the reference's own
code compiled from its headers). Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
package, and only as the checker or the timed CPU baseline -- never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libref.so"

POLICIES = {"int64_grid": 0, "int128_grid": 1, "exact_float": 2}

_i64, _p, _int = C.c_int64, C.c_void_p, C.c_int


def build(ref: bool = True) -> None:
    """make -C oracle (liboracle.so, and _ref/libref.so when the reference sources are present)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "all" if ref else "oracle"], check=True)


def _ptr(a):
    return None if a is None else a.ctypes.data


def _c(a, dt, cols=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=dt))
    return a.reshape(-1, cols) if cols else a


class _Lib:
    def __init__(self, path: Path, prefix: str):
        self.path = path
        self.prefix = prefix
        self._lib = None

    @property
    def lib(self) -> C.CDLL:
        if self._lib is None:
            if not self.path.exists():
                raise FileNotFoundError(f"{self.path} not built (make -C oracle)")
            self._lib = C.CDLL(str(self.path))
        return self._lib

    def available(self) -> bool:
        return self.path.exists()

    def fn(self, name, res=None):
        f = getattr(self.lib, f"{self.prefix}_{name}")
        f.restype = res
        return f


class Oracle(_Lib):
    """The CPU restatement (liboracle.so)."""

    def __init__(self):
        super().__init__(ORACLE_SO, "orc")

    def pair_half_area(self, x1, v0, v1, dtype=np.float64):
        s = "f32" if dtype == np.float32 else "f64"
        f = self.fn(f"pair_half_area_{s}", C.c_float if s == "f32" else C.c_double)
        f.argtypes = [_p, _p, _p]
        a, b, c = (_c(x, dtype) for x in (x1, v0, v1))
        return f(_ptr(a), _ptr(b), _ptr(c))

    def frac(self, pos, seg7, x0, dtype=np.float64, nthreads=0):
        s = "f32" if dtype == np.float32 else "f64"
        pos, seg, x = _c(pos, dtype, 3), _c(seg7, dtype, 7), _c(x0, dtype)
        out = np.empty(len(pos), dtype=dtype)
        f = self.fn(f"frac_batch_{s}")
        f.argtypes = [_p, _i64, _p, _i64, _p, _p, _int]
        f(_ptr(pos), len(pos), _ptr(seg), len(seg), _ptr(x), _ptr(out), nthreads)
        return out

    def frac_ext(self, pos, seg7, x0, nthreads=0):
        """Fractional term of double inputs computed in 80-bit extended precision (the arbiter of FP64 results)."""
        pos, seg, x = _c(pos, np.float64, 3), _c(seg7, np.float64, 7), _c(x0, np.float64)
        out = np.empty(len(pos), dtype=np.float64)
        f = self.fn("frac_batch_f64x")
        f.argtypes = [_p, _i64, _p, _i64, _p, _p, _int]
        f(_ptr(pos), len(pos), _ptr(seg), len(seg), _ptr(x), _ptr(out), nthreads)
        return out

    def naive_count(self, verts, idx, pos, direction, dtype=np.float64, nthreads=0):
        s = "f32" if dtype == np.float32 else "f64"
        v, i, pos, d = _c(verts, dtype, 3), _c(idx, np.int32), _c(pos, dtype, 3), _c(direction, dtype)
        out = np.empty(len(pos), dtype=np.int32)
        f = self.fn(f"naive_count_{s}")
        f.argtypes = [_p, _i64, _p, _i64, _p, _i64, _p, _p, _int]
        f(_ptr(v), len(v), _ptr(i), i.size, _ptr(pos), len(pos), _ptr(d), _ptr(out), nthreads)
        return out

    def full(self, verts, idx, pos, x0, seg7, dtype=np.float64, nthreads=0):
        s = "f32" if dtype == np.float32 else "f64"
        v, i, pos, x, seg = _c(verts, dtype, 3), _c(idx, np.int32), _c(pos, dtype, 3), _c(x0, dtype), _c(seg7, dtype, 7)
        out = np.empty(len(pos), dtype=dtype)
        f = self.fn(f"full_batch_{s}")
        f.argtypes = [_p, _i64, _p, _i64, _p, _i64, _p, _p, _i64, _p, _int]
        f(_ptr(v), len(v), _ptr(i), i.size, _ptr(pos), len(pos), _ptr(x), _ptr(seg), len(seg), _ptr(out), nthreads)
        return out

    def gwn_baseline(self, verts, idx, pos, dtype=np.float64, compensated=False, nthreads=0):
        s = "f32" if dtype == np.float32 else "f64"
        v, i, pos = _c(verts, dtype, 3), _c(idx, np.int32), _c(pos, dtype, 3)
        out = np.empty(len(pos), dtype=dtype)
        f = self.fn(f"gwn_baseline_{s}")
        f.argtypes = [_p, _i64, _p, _i64, _p, _i64, _p, _int, _int]
        f(_ptr(v), len(v), _ptr(i), i.size, _ptr(pos), len(pos), _ptr(out), int(compensated), nthreads)
        return out

    def boundary(self, verts, idx, dtype=np.float64):
        s = "f32" if dtype == np.float32 else "f64"
        v, i = _c(verts, dtype, 3), _c(idx, np.int32)
        out = np.empty((max(i.size, 1), 7), dtype=dtype)
        f = self.fn(f"boundary_{s}", _i64)
        f.argtypes = [_p, _i64, _p, _i64, _p, _i64]
        n = f(_ptr(v), len(v), _ptr(i), i.size, _ptr(out), i.size)
        return out[:n].copy()

    def preds(self, cases, ctype):
        """Batched robust predicates; cases (n, 12) = q|a|b|c. Returns edge_sign, num_sign, inside, w (n, 3)."""
        s = {np.int32: "i32", np.int64: "i64", np.float64: "f64"}[ctype]
        c = _c(cases, ctype, 12)
        n = len(c)
        es, ns, ins = (np.empty(n, dtype=np.int32) for _ in range(3))
        w = np.empty((n, 3), dtype=np.float64)
        f = self.fn(f"pred_batch_{s}")
        f.argtypes = [_p, _i64, _p, _p, _p, _p]
        f(_ptr(c), n, _ptr(es), _ptr(ns), _ptr(ins), _ptr(w))
        return es, ns, ins, w

    def edge_det(self, q, a, b, ctype):
        s = {np.int32: "i32", np.int64: "i64", np.float64: "f64"}[ctype]
        f = self.fn(f"edge_det_{s}", C.c_double)
        f.argtypes = [_p, _p, _p]
        q, a, b = (_c(x, ctype) for x in (q, a, b))
        return f(_ptr(q), _ptr(a), _ptr(b))

    def robust(self, verts, idx, seg7, pos, policy="exact_float", dtype=np.float64, nthreads=0):
        """RobustMeshGwn restatement: returns (fractional, count, eval) per position."""
        v, i, seg, pos = _c(verts, dtype, 3), _c(idx, np.int32), _c(seg7, dtype, 7), _c(pos, dtype, 3)
        fr = np.empty(len(pos), dtype=dtype)
        ev = np.empty(len(pos), dtype=dtype)
        cnt = np.empty(len(pos), dtype=np.int32)
        f = self.fn("robust_eval", _int)
        f.argtypes = [_int, _int, _p, _i64, _p, _i64, _p, _i64, _p, _i64, _p, _p, _p, _int]
        r = f(np.dtype(dtype).itemsize, POLICIES[policy], _ptr(v), len(v), _ptr(i), i.size, _ptr(seg), len(seg),
              _ptr(pos), len(pos), _ptr(fr), _ptr(cnt), _ptr(ev), nthreads)
        if r != 0:
            raise ValueError("orc_robust_eval failed")
        return fr, cnt, ev


class Reference(_Lib):
    """The reference's own code (oracle/_ref/libref.so), built from its headers in this container."""

    def __init__(self):
        super().__init__(REF_SO, "ref")

    # mode: 0 SinglethreadDispatcher, 1 OpenMPDispatcher, 2 PoolDispatcher
    def frac(self, pos, seg7, x0, dtype=np.float64, mode=0):
        s = "f32" if dtype == np.float32 else "f64"
        pos, seg, x = _c(pos, dtype, 3), _c(seg7, dtype, 7), _c(x0, dtype)
        out = np.empty(len(pos), dtype=dtype)
        f = self.fn(f"frac_batch_{s}")
        f.argtypes = [_p, _i64, _p, _i64, _p, _p, _int]
        f(_ptr(pos), len(pos), _ptr(seg), len(seg), _ptr(x), _ptr(out), mode)
        return out

    def naive_count(self, verts, idx, pos, direction, dtype=np.float64, mode=0):
        s = "f32" if dtype == np.float32 else "f64"
        v, i, pos, d = _c(verts, dtype, 3), _c(idx, np.int32), _c(pos, dtype, 3), _c(direction, dtype)
        out = np.empty(len(pos), dtype=np.int32)
        f = self.fn(f"naive_count_{s}")
        f.argtypes = [_p, _i64, _p, _i64, _p, _i64, _p, _p, _int]
        f(_ptr(v), len(v), _ptr(i), i.size, _ptr(pos), len(pos), _ptr(d), _ptr(out), mode)
        return out

    def full(self, verts, idx, pos, x0, seg7, dtype=np.float64, mode=0):
        s = "f32" if dtype == np.float32 else "f64"
        v, i, pos, x, seg = _c(verts, dtype, 3), _c(idx, np.int32), _c(pos, dtype, 3), _c(x0, dtype), _c(seg7, dtype, 7)
        out = np.empty(len(pos), dtype=dtype)
        f = self.fn(f"full_batch_{s}")
        f.argtypes = [_p, _i64, _p, _i64, _p, _i64, _p, _p, _i64, _p, _int]
        f(_ptr(v), len(v), _ptr(i), i.size, _ptr(pos), len(pos), _ptr(x), _ptr(seg), len(seg), _ptr(out), mode)
        return out

    def gwn_baseline(self, verts, idx, pos, dtype=np.float64):
        s = "f32" if dtype == np.float32 else "f64"
        v, i, pos = _c(verts, dtype, 3), _c(idx, np.int32), _c(pos, dtype, 3)
        out = np.empty(len(pos), dtype=dtype)
        f = self.fn(f"gwn_baseline_{s}")
        f.argtypes = [_p, _i64, _p, _i64, _p, _i64, _p]
        f(_ptr(v), len(v), _ptr(i), i.size, _ptr(pos), len(pos), _ptr(out))
        return out

    def boundary(self, verts, idx, dtype=np.float64):
        s = "f32" if dtype == np.float32 else "f64"
        v, i = _c(verts, dtype, 3), _c(idx, np.int32)
        out = np.empty((max(i.size, 1), 7), dtype=dtype)
        f = self.fn(f"boundary_{s}", _i64)
        f.argtypes = [_p, _i64, _p, _i64, _p, _i64]
        n = f(_ptr(v), len(v), _ptr(i), i.size, _ptr(out), i.size)
        return out[:n].copy()

    def pair_half_area(self, x1, v0, v1, dtype=np.float64):
        s = "f32" if dtype == np.float32 else "f64"
        f = self.fn(f"pair_half_area_{s}", C.c_float if s == "f32" else C.c_double)
        f.argtypes = [_p, _p, _p]
        a, b, c = (_c(x, dtype) for x in (x1, v0, v1))
        return f(_ptr(a), _ptr(b), _ptr(c))

    def preds(self, cases, ctype):
        s = {np.int32: "i32", np.int64: "i64", np.float64: "f64"}[ctype]
        c = _c(cases, ctype, 12)
        n = len(c)
        es, ns, ins = (np.empty(n, dtype=np.int32) for _ in range(3))
        w = np.empty((n, 3), dtype=np.float64)
        f = self.fn(f"pred_batch_{s}")
        f.argtypes = [_p, _i64, _p, _p, _p, _p]
        f(_ptr(c), n, _ptr(es), _ptr(ns), _ptr(ins), _ptr(w))
        return es, ns, ins, w

    def omp_threads(self) -> int:
        f = self.fn("omp_max_threads", _int)
        f.argtypes = []
        return f()


def sort_segments(seg7: np.ndarray) -> np.ndarray:
    """Canonical order of a boundary (for multiset comparisons): lexicographic over the 7 values."""
    seg7 = np.asarray(seg7)
    return seg7[np.lexsort(seg7.T[::-1])] if len(seg7) else seg7
