"""ctypes binding of libagwn.so (include/agwn.h).

The library is the HIP product; this module only declares argument types, converts numpy / torch buffers to
pointers and raises AgwnError on a nonzero status. There is no CPU fallback: if libagwn.so is missing, importing
antipodal_amd fails with ImportError.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().with_name("libagwn.so")
HEADER = Path(__file__).resolve().parent.parent / "include" / "agwn.h"

_i64 = C.c_int64
_p = C.c_void_p
_int = C.c_int


class AgwnError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"agwn status {code}: {msg}")
        self.code = code


# argument signatures; every entry point returns an int status except the informational ones
_SIG = {
    "agwn_last_error": (C.c_char_p, []),
    "agwn_version": (_int, []),
    "agwn_device_available": (_int, []),
    "agwn_frac_batch_f32": (_int, [_p, _i64, _p, _i64, _p, _p]),
    "agwn_frac_batch_f64": (_int, [_p, _i64, _p, _i64, _p, _p]),
    "agwn_frac_batch_f32_dev": (_int, [_p, _i64, _p, _i64, _p, _p, _p]),
    "agwn_frac_batch_f64_dev": (_int, [_p, _i64, _p, _i64, _p, _p, _p]),
    "agwn_frac_batch_f64_dev_noalloc": (_int, [_p, _i64, _p, _i64, _p, _p, _p]),
    "agwn_mesh_create_f32": (_int, [_p, _i64, _p, _i64, C.POINTER(_p)]),
    "agwn_mesh_create_f64": (_int, [_p, _i64, _p, _i64, C.POINTER(_p)]),
    "agwn_mesh_destroy": (_int, [_p]),
    "agwn_mesh_info": (_int, [_p, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_int)]),
    "agwn_int_batch_f32": (_int, [_p, _p, _i64, _p, _p]),
    "agwn_int_batch_f64": (_int, [_p, _p, _i64, _p, _p]),
    "agwn_int_batch_f32_dev": (_int, [_p, _p, _i64, _p, _p, _p]),
    "agwn_int_batch_f64_dev": (_int, [_p, _p, _i64, _p, _p, _p]),
    "agwn_full_batch_f32": (_int, [_p, _p, _i64, _p, _p, _i64, _p]),
    "agwn_full_batch_f64": (_int, [_p, _p, _i64, _p, _p, _i64, _p]),
    "agwn_full_batch_f32_dev": (_int, [_p, _p, _i64, _p, _p, _i64, _p, _p]),
    "agwn_full_batch_f64_dev": (_int, [_p, _p, _i64, _p, _p, _i64, _p, _p]),
    "agwn_robust_create": (_int, [_int, _int, _p, _i64, _p, _i64, _p, _i64, C.POINTER(_p)]),
    "agwn_robust_destroy": (_int, [_p]),
    "agwn_robust_info": (_int, [_p, C.POINTER(_int), C.POINTER(_int), C.POINTER(_i64), C.POINTER(_i64), _p,
                                C.POINTER(C.c_double)]),
    "agwn_robust_eval_batch": (_int, [_p, _p, _i64, _p]),
    "agwn_robust_fractional_batch": (_int, [_p, _p, _i64, _p]),
    "agwn_robust_count_batch": (_int, [_p, _p, _i64, _p]),
    "agwn_robust_eval_batch_dev": (_int, [_p, _p, _i64, _p, _p]),
    "agwn_robust_fractional_batch_dev": (_int, [_p, _p, _i64, _p, _p]),
    "agwn_robust_count_batch_dev": (_int, [_p, _p, _i64, _p, _p]),
    "agwn_grid_create": (_int, [C.POINTER(_p)]),
    "agwn_grid_destroy": (_int, [_p]),
    "agwn_grid_set_scene": (_int, [_p, _p, _i64, _p, _i64, _p, _i64]),
    "agwn_grid_eval": (_int, [_p, _p, _p, _p, _p, _int, _int, _int, _p]),
    "agwn_grid_eval_dev": (_int, [_p, _p, _p, _p, _p, _int, _int, _int, _p, _p]),
    "agwn_boundary_f32": (_int, [_p, _i64, _p, _i64, _p, _i64, C.POINTER(_i64)]),
    "agwn_boundary_f64": (_int, [_p, _i64, _p, _i64, _p, _i64, C.POINTER(_i64)]),
    "agwn_boundary_f32_dev": (_int, [_p, _i64, _p, _i64, _p, _i64, C.POINTER(_i64), _p]),
    "agwn_boundary_f64_dev": (_int, [_p, _i64, _p, _i64, _p, _i64, C.POINTER(_i64), _p]),
    "agwn_batch_create": (_int, [_i64, _p, _p, _p, _p, _p, _p, C.POINTER(_p)]),
    "agwn_batch_destroy": (_int, [_p]),
    "agwn_batch_tune": (_int, [_p, _p, _p, _p, _p, _p]),
    "agwn_batch_eval_dev": (_int, [_p, _p, _p, _p, _p, _p]),
    "agwn_batch_info": (_int, [_p, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_int),
                               C.POINTER(_int), _p, _p]),
    "agwn_batch_set_tuning": (_int, [_p, _int, _int]),
    "agwn_batch_frac_shape": (_int, [_int, C.POINTER(_int), C.POINTER(_int), C.POINTER(_int)]),
    "agwn_batch_leaf_size": (_int, [_int, C.POINTER(_int)]),
    "agwn_mlp_num_params": (_i64, []),
    "agwn_mlp_create": (_int, [_p, _p, C.c_float, C.c_uint64, C.POINTER(_p)]),
    "agwn_mlp_destroy": (_int, [_p]),
    "agwn_mlp_info": (_int, [_p, _p, _p, _p]),
    "agwn_mlp_get_params": (_int, [_p, _p]),
    "agwn_mlp_set_params": (_int, [_p, _p]),
    "agwn_mlp_train_dev": (_int, [_p, _p, _p, _i64, _int, _int, C.c_float, C.c_uint64, _p, _p]),
    "agwn_mlp_grad": (_int, [_p, _p, _p, _i64, _p, _p, C.POINTER(C.c_float)]),
    "agwn_mlp_predict_dev": (_int, [_p, _p, _i64, _p, _p]),
    "agwn_mlp_predict_grid_dev": (_int, [_p, _p, _p, _p, _p, _int, _int, _int, _p, _p]),
    "agwn_mlp_set_surface": (_int, [_p, _p, _i64, _p, _i64, C.c_float, C.c_float]),
    "agwn_mlp_near_surface_dev": (_int, [_p, _p, _i64, _p, _p]),
    "agwn_training_samples_dev": (_int, [_p, _p, _p, _i64, _p, _p, _i64, C.c_double, C.c_double, C.c_double,
                                         C.c_uint64, _p, _p]),
    "agwn_label_surrogate_dev": (_int, [_p, _p, _p, _i64, _p, _p, _i64, C.c_float, _i64, C.c_uint64, _p, _p, _p]),
    "agwn_grid_label_surrogate_dev": (_int, [_p, _p, _p, _p, _p, _p, _int, _int, _int, C.c_float, _i64, C.c_uint64,
                                             _p, _p, _p]),
}


def header_symbols() -> list[str]:
    """Names of every function include/agwn.h declares."""
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(agwn_[a-z0-9_]+)\s*\(", text)))


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    # torch-ROCm ships its own HIP and HSA runtimes under the same sonames (libamdhip64.so.7, libhsa-runtime64.so.1)
    # but NEEDs them by their unversioned names. Loaded first, they satisfy libagwn.so's dependencies and the process
    # keeps one runtime; loaded after libagwn.so they become a second HSA instance that cannot open the device, and
    # torch then reports no GPU. So torch, when present, is imported before the library is opened.
    try:
        import torch  # noqa: F401
    except ImportError:  # pragma: no cover - the C ABI does not need torch
        pass
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIG.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue  # reported by the symbol test, not here
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(status: int) -> None:
    if status != 0:
        raise AgwnError(status, lib.agwn_last_error().decode(errors="replace"))


def last_error() -> str:
    return lib.agwn_last_error().decode(errors="replace")


def device_available() -> bool:
    return bool(lib.agwn_device_available())


def ptr(a) -> int | None:
    """Raw address of a C-contiguous numpy array or torch tensor (None for None)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return a.data_ptr()
    raise TypeError(f"unsupported buffer type {type(a)!r}")


def sfx(dtype) -> str:
    dt = np.dtype(dtype) if not hasattr(dtype, "is_floating_point") else None
    if dt is not None:
        if dt == np.float32:
            return "f32"
        if dt == np.float64:
            return "f64"
    else:  # torch dtype
        import torch

        if dtype == torch.float32:
            return "f32"
        if dtype == torch.float64:
            return "f64"
    raise TypeError(f"unsupported scalar type {dtype}")
