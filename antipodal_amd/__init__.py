"""antipodal_amd -- the batch mesh generalized-winding-number path of antipodal, on MI355X (gfx950).

The product is libagwn.so (hand-written HIP kernels behind the C ABI in include/agwn.h); the C++20 plugin headers in
include/antipodal_hip/ put it behind the reference's Dispatcher / Intersector / RobustMeshGwn API. This Python module
is a thin host layer over the same C ABI (used by the tests and bench.py); names follow the reference's functions
(gwn_mesh.hh, gwn_mesh_robust.hh, math/common.hh).

Importing it without a built libagwn.so raises ImportError: there is no CPU fallback.
"""
from __future__ import annotations

from . import _lib
from ._lib import AgwnError, device_available, last_error
from .api import (
    HipGridGwn,
    HipMesh,
    HipMeshBatch,
    RobustMeshGwn,
    eval_gwnr_mesh_batch_robust,
    build_boundary_segments,
    eval_gwnr_mesh_batch,
    eval_gwnr_mesh_batch_fractional,
    eval_gwnr_mesh_batch_integer,
)
from .surrogate import GwnSurrogate, TrainingSampler, fit_surrogate

__all__ = [
    "AgwnError",
    "GwnSurrogate",
    "TrainingSampler",
    "HipGridGwn",
    "HipMesh",
    "HipMeshBatch",
    "RobustMeshGwn",
    "eval_gwnr_mesh_batch_robust",
    "build_boundary_segments",
    "device_available",
    "eval_gwnr_mesh_batch",
    "eval_gwnr_mesh_batch_fractional",
    "eval_gwnr_mesh_batch_integer",
    "fit_surrogate",
    "last_error",
]
