"""The allocation-free FP64 fractional entry point (agwn_frac_batch_f64_dev_noalloc) against the allocating one
(agwn_frac_batch_f64_dev), through the C ABI on the GPU.

Both run the same two kernels (boundary rotation into the pole frame, then the FP64 pair sum), the first with its
workspace in the library's device array and the second in a stream-ordered allocation, so their outputs must be
equal bit for bit. Two calls in a row, the second with a smaller boundary and fewer queries than the first, check
that nothing of the first call's workspace leaks into the second; a boundary above the stated maximum is rejected
on the host with a clear error before anything is launched.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from antipodal_amd import meshgen as mg

pytestmark = pytest.mark.gpu

MAX_SEG = 1 << 19  # the entry point's stated maximum boundary size


def _noalloc(lib, pos, seg, x0, out, stream):
    n = pos.shape[0]
    nseg = seg.shape[0]
    return lib.agwn_frac_batch_f64_dev_noalloc(pos.data_ptr(), n, seg.data_ptr() if nseg else None, nseg,
                                               x0.ctypes.data, out.data_ptr(), stream)


def test_noalloc_matches_allocating_entry_point_twice(gpu):
    import torch

    from antipodal_amd import _lib

    lib = _lib.lib
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    x0 = mg.example_x0(np.float64)
    # call 1: a holed torus (a few thousand boundary segments), 60k queries in its expanded box
    verts, idx = mg.holed_torus_mesh(seed=3, nu=240, nv=84, holes=40, hole_radius=0.08)
    seg1 = gpu.build_boundary_segments(verts, idx)
    lo, hi = mg.expanded_bbox(verts)
    pos1 = mg.uniform_points(5, 60_000, lo, hi)
    # call 2: the open icosphere cap (36 segments), 5000 queries -- smaller in both sizes
    cv, ci = mg.icosphere_cap_mesh()
    seg2 = gpu.build_boundary_segments(cv, ci)
    pos2 = mg.uniform_points(6, 5_000, [-1.5] * 3, [1.5] * 3)
    assert len(seg2) < len(seg1) and len(pos2) < len(pos1)
    for k, (pos, seg) in enumerate([(pos1, seg1), (pos2, seg2)]):
        print(f"case noalloc call {k + 1}: n={len(pos)} nseg={len(seg)}", flush=True)
        p = torch.from_numpy(np.ascontiguousarray(pos)).to(dev)
        s = torch.from_numpy(np.ascontiguousarray(seg)).to(dev)
        want = torch.full((len(pos),), float("nan"), dtype=torch.float64, device=dev)
        gpu.eval_gwnr_mesh_batch_fractional(p, x0, s, out=want)
        got = torch.full((len(pos),), float("nan"), dtype=torch.float64, device=dev)
        st = _noalloc(lib, p, s, x0, got, stream)
        assert st == 0, _lib.last_error()
        torch.cuda.synchronize()
        w, g = want.cpu().numpy(), got.cpu().numpy()
        assert np.all(np.isfinite(w))
        assert np.array_equal(g, w), f"call {k + 1}: {int((g != w).sum())} of {len(w)} values differ"


def test_noalloc_rejects_boundary_above_maximum(gpu):
    import torch

    from antipodal_amd import _lib

    lib = _lib.lib
    dev = torch.device("cuda:0")
    x0 = mg.example_x0(np.float64)
    pos = torch.zeros((4, 3), dtype=torch.float64, device=dev)
    seg = torch.zeros((1, 7), dtype=torch.float64, device=dev)
    out = torch.zeros(4, dtype=torch.float64, device=dev)
    # the size is checked on the host before any pointer is used or any kernel launched
    st = lib.agwn_frac_batch_f64_dev_noalloc(pos.data_ptr(), 4, seg.data_ptr(), MAX_SEG + 1, x0.ctypes.data,
                                             out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    assert st == 1  # AGWN_ERR_INVALID
    assert "maximum" in _lib.last_error()
    # at the maximum itself the size check passes (n = 0: nothing is launched, nothing is read)
    st = lib.agwn_frac_batch_f64_dev_noalloc(pos.data_ptr(), 0, seg.data_ptr(), MAX_SEG, x0.ctypes.data,
                                             out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    assert st == 0, _lib.last_error()
