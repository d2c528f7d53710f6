"""The C ABI (include/agwn.h) on the CPU: the library loads, exports every declared symbol, rejects bad arguments with
a status and a message (the reference asserts on span sizes, gwn_mesh.hh:74,134,209), and its host-only
build_boundary_segments matches the reference's multiset. No kernel is launched here."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import load_golden
from oracle import sort_segments
from antipodal_amd import meshgen as mg


def test_every_declared_symbol_is_exported(agwn):
    from antipodal_amd import _lib

    names = _lib.header_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(_lib.lib, n)]
    assert not missing, f"declared in include/agwn.h but not exported: {missing}"


def test_version_and_error_string(agwn):
    from antipodal_amd._lib import lib

    assert lib.agwn_version() >= 100
    assert isinstance(lib.agwn_last_error(), bytes)


def test_argument_errors_are_statuses(agwn):
    from antipodal_amd._lib import lib

    x0 = np.array([0.0, 0.0, 1.0])
    out = np.zeros(4)
    # negative count
    assert lib.agwn_frac_batch_f64(None, -1, None, 0, x0.ctypes.data, out.ctypes.data) == 1
    assert "negative" in agwn.last_error()
    # null positions with n > 0
    assert lib.agwn_frac_batch_f64(None, 4, None, 0, x0.ctypes.data, out.ctypes.data) == 1
    # null boundary with nseg > 0
    pos = np.zeros((4, 3))
    assert lib.agwn_frac_batch_f64(pos.ctypes.data, 4, None, 3, x0.ctypes.data, out.ctypes.data) == 1
    # null x0
    assert lib.agwn_frac_batch_f64(pos.ctypes.data, 4, None, 0, None, out.ctypes.data) == 1
    # n == 0 is a no-op success, without touching the device
    assert lib.agwn_frac_batch_f64(None, 0, None, 0, x0.ctypes.data, None) == 0
    # mesh: index count not a multiple of 3, index out of range
    v = np.zeros((3, 3))
    h = C.c_void_p()
    bad = np.array([0, 1], dtype=np.int32)
    assert lib.agwn_mesh_create_f64(v.ctypes.data, 3, bad.ctypes.data, 2, C.byref(h)) == 1
    oob = np.array([0, 1, 3], dtype=np.int32)
    assert lib.agwn_mesh_create_f64(v.ctypes.data, 3, oob.ctypes.data, 3, C.byref(h)) == 1
    assert "out of range" in agwn.last_error()
    # a null handle is rejected, not dereferenced
    cnt = np.zeros(1, dtype=np.int32)
    assert lib.agwn_int_batch_f64(None, pos.ctypes.data, 1, x0.ctypes.data, cnt.ctypes.data) == 1
    assert lib.agwn_mesh_info(None, None, None, None) == 1


def test_python_layer_size_checks(agwn):
    with pytest.raises(ValueError):
        agwn.eval_gwnr_mesh_batch_fractional(np.zeros((3, 3)), [0, 0, 1], np.zeros((0, 7)), out=np.zeros(2))
    with pytest.raises(ValueError):
        agwn.build_boundary_segments(np.zeros((3, 3)), np.array([0, 1], dtype=np.int32))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_boundary_matches_reference_multiset(agwn, dt):
    g = load_golden("c1_icosphere_cap")
    s = "f64" if dt == np.float64 else "f32"
    verts, idx = mg.icosphere_cap_mesh()
    got = agwn.build_boundary_segments(verts.astype(dt), idx)
    assert got.dtype == dt
    assert np.array_equal(sort_segments(got), sort_segments(g[f"seg_{s}"]))
    cube = load_golden("cube")
    got = agwn.build_boundary_segments(cube["verts"].astype(dt), cube["open_y"])
    assert np.array_equal(sort_segments(got), sort_segments(cube[f"boundary_ref_{s}"]))


def test_boundary_edge_cases(agwn, oracle):
    # closed mesh -> empty boundary; empty mesh -> empty boundary
    cube = load_golden("cube")
    assert len(agwn.build_boundary_segments(cube["verts"], cube["closed"])) == 0
    assert len(agwn.build_boundary_segments(np.zeros((0, 3)), np.zeros(0, dtype=np.int32))) == 0
    # non-manifold: three triangles on one edge, two traversing it 0->1 -> net weight 1 (common.hh:121-122);
    # a fourth fin traversing 0->1 again -> weight 2 (the surviving net count becomes the weight)
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=np.float64)
    idx = np.array([0, 1, 2, 1, 0, 3, 0, 1, 4, 0, 1, 5], dtype=np.int32)
    got = agwn.build_boundary_segments(v, idx)
    ref = oracle.boundary(v, idx)
    assert np.array_equal(sort_segments(got), sort_segments(ref))
    assert sorted(got[:, 6].tolist()) == sorted(ref[:, 6].tolist())
    assert 2.0 in got[:, 6]


def test_boundary_torus_matches_oracle(agwn, oracle):
    verts, idx = mg.holed_torus_mesh()
    got = agwn.build_boundary_segments(verts, idx)
    g = load_golden("torus_sub")
    assert len(got) == int(g["nseg_f64"])
    assert np.array_equal(sort_segments(got), sort_segments(oracle.boundary(verts, idx)))


def test_batch_tuning_space(agwn):
    """The batch autotuner's search space (host tables, no launch): 15 distinct fractional kernel shapes -- queries
    per lane x workgroup size x stream points per LDS stage -- and leaf-bucket sizes 1..16; indexes outside the space
    are errors."""
    from antipodal_amd.api import HipMeshBatch

    shapes = [HipMeshBatch.frac_shape(v) for v in range(HipMeshBatch.FRAC_SHAPES)]
    assert len(set(shapes)) == 15
    assert {s[2] for s in shapes} == {1, 2, 4} and {s[1] for s in shapes} == {128, 256}
    assert [HipMeshBatch.leaf_size(l) for l in range(HipMeshBatch.LEAF_SIZES)] == [1, 2, 4, 8, 16]
    with pytest.raises(agwn.AgwnError, match="out of range"):
        HipMeshBatch.frac_shape(15)
    with pytest.raises(agwn.AgwnError, match="out of range"):
        HipMeshBatch.leaf_size(-1)
