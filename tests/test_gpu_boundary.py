"""GPU build_boundary_segments (math/common.hh:123-174): the device extraction returns the host extraction's array
element for element (same order: sorted by (min index, max index)), and so the reference's multiset, which the host
extraction is checked against in test_capi.py and test_oracle.py."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden
from antipodal_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def both(agwn, v, idx, dt):
    import torch

    host = agwn.build_boundary_segments(np.asarray(v, dtype=dt), idx)
    vt = torch.from_numpy(np.ascontiguousarray(v, dtype=dt)).cuda()
    it = torch.from_numpy(np.ascontiguousarray(idx, dtype=np.int32)).cuda()
    dev = agwn.build_boundary_segments(vt, it).cpu().numpy()
    return host, dev


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_boundary_matches_host(gpu, dt):
    cases = {"c1": mg.icosphere_cap_mesh()}
    g = load_golden("cube")
    cases["cube_closed"] = (g["verts"], g["closed"])
    cases["cube_open_y"] = (g["verts"], g["open_y"])
    cases["torus"] = mg.holed_torus_mesh()
    # non-manifold: three triangles on one edge, one of them reversed; a repeated triangle; a degenerate one
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [2, 2, 2]], dtype=np.float64)
    cases["nonmanifold"] = (v, np.array([0, 1, 2, 1, 0, 3, 0, 1, 4, 0, 1, 2, 5, 5, 5], dtype=np.int32))
    for name, (v, idx) in cases.items():
        host, dev = both(gpu, v, idx, dt)
        assert host.shape == dev.shape, (name, host.shape, dev.shape)
        assert np.array_equal(host, dev), name
    assert len(both(gpu, *cases["cube_closed"], dt)[1]) == 0
    assert len(both(gpu, *cases["torus"], dt)[1]) == 31376


def test_boundary_edge_cases(gpu):
    import torch

    v = torch.zeros((3, 3), dtype=torch.float32, device="cuda")
    empty = gpu.build_boundary_segments(v, torch.zeros(0, dtype=torch.int32, device="cuda"))
    assert empty.shape == (0, 7)
    with pytest.raises(gpu.AgwnError, match="out of range"):
        gpu.build_boundary_segments(v, torch.tensor([0, 1, 3], dtype=torch.int32, device="cuda"))
    with pytest.raises(ValueError):
        gpu.build_boundary_segments(v, torch.tensor([0, 1], dtype=torch.int32, device="cuda"))


@pytest.mark.parametrize("s,dt", [("f64", np.float64), ("f32", np.float32)])
def test_boundary_matches_reference_golden(gpu, s, dt):
    """The device extraction against the reference's own build_boundary_segments output (c1_icosphere_cap.npz and
    cube.npz, written by the reference's code): the same multiset of oriented weighted segments (the reference's order
    is unordered_map order, so both are sorted before comparing)."""
    import torch

    from oracle import sort_segments

    c1 = load_golden("c1_icosphere_cap")
    cube = load_golden("cube")
    cases = [("c1", c1["verts"], c1["idx"], c1[f"seg_{s}"]),
             ("cube_open_y", cube["verts"], cube["open_y"], cube[f"boundary_ref_{s}"])]
    for name, v, idx, ref in cases:
        vt = torch.from_numpy(np.ascontiguousarray(v, dtype=dt)).cuda()
        it = torch.from_numpy(np.ascontiguousarray(idx, dtype=np.int32)).cuda()
        dev = gpu.build_boundary_segments(vt, it).cpu().numpy()
        assert dev.dtype == ref.dtype and dev.shape == ref.shape, (name, dev.shape, ref.shape)
        assert np.array_equal(sort_segments(dev), sort_segments(ref)), name
