"""GPU: the slice example end to end (antipodal_amd/slice.py, examples/main.cc:157-231) on the reference example's
mesh -- the full FP64 GWN at every grid point against the oracle (Oracle.full, gwn_mesh.hh:86-113) to 1e-12, the
cube's centre at 5/6, and the CLI's PNG against the colour map of the oracle's values (at most 1 level apart, where a
value sits on a rounding edge).
"""
from __future__ import annotations

import numpy as np
import pytest

from antipodal_amd import slice as sl
from antipodal_amd.obj import load_obj
from conftest import GOLDEN
from test_obj_slice import _read_png

pytestmark = pytest.mark.gpu

CUBE = GOLDEN / "rotated_cube_open.obj"


@pytest.mark.parametrize("normal,pos,res,extent", [((0, 1, 0), 0.5, 49, 1.2), ((-0.3, 0.2, 0.9), 0.3, 64, 2.0),
                                                   ((1, 0, 0), 0.9, 40, 1.5)])
def test_slice_matches_oracle(gpu, oracle, normal, pos, res, extent):
    v, idx = load_obj(CUBE)
    wn = sl.slice_gwn(v, idx, normal, pos, res, extent)
    assert wn.shape == (res, res) and wn.dtype == np.float64
    pts = sl.slice_points(v, normal, pos, res, extent)
    ref = oracle.full(v, idx, pts, sl.X0, gpu.build_boundary_segments(v, idx)).reshape(res, res)
    np.testing.assert_allclose(wn, ref, rtol=0, atol=1e-12)
    if res % 2 and pos == 0.5 and normal == (0, 1, 0):
        assert abs(wn[res // 2, res // 2] - 5.0 / 6.0) < 1e-6


def test_cli_writes_the_oracle_image(gpu, oracle, tmp_path, capsys):
    out = tmp_path / "slice.png"
    assert sl.main([str(CUBE), str(out), "32", "0.4", "0.2", "1", "0.1", "1.3"]) == 0
    log = capsys.readouterr().out
    assert "loaded 8 vertices, 10 triangles" in log and "open boundary: 4 segments" in log
    img = _read_png(out)
    assert img.shape == (32, 32, 3)
    v, idx = load_obj(CUBE)
    pts = sl.slice_points(v, (0.2, 1, 0.1), 0.4, 32, 1.3)
    ref = oracle.full(v, idx, pts, sl.X0, gpu.build_boundary_segments(v, idx)).reshape(32, 32)
    want = sl.colormap(ref[::-1])
    assert np.abs(img.astype(int) - want.astype(int)).max() <= 1
    assert len(np.unique(img.reshape(-1, 3), axis=0)) > 20  # a real picture, not one flat colour
