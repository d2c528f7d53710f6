"""CPU: OBJ ingest (antipodal_amd/obj.py, the reader of examples/main.cc:52-94) and the slice example's host side
(antipodal_amd/slice.py, examples/main.cc:157-231): plane basis, grid, colour map, PNG writer, CLI errors.

The fixture tests/golden/rotated_cube_open.obj is the reference example's own mesh (a rotated unit cube without its
+z face). Closed forms pin what the restated formulas could get wrong: the grid lies on the plane, is centred on the
in-plane bounding box and spaced side / res along a right-handed basis; the colour map hits its three anchors and
clamps; the PNG decodes (CRC, zlib) back to the pixels; the oracle's full GWN at the cube's centre is 5/6 (five of six
faces, each 1/6 of the sphere seen from the centre; 1e-6 because the file keeps six decimals).
"""
from __future__ import annotations

import struct
import zlib

import numpy as np
import pytest

from antipodal_amd import slice as sl
from antipodal_amd.obj import load_obj, parse_obj
from conftest import GOLDEN

CUBE = GOLDEN / "rotated_cube_open.obj"


def test_parse_obj_tokens_and_comments():
    text = "# c\nv 0 0 0\nvn 0 0 1\nvt 0 0\nv 1 0 0\nv 1 1 0\n\no name\nf 1/1 2//1 3/1/1\nv 0 1 0\nf -4 -2 -1\n"
    v, idx = parse_obj(text)
    assert v.dtype == np.float64 and v.shape == (4, 3)
    assert idx.dtype == np.int32
    # relative indices count back from the vertices read so far (4 here): -4 -> 0, -2 -> 2, -1 -> 3
    assert idx.tolist() == [0, 1, 2, 0, 2, 3]


def test_parse_obj_fans_polygons_from_the_first_corner():
    text = "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv -1 1 0\nf 1 2 3 4 5\nf 2 3\n"
    v, idx = parse_obj(text)
    assert idx.reshape(-1, 3).tolist() == [[0, 1, 2], [0, 2, 3], [0, 3, 4]]  # a 2-gon adds nothing


def test_load_obj_fixture_and_missing_file(agwn, tmp_path):
    v, idx = load_obj(CUBE)
    assert v.shape == (8, 3) and idx.size == 30
    np.testing.assert_allclose(v.mean(axis=0), 0.0, atol=1e-12)  # opposite corners are negatives of each other
    seg = agwn.build_boundary_segments(v, idx)
    assert len(seg) == 4 and np.all(seg[:, 6] == 1.0)  # the missing face's four edges
    with pytest.raises(OSError):
        load_obj(tmp_path / "absent.obj")


def test_plane_basis_is_right_handed_and_orthonormal():
    for n in [(0, 1, 0), (1, 0, 0), (0.95, 0.1, 0.0), (-0.3, 0.2, 0.9), (0, 0, -2)]:
        nn, u, w = sl.plane_basis(n)
        m = np.stack([u, w, nn])
        np.testing.assert_allclose(m @ m.T, np.eye(3), atol=1e-14)
        np.testing.assert_allclose(np.cross(u, w), nn, atol=1e-14)
        np.testing.assert_allclose(nn, np.asarray(n, float) / np.linalg.norm(n), atol=1e-15)
    # the default normal (0, 1, 0) gives u = +z, v = +x
    _, u, w = sl.plane_basis((0, 1, 0))
    assert u.tolist() == [0.0, 0.0, 1.0] and w.tolist() == [1.0, 0.0, 0.0]


@pytest.mark.parametrize("normal,pos,res,extent", [((0, 1, 0), 0.5, 33, 1.2), ((-0.3, 0.2, 0.9), 0.25, 40, 2.0)])
def test_slice_grid_geometry(normal, pos, res, extent):
    v, _ = load_obj(CUBE)
    p = sl.slice_points(v, normal, pos, res, extent)
    assert p.shape == (res * res, 3)
    n, u, w = sl.plane_basis(normal)
    d, a, b = v @ n, v @ u, v @ w
    np.testing.assert_allclose(p @ n, d.min() + pos * (d.max() - d.min()), atol=1e-14)
    side = extent * max(np.ptp(a), np.ptp(b))
    g = p.reshape(res, res, 3)
    # row j, column i: steps of side / res along u (columns) and v (rows)
    np.testing.assert_allclose((g[:, 1:] - g[:, :-1]) @ u, side / res, rtol=1e-12)
    np.testing.assert_allclose((g[:, 1:] - g[:, :-1]) @ w, 0.0, atol=1e-14)
    np.testing.assert_allclose((g[1:] - g[:-1]) @ w, side / res, rtol=1e-12)
    # centred on the in-plane bounding box, spanning side (1 - 1 / res) between the outer cell centres
    np.testing.assert_allclose([(p @ u).mean(), (p @ w).mean()], [0.5 * (a.min() + a.max()), 0.5 * (b.min() + b.max())],
                               atol=1e-12)
    np.testing.assert_allclose(np.ptp(p @ u), side * (1 - 1 / res), rtol=1e-12)


def test_degenerate_mesh_uses_unit_side():
    p = sl.slice_points(np.zeros((3, 3)), (0, 0, 1), 0.5, 4, 1.0)
    np.testing.assert_allclose(np.ptp(p, axis=0), [0.75, 0.75, 0.0])


def test_colormap_anchors_blend_and_clamp():
    c = sl.colormap(np.array([0.0, 1.0, -1.0, 2.5, -7.0, 0.5, -0.25, np.nextafter(0.0, -1.0)]))
    assert c.dtype == np.uint8
    assert c[0].tolist() == [47, 89, 148]
    assert c[1].tolist() == [243, 230, 118] and c[3].tolist() == [243, 230, 118]
    assert c[2].tolist() == [210, 85, 46] and c[4].tolist() == [210, 85, 46]
    assert c[5].tolist() == [145, 160, 133]  # 47 + 196 / 2 = 145, 89 + 141 / 2 = 159.5 -> 160, 148 - 30 / 2 = 133
    assert c[6].tolist() == [88, 88, 123]  # 47 + 163 / 4 = 87.75, 89 - 4 / 4 = 88, 148 - 102 / 4 = 122.5 -> 123
    assert c[7].tolist() == [47, 89, 148]


def _read_png(path):
    data = path.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    while pos < len(data):
        (ln,) = struct.unpack(">I", data[pos: pos + 4])
        tag, body = data[pos + 4: pos + 8], data[pos + 8: pos + 8 + ln]
        (crc,) = struct.unpack(">I", data[pos + 8 + ln: pos + 12 + ln])
        assert crc == zlib.crc32(tag + body) & 0xFFFFFFFF
        chunks.setdefault(tag, b"")
        chunks[tag] += body
        pos += 12 + ln
    assert pos == len(data) and b"IEND" in chunks
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[b"IHDR"])
    assert (depth, ctype, comp, filt, inter) == (8, 2, 0, 0, 0)
    raw = np.frombuffer(zlib.decompress(chunks[b"IDAT"]), dtype=np.uint8).reshape(h, 1 + 3 * w)
    assert np.all(raw[:, 0] == 0)
    return raw[:, 1:].reshape(h, w, 3)


def test_png_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (7, 11, 3), dtype=np.uint8)
    sl.write_png(tmp_path / "x.png", img)
    np.testing.assert_array_equal(_read_png(tmp_path / "x.png"), img)


def test_cli_errors(tmp_path, capsys):
    assert sl.main([]) == 1
    assert "usage" in capsys.readouterr().err
    assert sl.main([str(CUBE), str(tmp_path / "o.png"), "0"]) == 1
    assert "resolution must be positive" in capsys.readouterr().err
    assert sl.main([str(tmp_path / "absent.obj"), str(tmp_path / "o.png"), "8"]) == 1
    assert "could not open" in capsys.readouterr().err
    (tmp_path / "empty.obj").write_text("v 0 0 0\n")
    assert sl.main([str(tmp_path / "empty.obj"), str(tmp_path / "o.png"), "8"]) == 1
    assert "no triangles" in capsys.readouterr().err


def test_oracle_centre_value_is_five_sixths(oracle, agwn):
    v, idx = load_obj(CUBE)
    seg = agwn.build_boundary_segments(v, idx)
    p = sl.slice_points(v, (0, 1, 0), 0.5, 49, 1.2)
    assert np.array_equal(p[24 * 49 + 24], [0.0, 0.0, 0.0])
    w = oracle.full(v, idx, p[24 * 49 + 24: 24 * 49 + 25], sl.X0, seg)
    assert abs(w[0] - 5.0 / 6.0) < 1e-6
