"""Slice image of a mesh's generalized winding number, end to end on the GPU path (examples/main.cc:157-231).

    python -m antipodal_amd.slice <input.obj> <output.png> <resolution> [slice_pos=0.5] [nx ny nz = 0 1 0] [extent=1.2]

The mesh is read with load_obj, its boundary extracted (build_boundary_segments), and the full GWN (fractional +
integer term, FP64, x0 = normalize(0.31, 0.57, 0.76)) evaluated by HipMesh at a res x res grid on the plane
dot(x, n) = dmin + slice_pos (dmax - dmin), where [dmin, dmax] is the mesh's extent along n. The grid is centred on
the mesh's in-plane bounding box, with side extent x its larger in-plane extent, in the in-plane basis
u = normalize(h x n), v = n x u (h = x unless n is within ~25 degrees of x, then y). Colours blend from #2F5994 at
w = 0 toward #F3E676 at w = 1 and #D2552E at w = -1 (clamped), row 0 at the top (largest v).
"""
from __future__ import annotations

import struct
import sys
import zlib

import numpy as np

X0 = np.array([0.31, 0.57, 0.76]) / np.linalg.norm([0.31, 0.57, 0.76])


def plane_basis(normal):
    n = np.asarray(normal, dtype=np.float64)
    n = n / np.linalg.norm(n)
    h = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    u = np.cross(h, n)
    u /= np.linalg.norm(u)
    return n, u, np.cross(n, u)


def slice_points(verts, normal=(0.0, 1.0, 0.0), slice_pos=0.5, res=256, extent=1.2) -> np.ndarray:
    """(res * res, 3) float64 query points, point [j * res + i] at column i, row j of the plane grid."""
    v = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
    n, u, w = plane_basis(normal)
    d, a, b = v @ n, v @ u, v @ w
    off = d.min() + slice_pos * (d.max() - d.min())
    ac, bc = 0.5 * (a.min() + a.max()), 0.5 * (b.min() + b.max())
    larger = max(a.max() - a.min(), b.max() - b.min())
    side = extent * (larger if larger > 0 else 1.0)
    t = (np.arange(res) + 0.5) / res - 0.5
    au = ac + t * side
    bv = bc + t * side
    return (off * n[None, None, :] + au[None, :, None] * u[None, None, :] + bv[:, None, None] * w[None, None, :]) \
        .reshape(-1, 3)


def colormap(wn: np.ndarray) -> np.ndarray:
    """uint8 RGB of winding numbers: #2F5994 at 0, toward #F3E676 at 1 and #D2552E at -1, clamped to [-1, 1]."""
    c_out = np.array([47.0, 89.0, 148.0])
    c_in = np.array([243.0, 230.0, 118.0])
    c_neg = np.array([210.0, 85.0, 46.0])
    w = np.clip(np.asarray(wn, dtype=np.float64), -1.0, 1.0)[..., None]
    target = np.where(w >= 0, c_in, c_neg)
    rgb = c_out + (target - c_out) * np.abs(w)
    return np.clip(np.floor(rgb + 0.5), 0, 255).astype(np.uint8)


def write_png(path, rgb: np.ndarray) -> None:
    """8-bit RGB PNG (filter 0 rows, one zlib stream)."""
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
    h, w, _ = rgb.shape
    raw = b"".join(b"\x00" + rgb[r].tobytes() for r in range(h))

    def chunk(tag: bytes, data: bytes) -> bytes:
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)) + \
        chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(png)


def slice_gwn(verts, idx, normal=(0.0, 1.0, 0.0), slice_pos=0.5, res=256, extent=1.2, log=None) -> np.ndarray:
    """(res, res) full GWN on the slice grid (row j = grid row j, bottom to top), evaluated by HipMesh in FP64."""
    from .api import HipMesh, build_boundary_segments

    v = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
    i = np.asarray(idx, dtype=np.int32).reshape(-1)
    pts = slice_points(v, normal, slice_pos, res, extent)
    seg = build_boundary_segments(v, i)
    if log is not None:
        log(f"open boundary: {len(seg)} segments")
    with HipMesh(v, i, dtype=np.float64) as mesh:
        wn = mesh.full(pts, X0, seg if len(seg) else None)
    return np.asarray(wn).reshape(res, res)


def main(argv=None) -> int:
    from .obj import load_obj

    a = list(sys.argv[1:] if argv is None else argv)
    if len(a) < 3:
        print("usage: python -m antipodal_amd.slice <input.obj> <output.png> <resolution> [slice_pos=0.5] "
              "[nx ny nz = 0 1 0] [extent=1.2]", file=sys.stderr)
        return 1
    res = int(a[2])
    if res <= 0:
        print("error: resolution must be positive", file=sys.stderr)
        return 1
    slice_pos = float(a[3]) if len(a) > 3 else 0.5
    normal = (float(a[4]), float(a[5]), float(a[6])) if len(a) > 6 else (0.0, 1.0, 0.0)
    extent = float(a[7]) if len(a) > 7 else 1.2
    try:
        verts, idx = load_obj(a[0])
    except OSError:
        print(f"error: could not open '{a[0]}'", file=sys.stderr)
        return 1
    if len(verts) == 0 or len(idx) == 0:
        print(f"error: no triangles loaded from '{a[0]}'", file=sys.stderr)
        return 1
    print(f"loaded {len(verts)} vertices, {len(idx) // 3} triangles")
    wn = slice_gwn(verts, idx, normal, slice_pos, res, extent, log=print)
    write_png(a[1], colormap(wn[::-1]))
    print(f"wrote {res}x{res} slice to '{a[1]}'")
    return 0


if __name__ == "__main__":
    sys.exit(main())
