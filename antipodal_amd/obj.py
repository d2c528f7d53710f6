"""Wavefront OBJ ingest for the GPU path (the reader of examples/main.cc:52-94).

Only `v x y z` and `f ...` lines are read; every other line is ignored. A face token is `v`, `v/vt`, `v//vn` or
`v/vt/vn` -- the index before the first '/' is kept, 1-based, or relative to the vertices read so far when negative --
and a polygon is fan-triangulated from its first corner: (f0, f_{i-1}, f_i) for i >= 2. Not a general importer.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np


def parse_obj(text: str):
    """(vertices float64 (n, 3), triangle indices int32 (3 m,)) of OBJ text."""
    verts: list[tuple[float, float, float]] = []
    idx: list[int] = []
    for line in text.splitlines():
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "v":
            verts.append((float(parts[1]), float(parts[2]), float(parts[3])))
        elif parts[0] == "f":
            face = []
            for tok in parts[1:]:
                v = tok.split("/", 1)[0]
                if not v:
                    continue
                k = int(v)
                face.append(k - 1 if k > 0 else len(verts) + k)
            for i in range(2, len(face)):
                idx += [face[0], face[i - 1], face[i]]
    return np.asarray(verts, dtype=np.float64).reshape(-1, 3), np.asarray(idx, dtype=np.int32)


def load_obj(path):
    """parse_obj of a file; raises FileNotFoundError when it cannot be opened."""
    return parse_obj(Path(path).read_text())
