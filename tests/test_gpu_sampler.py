"""The surrogate's training sampler on the device (agwn_training_samples_dev) against its numpy restatement: the same
counter-based SplitMix64 streams (meshgen.uniform01(seed, c, n), c = 0..5), the same float64 formulas, rounded to
float32. Bars: at least 99.9 % of the coordinates bit-identical, the rest within 2 float32 ulp of the coordinate scale
(sqrt and pow may round differently on the device); the near/uniform split, the offsets' range and the box follow the
definition."""
from __future__ import annotations

import numpy as np
import pytest

from antipodal_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def restated(verts, idx, lo, hi, n, near_frac, band, seed, power):
    v = np.asarray(verts, dtype=np.float32).astype(np.float64).reshape(-1, 3)
    tri = np.asarray(idx, dtype=np.int64).reshape(-1, 3)
    a, b, c = v[tri[:, 0]], v[tri[:, 1]], v[tri[:, 2]]
    e1, e2 = b - a, c - a
    nr = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1], e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                   e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], axis=1)
    cdf = np.cumsum(np.linalg.norm(np.cross(e1, e2), axis=1))
    u = [mg.uniform01(seed, k, n) for k in range(6)]
    near = u[0] < near_frac
    k = np.minimum(np.searchsorted(cdf, u[1] * cdf[-1], side="right"), len(tri) - 1)
    ln = np.sqrt((nr[:, 0] * nr[:, 0] + nr[:, 1] * nr[:, 1]) + nr[:, 2] * nr[:, 2])
    nrm = np.where(ln[:, None] > 0, nr / np.where(ln > 0, ln, 1.0)[:, None], nr)
    r1, r2 = np.sqrt(u[2]), u[3]
    w0, w1, w2 = 1.0 - r1, r1 * (1.0 - r2), r1 * r2
    off = (band * np.where(u[4] < 0.5, -1.0, 1.0)) * u[5] ** power
    pn = ((w0[:, None] * a[k] + w1[:, None] * b[k]) + w2[:, None] * c[k]) + off[:, None] * nrm[k]
    lo64, hi64 = np.asarray(lo, np.float32).astype(np.float64), np.asarray(hi, np.float32).astype(np.float64)
    pu = lo64 + (hi64 - lo64) * np.stack([u[1], u[2], u[3]], axis=1)
    return np.where(near[:, None], pn, pu).astype(np.float32), near, off


@pytest.mark.parametrize("power", [1.0, 2.0])
def test_sampler_matches_restatement(gpu, power):
    v, idx = mg.icosphere_cap_mesh()
    lo, hi = mg.expanded_bbox(v, 0.10)
    n, near_frac, band, seed = 100_000, 0.7, 0.05, 11
    smp = gpu.TrainingSampler(v.astype(np.float32), idx, lo, hi)
    got = smp(n, near_frac=near_frac, band=band, seed=seed, power=power).cpu().numpy()
    want, near, off = restated(v, idx, lo.astype(np.float32), hi.astype(np.float32), n, near_frac, band, seed, power)
    eq = np.mean(got == want)
    scale = float(np.max(np.abs(want)))
    d = float(np.max(np.abs(got - want)))
    print(f"power {power}: equal {eq:.6f}, max |d| {d:.3e}, near {near.mean():.4f}")
    assert eq >= 0.999 and d <= 2 * np.finfo(np.float32).eps * scale
    assert abs(near.mean() - near_frac) < 0.01
    assert np.all(np.abs(off) <= band)
    u = ~near
    assert np.all(got[u] >= lo.astype(np.float32)) and np.all(got[u] <= hi.astype(np.float32))
    # another seed gives other points; the same seed the same ones
    again = smp(n, near_frac=near_frac, band=band, seed=seed, power=power).cpu().numpy()
    other = smp(n, near_frac=near_frac, band=band, seed=seed + 1, power=power).cpu().numpy()
    assert np.array_equal(again, got) and np.mean(other == got) < 0.01


def test_sampler_rejects_bad_parameters(gpu):
    v, idx = mg.icosphere_cap_mesh()
    lo, hi = mg.expanded_bbox(v, 0.10)
    smp = gpu.TrainingSampler(v.astype(np.float32), idx, lo, hi)
    with pytest.raises(gpu.AgwnError, match="sampling parameters"):
        smp(16, near_frac=1.5)
    with pytest.raises(gpu.AgwnError, match="sampling parameters"):
        smp(16, power=0.0)
    with pytest.raises(gpu.AgwnError, match="box"):
        gpu.TrainingSampler(v.astype(np.float32), idx, hi, lo)(16)
