"""Host-side logic of the surrogate's band calibration (no GPU): delta = factor x the worst wrong-side margin of the
validation points, plus a margin; points flagged as exact-path disagreements are left out."""
from __future__ import annotations

import numpy as np

from antipodal_amd.surrogate import calibrate_delta


def test_delta_from_worst_wrong_margin():
    pred = np.array([0.9, 0.1, 0.55, 0.3, 0.7], dtype=np.float32)
    exact = np.array([1.0, 0.0, 0.2, 0.9, 0.8])
    # wrong: 0.55 vs 0.2 (margin 0.05), 0.3 vs 0.9 (margin 0.2)
    assert np.isclose(calibrate_delta(pred, exact, 2.0, 0.02), 2 * 0.2 + 0.02)
    assert calibrate_delta(pred[:2], exact[:2], 2.0, 0.02) == 0.02  # nothing wrong


def test_delta_skips_flagged_points():
    pred = np.array([0.02, 0.55, 0.95], dtype=np.float32)
    exact = np.array([0.98, 0.2, 1.0])  # the first: an exact value off by one (grazing ray), margin 0.48
    skip = np.array([True, False, False])
    assert np.isclose(calibrate_delta(pred, exact, 2.0, 0.02), 2 * 0.48 + 0.02)
    assert np.isclose(calibrate_delta(pred, exact, 2.0, 0.02, skip), 2 * float(np.float32(0.55) - 0.5) + 0.02)
