"""The FP64 arctangent polynomial of csrc/agwn_math.hpp (atan2_f64) against the Remez fit that produced it
(csrc/atan_fit.py): the header's coefficients are the fit's, rounded to double, and the polynomial evaluated as the
device does (double Horner with fused multiply-adds) stays within 1.7 ulp of atan over the reduced range |t| <= 0.2.
The FP32 polynomial of atan2_f32 stays within its stated 1.1e-7 on [0, 1]."""
from __future__ import annotations

import importlib.util
import re
from pathlib import Path

import numpy as np

CSRC = Path(__file__).resolve().parents[1] / "antipodal_amd" / "csrc"


def _fit():
    spec = importlib.util.spec_from_file_location("atan_fit", CSRC / "atan_fit.py")
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_fp64_coefficients_are_the_remez_fit():
    fit = _fit()
    text = (CSRC / "agwn_math.hpp").read_text()
    body = text[text.index("atan2_f64"):]
    got = [float(re.search(rf"\bc{k} = ([-0-9.e]+)", body).group(1)) for k in range(8)]
    c, E = fit.remez(7)
    want = [float(x) for x in c]
    assert got == want
    assert E < 2e-17  # levelled relative error, below a tenth of an ulp


def test_fp64_polynomial_error_bound():
    fit = _fit()
    c, _ = fit.remez(7)
    t = np.linspace(0.0, 0.2, 100001)
    v = fit.fma_horner_double([float(x) for x in c], t).astype(np.longdouble)
    ref = np.arctan(t.astype(np.longdouble))
    rel = np.abs(v[1:] - ref[1:]) / ref[1:]
    assert float(rel.max()) < 1.7 * 2.0**-53


def test_fp32_polynomial_error_bound():
    assert _fit().check_fp32() <= 1.1e-7
