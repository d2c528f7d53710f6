"""Shared test fixtures. GPU tests are marked `gpu`; everything else runs on the CPU."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def load_golden(name: str) -> dict:
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle

    o = Oracle()
    if not o.available():
        import oracle as pkg

        pkg.build(ref=False)
    return o


@pytest.fixture(scope="session")
def reference():
    """The reference's own code (oracle/_ref/libref.so) where it was built; tests using it skip otherwise."""
    from oracle import Reference

    r = Reference()
    if not r.available():
        pytest.skip("oracle/_ref/libref.so not built (reference sources absent)")
    return r


@pytest.fixture(scope="session")
def agwn():
    """The product (libagwn.so) -- import fails loudly when it is not built."""
    import antipodal_amd

    return antipodal_amd


@pytest.fixture(scope="session")
def gpu(agwn):
    if not agwn.device_available():
        raise RuntimeError(f"GPU test without a usable HIP device: {agwn.last_error()}")
    return agwn
