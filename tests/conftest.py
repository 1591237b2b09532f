"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on CPU."""
from __future__ import annotations

import importlib
import json
import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent
GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(REPO / "oracle"))
sys.path.insert(0, str(GOLDEN))
sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def O():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def W():
    return importlib.import_module("sidon-autocorrelation-bound_amd.workloads")


@pytest.fixture(scope="session")
def certs():
    out = {}
    for n in (512, 4096, 8400):
        c = json.loads((GOLDEN / f"cert-{n}.json").read_text())
        c["a"] = np.array(c["weights"], dtype=np.float64)
        out[n] = c
    return out


@pytest.fixture(scope="session")
def S():
    """The HIP package (GPU tests only)."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected but no GPU is visible")
    mod = importlib.import_module("sidon-autocorrelation-bound_amd")
    mod.init()
    return mod


def ulp_diff32(x, y):
    xi = np.asarray(x, np.float32).view(np.int32).astype(np.int64)
    yi = np.asarray(y, np.float32).view(np.int32).astype(np.int64)
    return np.abs(xi - yi)
