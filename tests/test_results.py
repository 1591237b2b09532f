"""The witnesses this build's search found (results/), re-checked on the CPU by
the oracle: the certificate rules of the reference's verify.py:81-100 (n
matches, no negative height, stored F = the fp64 hostF of the weights to
1e-12) and the long-double verifier of tools/probe.cpp --verify on the
saved-vector file, whose rigorous bound must not exceed the stated one."""
from __future__ import annotations

import json

import numpy as np
import pytest

from conftest import REPO

RESULTS = sorted((REPO / "results").glob("sidon-*.json"))


@pytest.mark.parametrize("path", RESULTS, ids=[p.stem for p in RESULTS])
def test_witness_verifies(O, path):
    cert = json.loads(path.read_text())
    a = np.array(cert["weights"], dtype=np.float64)
    assert a.size == cert["n"] and (a >= 0).all()
    F = O.F_hostorder(a)
    assert abs(F - cert["F"]) <= 1e-12 * F
    v = O.verify_file(str(path.with_suffix(".txt")))
    assert v["rc"] == 0 and v["n"] == cert["n"]
    bound = v["F"] * (1 + v["relBound"])                      # what --verify proves: C <= bound
    assert abs(v["F"] - F) <= 1e-13 * F
    assert bound < 1.5039722   # each is below the reference's n = 4096 certificate


def test_best_witness_beats_the_reference_certificate(O, certs):
    """The n = 8192 witness is below the reference's best certificate (n = 8400)."""
    best = json.loads((REPO / "results" / "sidon-8192.json").read_text())
    ref = certs[8400]
    assert O.F_hostorder(np.array(best["weights"])) < O.F_hostorder(ref["a"])
