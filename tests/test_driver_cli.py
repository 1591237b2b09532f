"""The driver's command line against the reference's (src/main.cu:41-82,
:357-381): every reference flag is accepted, unknown options exit 1 with a
message, bad values are refused before any device work.  Runs without a GPU
(argument handling happens before the device is touched)."""
from __future__ import annotations

import subprocess

import pytest

from conftest import REPO

DRIVER = REPO / "sidon-autocorrelation-bound_amd" / "sidon_search"
REFERENCE_FLAGS = ["--ladder", "--pop", "--islands", "--elite", "--ls-iters", "--cold-iters",
                   "--pmax", "--fw-rounds", "--fw-iters", "--gens", "--seconds", "--seed", "--out",
                   "--resume", "--polish-iters", "--polish-top", "--selftest", "--bench",
                   "--quiet", "-h"]

pytestmark = pytest.mark.skipif(not DRIVER.exists(), reason="driver not built (make -C "
                                "sidon-autocorrelation-bound_amd)")


def run(*args):
    return subprocess.run([str(DRIVER), *args], capture_output=True, text=True, timeout=60)


def test_help_lists_every_reference_flag():
    for flag in ("-h", "--help"):
        r = run(flag)
        assert r.returncode == 0
        for f in REFERENCE_FLAGS:
            assert f in r.stdout, f


@pytest.mark.parametrize("args,msg", [
    (["--bogus"], "unknown option --bogus"),
    (["--ladder", "--pop", "32"], "bad list for --ladder"),     # a list is required
    (["--ladder", "64,0"], "bad list for --ladder"),
    (["--search-graph", "--ladder", "64,100"], "twice the one before"),
    (["--levels", "x"], "bad list for --ladder"),               # alias spelling
    (["--pop", "100", "--islands", "8"], "pop must be divisible by islands"),
    (["--gens", "-1"], "must not be negative"),
    (["--seconds", "-5"], "must not be negative"),
    (["--seed"], "needs a value"),
    (["--fw-ladder", "--ladder", "1050,2000"], "twice the one before"),
])
def test_bad_arguments_exit_1(args, msg):
    r = run(*args)
    assert r.returncode == 1
    assert msg in r.stderr
