"""The C++ host driver (sidon_search, the counterpart of the reference's
sidon.exe host loop): its device self-test, the policy search (three anneal
tiers, support signatures, tabu / operator weights, all fourteen operators,
n -> 2n splits) and the files it writes, checked with the oracle's verifier
(tools/probe.cpp --verify semantics) and hostF."""
from __future__ import annotations

import json
import subprocess

import numpy as np
import pytest

from conftest import REPO

pytestmark = pytest.mark.gpu

DRIVER = REPO / "sidon-autocorrelation-bound_amd" / "sidon_search"


def run(args, timeout=150):
    print("[driver]", " ".join(args), flush=True)
    out = subprocess.run([str(DRIVER)] + args, capture_output=True, text=True, timeout=timeout)
    assert out.returncode == 0, out.stderr[-2000:]
    return out.stdout


def test_selftest():
    out = run(["--selftest"])
    assert "FAIL" not in out and "OK" in out


STAT_KEYS = ["t", "n", "gen", "best", "lvl", "mean", "evals", "lsiters", "fw", "cells", "pflop",
             "peaks", "nnz"]   # the reference's progress record, main.cu:778-781


def stat_lines(out):
    recs = []
    for line in out.splitlines():
        if line.startswith("STAT "):
            kv = dict(f.split("=", 1) for f in line.split()[1:])
            assert list(kv) == STAT_KEYS
            recs.append(kv)
    return recs


def test_policy_search_improves_and_writes_verifiable_files(O, tmp_path):
    """The reference's command line (no mode flag): FW rounds on by default,
    --gens caps each level, the time budget is generous."""
    base = tmp_path / "best"
    out = run(["--pop", "64", "--islands", "4", "--elite", "2", "--ladder", "128,256",
               "--ls-iters", "60", "--cold-iters", "240", "--fw-iters", "200", "--gens", "12",
               "--seconds", "60", "--polish-iters", "1200", "--out", str(base), "--seed", "7"])
    assert out.startswith("GPU: ")
    recs = [json.loads(line) for line in out.splitlines() if line.startswith("{")]
    stats = [r["stat"] for r in recs if "stat" in r]
    levels = [r for r in recs if "level" in r]
    final = recs[-1]
    assert [r["level"] for r in levels] == [128, 256]
    assert all(r["generations"] == 12 for r in levels)
    # STAT records: six FW rounds a generation for every candidate, before the record
    st = stat_lines(out)
    assert len(st) == len(stats) and st
    for r in st:
        done = int(r["gen"]) + 1 + (12 if r["n"] == "256" else 0)
        assert int(r["fw"]) == 6 * 64 * done
    pf = [float(r["pflop"]) for r in st]
    assert pf == sorted(pf) and pf[0] > 0
    assert len(levels[0]["op_weights"]) == 14 and levels[0]["cells"] > 1
    assert stats[0]["cold"] > 0                       # fresh seeds take the cold tier
    polish = [r["polish"] for r in recs if "polish" in r]
    assert polish and polish[0]["best_F"] <= levels[-1]["best_F"] + 1e-12   # fp64 polish
    assert final["best_F"] < 1.56 and final["n"] in (128, 256)
    # the saved vector verifies, and its stored F is the hostF of its weights
    v = O.verify_file(str(base) + ".txt")
    assert v["rc"] == 0 and v["n"] == final["n"]
    a = np.array(json.loads((tmp_path / "best.json").read_text())["weights"])
    assert a.size == final["n"] and (a >= 0).all()
    assert O.F_hostorder(a) == pytest.approx(final["host_recheck"], rel=0, abs=1e-14)  # %.15f
    assert abs(final["best_F"] - final["host_recheck"]) <= 1e-12 * final["best_F"]


def test_graph_search_runs(tmp_path):
    out = run(["--search-graph", "--P", "64", "--islands", "4", "--levels", "128,256",
               "--iters", "50", "--gens", "3", "--seconds", "30", "--out", str(tmp_path / "g")])
    recs = [json.loads(line) for line in out.splitlines() if line.startswith("{")]
    assert [r["generations"] for r in recs if "level" in r] == [3, 3]
    final = recs[-1]
    assert final["best_F"] < 1.7 and (tmp_path / "g.json").exists()


def test_resume_adopts_the_vector_and_drops_coarser_levels(O, certs, tmp_path):
    a = certs[512]["a"]
    src = tmp_path / "seed.txt"
    O.write_vec(str(src), a, O.F_hostorder(a))
    base = tmp_path / "res"
    out = run(["--search", "--P", "64", "--islands", "4", "--levels", "256,512,1024",
               "--iters", "60", "--cold-iters", "240", "--fw-rounds", "0", "--gens", "3",
               "--seconds", "60", "--polish-iters", "600", "--quiet",
               "--resume", str(src), "--out", str(base)])
    assert "STAT" not in out and '"stat"' not in out    # --quiet
    recs = [json.loads(line) for line in out.splitlines() if line.startswith("{")]
    assert recs[0]["resumed"] == str(src) and recs[0]["levels_dropped"] == 1
    assert recs[0]["F"] == pytest.approx(O.F_hostorder(a), rel=0, abs=1e-14)
    assert [r["level"] for r in recs if "level" in r] == [512, 1024]
    final = recs[-1]
    assert final["host_recheck"] <= O.F_hostorder(a) + 1e-14   # never worse than it was handed
    v = O.verify_file(str(base) + ".txt")
    assert v["rc"] == 0 and v["n"] == final["n"]


def test_unreadable_resume_is_ignored(tmp_path):
    """As the reference (main.cu:424-426): a missing --resume file is reported
    and the search runs without it."""
    out = subprocess.run([str(DRIVER), "--pop", "32", "--islands", "4", "--ladder", "64,128",
                          "--ls-iters", "20", "--cold-iters", "60", "--fw-rounds", "0",
                          "--gens", "2", "--seconds", "30", "--polish-iters", "0",
                          "--resume", str(tmp_path / "missing.txt"), "--out", str(tmp_path / "r")],
                         capture_output=True, text=True, timeout=150)
    assert out.returncode == 0, out.stderr[-2000:]
    assert "could not read" in out.stderr and "ignoring" in out.stderr
    assert '"resumed"' not in out.stdout and (tmp_path / "r.json").exists()


def test_uneven_islands_graph():
    """Uneven islands in small: uneven lengths, generations replayed from a HIP graph."""
    out = run(["--uneven", "--P", "64", "--islands", "4", "--iters", "40", "--gens", "2",
               "--lengths", "128,256,600", "--no-tune"])
    rec = json.loads(out.strip().splitlines()[-1])
    assert rec["generations"] == 2 and rec["ms_per_generation"] > 0
    assert rec["candidate_steps_per_s"] > 0 and 1.0 < rec["best_F"] < 2.5


def test_ladder_splits_preserve_F_exactly():
    """The Frank-Wolfe ladder in small: Frank-Wolfe rounds per level, every n -> 2n split keeps
    the correctly rounded F bit for bit (and certifies the rounding)."""
    out = run(["--fw-ladder", "--pop", "32", "--ladder", "256,512,1024", "--fw-rounds", "1",
               "--fw-iters", "200"])
    recs = [json.loads(line) for line in out.splitlines() if line.startswith("{")]
    splits = [r for r in recs if "split" in r]
    levels = [r for r in recs if "level" in r]
    assert [r["level"] for r in levels] == [256, 512, 1024]
    assert len(splits) == 2
    for r in splits:
        assert r["F_bit_exact"] == r["of"] == 32 and r["ambiguous"] == 0
    for r in levels:
        assert r["mean_F_after"] <= r["mean_F_before"] + 1e-12
        assert r["executed_block_steps"] > 0 and 0 < r["executed_tflops"] < 157.3


def test_bench_mode():
    """--bench: the reference's micro-benchmark lengths (main.cu:296) plus 8400."""
    out = run(["--bench"])
    recs = [json.loads(line)["bench"] for line in out.splitlines() if line.startswith("{")]
    assert [r["n"] for r in recs] == [128, 256, 512, 1024, 2048, 4096, 8400]
    for r in recs:
        assert r["P"] == max(32, min(512, (4 << 20) // r["n"]))
        assert r["ls_ms"] > 0 and r["eval_f64_ms"] > 0
        assert 0 < r["tflops_executed"] < 157.3


def test_search_any_ladder_ratio_and_fresh_seeds(O, tmp_path):
    """A ladder that does not double (src/main.cu:491-501): each level carries
    the population by nearest-neighbour resampling, the last 20 % of the
    slots start from fresh host seeds, and the output verifies."""
    base = tmp_path / "lad"
    out = run(["--pop", "64", "--islands", "4", "--ladder", "600,1000,2100", "--ls-iters", "40",
               "--cold-iters", "120", "--fw-rounds", "1", "--fw-iters", "50", "--gens", "2",
               "--seconds", "60", "--polish-iters", "0", "--out", str(base), "--seed", "3"])
    recs = [json.loads(line) for line in out.splitlines() if line.startswith("{")]
    assert [r["level"] for r in recs if "level" in r] == [600, 1000, 2100]
    carries = [r for r in recs if "carry" in r]
    assert [c["carry"] for c in carries] == ["600 -> 1000", "1000 -> 2100"]
    assert all(c["form"] == "nearest-neighbour resample" and c["fresh"] == 12 for c in carries)
    stats = [r["stat"] for r in recs if "stat" in r]
    assert stats[0]["cold"] > 0
    final = recs[-1]
    v = O.verify_file(str(base) + ".txt")
    assert v["rc"] == 0 and v["n"] == final["n"]
    assert abs(final["best_F"] - final["host_recheck"]) <= 1e-12 * final["best_F"]


def test_search_past_the_on_chip_length(O, tmp_path):
    """A ladder into the lengths whose correlation output lives in global
    memory (6608 -> 13216 by the cell split): the search runs there with its
    moments' buffer as the launch's scratch, Frank-Wolfe is skipped where it
    does not fit (src/main.cu:633-639), and the output verifies."""
    base = tmp_path / "long"
    out = run(["--pop", "32", "--islands", "4", "--elite", "2", "--ladder", "6608,13216",
               "--ls-iters", "20", "--cold-iters", "40", "--fw-rounds", "1", "--fw-iters", "50",
               "--gens", "2", "--seconds", "60", "--polish-iters", "200", "--polish-top", "2",
               "--out", str(base), "--seed", "5"])
    recs = [json.loads(line) for line in out.splitlines() if line.startswith("{")]
    assert [r["level"] for r in recs if "level" in r] == [6608, 13216]
    final = recs[-1]
    assert final["n"] == 13216
    v = O.verify_file(str(base) + ".txt")
    assert v["rc"] == 0 and v["n"] == 13216
    assert abs(final["best_F"] - final["host_recheck"]) <= 1e-12 * final["best_F"]
