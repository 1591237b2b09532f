"""The island search policy (src/main.cu:461-472, :598-609, :650-769) behind
sidon_policy_*: elitism, tournaments inside islands, ring migration, the
tabu override on over-visited support cells, stagnation restarts, credit
assignment of the operator weights, best re-seeding and anneal tiers.  Host
code only; runs without a GPU."""
from __future__ import annotations

import importlib

import numpy as np
import pytest

S = importlib.import_module("sidon-autocorrelation-bound_amd")
DIVERS = {5, 7, 8, 9, 10, 11, 12}


def island_ranks(fit, NI):
    PI = fit.size // NI
    return [sorted(range(i * PI, (i + 1) * PI), key=lambda p: (fit[p], p)) for i in range(NI)]


def test_elites_tournaments_and_strengths():
    P, NI, elite = 64, 4, 2
    pol = S.Policy(P, NI, elite, 7)
    rng = np.random.default_rng(1)
    fit = rng.uniform(1.5, 1.6, size=P)
    fit[3] = fit[5]                                    # a tie, ordered by slot
    sig = rng.integers(0, 2**63, size=P, dtype=np.uint64)
    pol.observe(fit, sig)
    pl = pol.plan(5)
    ranks = island_ranks(fit, NI)
    PI = P // NI
    for i in range(NI):
        for r in range(PI):
            slot = i * PI + r
            if r < elite:
                assert pl["p1"][slot] == pl["p2"][slot] == ranks[i][r]
                assert pl["op"][slot] == 20 and pl["strength"][slot] == 0
            else:
                assert i * PI <= pl["p1"][slot] < (i + 1) * PI   # parents from the island
                assert i * PI <= pl["p2"][slot] < (i + 1) * PI
                assert 0 <= pl["op"][slot] < 14
                warm = min(1.0, 5 / 40)
                s = pl["strength"][slot] / (0.5 + warm)
                assert 0.03 - 1e-6 <= s <= 0.38 + 1e-6
    assert pl["mig_src"].size == 0 and pl["reseed"] == -1


def test_tournaments_prefer_the_fitter():
    P, NI = 256, 1
    pol = S.Policy(P, NI, 0, 3)
    fit = np.linspace(1.5, 2.5, P)
    pol.observe(fit, np.arange(P, dtype=np.uint64))
    pl = pol.plan(1)
    # the better of two uniform draws: the mean rank is about P/3
    assert pl["p1"].mean() < 0.4 * P and pl["p2"].mean() < 0.4 * P


def test_ring_migration_every_12_generations():
    P, NI = 32, 4
    pol = S.Policy(P, NI, 1, 5)
    rng = np.random.default_rng(2)
    fit = rng.uniform(1.5, 1.6, size=P)
    pol.observe(fit, np.arange(P, dtype=np.uint64))
    ranks = island_ranks(fit, NI)
    assert pol.plan(0)["mig_src"].size == 0
    assert pol.plan(11)["mig_src"].size == 0
    pl = pol.plan(12)
    assert pl["mig_src"].tolist() == [ranks[i][0] for i in range(NI)]
    assert pl["mig_dst"].tolist() == [ranks[(i + 1) % NI][-1] for i in range(NI)]


def test_tabu_forces_diversifying_operators():
    P = 64
    pol = S.Policy(P, 1, 0, 11)
    fit = np.full(P, 2.0)
    fit[-1] = 1.5                                       # the incumbent, its own cell
    sig = np.full(P, 77, np.uint64)
    sig[-1] = 99
    for _ in range(2):                                  # cell 77 revisited, never improved
        pol.observe(fit, sig)
    pl = pol.plan(3)
    crowded = sig[pl["p1"]] == 77
    assert crowded.sum() > P // 2
    assert set(pl["op"][crowded].tolist()) <= DIVERS
    assert pol.stats()["cells"] == 2


def test_competitive_cells_are_not_tabu():
    P = 64
    pol = S.Policy(P, 1, 0, 11)
    fit = np.full(P, 1.5 * (1 + 1e-4))                  # within 2e-4 of the incumbent
    fit[-1] = 1.5
    sig = np.full(P, 77, np.uint64)
    sig[-1] = 99
    for _ in range(3):
        pol.observe(fit, sig)
    ops = np.concatenate([pol.plan(g)["op"] for g in range(1, 9)])
    # weights start at 1 (2.5, 1.8, 1.8 for 11, 12, 13): ops outside DIVERS keep appearing
    assert np.isin(ops, [0, 1, 2, 3, 4, 6, 13]).mean() > 0.3


def test_stagnation_brings_restarts():
    P = 512
    pol = S.Policy(P, 1, 0, 13)
    fit = np.random.default_rng(4).uniform(1.5, 1.6, size=P)
    sig = np.arange(P, dtype=np.uint64)
    for _ in range(20):                                 # no improvement after the first
        pol.observe(fit, sig)
    assert pol.stats()["since_improved"] == 19
    frac9 = (pol.plan(30)["op"] == 9).mean()
    fresh = S.Policy(P, 1, 0, 13)
    fresh.observe(fit, sig)
    frac9_fresh = (fresh.plan(30)["op"] == 9).mean()
    assert frac9 > frac9_fresh + 0.05                   # ~10 % extra restarts


def test_credit_moves_weights_by_the_rule():
    P = 1024
    pol = S.Policy(P, 1, 0, 17)
    fit = np.random.default_rng(5).uniform(1.5, 1.6, size=P)
    pol.observe(fit, np.arange(P, dtype=np.uint64))
    w0 = pol.stats()["weights"]
    pl = pol.plan(5)
    used = np.bincount(pl["op"], minlength=14)
    pol.credit(np.full(P, -1.0))                        # every child beats its parent
    w1 = pol.stats()["weights"]
    for o in range(14):
        rate = 1.0 if used[o] > 20 else 0.25
        want = min(4.0, max(0.05, 0.85 * w0[o] + 0.15 * (0.15 + 3.0 * rate)))
        assert w1[o] == pytest.approx(want, rel=1e-12)
    fresh = S.Policy(P, 1, 0, 17)
    fresh.observe(fit, np.arange(P, dtype=np.uint64))
    fresh.plan(5)
    fresh.credit(np.full(P, 9.0))                       # no child wins: rate 0 where used
    w3 = fresh.stats()["weights"]
    for o in range(14):
        rate = 0.0 if used[o] > 20 else 0.25
        assert w3[o] == pytest.approx(0.85 * w0[o] + 0.15 * (0.15 + 3.0 * rate), rel=1e-12)


def test_reseed_slot_and_best_tracking():
    P, NI = 32, 4
    pol = S.Policy(P, NI, 1, 19)
    fit = np.random.default_rng(6).uniform(1.5, 1.6, size=P)
    bi, imp = pol.observe(fit, np.arange(P, dtype=np.uint64))
    assert bi == int(np.argmin(fit)) and imp
    assert pol.plan(9)["reseed"] == island_ranks(fit, NI)[0][-1]
    assert pol.plan(8)["reseed"] == -1
    _, imp = pol.observe(fit, np.arange(P, dtype=np.uint64))
    assert not imp
    f2 = fit.copy()
    f2[7] = fit.min() - 1e-9
    bi, imp = pol.observe(f2, np.arange(P, dtype=np.uint64))
    assert bi == 7 and imp and pol.stats()["best"] == f2[7]
    pol.new_level()
    assert pol.stats()["since_improved"] == 0 and pol.stats()["level_best"] == 1e30


def test_plans_are_reproducible_by_seed():
    P = 64
    fit = np.random.default_rng(7).uniform(1.5, 1.6, size=P)
    sig = np.arange(P, dtype=np.uint64)
    plans = []
    for seed in (21, 21, 22):
        pol = S.Policy(P, 4, 2, seed)
        pol.observe(fit, sig)
        plans.append(pol.plan(3))
    for k in ("p1", "p2", "op", "strength"):
        np.testing.assert_array_equal(plans[0][k], plans[1][k])
    assert any(not np.array_equal(plans[0][k], plans[2][k]) for k in ("p1", "op"))


def test_anneal_tiers():
    P = 100
    cur = np.array([9, 11] * 20 + [5, 7, 8, 10, 12, 13] * 5 + [2, 20, 0, 1, 3, 4, 6] * 4
                   + [2, 20], np.int32)
    assert cur.size == P
    tier, nc, nm = S.Policy.tiers(cur)
    assert nc == 18                                     # 18 % of P, first come
    assert (tier[:40] == 0).sum() == 18 and (tier[:40] == 1).sum() == 22
    assert (tier[40:70] == 1).all() and (tier[70:] == 2).all()
    assert nm == 22 + 30


def test_bad_arguments():
    with pytest.raises(S.SidonError):
        S.Policy(10, 3, 1, 0)                          # P not divisible by islands
    with pytest.raises(S.SidonError):
        S.Policy(8, 2, 5, 0)                           # more elites than an island holds
    pol = S.Policy(8, 2, 1, 0)
    with pytest.raises(S.SidonError):
        pol.plan(0)                                    # nothing observed yet
