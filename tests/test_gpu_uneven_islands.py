"""Uneven islands at full size: 512 candidates with n drawn uniformly from
{600, 1200, 2100, 4096, 8400}, 8 islands of 64, a generation = 200 fused
projected-Adam steps + log-normal mutation (sigma = 0.05) + binary tournaments
on fp64 F, replayed from the captured HIP graphs.

The same generations are run eagerly, one entry point at a time with a
read-back after each stage, and every stage is checked against the oracle
applied to that stage's GPU input (reference: the generation loop body of
src/main.cu:586-785):
  * local search: one candidate of every length class against the fp64
    oracle's 200-step trajectory (kernel semantics): F to 1.5e-3 (the fp32
    and fp64 trajectories part once a direction passes through zero; the
    per-height bound is the 10-step test's);
  * fp64 fitness: equal to hostF for all 512 candidates, bit for bit;
  * selection: the plan (parents, operators, strengths) and the migrated
    fitness equal the oracle's plan bit for bit, and migrated vectors and
    lengths moved;
  * variation: every offspring within 1 ulp of the oracle operator (exp
    enters operator 2), clones bit-exact, lengths inherited from parent 1.
The graph replays must equal the eager chain bit for bit.  Generation 12 is
the first with a ring migration (every 12 generations, main.cu:684-692), so
13 generations are run and generations 0 and 12 are checked stage by stage.
"""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

from conftest import ulp_diff32

pytestmark = pytest.mark.gpu

P, NI, ELITE, MIG, ITERS, SIGMA, SEED = 512, 8, 2, 12, 200, 0.05, 20260905
SCHED = dict(p_start=4.0, p_end=1e4, lr_start=0.085, lr_end=0.0015)
GENS = 13
CHECKED = (0, 12)


def dev(x, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device="cuda")


@pytest.fixture(scope="module")
def start(W):
    ns = W.uneven_lengths(P, seed=SEED)
    stride = int(ns.max())
    pop = np.zeros((P, stride), np.float32)
    for p, n in enumerate(ns):
        pop[p, :n] = W.structured_prior(1, int(n), seed=SEED + p)[0]
    return pop, ns


def test_uneven_lengths(start, W):
    _, ns = start
    assert set(ns.tolist()) == set(W.UNEVEN_NS) and ns.max() == 8400


class Eager:
    """The generation's entry points one at a time (the sequence the graph
    captures), with the population read back after every stage."""

    def __init__(self, S, pop, ns):
        self.S = S
        self.pop = [dev(pop), torch.zeros_like(dev(pop))]
        self.ns = [dev(ns, torch.int32), dev(ns, torch.int32)]
        self.nmax = int(ns.max())
        i32 = dict(dtype=torch.int32, device="cuda")
        self.order = torch.zeros(P, **i32)
        self.p1, self.p2, self.op = (torch.zeros(P, **i32) for _ in range(3))
        self.st = torch.zeros(P, dtype=torch.float32, device="cuda")
        self.gen = torch.zeros(1, dtype=torch.int64, device="cuda")
        self.ws = torch.empty(S.ragged_workspace_bytes(P, int(ns.max())), dtype=torch.uint8,
                              device="cuda")
        self.parity = 0

    def generation(self, record: bool):
        S, k = self.S, self.parity
        src, dst, ns_src, ns_dst = self.pop[k], self.pop[1 - k], self.ns[k], self.ns[1 - k]
        rec = {}
        if record:
            rec["ns"] = ns_src.cpu().numpy()
            rec["before_ls"] = src.cpu().numpy()
            rec["gen"] = int(self.gen.item())
        S.order_by_length(ns_src, self.order)
        S.local_search_ragged(src, ns_src, self.nmax, ITERS, order=self.order, workspace=self.ws,
                              **SCHED)
        if record:
            rec["pair_timeouts"] = S.pair_timeouts(self.ws)
        if record:
            rec["after_ls"] = src.cpu().numpy()
        fit, nact = S.eval_f64_ragged(src, ns_src, self.nmax)
        if record:
            rec["fit"] = fit.cpu().numpy()
        S.select(fit, src, ns_src, NI, ELITE, MIG, SIGMA, SEED, self.gen, self.p1, self.p2,
                 self.op, self.st)
        if record:
            rec.update(fit_sel=fit.cpu().numpy(), after_sel=src.cpu().numpy(),
                       ns_sel=ns_src.cpu().numpy(), p1=self.p1.cpu().numpy(),
                       p2=self.p2.cpu().numpy(), op=self.op.cpu().numpy(),
                       st=self.st.cpu().numpy())
        S.variation_ragged(src, ns_src, dst, ns_dst, self.p1, self.p2, self.op, self.st, SEED,
                           self.gen)
        S.advance_gen(self.gen, 1)
        if record:
            rec["child"] = dst.cpu().numpy()
            rec["ns_child"] = ns_dst.cpu().numpy()
        self.fit = fit
        self.parity ^= 1
        return rec

    def population(self):
        return self.pop[self.parity], self.ns[self.parity]


@pytest.fixture(scope="module")
def runs(S, start):
    pop, ns = start
    print("[uneven] eager generations, stage by stage", flush=True)
    eager = Eager(S, pop, ns)
    recs = {}
    for g in range(GENS):
        recs[g] = eager.generation(record=g in CHECKED)
        torch.cuda.synchronize()
        print(f"   generation {g} done", flush=True)
    print("[uneven] graph replays", flush=True)
    graph = S.Generation(dev(pop), dev(ns, torch.int32), islands=NI, iters=ITERS, sigma=SIGMA,
                         elite=ELITE, migrate_every=MIG, seed=SEED, graph=True, **SCHED)
    for g in range(GENS):
        graph.run(1)
        torch.cuda.synchronize()
        _ = float(graph.pop_a.sum().item())      # eager work between replays
    return eager, graph, recs


def test_uneven_graph_equals_eager(runs):
    """(Generation 0 splits its longest candidates over two workgroups when
    the makespan model says so; no pair wait gives up.)"""
    eager, graph, recs = runs
    assert all(r["pair_timeouts"] == 0 for r in recs.values() if r)
    gp, gn = graph.population()
    ep, en = eager.population()
    assert graph.gen.item() == eager.gen.item() == GENS
    assert torch.equal(gn, en)
    assert torch.equal(gp, ep)
    assert torch.equal(graph.fit, eager.fit)


def _classes(ns):
    """One slot of every length class (the first in slot order)."""
    return [int(np.flatnonzero(ns == n)[0]) for n in sorted(set(ns.tolist()))]


def test_uneven_local_search_vs_oracle(runs, O):
    """200 steps of the anneal, fp32 on the GPU against the fp64 oracle from
    the same start, one candidate of every length class in generation 0."""
    _, _, recs = runs
    r = recs[0]
    slots = _classes(r["ns"])
    starts = [r["before_ls"][s, :r["ns"][s]].astype(np.float64) for s in slots]
    with ThreadPoolExecutor(len(slots)) as ex:
        ref = list(ex.map(lambda a: O.local_search_gmin(
            a, ITERS, SCHED["p_start"], SCHED["p_end"], SCHED["lr_start"], SCHED["lr_end"]),
            starts))
    worst = 0.0
    for s, (want, Fw, _, _, gmin) in zip(slots, ref):
        n = int(r["ns"][s])
        got = r["after_ls"][s, :n].astype(np.float64)
        Fg = O.F_hostorder(got)
        F0 = O.F_hostorder(starts[slots.index(s)])
        rel = abs(Fg - Fw) / Fw
        worst = max(worst, rel)
        print(f"   n={n}: F start {F0:.9f}, gpu {Fg:.9f}, oracle {Fw:.9f} (rel {rel:.2e})")
        assert Fg < F0 and Fw < F0
        assert (got >= 0).all() and abs(got.sum() - 1) < 1e-5
    # 200 fp32 steps of the anneal to p = 1e4 are not a per-height match for
    # fp64 (a direction that passes through zero takes either Adam sign, and
    # the trajectories part after that); their F must agree to UNEVEN_F_BOUND
    print(f"   worst F difference {worst:.2e}")
    assert worst < UNEVEN_F_BOUND


# 200 fp32 steps against fp64: about 3x the worst F difference measured on
# MI355X (4.5e-4, n = 8400)
UNEVEN_F_BOUND = 1.5e-3


@pytest.mark.parametrize("g", CHECKED)
def test_uneven_fitness_is_hostF(runs, O, g):
    """fp64 F of all 512 local-search results equals hostF bit for bit."""
    _, _, recs = runs
    r = recs[g]
    rows = [r["after_ls"][p, :r["ns"][p]].astype(np.float64) for p in range(P)]
    with ThreadPoolExecutor(8) as ex:
        want = np.array(list(ex.map(O.F_hostorder, rows)))
    np.testing.assert_array_equal(r["fit"], want)


@pytest.mark.parametrize("g", CHECKED)
def test_uneven_selection_plan(runs, O, g):
    _, _, recs = runs
    r = recs[g]
    want = O.select_plan(r["fit"], NI, ELITE, SEED, r["gen"], migrate_every=MIG, sigma=SIGMA)
    for name, key in (("p1", "p1"), ("p2", "p2"), ("op", "op"), ("strength", "st"),
                      ("fit", "fit_sel")):
        np.testing.assert_array_equal(r[key], want[name], err_msg=name)
    migrated = [(s, d) for s, d in zip(want["mig_src"], want["mig_dst"]) if s >= 0]
    assert bool(migrated) == (r["gen"] > 0 and r["gen"] % MIG == 0)
    for s, d in migrated:
        assert r["ns_sel"][d] == r["ns"][s]
        n = int(r["ns"][s])
        np.testing.assert_array_equal(r["after_sel"][d, :n], r["after_ls"][s, :n])
    moved = {d for _, d in migrated}
    for p in range(P):
        if p not in moved:
            np.testing.assert_array_equal(r["after_sel"][p], r["after_ls"][p])


@pytest.mark.parametrize("g", CHECKED)
def test_uneven_offspring(runs, O, W, g):
    _, _, recs = runs
    r = recs[g]
    src, ns = r["after_sel"], r["ns_sel"]
    for j in range(P):
        a = int(r["p1"][j])
        n = int(ns[a])
        assert r["ns_child"][j] == n, j
        child = r["child"][j, :n]
        assert (child >= 0).all() and abs(child.astype(np.float64).sum() - 1) < 1e-5
        if r["op"][j] == 20:        # elite clone: the parent, renormalised
            want = O.variation2(src[a, :n], None, 20, 0.0, SEED, r["gen"], j)
            np.testing.assert_array_equal(child, want, err_msg=f"clone {j}")
            continue
        want = O.variation2(src[a, :n], None, int(r["op"][j]), float(r["st"][j]), SEED,
                            r["gen"], j)
        d = ulp_diff32(child, want)
        assert d.max() <= 1 and (d > 0).mean() < 1e-3, (j, int(d.max()))
    assert set(r["ns_child"].tolist()) <= set(W.UNEVEN_NS)
    # each island clones its two elites, ranked before the migration (the
    # reference ranks once, src/main.cu:677-706)
    fit = r["fit"]
    for i in range(NI):
        members = range(i * (P // NI), (i + 1) * (P // NI))
        best = sorted(members, key=lambda p: (fit[p], p))[:ELITE]
        assert sorted(r["p1"][i * (P // NI):i * (P // NI) + ELITE].tolist()) == sorted(best)
