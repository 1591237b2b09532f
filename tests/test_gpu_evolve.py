"""Mutation, selection, migration and the device generation counter vs the
oracle (Philox streams are bit-identical; fp64 exp/log may differ by an ulp)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import ulp_diff32

pytestmark = pytest.mark.gpu


def dev(x, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device="cuda")


def test_variation_matches_oracle(S, O, W):
    print("[evolve] variation", flush=True)
    P, n = 16, 1200
    pop = W.structured_prior(P, n, seed=4)
    out = torch.zeros(P, n, dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(1)
    p1 = rng.integers(0, P, size=P).astype(np.int32)
    op = np.where(rng.uniform(size=P) < 0.7, 2, 20).astype(np.int32)
    st = np.full(P, 0.05, np.float32)
    S.variation(dev(pop), out, dev(p1, torch.int32), dev(p1, torch.int32), dev(op, torch.int32),
                dev(st), 20260905, 7)
    got = out.cpu().numpy()
    for j in range(P):
        want = O.variation(pop[p1[j]], int(op[j]), 0.05, 20260905, 7, j)
        d = ulp_diff32(got[j], want)
        assert d.max() <= 1 and (d > 0).mean() < 1e-3


# operators with a library exp / log / cos in them may differ by an ulp
TRANSCENDENTAL = {2, 3, 4, 9, 11}


@pytest.mark.parametrize("n", [1200, 8400])
def test_all_operators_match_oracle(S, O, W, n):
    """Each of the fourteen operators (and clone) on two parents, GPU vs the
    oracle: bit for bit where only + - * / and comparisons are involved."""
    print(f"[evolve] operators n={n}", flush=True)
    ops = list(range(14)) + [20]
    reps = 3
    P = len(ops) * reps
    pop = W.structured_prior(P, n, seed=8).astype(np.float32)
    rng = np.random.default_rng(3)
    p1 = rng.integers(0, P, size=P).astype(np.int32)
    p2 = rng.integers(0, P, size=P).astype(np.int32)
    op = np.repeat(np.array(ops, np.int32), reps)
    st = rng.uniform(0.03, 0.38, size=P).astype(np.float32)
    out = torch.zeros(P, n, dtype=torch.float32, device="cuda")
    S.variation(dev(pop), out, dev(p1, torch.int32), dev(p2, torch.int32), dev(op, torch.int32),
                dev(st), 20260905, 11)
    got = out.cpu().numpy()
    for j in range(P):
        want = O.variation2(pop[p1[j]], pop[p2[j]], int(op[j]), float(st[j]), 20260905, 11, j)
        if op[j] in TRANSCENDENTAL:
            d = ulp_diff32(got[j], want)
            assert d.max() <= 1 and (d > 0).mean() < 1e-3, (op[j], j)
        else:
            np.testing.assert_array_equal(got[j], want, err_msg=f"op {op[j]} slot {j}")


def test_variation_ragged_crossover_lengths(S, O):
    """Uneven lengths: the offspring keeps p1's length; a second parent of
    another length is replaced by the first."""
    print("[evolve] ragged crossover", flush=True)
    rng = np.random.default_rng(9)
    ns = np.array([600, 600, 900, 600], np.int32)
    stride = 1000
    pop = np.zeros((4, stride), np.float32)
    for p, n in enumerate(ns):
        pop[p, :n] = rng.uniform(0.5, 1.5, size=n)
    p1 = np.array([0, 0, 2, 3, 1, 2], np.int32)
    p2 = np.array([1, 2, 0, 1, 3, 2], np.int32)
    op = np.array([0, 0, 1, 8, 8, 0], np.int32)
    st = np.full(6, 0.1, np.float32)
    out = torch.zeros(6, stride, dtype=torch.float32, device="cuda")
    ns_out = torch.zeros(6, dtype=torch.int32, device="cuda")
    gd = dev(np.array([5], np.uint64).view(np.int64), torch.int64)
    S.variation_ragged(dev(pop), dev(ns, torch.int32), out, ns_out, dev(p1, torch.int32),
                       dev(p2, torch.int32), dev(op, torch.int32), dev(st), 20260905, gd)
    got, gl = out.cpu().numpy(), ns_out.cpu().numpy()
    for j in range(6):
        n = ns[p1[j]]
        assert gl[j] == n
        B = pop[p2[j], :n] if ns[p2[j]] == n else None
        want = O.variation2(pop[p1[j], :n], B, int(op[j]), 0.1, 20260905, 5, j)
        np.testing.assert_array_equal(got[j, :n], want, err_msg=f"slot {j}")


@pytest.mark.parametrize("n", [1, 17, 256, 257, 1200, 8400])
def test_support_sig_matches_oracle(S, O, W, n):
    """kSupportSig's signature and support size, bit for bit (the fp32 mean in
    the reference kernel's summation order decides the support)."""
    print(f"[evolve] support signature n={n}", flush=True)
    rng = np.random.default_rng(n)
    rows = [rng.uniform(size=n), W.structured_prior(1, n, seed=n)[0],
            np.where(rng.uniform(size=n) < 0.4, 0.0, rng.lognormal(0, 2, size=n)),
            np.zeros(n), np.ones(n)]
    pop = np.stack(rows).astype(np.float32)
    for tol in (0.02, 0.5):
        sig, nnz = S.support_sig(dev(pop), tol)
        sig = sig.cpu().numpy().view(np.uint64)
        nnz = nnz.cpu().numpy()
        for p in range(pop.shape[0]):
            ws, wn = O.support_sig(pop[p], tol)
            assert int(sig[p]) == ws and int(nnz[p]) == wn, (p, tol)


def test_select_matches_oracle_and_migrates(S, O, W):
    print("[evolve] select", flush=True)
    P, NI, n = 512, 8, 64
    rng = np.random.default_rng(2)
    fit = rng.uniform(1.5, 1.6, size=P)
    fit[5] = fit[9]                      # a tie, broken by index
    pop = rng.uniform(size=(P, n)).astype(np.float32)
    for gen in (11, 12):
        fd, pd = dev(fit, torch.float64), dev(pop)
        gd = dev(np.array([gen], np.uint64).view(np.int64), torch.int64)
        outs = [torch.zeros(P, dtype=torch.int32, device="cuda") for _ in range(3)]
        stg = torch.zeros(P, dtype=torch.float32, device="cuda")
        S.select(fd, pd, None, NI, 2, 12, 0.05, 20260905, gd, outs[0], outs[1], outs[2], stg)
        want = O.select_plan(fit, NI, 2, 20260905, gen, migrate_every=12, sigma=0.05)
        for k, name in enumerate(("p1", "p2", "op")):
            np.testing.assert_array_equal(outs[k].cpu().numpy(), want[name])
        np.testing.assert_array_equal(stg.cpu().numpy(), want["strength"])
        np.testing.assert_array_equal(fd.cpu().numpy(), want["fit"])
        got_pop = pd.cpu().numpy()
        for i in range(NI):
            if want["mig_src"][i] >= 0:
                np.testing.assert_array_equal(got_pop[want["mig_dst"][i]], pop[want["mig_src"][i]])


def test_generation_counter_and_graph_replay(S, W):
    """A generation (select -> variation -> advance) captured as a HIP graph and
    replayed several times matches the same sequence run eagerly."""
    print("[evolve] graph replay", flush=True)
    P, NI, n = 64, 4, 600
    base = W.structured_prior(P, n, seed=6)

    def state():
        return dict(pop=dev(base), nxt=torch.zeros(P, n, dtype=torch.float32, device="cuda"),
                    fit=dev(np.linspace(1.6, 1.5, P), torch.float64),
                    gen=torch.zeros(1, dtype=torch.int64, device="cuda"),
                    p1=torch.zeros(P, dtype=torch.int32, device="cuda"),
                    p2=torch.zeros(P, dtype=torch.int32, device="cuda"),
                    op=torch.zeros(P, dtype=torch.int32, device="cuda"),
                    st=torch.zeros(P, dtype=torch.float32, device="cuda"),
                    ns=dev(np.full(P, n, np.int32), torch.int32),
                    ns2=dev(np.full(P, n, np.int32), torch.int32))

    def generation(s, src, dst, ns_src, ns_dst):
        S.select(s["fit"], s[src], s[ns_src], NI, 2, 3, 0.05, 99, s["gen"], s["p1"], s["p2"],
                 s["op"], s["st"])
        S.variation_ragged(s[src], s[ns_src], s[dst], s[ns_dst], s["p1"], s["p2"], s["op"],
                           s["st"], 99, s["gen"])
        S.advance_gen(s["gen"], 1)

    eager = state()
    for g in range(6):
        generation(eager, *(("pop", "nxt", "ns", "ns2") if g % 2 == 0 else
                            ("nxt", "pop", "ns2", "ns")))
    torch.cuda.synchronize()

    cap = state()
    stream = torch.cuda.Stream()
    graphs = []
    with torch.cuda.stream(stream):
        for parity in range(2):
            gph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(gph, stream=stream):
                generation(cap, *(("pop", "nxt", "ns", "ns2") if parity == 0 else
                                  ("nxt", "pop", "ns2", "ns")))
            graphs.append(gph)
    torch.cuda.synchronize()
    # capturing does not execute; reset the state the captures reference
    cap["gen"].zero_()
    for g in range(6):
        graphs[g % 2].replay()
        torch.cuda.synchronize()
        _ = cap["pop"].sum().item()       # eager work between replays
    assert cap["gen"].item() == 6 == eager["gen"].item()
    assert torch.equal(cap["pop"], eager["pop"]) and torch.equal(cap["nxt"], eager["nxt"])


def test_generation_graph_matches_eager(S, W):
    """Generations replayed from the captured HIP graphs equal the same
    sequence enqueued eagerly, bit for bit, with uneven lengths (the lengths
    travel with the vectors)."""
    print("[evolve] generation graph vs eager", flush=True)
    P, NI, stride = 64, 4, 320
    ns = np.array([64, 128, 200, 320] * 16, np.int32)
    pop = np.zeros((P, stride), np.float32)
    for p, n in enumerate(ns):
        pop[p, :n] = W.structured_prior(1, int(n), seed=200 + p)[0]
    kw = dict(islands=NI, iters=6, sigma=0.05, elite=2, migrate_every=3, seed=77)
    g = S.Generation(dev(pop), dev(ns, torch.int32), graph=True, **kw)
    e = S.Generation(dev(pop), dev(ns, torch.int32), graph=False, **kw)
    for rnd in range(3):
        g.run(3)
        e.run(3)
        torch.cuda.synchronize()
        _ = float(g.pop_a.sum().item())        # eager work between replays
        gp, gn = g.population()
        ep, en = e.population()
        assert g.gen.item() == e.gen.item() == 3 * (rnd + 1)
        assert torch.equal(gn, en) and torch.equal(gp, ep)
        assert torch.equal(g.fit, e.fit)
    gp, gn = g.population()
    lens = gn.cpu().numpy()
    assert set(lens.tolist()) <= {64, 128, 200, 320}
    arr = gp.cpu().numpy()
    for p in range(P):
        row = arr[p, :lens[p]].astype(np.float64)
        assert (row >= 0).all() and abs(row.sum() - 1) < 1e-5
    g.close()


def test_autotuner_picks_a_configuration(S):
    """The tuner times every configuration that fits n, keeps the built-in one
    unless another beats it by more than 1%, and leaves its choice as n's rule."""
    print("[evolve] autotuner", flush=True)
    builtin = S.get_ls_tile(600)
    tile, ms, table = S.tune_local_search(600, P=16, iters=5, repeats=1)
    try:
        assert ms > 0 and tile in table and table[tile] == ms
        assert len(table) == S.LS_TILES        # every configuration fits n = 600
        assert min(table.values()) <= ms
        if tile != builtin:
            assert ms < 0.99 * table[builtin]
        assert S.get_ls_tile(600) == tile
    finally:
        S.set_ls_tile(600, 0)
    assert S.get_ls_tile(600) == builtin
