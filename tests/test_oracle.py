"""CPU oracle pinned against the reference's golden data, the reference's own
CPU verifier/twin (built from tools/probe.cpp), closed forms and library
routines.  No GPU."""
from __future__ import annotations

import json
import math
import subprocess
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN


# ---- F: closed forms (self-test KATs of src/main.cu:196-226) -------------

@pytest.mark.parametrize("n", [4, 8, 37, 128])
def test_constant_is_two(O, n):
    a = np.ones(n)
    assert O.F_hostorder(a) == 2.0
    assert O.F_exact(a)["F"] == 2.0


@pytest.mark.parametrize("a,want", [((1, 2, 2, 1), 80.0 / 36.0), ((1, 0, 0, 1), 4.0),
                                    ((0, 1, 0, 0), 8.0)])
def test_small_kats(O, a, want):
    a = np.array(a, dtype=float)
    assert abs(O.F_hostorder(a) - want) <= 1e-15 * want
    assert abs(O.F_gather(a)[0] - want) <= 1e-15 * want
    assert O.F_exact(a)["F"] == float(Fraction(2 * 4) * Fraction(int(np.convolve(a, a).max()))
                                      / Fraction(int(a.sum())) ** 2)
    assert abs(float(O.verify_ld(a)["F"]) - want) <= 1e-15 * want


def test_exact_F_against_rationals(O):
    """Brute force: exact rational F on small dyadic inputs == F_exact (correctly rounded)."""
    rng = np.random.default_rng(3)
    for _ in range(40):
        n = int(rng.integers(2, 24))
        a = rng.integers(0, 1 << 20, size=n).astype(np.float64) * 2.0 ** -17
        fr = [Fraction(x) for x in a]
        b = [sum((fr[i] * fr[k - i] for i in range(max(0, k - n + 1), min(k, n - 1) + 1)),
                 Fraction(0)) for k in range(2 * n - 1)]
        M, S = max(b), sum(fr, Fraction(0))
        if S == 0:
            continue
        ex = O.F_exact(a)
        assert ex["M"] == float(M) and ex["S"] == float(S)
        assert ex["F"] == 2.0 * n * float(M) / (float(S) * float(S))


def test_fsum_matches_math_fsum(O):
    rng = np.random.default_rng(5)
    for _ in range(50):
        x = rng.standard_normal(int(rng.integers(1, 400))) * 10.0 ** rng.integers(-20, 20, 1)
        assert O.fsum(x) == math.fsum(x)


def test_exact_b_matches_fsum_of_products(O):
    rng = np.random.default_rng(7)
    a = rng.uniform(size=61).astype(np.float32).astype(np.float64)
    ex = O.F_exact(a, want_b=True)
    n = a.size
    for k in range(2 * n - 1):
        terms = [a[i] * a[k - i] for i in range(max(0, k - n + 1), min(k, n - 1) + 1)]
        assert ex["b"][k] == math.fsum(terms)  # fp32 x fp32 products are exact in fp64


def test_scale_invariance(O):
    rng = np.random.default_rng(11)
    a = rng.uniform(size=40)
    assert abs(O.F_hostorder(a * 37.5) - O.F_hostorder(a)) <= 1e-12 * O.F_hostorder(a)


# ---- certificates (golden data of the reference) --------------------------

@pytest.mark.parametrize("n", [512, 4096, 8400])
def test_certificates_hostorder_bitwise(O, certs, n):
    c = certs[n]
    a = c["a"]
    assert a.size == c["n"] and (a >= 0).all()
    assert int((a != 0).sum()) == c["non_zero_cells"] == c["ladder"]["non_zero_cells"]
    assert np.array_equal(a.astype(np.float32).astype(np.float64), a)  # fp32-exact weights
    F = O.F_hostorder(a)
    assert F == c["F"]                         # bit for bit with the stored value
    assert F == O.F_gather(a)[0]
    assert abs(F - float(c["ladder"]["F"])) <= 1e-12  # ladder.csv rounds to 16 digits
    assert abs(O.F_exact(a)["F"] - F) <= 4e-15 * F


@pytest.mark.parametrize("n", [512, 4096, 8400])
def test_certificates_probe_verify(O, certs, n, tmp_path):
    """The oracle's long-double verifier prints exactly what the reference's
    tools/probe.cpp --verify printed for the same file (fixture)."""
    want = json.loads((GOLDEN / "probe_verify.json").read_text())[str(n)]
    p = tmp_path / f"sidon-{n}.txt"
    O.write_vec(p, certs[n]["a"], certs[n]["F"])
    rc, got = O.verify_file_str(p)
    assert rc == 0 and want["exit"] == 0
    assert got == [want["S"], want["M"], str(want["kmax"]), str(want["nActive"]), want["F"],
                   want["relBound"], want["F_upper"]]
    v = O.verify_ld(certs[n]["a"])
    assert abs(float(v["F"]) - certs[n]["F"]) <= 3e-15 * certs[n]["F"]


def test_verify_rejects_negative_and_length_mismatch(O, tmp_path):
    p = tmp_path / "neg.txt"
    O.write_vec(p, [0.5, -0.1, 0.3], 1.0)
    assert O.verify_file(p)["rc"] == 1
    q = tmp_path / "len.txt"
    q.write_text("n 4\nF 1\n0.1\n0.2\n0.3\n")
    assert O.verify_file(q)["rc"] == 3


def test_refine_preserves_exact_F_bitwise(O, W):
    """Cell split n -> 2n keeps F bit for bit under correctly rounded b and S,
    and only to ~1e-15 under plain sequential fp64 (src/main.cu:208-219)."""
    rng = np.random.default_rng(19)
    for n in (23, 34, 45, 128, 512):
        a = rng.uniform(size=n)
        a[rng.uniform(size=n) < 0.3] = 0.0
        r = O.refine(a)
        assert np.array_equal(r[0::2], a) and np.array_equal(r[1::2], a)
        assert O.F_exact(r)["F"] == O.F_exact(a)["F"]
        assert abs(O.F_hostorder(r) - O.F_hostorder(a)) <= 1e-12 * O.F_hostorder(a)
    a = W.structured_prior(2, 1050, dtype=np.float32).astype(np.float64)
    for row in a:
        r2 = O.refine(O.refine(row))
        assert O.F_exact(r2)["F"] == O.F_exact(row)["F"]


# ---- local search: bit-for-bit against the reference's CPU twin ------------

def test_local_search_matches_reference_probe_opt(O):
    """tools/probe.cpp --opt (compiled from the reference source) wrote these
    results; the oracle rebuilds the same starts (mt19937_64 + libstdc++
    uniform) and must reproduce F and the vector bit for bit."""
    from mt64 import probe_opt_start
    runs = json.loads((GOLDEN / "probe_opt.json").read_text())
    for run in runs:
        best = None
        for r in range(run["restarts"]):
            a = probe_opt_start(run["n"], run["seed"], r)
            S = 0.0
            for x in a:
                S += x
            a = np.array([x / S for x in a])
            out, F, *_ = O.local_search(a, run["iters"], run["pA"], run["pB"], run["lrA"],
                                        run["lrB"], mode=run["mode"], probe_semantics=True)
            if best is None or F < best[0]:
                best = (F, out)
        assert repr(best[0]) == repr(float(run["F"])) or best[0] == float(run["F"])
        assert np.array_equal(best[1], np.array(run["a"]))


def test_local_search_kernel_semantics_first_step(O):
    """First Adam step from zero moments: step = (lr/n) sign(g) (closed form)."""
    rng = np.random.default_rng(23)
    n = 64
    a = rng.uniform(size=n)
    a /= a.sum()
    lr = 1e-3
    out, F, m, v, done = O.local_search(a, 1, 8.0, 8.0, lr, lr)
    assert done == 1
    b = O.autoconv(a)
    M = b.max()
    c, R = O.lp_weights(b, M, 8.0)
    g = O.corr(np.concatenate([c, np.zeros(n)]), a) / R - M
    x = np.maximum(a - (lr / n) * np.sign(g), 0.0)
    np.testing.assert_allclose(out, x / x.sum(), rtol=1e-12, atol=0)
    np.testing.assert_allclose(m, 0.1 * g, rtol=1e-15)


def test_local_search_gmin_records_the_direction(O):
    """The gmin record is min over the steps of |g_i| / M: with a constant
    schedule the per-step directions come from the first moments of 1- and
    2-step runs (m1 = 0.1 g1, m2 = 0.9 m1 + 0.1 g2), and M from the heights
    each step starts from.  The trajectory is the plain run's."""
    rng = np.random.default_rng(41)
    a = rng.uniform(size=90)
    a /= a.sum()
    kw = (8.0, 8.0, 0.02, 0.02)
    one, _, m1, _, gmin1 = O.local_search_gmin(a, 1, *kw)
    two, _, m2, _, gmin2 = O.local_search_gmin(a, 2, *kw)
    assert np.array_equal(two, O.local_search(a, 2, *kw)[0])
    g1 = m1 / (1 - 0.9)
    g2 = (m2 - 0.9 * m1) / (1 - 0.9)
    M1 = np.convolve(a, a).max()
    M2 = np.convolve(one, one).max()
    np.testing.assert_allclose(gmin1, np.abs(g1) / M1, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(gmin2, np.minimum(np.abs(g1) / M1, np.abs(g2) / M2),
                               rtol=1e-9, atol=1e-12)


def test_local_search_improves(O):
    """Self-test 8 (src/main.cu:257-288) on the fp64 oracle."""
    rng = np.random.default_rng(5)
    for _ in range(3):
        a = rng.uniform(0.1, 1.0, size=128)
        F0 = O.F_hostorder(a)
        out, F1, *_ = O.local_search(a, 800, 4.0, 300.0, 0.06, 0.002)
        assert F1 < F0 and (out >= 0).all() and abs(out.sum() - 1) < 1e-12


# ---- gradient ---------------------------------------------------------------

@pytest.mark.parametrize("p", [8.0, 64.0, 1024.0])
def test_grad_against_long_double_central_differences(O, p):
    rng = np.random.default_rng(int(p))
    n = 96
    a = rng.uniform(0.2, 1.0, size=n)
    a /= a.sum()
    J, g = O.Jp_grad(a, p)
    Jl, gl = O.Jp_grad_ld(a, p)
    assert abs(J - float(Jl)) <= 1e-13 * J
    al = a.astype(np.longdouble)
    for i in rng.choice(n, 12, replace=False):
        h = np.longdouble(1e-7) * al[i]
        ap, am = al.copy(), al.copy()
        ap[i] += h
        am[i] -= h
        fd = (O.Jp_grad_ld(ap, p)[0] - O.Jp_grad_ld(am, p)[0]) / (2 * h)
        assert abs(float(fd) - g[i]) <= 2e-6 * np.abs(g).max()
        assert abs(float(gl[i]) - g[i]) <= 1e-10 * np.abs(g).max()


def test_Jp_tends_to_F(O):
    rng = np.random.default_rng(2)
    a = rng.uniform(size=80)
    F = O.F_hostorder(a)
    J = O.Jp(a, 1e6)
    assert F <= J <= F * (1 + 1e-4)
    assert abs(O.Jp(a, 1.0) - 2 * 80 * np.convolve(a, a).sum() / a.sum() ** 2) < 1e-12 * J


# ---- Frank-Wolfe / PDHG ------------------------------------------------------

def test_pdhg_lp_properties(O):
    """PDHG on the LP max sum g s.t. (a*g)_k <= M, g >= 0: g stays >= 0, the
    constraint violation shrinks, and sum g exceeds 1 (g = a is feasible)."""
    rng = np.random.default_rng(31)
    n = 48
    a = rng.uniform(size=n)
    a /= a.sum()
    M = np.convolve(a, a).max()
    viol = []
    for iters in (50, 400, 3000):
        g, lam = O.pdhg(a, M, iters)
        assert (g >= 0).all() and (lam >= 0).all()
        viol.append(max(0.0, (np.convolve(a, g) - M).max() / M))
    assert viol[-1] < viol[0] and viol[-1] < 5e-3
    assert g.sum() > 1.0


def test_pdhg_reaches_the_lp_optimum(O):
    """After 4000 iterations PDHG's sum g is within 1e-3 of the LP optimum
    max sum g s.t. (a*g)_k <= M, g >= 0 that scipy's HiGHS solver finds
    (an extrapolation or step-order mistake leaves it far off)."""
    from scipy.optimize import linprog
    rng = np.random.default_rng(31)
    for n in (24, 48):
        a = rng.uniform(size=n)
        a /= a.sum()
        M = np.convolve(a, a).max()
        A = np.zeros((2 * n - 1, n))
        for i in range(n):
            A[i:i + n, i] = a
        res = linprog(-np.ones(n), A_ub=A, b_ub=np.full(2 * n - 1, M), bounds=(0, None),
                      method="highs")
        assert res.status == 0
        opt = -res.fun
        g, _ = O.pdhg(a, M, 4000)
        assert opt >= 1.0 - 1e-12          # g = a is feasible
        assert abs(g.sum() - opt) < 1e-3 * opt, (n, g.sum(), opt)


def test_frank_wolfe_trace(O):
    """The traced Frank-Wolfe round is the plain one, its direction is the
    normalised PDHG solution, and the accepted step is the line search's
    minimum."""
    rng = np.random.default_rng(43)
    n = 80
    a = rng.uniform(size=n)
    a /= a.sum()
    out, acc, M1 = O.frank_wolfe(a, 1, 200, 12, 0.30)
    out2, acc2, M2, g, lsM = O.frank_wolfe_trace(a, 1, 200, 12, 0.30)
    assert np.array_equal(out, out2) and acc == acc2 == 1 and M1 == M2
    # (the oracle rescales a by 1/sum a first: equal to rounding)
    gp, _ = O.pdhg(a, np.convolve(a, a).max(), 200)
    np.testing.assert_allclose(g, gp / gp.sum(), rtol=1e-12, atol=1e-16)
    s = int(np.argmin(lsM))
    assert M1 == lsM[s]
    t = 0.30 * (s + 1) / 12
    np.testing.assert_allclose(out, (1.0 - t) * a + t * g, rtol=1e-12, atol=1e-16)
    for k in range(12):
        tt = 0.30 * (k + 1) / 12
        h = (1.0 - tt) * a + tt * g
        assert abs(np.convolve(h, h).max() - lsM[k]) <= 1e-15


def test_frank_wolfe_decreases_max(O):
    rng = np.random.default_rng(37)
    n = 64
    a = rng.uniform(size=n)
    a /= a.sum()
    M0 = np.convolve(a, a).max()
    out, acc, M1 = O.frank_wolfe(a, 3, 300, 12, 0.30)
    assert acc >= 1 and M1 < M0
    assert abs(out.sum() - 1.0) < 1e-12 and (out >= 0).all()
    assert abs(np.convolve(out, out).max() - M1) <= 1e-15


# ---- RNG, variation, selection ---------------------------------------------

def test_philox_known_answers(O):
    """Philox4x32-10 known-answer vectors (Random123 distribution)."""
    kat = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
           ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF),
            (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
           ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
            (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for c, k, want in kat:
        assert tuple(int(x) for x in O.philox4x32_10(c, k)) == want


def test_normals_are_standard(O):
    z = np.array([O.normal(20260905, 3, 7, i) for i in range(40000)])
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02
    assert abs((z ** 3).mean()) < 0.05 and abs((z ** 4).mean() - 3) < 0.1


def test_variation_lognormal_and_clone(O):
    rng = np.random.default_rng(41)
    A = rng.uniform(size=4096).astype(np.float32)
    A /= A.sum()
    out = O.variation(A, 2, 0.05, 20260905, 1, 3)
    assert out.dtype == np.float32 and (out >= 0).all()
    assert abs(float(out.astype(np.float64).sum()) - 1) < 1e-6
    r = np.log(out.astype(np.float64) / A)
    r -= r.mean()
    assert abs(r.std() - 0.05) < 0.003
    clone = O.variation(A, 20, 0.0, 20260905, 1, 3)
    S = math.fsum(A.astype(np.float64))
    np.testing.assert_array_equal(clone, (A.astype(np.float64) * (1.0 / S)).astype(np.float32))


def test_select_plan_semantics(O):
    rng = np.random.default_rng(43)
    P, NI, PI, elite = 64, 4, 16, 2
    fit = rng.uniform(1.5, 2.0, size=P)
    for gen in (5, 12):
        plan = O.select_plan(fit, NI, elite, 20260905, gen, migrate_every=12, sigma=0.05)
        f2 = plan["fit"]
        for i in range(NI):
            members = np.arange(i * PI, (i + 1) * PI)
            ranked = sorted(members, key=lambda x: (fit[x], x))
            for r in range(PI):
                s = i * PI + r
                if r < elite:
                    assert plan["p1"][s] == ranked[r] and plan["op"][s] == 20
                else:
                    assert plan["op"][s] == 2 and plan["strength"][s] == np.float32(0.05)
                    assert i * PI <= plan["p1"][s] < (i + 1) * PI
                    assert i * PI <= plan["p2"][s] < (i + 1) * PI
        if gen % 12 == 0:
            for i in range(NI):
                src, dst = plan["mig_src"][i], plan["mig_dst"][i]
                assert src == min(range(i * PI, (i + 1) * PI), key=lambda x: (fit[x], x))
                j = (i + 1) % NI
                assert dst == max(range(j * PI, (j + 1) * PI), key=lambda x: (fit[x], x))
                assert f2[dst] == fit[src]
        else:
            assert (plan["mig_src"] == -1).all() and np.array_equal(f2, fit)
    # tournament pressure: winners are better than a random member on average
    plan = O.select_plan(fit, NI, elite, 7, 1, migrate_every=0)
    nonelite = plan["op"] == 2
    assert fit[plan["p1"][nonelite]].mean() < fit.mean()
