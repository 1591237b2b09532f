"""Frank-Wolfe / Chambolle-Pock kernel vs the fp64 oracle restatement of
src/kernels.cuh:408-532, and the LP properties the method guarantees."""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def dev(x, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device="cuda")


@pytest.mark.parametrize("n", [64, 300, 1050])
def test_frank_wolfe_vs_oracle(S, O, W, n):
    print(f"[fw] n={n}", flush=True)
    a = W.structured_prior(2, n, seed=n).astype(np.float64)
    t = dev(a)
    S.frank_wolfe(t, 2, 200, 12, 0.30)
    got = t.cpu().numpy().astype(np.float64)
    for p in range(2):
        start = a[p].astype(np.float32).astype(np.float64)
        want, acc, M1 = O.frank_wolfe(start, 2, 200, 12, 0.30)
        M0 = np.convolve(start / start.sum(), start / start.sum()).max()
        Mg = np.convolve(got[p], got[p]).max()
        assert (got[p] >= 0).all() and abs(got[p].sum() - 1) < 1e-5
        assert Mg <= M0 * (1 + 1e-6)                  # never worse than the start
        if acc > 0:
            assert abs(Mg - M1) < 2e-3 * M1
            assert np.abs(got[p] - want).max() < 5e-3 * want.max()


# the longest candidate the Frank-Wolfe workgroup's LDS holds with 20-float
# segments, and with the plain 16-float layout
FW_NARROW_MAX_N = 8720
FW_MAX_N = 9856


@pytest.mark.parametrize("n", [4200, 8400, FW_NARROW_MAX_N, FW_NARROW_MAX_N + 1, FW_MAX_N])
def test_frank_wolfe_vs_oracle_large_n(S, O, W, n):
    """The ladder's levels 4200 and 8400: one Frank-Wolfe round of 50
    Chambolle-Pock iterations against the fp64 oracle, per height.  The GPU's
    line search compares fp32 maxima: when the oracle's best two trial steps
    are within 1e-5 of each other either choice is right, and the GPU result
    must match the oracle's update for one of them."""
    print(f"[fw] n={n} vs oracle", flush=True)
    a = W.structured_prior(2, n, seed=n + 1).astype(np.float32).astype(np.float64)
    t = dev(a)
    S.frank_wolfe(t, 1, 50, 12, 0.30)
    got = t.cpu().numpy().astype(np.float64)
    with ThreadPoolExecutor(2) as ex:
        ref = list(ex.map(lambda x: O.frank_wolfe_trace(x, 1, 50, 12, 0.30), list(a)))
    for p in range(2):
        want, acc, M1, g, lsM = ref[p]
        start = a[p] / a[p].sum()
        order = np.argsort(lsM)
        ts = 0.30 * (order[:2] + 1) / 12
        alts = [(1 - tt) * start + tt * g for tt in ts] if acc else [start]
        tie = acc and (lsM[order[1]] - lsM[order[0]]) / lsM[order[0]] < 1e-5
        if acc:
            np.testing.assert_allclose(alts[0], want, rtol=1e-12, atol=1e-18)
        errs = [float(np.abs(got[p] - alt).max() / alt.max()) for alt in (alts if tie else
                                                                            alts[:1])]
        print(f"   row {p}: per-height err {min(errs):.2e} (line-search tie: {tie})")
        assert min(errs) < 5e-7, errs      # measured up to 1.4e-7
        assert (got[p] >= 0).all() and abs(got[p].sum() - 1) < 1e-5


def test_frank_wolfe_large_n_improves(S, O, W):
    print("[fw] n=8400 improves", flush=True)
    a = W.structured_prior(2, 8400, seed=3)
    t = dev(a)
    F0, _ = S.eval_f64(t)
    S.frank_wolfe(t, 1, 100, 12, 0.30)
    F1, _ = S.eval_f64(t)
    assert (F1 <= F0 * (1 + 1e-7)).all() and (F1 < F0).any()


def test_skip_and_ladder(S, O, W):
    """Cell split up the ladder 1050 -> 2100 -> 4200 -> 8400 keeps F bit for bit
    (correctly rounded), and FW at each level only lowers it."""
    print("[fw] ladder", flush=True)
    t = dev(W.structured_prior(3, 1050, seed=12))
    skip = dev(np.array([0, 1, 0]), torch.int32)
    for level in range(4):
        Fe0, *_ , amb0 = S.eval_exact(t)
        before = t.clone()
        S.frank_wolfe(t, 1, 50, 12, 0.30, skip=skip)
        assert torch.equal(t[1], before[1])
        Fe1, *_ = S.eval_exact(t)
        assert (Fe1 <= Fe0 * (1 + 1e-7)).all()
        if level < 3:
            r = S.refine(t)
            Fr, *_, amb = S.eval_exact(r)
            assert (amb == 0).all()
            assert torch.equal(Fr, Fe1)
            t = r


@pytest.mark.parametrize("n", [FW_MAX_N + 1, 13216])
def test_frank_wolfe_past_the_lds_is_refused(S, n):
    """Past FW_MAX_N the solve returns SIDON_ERR_LDS without a launch
    (src/main.cu:173's "does not fit"), also at lengths the local search
    still runs (13216), so a driver skips Frank-Wolfe there."""
    assert S.lds_bytes(3, FW_MAX_N, 1024) <= 160 * 1024 < S.lds_bytes(3, n, 1024)
    t = torch.full((1, n), 1.0 / n, dtype=torch.float32, device="cuda")
    with pytest.raises(S.SidonError) as e:
        S.frank_wolfe(t, 1, 5)
    assert e.value.status == S.SIDON_ERR_LDS
