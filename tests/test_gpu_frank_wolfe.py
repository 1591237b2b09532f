"""Frank-Wolfe / Chambolle-Pock kernel vs the fp64 oracle restatement of
src/kernels.cuh:408-532, and the LP properties the method guarantees."""
from __future__ import annotations

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
