"""fp64 polish spread over the device (sidon_polish_f64, the reference's polish
src/main.cu:798-846 at any n): the same step as the fused fp64 local search
(checked against it where it fits, n <= 4096) and against the fp64 oracle at
n = 8400, where the fused kernel does not fit one workgroup.  Sums run in
another order than either, so agreement is to fp64 rounding, not bitwise."""
from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def dev(x, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device="cuda")


def start(W, K, n, seed):
    a = W.structured_prior(K, n, seed=seed).astype(np.float64)
    return a / a.sum(axis=1, keepdims=True)


@pytest.mark.parametrize("n,mode", [(600, 0), (2100, 0), (4096, 0), (1200, 1)])
def test_polish_matches_fused_fp64(S, W, n, mode):
    print(f"[polish] vs fused fp64 n={n} mode={mode}", flush=True)
    a = start(W, 3, n, n)
    sched = (200.0, 6000.0, 0.01, 0.0002)
    t1, t2 = dev(a), dev(a)
    m1, v1 = torch.zeros_like(t1), torch.zeros_like(t1)
    m2, v2 = torch.zeros_like(t2), torch.zeros_like(t2)
    S.local_search(t1, 12, *sched, mom=m1, vel=v1, mode=mode)
    S.polish_f64(t2, 12, *sched, mode=mode, mom=m2, vel=v2)
    torch.cuda.synchronize()
    x1, x2 = t1.cpu().numpy(), t2.cpu().numpy()
    assert (x2 >= 0).all() and np.allclose(x2.sum(axis=1), 1.0, rtol=0, atol=1e-13)
    np.testing.assert_allclose(x2, x1, rtol=0, atol=1e-10 * x1.max())
    if mode == 0:
        np.testing.assert_allclose(m2.cpu().numpy(), m1.cpu().numpy(), rtol=0,
                                   atol=1e-8 * np.abs(m1.cpu().numpy()).max())


@pytest.mark.parametrize("n", [1, 7, 17, 255, 257, 1000])
def test_polish_ragged_lengths_match_fused_fp64(S, n):
    """Lengths that end inside a 16-segment, a 256-output tile or a chunk group;
    uniform-random starts with exact zeros (not the structured prior)."""
    print(f"[polish] ragged n={n}", flush=True)
    rng = np.random.default_rng(n)
    a = rng.random((2, n)) * (rng.random((2, n)) > 0.3)
    a[:, 0] += 0.1                                        # no all-zero row
    a /= a.sum(axis=1, keepdims=True)
    sched = (200.0, 6000.0, 0.01, 0.0002)
    t1, t2 = dev(a), dev(a)
    S.local_search(t1, 15, *sched)
    S.polish_f64(t2, 15, *sched)
    torch.cuda.synchronize()
    x1, x2 = t1.cpu().numpy(), t2.cpu().numpy()
    assert (x2 >= 0).all() and np.allclose(x2.sum(axis=1), 1.0, rtol=0, atol=1e-13)
    np.testing.assert_allclose(x2, x1, rtol=0, atol=1e-10 * x1.max())


def test_polish_at_8400_matches_oracle(S, O, W):
    print("[polish] n=8400 vs oracle", flush=True)
    n = 8400
    a = start(W, 2, n, 5)
    sched = (200.0, 600.0, 0.01, 0.004)
    t = dev(a)
    S.polish_f64(t, 3, *sched)
    got = t.cpu().numpy()
    for p in range(2):
        want, F, m, v, done = O.local_search(a[p], 3, *sched)
        assert done == 3
        np.testing.assert_allclose(got[p], want, rtol=0, atol=1e-10 * want.max())


def test_polish_improves_F_at_8400_and_is_deterministic(S, W):
    print("[polish] n=8400 F improves, deterministic", flush=True)
    n = 8400
    a = start(W, 4, n, 9)
    t1, t2 = dev(a), dev(a)
    F0, _ = S.eval_f64(t1)
    S.polish_f64(t1, 40, 200.0, 6000.0, 0.01, 0.0005)
    S.polish_f64(t2, 40, 200.0, 6000.0, 0.01, 0.0005)
    F1, _ = S.eval_f64(t1)
    assert torch.equal(t1, t2)
    assert (F1 < F0).all()


def test_polish_arguments(S, W):
    t = dev(start(W, 2, 300, 1))
    small = torch.empty(16, dtype=torch.uint8, device="cuda")
    with pytest.raises(S.SidonError):
        S.polish_f64(t, 3, 200.0, 600.0, 0.01, 0.004, workspace=small)
    with pytest.raises(S.SidonError):
        S.polish_f64(t, 3, 0.5, 600.0, 0.01, 0.004)          # p < 1
    before = t.clone()
    S.polish_f64(t, 0, 200.0, 600.0, 0.01, 0.004)             # no steps: untouched
    assert torch.equal(t, before)
