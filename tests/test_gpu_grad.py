"""Backward kernel of J_p (recomputes b) through the C ABI and the
torch.autograd.Function harness: vs the fp64 / long-double oracle gradient and
central finite differences, including five planted near-tie peaks."""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def dev(x, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device="cuda")


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def planted_rows(W):
    """The adjoint check's batch: 64 structured-prior rows at n = 8400, each with five
    b_k planted within 1e-6 (relative, fp64, after rounding to fp32) of the
    max; rows where planting reached fewer than five ties are passed over."""
    base = W.structured_prior(96, 8400, seed=77)
    rows = []
    for r in range(base.shape[0]):
        a, near = W.plant_peaks(base[r], seed=r)
        if near.size >= 5:
            rows.append(a)
        if len(rows) == 64:
            break
    assert len(rows) == 64
    return np.stack(rows)


# per-height bounds on |g - g_ref| / max|g_ref|, about 3x the worst error
# measured over the 64 rows on MI355X (2.1e-6, 4.1e-5, 6.4e-4): fp32 b carries
# ~1e-7 relative error, which moves (b/M)^(p-1) near the planted ties by
# ~(p-1) 1e-7
ADJOINT_BOUND = {8.0: 7e-6, 64.0: 1.2e-4, 1024.0: 2e-3}


@pytest.mark.parametrize("p", [8.0, 64.0, 1024.0])
def test_adjoint_planted_ties_n8400(S, O, planted_rows, p):
    """Adjoint check at n = 8400, batch 64, through the autograd harness: J_p to
    1e-6 and every gradient component against the fp64 oracle, per height."""
    print(f"[grad] n=8400 batch 64 p={p}", flush=True)
    A = planted_rows
    t = dev(A).requires_grad_(True)
    J = S.lp_objective(t, p)
    J.sum().backward()
    g = t.grad.cpu().numpy().astype(np.float64)
    Jg = J.detach().cpu().numpy()
    with ThreadPoolExecutor(8) as ex:
        ref = list(ex.map(lambda a: O.Jp_grad(a, p), list(A)))
    worst = 0.0
    for r, (Jr, gr) in enumerate(ref):
        assert abs(Jg[r] - Jr) < 1e-6 * Jr, (r, Jg[r], Jr)
        e = float(np.abs(g[r] - gr).max() / np.abs(gr).max())
        worst = max(worst, e)
        assert e < ADJOINT_BOUND[p], (r, e)
    print(f"   worst per-height error over 64 rows: {worst:.3e} of max|g|")


@pytest.mark.parametrize("n", [1, 2, 3, 7, 15, 16, 17, 31, 33, 63, 65, 255, 257, 1000, 2047,
                               4095, 9297, 12177, 13216])
def test_lp_grad_fp32_ragged_lengths(S, O, n):
    """fp32 J_p and gradient at p = 8 at every tile edge the kernel has (the
    symmetric autoconvolution's zone and finish, Hankel tile tails), per
    height against the fp64 oracle; 9297, 12177 and 13216 with the
    correlation accumulated in the gradient rows (24- and 16-float layouts)."""
    print(f"[grad] fp32 ragged n={n}", flush=True)
    rng = np.random.default_rng(n)
    A = rng.uniform(0.1, 1.0, size=(3, n))
    A[1][rng.uniform(size=n) < 0.33] = 0.0
    A[1][0] = 1.0
    A[2] = rng.lognormal(0.0, 1.0, size=n)
    A = A.astype(np.float32).astype(np.float64)
    J, g = S.lp_grad(dev(A), 8.0)
    J = J.cpu().numpy().astype(np.float64)
    g = g.cpu().numpy().astype(np.float64)
    for r in range(3):
        Jr, gr = O.Jp_grad(A[r], 8.0)
        assert abs(J[r] - Jr) < 2e-6 * Jr, (r, J[r], Jr)
        # the natural size of a component: scale * M / S (the two terms of
        # G_i/R - M/S are both of that order)
        S_ = A[r].sum()
        M = np.convolve(A[r], A[r]).max()
        nat = 4.0 * n * Jr / (2.0 * n * M) * M / S_   # = 4 n R^{1/p} M / S^3
        e = float(np.abs(g[r] - gr).max() / nat)
        print(f"   row {r}: per-height err {e:.2e}")
        assert e < 6e-6, (r, e)            # measured up to 1.8e-6 (n = 4095)


def _dense_steps(n: int):
    """Matrix block steps of one pass when every weight c_k, k < 2n-1, is
    non-zero: (the symmetric autoconvolution's tiles, block offsets
    max(0, 16j-N16+1) <= d <= min(16j+15, N16, 8j+7); the correlation steps
    e in [16j, 16j+N16+15) whose weight window c[16e .. 16e+30] reaches below
    2n-1, i.e. e <= (2n-2) // 16)."""
    N16 = (n + 15) // 16
    NTc = (((2 * n - 1) + 15) // 16 + 15) // 16
    conv = sum(max(0, min(16 * j + 15, N16, 8 * j + 7) - max(0, 16 * j - N16 + 1) + 1)
               for j in range(NTc))
    last = (2 * n - 2) // 16
    corr = sum(max(0, min(16 * j + N16 + 14, last) - 16 * j + 1) for j in range((N16 + 15) // 16))
    return conv, corr


@pytest.mark.parametrize("n", [600, 8400, 13216])
def test_grad_block_step_count_dense(S, n):
    """Strictly positive heights at p = 4: every weight is non-zero, so the
    gradient kernel runs the dense schedule of one autoconvolution and one
    correlation a row (13216: the correlation in the gradient rows), and J
    alone the autoconvolution; the counts feed the bench's executed work."""
    print(f"[grad] block-step count n={n}", flush=True)
    rng = np.random.default_rng(n)
    a = dev(rng.uniform(0.5, 1.0, size=(3, n)))
    conv, corr = _dense_steps(n)
    S.ls_block_steps(reset=True)
    S.lp_grad(a, 4.0)
    torch.cuda.synchronize()
    assert S.ls_block_steps(reset=True) == 3 * (conv + corr)
    S.lp_grad(a, 4.0, want_grad=False)
    torch.cuda.synchronize()
    assert S.ls_block_steps(reset=True) == 3 * conv


def test_grad_fp64_kernel_and_fd(S, O, W):
    """fp64 backward kernel vs the long-double oracle gradient and central
    finite differences of the long-double J (step 1e-6 relative)."""
    print("[grad] fp64 + finite differences", flush=True)
    a, _ = W.plant_peaks(W.structured_prior(1, 700, seed=5)[0], seed=2)
    for p in (8.0, 64.0, 1024.0):
        J, g = S.lp_grad(dev(a[None, :], torch.float64), p)
        g = g.cpu().numpy()[0]
        Jl, gl = O.Jp_grad_ld(a, p)
        assert abs(J.item() - float(Jl)) < 1e-12 * float(Jl)
        assert rel(g, gl.astype(np.float64)) < 1e-9
        al = a.astype(np.longdouble)
        for i in np.random.default_rng(1).choice(np.flatnonzero(a > 0), 6, replace=False):
            h = np.longdouble(1e-6) * al[i]
            ap, am = al.copy(), al.copy()
            ap[i] += h
            am[i] -= h
            fd = float((O.Jp_grad_ld(ap, p)[0] - O.Jp_grad_ld(am, p)[0]) / (2 * h))
            assert abs(fd - g[i]) < 1e-5 * np.abs(g).max()


def test_autograd_gradcheck_small(S):
    print("[grad] torch gradcheck", flush=True)
    rng = np.random.default_rng(3)
    a = dev(rng.uniform(0.2, 1.0, size=(2, 40)), torch.float64).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda x: S.lp_objective(x, 16.0), (a,), eps=1e-7,
                                    atol=1e-6, rtol=1e-5)
