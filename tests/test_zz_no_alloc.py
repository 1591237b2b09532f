"""sidon_polish_f64_noalloc (the fp64 polish on the library's static scratch
array) against sidon_polish_f64 with a caller workspace, through the C ABI:
the same kernels, so heights and moments must agree bit for bit.  Two calls
in a row, the second with a smaller input (fewer candidates, shorter rows),
so the second runs on a scratch array the first left dirty.  The size limit
is rejected before any device work."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _run_workspace(S, pop, iters, sched, mode):
    L = S.lib()
    K, n = pop.shape
    out = pop.clone()
    mom, vel = torch.empty_like(pop), torch.empty_like(pop)
    need = L.sidon_polish_f64_workspace(n, K)
    ws = torch.empty(need, dtype=torch.uint8, device=pop.device)
    rc = L.sidon_polish_f64(out.data_ptr(), mom.data_ptr(), vel.data_ptr(), n, K, iters, *sched,
                            0.9, 0.999, mode, ws.data_ptr(), need,
                            torch.cuda.current_stream().cuda_stream)
    assert rc == S.SIDON_OK, L.sidon_last_error()
    return out, mom, vel


def _run_noalloc(S, pop, iters, sched, mode):
    L = S.lib()
    K, n = pop.shape
    out = pop.clone()
    mom, vel = torch.empty_like(pop), torch.empty_like(pop)
    rc = L.sidon_polish_f64_noalloc(out.data_ptr(), mom.data_ptr(), vel.data_ptr(), n, K, iters,
                                    *sched, 0.9, 0.999, mode,
                                    torch.cuda.current_stream().cuda_stream)
    assert rc == S.SIDON_OK, L.sidon_last_error()
    return out, mom, vel


@pytest.mark.parametrize("K,n,mode", [(6, 4096, 0), (2, 1000, 1)])
def test_noalloc_polish_matches_workspace_polish(S, W, K, n, mode):
    print(f"[no-alloc] polish K={K} n={n} mode={mode}", flush=True)
    sched = (8.0, 2e3, 0.02, 0.004)
    big = torch.as_tensor(W.structured_prior(8, 8400, seed=3).astype(np.float64), device="cuda")
    small = torch.as_tensor(W.structured_prior(K, n, seed=K + n).astype(np.float64),
                            device="cuda")
    for pop in (big, small):            # second call: smaller input on dirty scratch
        ref = _run_workspace(S, pop, 30, sched, mode)
        got = _run_noalloc(S, pop, 30, sched, mode)
        torch.cuda.synchronize()
        for r, g in zip(ref, got):
            assert torch.equal(r, g)
        assert not torch.equal(ref[0], pop)


def test_noalloc_polish_rejects_oversize(S):
    print("[no-alloc] size limit", flush=True)
    L = S.lib()
    fake = C.c_void_p(0x1000)
    assert L.sidon_polish_f64_workspace(8400, 15) <= 32 << 20 < L.sidon_polish_f64_workspace(8400,
                                                                                            16)
    rc = L.sidon_polish_f64_noalloc(fake, None, None, 8400, 16, 10, 8.0, 8.0, 0.01, 0.01, 0.9,
                                    0.999, 0, None)
    assert rc == S.SIDON_ERR_ARG
    assert b"32" in L.sidon_last_error() or b"scratch" in L.sidon_last_error()
