"""The index bookkeeping of the symmetric MFMA autoconvolution, emulated on the
CPU in exact integer arithmetic and checked against np.convolve.

The search kernel forms b_k = 2 sum_{i < k-i} a_i a_{k-i} + [k even] a_{k/2}^2
on the matrix cores: output segment K (outputs 16K .. 16K+15) takes Toeplitz
blocks d in [lo_j, hi_j] of its tile j = K / 16 (lo_j = max(0, 16j - N16 + 1),
hi_j = min(16j + 15, N16, 8j + 7)), block d adding
sum_s a[16d + t - s] a[16(K - d) + s] to row t, but only while d <
dp(K) = ceil(K/2) (past that the column reads zeros); the finish then adds
the products 16(dp - 1) + t < i < k/2 (at most 15, ascending i), doubles,
and adds the square term.  This test restates exactly those index ranges
(the kernel's autoconv_drange, the zone's live test and autoconv_finish /
autoconv_finish_lane) and checks that every product a_i a_j is counted once:
integer heights make every sum exact, so any dropped or doubled product or
off-by-one range shows as a mismatch.  The emulation also tries the two
mistakes the design guards against (a zone without the live mask, a finish
without the square term) and checks that they do fail.
"""
from __future__ import annotations

import numpy as np
import pytest


def emulate(a: np.ndarray, live_mask: bool = True, square: bool = True) -> np.ndarray:
    n = a.size
    N16 = (n + 15) // 16
    NK = (2 * n - 1 + 15) // 16          # output segments
    NT = (NK + 15) // 16                 # tiles of 16 segments
    ext = np.zeros(16 * (N16 + 2 * 16) + 64, dtype=np.int64)
    off = 16 * 16                        # a[x] at ext[x + off], zero outside [0, n)
    ext[off:off + n] = a
    A = lambda x: ext[x + off]           # noqa: E731

    part = np.zeros(16 * 16 * NT, dtype=np.int64)
    t = np.arange(16)
    s = np.arange(16)
    for j in range(NT):
        lo = max(0, 16 * j - N16 + 1)
        hi = min(16 * j + 15, N16, 8 * j + 7)
        for K in range(16 * j, 16 * j + 16):
            dp = (K + 1) >> 1
            for d in range(lo, hi + 1):
                if live_mask and d >= dp:
                    continue                      # the column reads the zero segment
                H = A(16 * d + t[:, None] - s[None, :])   # H_d[t][s] = a[16d + t - s]
                B = A(16 * (K - d) + s)                   # a[16(K - d) + s]
                part[16 * K + t] += H @ B
    b = np.zeros(2 * n - 1, dtype=np.int64)
    for k in range(2 * n - 1):
        K, row = k >> 4, k & 15
        dp = (K + 1) >> 1
        i0 = 16 * (dp - 1) + row + 1
        cnt = ((k + 1) >> 1) - i0
        r = part[k]
        for u in range(15):
            if u < cnt:
                r += A(i0 + u) * A(k - i0 - u)
        v = 2 * r
        if square and not (k & 1):
            v += A(k >> 1) * A(k >> 1)
        b[k] = v
    return b


def emulate_lane_finish(a: np.ndarray) -> np.ndarray:
    """autoconv_finish_lane's form: a lane (column K, row group kk) reads the
    eighteen heights a_{i0 .. i0+17} and the fifteen a_{16(K-dp)+15-u} once and
    finishes its four outputs (rows 4kk + r) from them; the matrix share is
    taken from emulate()'s range rules (a checked one) by subtracting the
    finish of that one."""
    n = a.size
    full = np.convolve(a, a)
    out = np.zeros(2 * n - 1, dtype=np.int64)
    ext = np.zeros(4 * n + 64, dtype=np.int64)
    off = 32
    ext[off:off + n] = a
    A = lambda x: ext[x + off]           # noqa: E731
    NK = (2 * n - 1 + 15) // 16
    for K in range(NK):
        dp = (K + 1) >> 1
        for kk in range(4):
            i0 = 16 * (dp - 1) + 4 * kk + 1
            xv = [A(i0 + v) for v in range(18)]
            yv = [A(16 * (K - dp) + 15 - u) for u in range(15)]
            for r in range(4):
                k = 16 * K + 4 * kk + r
                if k >= 2 * n - 1:
                    continue
                cnt = ((k + 1) >> 1) - (i0 + r)
                # the matrix share: the pairs i < j, i <= 16(dp - 1) + row
                share = 0
                for i in range(0, min(k, 16 * (dp - 1) + 4 * kk + r) + 1):
                    if i < k - i:
                        share += A(i) * A(k - i)
                s_ = share + sum(xv[r + u] * yv[u] for u in range(15) if u < cnt)
                v = 2 * s_ + (A(k >> 1) ** 2 if not (k & 1) else 0)
                out[k] = v
    np.testing.assert_array_equal(out, full)
    return out


LENGTHS = [1, 2, 3, 7, 15, 16, 17, 31, 33, 63, 65, 255, 257, 1000]


@pytest.mark.parametrize("n", LENGTHS)
def test_symmetric_split_counts_every_product_once(n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 10, size=n).astype(np.int64)
    a[rng.uniform(size=n) < 0.3] = 0
    np.testing.assert_array_equal(emulate(a), np.convolve(a, a))


@pytest.mark.parametrize("n", [33, 257])
def test_lane_finish_matches_convolution(n):
    rng = np.random.default_rng(7 + n)
    emulate_lane_finish(rng.integers(0, 10, size=n).astype(np.int64))


@pytest.mark.parametrize("n", [17, 65, 257])
def test_the_checks_catch_the_mistakes(n):
    """Without the zone's live mask the products i > j of the zone are added a
    second time; without the square term every even output is short."""
    a = np.random.default_rng(3).integers(1, 10, size=n).astype(np.int64)
    want = np.convolve(a, a)
    assert not np.array_equal(emulate(a, live_mask=False), want)
    assert not np.array_equal(emulate(a, square=False), want)
