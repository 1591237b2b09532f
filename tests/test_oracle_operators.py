"""The oracle's fourteen mutation operators (src/kernels.cuh:614-787) and the
support signature (src/kernels.cuh:803-849), pinned by what each operator is
rather than by its formula: a splice is one contiguous segment of the second
parent, a reflection reverses, a shift moves every height by one offset, a
sparsify zeroes exactly the heights below one threshold, and so on, each up
to the renormalisation.  Parents are random and continuous, so accidental
agreement is impossible."""
from __future__ import annotations

import math

import numpy as np
import pytest

SEED = 20260905
N = 1000


def parents(k=0):
    rng = np.random.default_rng(100 + k)
    A = rng.uniform(0.5, 1.5, size=N).astype(np.float32)
    B = rng.uniform(0.5, 1.5, size=N).astype(np.float32)
    return A, B


def close(x, y, rtol=2e-6):
    return np.abs(np.asarray(x, np.float64) - np.asarray(y, np.float64)) <= rtol * np.abs(y)


def one_run(mask):
    """True if the True entries of mask form one contiguous run."""
    idx = np.flatnonzero(mask)
    return idx.size > 0 and idx[-1] - idx[0] + 1 == idx.size


def common_scale(out, X):
    """The scale c with out = c X on the entries of X that are not zero (the
    median ratio), and the mask of entries where it holds."""
    m = X != 0
    c = float(np.median(out[m].astype(np.float64) / X[m]))
    return c, close(out, c * X.astype(np.float64)) & m


def check_sum(out):
    assert out.dtype == np.float32 and (out >= 0).all()
    assert abs(math.fsum(out.astype(np.float64)) - 1.0) < N * 2.0 ** -24


SLOTS = range(12)


def test_params_are_uniform_and_keyed(O):
    u = np.stack([O.op_params(SEED, g, s) for g in range(8) for s in range(64)])
    assert (u > 0).all() and (u <= 1).all()
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.005
    assert not np.array_equal(O.op_params(SEED, 1, 2), O.op_params(SEED, 1, 3))
    assert not np.array_equal(O.op_params(SEED, 1, 2), O.op_params(SEED, 2, 2))
    np.testing.assert_array_equal(O.op_params(SEED, 5, 9), O.op_params(SEED, 5, 9))


@pytest.mark.parametrize("slot", SLOTS)
def test_op0_blend(O, slot):
    A, B = parents(slot)
    out = O.variation2(A, B, 0, 0.1, SEED, 3, slot)
    check_sum(out)
    # out = c (lam A + (1 - lam) B), clamped at 0: fit c and c*lam on the positive part
    pos = out > 0
    M = np.stack([B[pos], (A - B)[pos]], 1).astype(np.float64)
    (c, clam), *_ = np.linalg.lstsq(M, out[pos].astype(np.float64), rcond=None)
    lam = clam / c
    assert -0.3 - 1e-6 <= lam <= 1.3 + 1e-6
    X = lam * A.astype(np.float64) + (1 - lam) * B.astype(np.float64)
    assert np.all(close(out[pos], c * X[pos], 1e-5))
    assert np.all(X[~pos] <= 1e-6)


@pytest.mark.parametrize("op", [1, 8])
@pytest.mark.parametrize("slot", SLOTS)
def test_op1_op8_splice(O, op, slot):
    A, B = parents(slot)
    out = O.variation2(A, B, op, 0.1, SEED, 3, slot)
    check_sum(out)
    # one scale for both parents: out = c A outside the segment, c B inside
    cA, okA = common_scale(out, A)
    okB = close(out, cA * B.astype(np.float64))
    seg = okB & ~okA
    assert np.all(okA | okB)
    if seg.any():
        assert one_run(seg)
        if op == 8:
            assert seg.sum() <= N // 2 + 1
    else:
        assert op == 1 or np.all(okA)  # an empty segment leaves A


@pytest.mark.parametrize("slot", SLOTS)
def test_op3_smooth(O, slot):
    A, _ = parents(slot)
    sg = 0.1
    out = O.variation2(A, None, 3, sg, SEED, 3, slot)
    check_sum(out)
    r = out.astype(np.float64) / A
    c = r.mean()
    assert np.all(np.abs(r / c - 1) <= 6 * sg + 1e-6)          # 1 + 3 sg m, |m| <= 1
    assert np.abs(np.diff(r / c)).max() < 3 * sg * 2 * np.pi * 13 * 2 / N  # few low modes


@pytest.mark.parametrize("slot", SLOTS)
def test_op4_additive(O, slot):
    A, _ = parents(slot)
    sg = 0.05
    out = O.variation2(A, None, 4, sg, SEED, 3, slot)
    check_sum(out)
    z = np.array([O.normal(SEED, 3, slot, i) for i in range(N)], np.float64)
    mean = math.fsum(A.astype(np.float64)) / N
    X = A + sg * mean * z
    c, ok = common_scale(out, X)
    assert ok.mean() > 0.999


@pytest.mark.parametrize("slot", SLOTS)
def test_op5_sparsify(O, slot):
    A, _ = parents(slot)
    out = O.variation2(A, None, 5, 0.1, SEED, 3, slot)
    check_sum(out)
    zero = out == 0
    c, ok = common_scale(out, A)
    assert np.all(ok | zero)
    mean = A.astype(np.float64).mean()
    if zero.any():
        assert A[zero].max() < A[~zero].min()            # one threshold
        assert A[zero].max() < 0.95 * mean + 1e-6
    if (~zero).any() and zero.any():
        assert A[~zero].min() >= 0.05 * mean - 1e-6


@pytest.mark.parametrize("slot", SLOTS)
def test_op6_reflect_or_symmetrise(O, slot):
    A, _ = parents(slot)
    out = O.variation2(A, None, 6, 0.1, SEED, 3, slot)
    check_sum(out)
    c, ok = common_scale(out, A[::-1])
    if ok.all():
        return                                           # hard reflection
    np.testing.assert_array_equal(out, out[::-1])        # symmetrised, bit for bit
    c, ok = common_scale(out, 0.5 * (A.astype(np.float64) + A[::-1]))
    assert ok.all()


@pytest.mark.parametrize("slot", SLOTS)
def test_op7_shift(O, slot):
    A, _ = parents(slot)
    out = O.variation2(A, None, 7, 0.1, SEED, 3, slot)
    check_sum(out)
    found = []
    for sh in range(-N // 8, N // 8 + 1):
        X = np.zeros(N)
        if sh >= 0:
            X[sh:] = A[:N - sh]
        else:
            X[:N + sh] = A[-sh:]
        c = out[X > 0].astype(np.float64) / X[X > 0]
        if c[0] > 0 and np.all(np.abs(c / c[0] - 1) < 2e-6) and np.all(out[X == 0] == 0):
            found.append(sh)
    assert len(found) == 1


@pytest.mark.parametrize("slot", SLOTS)
def test_op9_op11_ignore_parent(O, slot):
    A, B = parents(slot)
    for op in (9, 11):
        o1 = O.variation2(A, None, op, 0.1, SEED, 3, slot)
        o2 = O.variation2(B, None, op, 0.1, SEED, 3, slot)
        check_sum(o1)
        np.testing.assert_array_equal(o1, o2)
    s = O.variation2(A, None, 11, 0.1, SEED, 3, slot)
    assert int(np.argmax(s)) in (0, N - 1)               # the endpoint atom
    z = np.flatnonzero(s == 0)
    assert z.size > 0.04 * N and one_run(s == 0)         # the hard gap


@pytest.mark.parametrize("slot", SLOTS)
def test_op10_spike(O, slot):
    A, _ = parents(slot)
    out = O.variation2(A, None, 10, 0.1, SEED, 3, slot)
    check_sum(out)
    c, ok = common_scale(out, A)
    win = ~ok
    assert one_run(win) and win.sum() <= 2 * (1 + int(0.03 * N)) + 1
    amp = out[win].astype(np.float64) / c - A[win]
    assert np.all(np.abs(amp / amp[0] - 1) < 1e-4)       # one constant amplitude
    mean = A.astype(np.float64).mean()
    assert 1 - 1e-4 <= amp[0] / mean <= 7 + 1e-4


@pytest.mark.parametrize("slot", SLOTS)
def test_op12_gap(O, slot):
    A, _ = parents(slot)
    out = O.variation2(A, None, 12, 0.1, SEED, 3, slot)
    check_sum(out)
    c, ok = common_scale(out, A)
    win = ~ok
    assert one_run(win) and win.sum() <= 1 + int(0.145 * N)
    if (out[win] == 0).all():
        return                                           # carved
    fill = out[win].astype(np.float64) / c - A[win]
    mean = A.astype(np.float64).mean()
    assert np.all(np.abs(fill / fill[0] - 1) < 1e-4) and 0.2 - 1e-4 <= fill[0] / mean <= 2 + 1e-4


@pytest.mark.parametrize("slot", SLOTS)
def test_op13_endpoint_atom(O, slot):
    A, _ = parents(slot)
    out = O.variation2(A, None, 13, 0.1, SEED, 3, slot)
    check_sum(out)
    c, ok = common_scale(out, A)
    ch = np.flatnonzero(~ok)
    wmax = 1 + int(0.012 * N)
    assert ch.size <= wmax
    if ch.size:
        left = ch[0] == 0
        d = ch if left else N - 1 - ch
        assert d.max() < wmax
        delta = (out[ch].astype(np.float64) / c - A[ch]) * (1 + d)
        live = out[ch] > 0
        if live.sum() > 1:
            assert np.all(np.abs(delta[live] / delta[live][0] - 1) < 1e-3)


def test_operators_renormalise_and_clone(O):
    A, B = parents(0)
    for op in list(range(14)) + [20, -1]:
        check_sum(O.variation2(A, B, op, 0.1, SEED, 3, 0))
    S = math.fsum(A.astype(np.float64))
    np.testing.assert_array_equal(O.variation2(A, B, 20, 0.1, SEED, 3, 0),
                                  (A.astype(np.float64) * (1.0 / S)).astype(np.float32))
    np.testing.assert_array_equal(O.variation2(A, None, 2, 0.05, SEED, 1, 3),
                                  O.variation(A, 2, 0.05, SEED, 1, 3))
    z = np.zeros(N, np.float32)                          # zero sum -> uniform
    np.testing.assert_array_equal(O.variation2(z, None, 5, 0.1, SEED, 3, 0),
                                  np.full(N, 1.0 / N, np.float32))


# ------------------------------------------------------------ signature ----

def xor_mix(idx, O):
    h = 0
    for i in idx:
        h ^= O.sig_mix(int(i))
    return h


def test_sig_mix_zero_and_spread(O):
    assert O.sig_mix(0) == 0                             # 0 * odd constant, shifts of 0
    vals = {O.sig_mix(i) for i in range(1, 5000)}
    assert len(vals) == 4999
    bits = np.array([bin(v).count("1") for v in vals])
    assert abs(bits.mean() - 32) < 0.5


def test_support_sig_threshold_is_tol_times_mean(O):
    rng = np.random.default_rng(5)
    n = 3000
    a = rng.uniform(1.0, 2.0, size=n).astype(np.float32)
    mean = a.astype(np.float64).mean()
    lo = rng.choice(n, 300, replace=False)
    a[lo[:150]] = 0.5 * 0.02 * mean                      # clearly below mean * tol
    a[lo[150:]] = 2.0 * 0.02 * mean                      # clearly above
    mean = a.astype(np.float64).mean()
    sup = np.flatnonzero(a > 0.02 * mean)
    assert sup.size == n - 150
    sig, nnz = O.support_sig(a, 0.02)
    assert nnz == sup.size and sig == xor_mix(sup, O)
    sig5, nnz5 = O.support_sig(a, 0.5)                   # another tol moves the cut
    assert nnz5 == n - 300 and sig5 == xor_mix(np.flatnonzero(a > 0.5 * mean), O)


def test_support_sig_depends_only_on_the_support(O):
    rng = np.random.default_rng(6)
    n = 1200
    a = rng.uniform(1.0, 2.0, size=n).astype(np.float32)
    a[rng.uniform(size=n) < 0.3] = 0
    b = a.copy()
    nz = np.flatnonzero(b)
    b[nz] = rng.permutation(b[nz])                       # same support, other values
    assert O.support_sig(a) == O.support_sig(b)
    # disjoint supports combine by XOR
    s1 = np.zeros(n, np.float32); s1[:400] = 1
    s2 = np.zeros(n, np.float32); s2[400:900] = 1
    h1, _ = O.support_sig(s1); h2, _ = O.support_sig(s2)
    h12, c12 = O.support_sig(s1 + s2)
    assert h12 == h1 ^ h2 and c12 == 900
    assert O.support_sig(np.zeros(n, np.float32)) == (0, 0)
