"""Seeded synthetic inputs for the benchmark configurations (numpy PCG64).

Every sampler is the build's own; the default seed 20260905 is the reference
search's default seed (src/main.cu:54) and is recorded with every result.

  * plain small vectors: a_i = 0 with probability 0.33, else U(0,1);
  * structured prior: endpoint atom a_0 = 5 x body mean, zeros on cells
    [1, 0.1 n), body a_i ~ LogNormal(0,1) (1 + 0.5 i/n) with 33 % exact zeros,
    renormalised to sum 1 (the shape every published record shares,
    src/kernels.cuh:570-575);
  * planted near-ties: a structured-prior vector adjusted by a few Newton steps
    on a handful of cells so that several b_k sit within 1e-6 of the max;
  * uneven lengths drawn uniformly from a set.
"""
from __future__ import annotations

import numpy as np

SEED = 20260905
UNEVEN_NS = (600, 1200, 2100, 4096, 8400)


def plain_vectors(P: int = 64, n: int = 512, seed: int = SEED) -> np.ndarray:
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.0, 1.0, size=(P, n))
    a[rng.uniform(size=(P, n)) < 0.33] = 0.0
    return a


def structured_prior(P: int, n: int, seed: int = SEED, dtype=np.float32) -> np.ndarray:
    rng = np.random.default_rng(seed)
    out = np.zeros((P, n), dtype=np.float64)
    i = np.arange(n, dtype=np.float64)
    gap_end = max(1, int(0.1 * n))
    for p in range(P):
        body = rng.lognormal(0.0, 1.0, size=n) * (1.0 + 0.5 * i / n)
        body[rng.uniform(size=n) < 0.33] = 0.0
        body[:gap_end] = 0.0
        nz = body[gap_end:]
        mean = nz.mean() if nz.size else 1.0
        body[0] = 5.0 * mean
        out[p] = body / body.sum()
    return out.astype(dtype)


def autoconv64(a: np.ndarray) -> np.ndarray:
    """fp64 autoconvolution via numpy (test-data preparation only)."""
    return np.convolve(a, a)


def _b_at(a: np.ndarray, k: int) -> float:
    """b_k = sum_i a_i a_{k-i} alone (O(n))."""
    n = a.size
    lo, hi = max(0, k - n + 1), min(k, n - 1)
    return float(np.dot(a[lo:hi + 1], a[k - hi:k - lo + 1][::-1]))


def plant_peaks(a: np.ndarray, n_peaks: int = 5, spread: float = 1e-7, cells: int = 48,
                seed: int = SEED, round_fp32: bool = True, tol: float = 1e-6):
    """Adjust a (fp64 copy) so that at least n_peaks values b_k lie within tol of
    max b (relative), checked in fp64 on the returned (optionally fp32-rounded)
    values.  Returns (a_new, peak_indices)."""
    rng = np.random.default_rng(seed)
    a = np.array(a, dtype=np.float64)
    n = a.size
    b = autoconv64(a)
    ks = np.argsort(b)[::-1][:n_peaks]
    for _ in range(6):
        # bring the top values down to the level of the n_peaks-th largest
        # (lowering b can push no other value above the planted level)
        L = b[ks[-1]]
        target = L * (1.0 - spread * np.arange(n_peaks))
        # free cells: for every target k, cells i with a_i > 0 and a_{k-i} > 0
        # (so that db_k/da_i = 2 a_{k-i} is non-zero for that row)
        free = set()
        per = max(2, cells // n_peaks)
        for k in ks:
            i = np.arange(max(0, k - n + 1), min(k, n - 1) + 1)
            ok = i[(a[i] > 0) & (a[k - i] > 0)]
            if ok.size:
                free.update(rng.choice(ok, size=min(per, ok.size), replace=False).tolist())
        free = np.array(sorted(free))
        for _ in range(12):
            r = target - np.array([_b_at(a, k) for k in ks])
            if np.max(np.abs(r) / L) < 1e-14:
                break
            J = np.zeros((n_peaks, free.size))
            for row, k in enumerate(ks):
                idx = k - free
                ok = (idx >= 0) & (idx < n)
                J[row, ok] = 2.0 * a[idx[ok]]
            delta = J.T @ np.linalg.lstsq(J @ J.T, r, rcond=None)[0]
            a[free] = np.maximum(a[free] + delta, 0.0)
        b = autoconv64(a)
        if b.max() <= target[0] * (1.0 + 1e-15):
            break
        # another k rose above the planted level: plant on the new top values
        ks = np.argsort(b)[::-1][:n_peaks]
    if round_fp32:
        a = a.astype(np.float32).astype(np.float64)
    b = autoconv64(a)
    M = b.max()
    near = np.flatnonzero(b >= M * (1.0 - tol))
    return a, near


def uneven_lengths(P: int = 512, choices=UNEVEN_NS, seed: int = SEED) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.choice(np.asarray(choices, dtype=np.int32), size=P).astype(np.int32)
