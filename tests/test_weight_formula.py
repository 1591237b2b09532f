"""The fp32 search weight's log2(1 + d) (csrc/common.hpp lp_weight_fast).

The kernel forms log2(1 + d), -1 < d <= 0, from u = fl32(1 + d) as
log2(u) + (d - (u - 1)) * log2(e): u - 1 is exact, so d - (u - 1) is the
rounding error of 1 + d, and the correction restores what log2(u) alone loses
for small |d| (the reason the reference uses log1p, src/kernels.cuh:35-44).
Here the same float32 steps run in numpy with a correctly rounded float32
log2 standing in for the hardware one, against float64 log1p; dropping the
correction or flipping its sign fails the bound by orders of magnitude.
"""
import numpy as np

LOG2E = np.float32(1.44269504088896341)


def _kernel_log2_1p(d: np.ndarray) -> np.ndarray:
    d = d.astype(np.float32)
    u = (np.float32(1.0) + d).astype(np.float32)
    l2u = np.log2(u.astype(np.float64)).astype(np.float32)  # correctly rounded fp32 log2
    e = (d - (u - np.float32(1.0))).astype(np.float32)
    return (l2u + e * LOG2E).astype(np.float32)


def _samples() -> np.ndarray:
    rng = np.random.default_rng(20260905)
    parts = [-(2.0 ** -rng.uniform(1, 40, 200000)), -rng.uniform(0, 1, 200000),
             -(1.0 - 2.0 ** -rng.uniform(1, 20, 50000))]
    d = np.concatenate(parts).astype(np.float32)
    return d[(d > -1.0) & (d < 0.0)]


def test_log2_1p_within_three_ulp_of_fp64():
    d = _samples()
    got = _kernel_log2_1p(d).astype(np.float64)
    want = np.log1p(d.astype(np.float64)) / np.log(2.0)
    rel = np.abs(got - want) / np.abs(want)
    assert rel.max() <= 3.0 * 2.0 ** -24, rel.max()


def test_correction_term_is_what_keeps_small_d_accurate():
    d = _samples()
    d = d[np.abs(d) < 1e-4]
    want = np.log1p(d.astype(np.float64)) / np.log(2.0)
    u = (np.float32(1.0) + d).astype(np.float32)
    plain = np.log2(u.astype(np.float64)).astype(np.float32).astype(np.float64)
    # without the correction the error at small |d| is far above float32 rounding
    assert (np.abs(plain - want) / np.abs(want)).max() > 1e-3
    flipped = (np.log2(u.astype(np.float64)).astype(np.float32)
               - (d - (u - np.float32(1.0))) * LOG2E).astype(np.float64)
    assert (np.abs(flipped - want) / np.abs(want)).max() > 1e-3


def test_exact_at_zero_and_flush_bound():
    assert _kernel_log2_1p(np.array([0.0], np.float32))[0] == 0.0
    # the weight 2^((p-1) log2(1+d)) flushes below 2^-126; the prefilter
    # b < 0.999 M exp(-126 ln2 / (p-1)) only ever drops weights that flush
    for pm1 in (1.0, 3.0, 63.0, 999.0, 9999.0):
        cut_ratio = 0.999 * np.exp(-87.33654475055310898657 / pm1)
        d = np.float32(cut_ratio - 1.0)
        if not d > -1.0:  # the kernel returns a zero weight before any logarithm
            continue
        y = np.float32(pm1) * _kernel_log2_1p(np.array([d], np.float32))[0]
        assert y < -126.0
