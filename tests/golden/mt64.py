"""std::mt19937_64 (Matsumoto & Nishimura 64-bit Mersenne Twister, the C++11
standard's parameters) and libstdc++'s uniform_real_distribution<double>(0,1)
over it (generate_canonical with one 64-bit draw: x / 2^64, clamped below 1).

Used only to rebuild the start vectors that tools/probe.cpp --opt draws
(probe.cpp:241-267), so the oracle's local search can be compared with the
reference binary's output on the same starts."""
from __future__ import annotations

import math

_MASK = (1 << 64) - 1


class MT19937_64:
    n, m = 312, 156

    def __init__(self, seed: int):
        self.mt = [0] * self.n
        self.mt[0] = seed & _MASK
        for i in range(1, self.n):
            prev = self.mt[i - 1]
            self.mt[i] = (6364136223846793005 * (prev ^ (prev >> 62)) + i) & _MASK
        self.idx = self.n

    def _twist(self):
        mt = self.mt
        upper, lower = 0xFFFFFFFF80000000, 0x7FFFFFFF
        for i in range(self.n):
            x = (mt[i] & upper) | (mt[(i + 1) % self.n] & lower)
            xa = x >> 1
            if x & 1:
                xa ^= 0xB5026F5AA96619E9
            mt[i] = mt[(i + self.m) % self.n] ^ xa
        self.idx = 0

    def __call__(self) -> int:
        if self.idx >= self.n:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & _MASK

    def uniform01(self) -> float:
        r = float(self()) / 18446744073709551616.0
        return r if r < 1.0 else math.nextafter(1.0, 0.0)


def probe_opt_start(n: int, seed: int, r: int) -> list[float]:
    """Start vector of restart r of probe.cpp --opt (kinds 0, 1, 3; kind 2 uses
    normal_distribution and is not rebuilt here).  Not yet normalised."""
    rng = MT19937_64((seed * 1000003 + r) & _MASK)
    U = rng.uniform01
    kind = r % 4
    if kind == 0:
        nf = 2 + int(U() * 6)
        fr, ph = [0.0] * nf, [0.0] * nf
        for f in range(nf):
            fr[f] = 1 + U() * 10
            ph[f] = U() * 6.283185307
        base = 0.2 + 0.9 * U()
        amp = 0.3 + 1.5 * U()
        a = []
        for i in range(n):
            m = 0.0
            for f in range(nf):
                m += math.cos(6.283185307 * fr[f] * i / n + ph[f])
            a.append(max(0.0, base + amp * m / nf))
        return a
    if kind == 1:
        return [U() for _ in range(n)]
    if kind == 3:
        out = []
        for _ in range(n):
            out.append(U() if U() < 0.5 else 0.0)
        return out
    raise ValueError("kind 2 (normal_distribution) is not rebuilt")
