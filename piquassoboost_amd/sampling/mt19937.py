"""The 32-bit Mersenne Twister of C++'s std::mt19937 (Matsumoto and Nishimura,
ACM TOMACS 8 (1998); parameters of [rand.predef] in the C++ standard), and the
two draws the samplers make from it.

The reference's samplers share one std::mt19937 seeded with the user's seed
(samplingHelperFunctions.cpp:56-75): a uniform double rng() / rng.max() and an
index rng() % n.  Reproducing the generator bit for bit reproduces their
seeded samples.
"""

from __future__ import annotations

_N, _M = 624, 397
_MATRIX_A = 0x9908B0DF
_UPPER, _LOWER = 0x80000000, 0x7FFFFFFF
MAX = 0xFFFFFFFF


class MT19937:
    def __init__(self, seed: int = 5489):
        self.seed(seed)

    def seed(self, value: int) -> None:
        """std::mt19937::seed: x0 = value mod 2^32, x_i = 1812433253 (x_{i-1} ^ (x_{i-1} >> 30)) + i."""
        mt = [0] * _N
        mt[0] = int(value) & MAX
        for i in range(1, _N):
            mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & MAX
        self._mt = mt
        self._idx = _N

    def _twist(self) -> None:
        # in place and in order: entries below i are already the new state
        mt = self._mt
        for i in range(_N):
            y = (mt[i] & _UPPER) | (mt[(i + 1) % _N] & _LOWER)
            mt[i] = mt[(i + _M) % _N] ^ (y >> 1) ^ (_MATRIX_A if y & 1 else 0)
        self._idx = 0

    def __call__(self) -> int:
        if self._idx >= _N:
            self._twist()
        y = self._mt[self._idx]
        self._idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & MAX

    def random_double_0_1(self) -> float:
        """samplingHelperFunctions.cpp:62-64"""
        return float(self()) / float(MAX)

    def random_index(self, upper_bound_exclusive: int) -> int:
        """samplingHelperFunctions.cpp:66-72"""
        if upper_bound_exclusive == 0:
            return 0
        return self() % upper_bound_exclusive
