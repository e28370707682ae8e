"""Exact BBFG/Glynn sums in integer arithmetic (test infrastructure).

Every double is an integer times a power of two, so after one common scaling
the column sums and products of

    perm = 2^-(R-1) sum_{delta in {+-1}^R, delta_0 = +1} (prod delta) prod_j sum_i delta_i a_ij

are Python integers and the sum is exact.  Rounding the exact value once to
the nearest double gives what an exact-arithmetic calculator returns -- the
reference's MPFR calculators (BBFGPermanentCalculator.cpp:100-107 with
ComplexInf; InfinitePrecisionComplex.h keeps every sum and product exact and
converts with mpfr_get_d, round to nearest).  Independent of the oracle: this
is the definition evaluated without rounding.
"""

from __future__ import annotations

from fractions import Fraction

import numpy as np


def _scaled(A):
    """Integer matrices (re, im) and exponent e with A = (re + i im) * 2^e."""
    vals = [float(x) for x in np.concatenate([A.real.ravel(), A.imag.ravel()])]
    den = max(x.as_integer_ratio()[1] for x in vals)  # a power of two
    e = -(den.bit_length() - 1)

    def to_int(x):
        num, d = float(x).as_integer_ratio()
        return num * (den // d)

    re = [[to_int(A[i, j].real) for j in range(A.shape[1])] for i in range(A.shape[0])]
    im = [[to_int(A[i, j].imag) for j in range(A.shape[1])] for i in range(A.shape[0])]
    return re, im, e


def _round(n: int, k: int) -> float:
    """n * 2^k rounded to the nearest double (ties to even)."""
    if n == 0:
        return 0.0
    if k >= 0:
        return float(n << k)
    return float(Fraction(n, 1 << (-k)))  # int / int true division rounds correctly


def exact_sum(A):
    """The BBFG sum of a rows x cols complex matrix (rows <= cols + 1), exact,
    as the integer pair (S_re, S_im) and the exponent k with sum = S * 2^k."""
    A = np.asarray(A, dtype=np.complex128)
    R, C = A.shape
    re, im, e = _scaled(A)
    cr = [sum(re[i][j] for i in range(R)) for j in range(C)]
    ci = [sum(im[i][j] for i in range(R)) for j in range(C)]
    delta = [1] * R
    sign = 1
    tr = ti = 0
    for t in range(1 << (R - 1)):
        if t:
            b = (t & -t).bit_length() - 1
            i = R - 1 - b
            delta[i] = -delta[i]
            sign = -sign
            f = 2 * delta[i]
            for j in range(C):
                cr[j] += f * re[i][j]
                ci[j] += f * im[i][j]
        pr, pi = 1, 0
        for j in range(C):
            pr, pi = pr * cr[j] - pi * ci[j], pr * ci[j] + pi * cr[j]
        tr += sign * pr
        ti += sign * pi
    return tr, ti, C * e - (R - 1)


def exact_bbfg(A) -> complex:
    """Correctly rounded BBFG permanent of A (rows <= cols + 1)."""
    tr, ti, k = exact_sum(A)
    return complex(_round(tr, k), _round(ti, k))


def expand(A, row_mult, col_mult):
    """Rows and columns repeated by their multiplicities (the matrix whose
    permanent the repeated-row calculators compute)."""
    A = np.asarray(A, dtype=np.complex128)
    ri = np.repeat(np.arange(A.shape[0]), np.asarray(row_mult, dtype=np.int64))
    ci = np.repeat(np.arange(A.shape[1]), np.asarray(col_mult, dtype=np.int64))
    return A[np.ix_(ri, ci)]


def exact_repeated(A, row_mult, col_mult) -> complex:
    """Correctly rounded permanent of the expanded matrix."""
    E = expand(A, row_mult, col_mult)
    if E.shape[0] == 0:
        return 1.0 + 0.0j
    return exact_bbfg(E)
