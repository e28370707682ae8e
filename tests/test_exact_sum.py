"""The exact-sum checker (tests/exact_sum.py) against independent exact
definitions: the permutation expansion in rational arithmetic, the exact
integer permanent, and the identities of the BBFG sum.  Also pins the
oracle's binary128 sweep to the correctly rounded value on small cases, so it
can stand in for the exact sum at sizes Python integers are too slow for."""

import itertools
from fractions import Fraction

import numpy as np
import pytest

import oracle
from conftest import random_complex
from exact_sum import _round, exact_bbfg, exact_repeated, exact_sum, expand


def frac_sum(A):
    tr, ti, k = exact_sum(A)
    s = Fraction(2) ** k
    return Fraction(tr) * s, Fraction(ti) * s


def perm_fractions(A):
    """sum over permutations of prod a_{i, sigma(i)} in exact rationals"""
    n = A.shape[0]
    re = [[Fraction(float(A[i, j].real)) for j in range(n)] for i in range(n)]
    im = [[Fraction(float(A[i, j].imag)) for j in range(n)] for i in range(n)]
    tr = ti = Fraction(0)
    for p in itertools.permutations(range(n)):
        pr, pi = Fraction(1), Fraction(0)
        for i in range(n):
            a, b = re[i][p[i]], im[i][p[i]]
            pr, pi = pr * a - pi * b, pr * b + pi * a
        tr += pr
        ti += pi
    return tr, ti


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_exact_sum_equals_permutation_expansion(n):
    rng = np.random.default_rng(500 + n)
    A = random_complex(rng, n, n) * 2.0 ** rng.integers(-30, 30, size=(n, n))
    assert frac_sum(A) == perm_fractions(A)


def test_rows_equal_cols_plus_one_is_the_permanent_without_row_0():
    rng = np.random.default_rng(7)
    for c in (1, 2, 3, 5):
        A = random_complex(rng, c + 1, c)
        assert frac_sum(A) == perm_fractions(A[1:, :])


def test_integer_matrices_match_the_integer_permanent():
    rng = np.random.default_rng(8)
    for n in (3, 7, 10):
        A = rng.integers(-3, 4, size=(n, n))
        tr, ti, k = exact_sum(A.astype(float))
        assert ti == 0
        assert Fraction(tr) * Fraction(2) ** k == oracle.int_permanent(A)


def test_planted_mistakes_are_caught():
    """The permutation check separates the sum from plausible slips: the walk
    without the (prod delta) sign, and a transposed operand on a rectangular
    matrix."""
    rng = np.random.default_rng(9)
    A = random_complex(rng, 4, 4)
    ref = perm_fractions(A)
    # (prod delta) dropped: sum_delta prod_j c_j(delta) over delta_0 = +1
    n = 4
    unsigned = [Fraction(0), Fraction(0)]
    for bits in range(1 << (n - 1)):
        d = [1] + [(-1) ** ((bits >> b) & 1) for b in range(n - 1)]
        pr, pi = Fraction(1), Fraction(0)
        for j in range(n):
            cr = sum(Fraction(d[i]) * Fraction(float(A[i, j].real)) for i in range(n))
            cim = sum(Fraction(d[i]) * Fraction(float(A[i, j].imag)) for i in range(n))
            pr, pi = pr * cr - pi * cim, pr * cim + pi * cr
        unsigned[0] += pr / 2 ** (n - 1)
        unsigned[1] += pi / 2 ** (n - 1)
    assert frac_sum(A) == ref
    assert tuple(unsigned) != ref
    B = random_complex(rng, 3, 2)
    assert frac_sum(B) != frac_sum(B.T)  # rows = cols + 1 is not transpose invariant


def test_round_half_even_and_powers():
    assert _round(3, -1) == 1.5
    assert _round((1 << 53) + 1, 0) == float(1 << 53)  # tie to even
    assert _round((1 << 53) + 3, 0) == float((1 << 53) + 4)
    assert _round(-5, -2) == -1.25
    assert _round(0, -100) == 0.0


def test_repeated_is_the_expanded_permanent():
    rng = np.random.default_rng(10)
    U = random_complex(rng, 4, 4)
    rows, cols = [2, 0, 1, 1], [1, 1, 2, 0]
    E = expand(U, rows, cols)
    assert E.shape == (4, 4)
    assert exact_repeated(U, rows, cols) == exact_bbfg(E)


@pytest.mark.parametrize("n", [3, 6, 9, 11])
def test_oracle_quad_rounds_to_the_exact_value(n):
    """The binary128 sweep of the oracle, rounded to double, is the correctly
    rounded sum on these cases (113-bit significands, no cancellation beyond
    a few bits)."""
    rng = np.random.default_rng(600 + n)
    A = random_complex(rng, n, n)
    assert oracle.bbfg(A, "quad") == exact_bbfg(A)
