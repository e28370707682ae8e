"""Pins the CPU oracle (CPU only).

Every restated reference function is checked against things that do not share
its code: the reference's own known-answer vectors, a permutation-sum brute
force, Ryser's formula, exact integer permanents and closed forms.  Each check
is chosen so that a plausible slip (dropped term, wrong sign or index,
transposed operand, wrong anchor or binomial weight) fails it.
"""

import math

import numpy as np
import pytest

import oracle
from conftest import random_complex, rel_err

PLAIN_PRECS = ["double", "long double", "dd", "quad"]


def test_kat_repeated(kat_cases):
    """unittests/test_permanent.py:47-251 through the restated BBFG repeated
    (lib 7 and 8 arithmetic) and the restated Glynn repeated (lib 1 / 6)."""
    for c in kat_cases:
        for prec in ("double", "long double", "dd"):
            v = oracle.bbfg_repeated(c["matrix"], c["input"], c["output"], prec)
            assert rel_err(v, c["value"]) < 1e-12, (c["name"], prec, v)
        for prec in ("double", "long double"):
            v = oracle.glynn_repeated(c["matrix"], c["input"], c["output"], prec)
            assert rel_err(v, c["value"]) < 1e-12, (c["name"], prec, v)


def test_kat_plain_6x6(kat_cases):
    c = [k for k in kat_cases if k["name"] == "no_repetition_6x6"][0]
    for prec in PLAIN_PRECS:
        assert rel_err(oracle.bbfg(c["matrix"], prec), c["value"]) < 1e-12
        assert rel_err(oracle.glynn_recursive(c["matrix"], prec), c["value"]) < 1e-12
        assert rel_err(oracle.glynn_simple(c["matrix"], prec), c["value"]) < 1e-12
    assert rel_err(oracle.bruteforce(c["matrix"]), c["value"]) < 1e-12
    assert rel_err(oracle.ryser(c["matrix"]), c["value"]) < 1e-12


@pytest.mark.parametrize("n", range(1, 10))
def test_plain_vs_bruteforce(n):
    rng = np.random.default_rng(100 + n)
    A = random_complex(rng, n, n)
    ref = oracle.bruteforce(A, "quad")
    for prec in PLAIN_PRECS:
        for nj in (1, 3, 64):
            assert rel_err(oracle.bbfg(A, prec, njobs=nj), ref) < 1e-12, (prec, nj)
    assert rel_err(oracle.glynn_recursive(A, "long double"), ref) < 1e-12
    assert rel_err(oracle.glynn_recursive(A, "double"), ref) < 1e-12
    assert rel_err(oracle.glynn_simple(A, "long double"), ref) < 1e-12
    assert rel_err(oracle.ryser(A, "dd"), ref) < 1e-12


def test_ryser_vs_bbfg_larger():
    rng = np.random.default_rng(7)
    for n in (12, 15):
        A = random_complex(rng, n, n)
        assert rel_err(oracle.ryser(A, "dd"), oracle.bbfg(A, "dd")) < 1e-13


def test_closed_forms():
    for n in range(1, 13):
        J = np.ones((n, n))
        assert oracle.bbfg(J, "long double") == math.factorial(n)
        assert oracle.int_permanent(J.astype(np.int64)) == math.factorial(n)
        assert oracle.bbfg(np.eye(n), "double") == 1.0
        d = np.arange(1, n + 1) * (1 + 0.5j)
        assert rel_err(oracle.bbfg(np.diag(d), "dd"), np.prod(d)) < 1e-14
    # derangements !n = perm(J - I)
    for n in range(1, 14):
        D = np.ones((n, n), dtype=np.int64) - np.eye(n, dtype=np.int64)
        der = round(math.factorial(n) * sum((-1) ** k / math.factorial(k) for k in range(n + 1)))
        assert oracle.int_permanent(D) == der
        assert oracle.bbfg(D, "long double").real == der
    # 2x2: ad + bc
    A = np.array([[1 + 2j, 3 - 1j], [0.5j, -2 + 1j]])
    assert rel_err(oracle.bbfg(A, "dd"), A[0, 0] * A[1, 1] + A[0, 1] * A[1, 0]) < 1e-15


def test_invariances():
    rng = np.random.default_rng(11)
    n = 8
    A = random_complex(rng, n, n)
    p = oracle.bbfg(A, "dd")
    P = rng.permutation(n)
    Q = rng.permutation(n)
    assert rel_err(oracle.bbfg(A[P][:, Q], "dd"), p) < 1e-13
    assert rel_err(oracle.bbfg(A.T, "dd"), p) < 1e-13
    c = 0.7 - 0.3j
    assert rel_err(oracle.bbfg(c * A, "dd"), c ** n * p) < 1e-13
    Z = A.copy()
    Z[3] = 0
    # a zero row cancels the Glynn sum exactly in exact arithmetic; dd leaves
    # rounding far below the permanent's scale
    assert abs(oracle.bbfg(Z, "dd")) < 1e-25 * np.prod(np.abs(A).sum(axis=0))
    assert oracle.int_permanent(np.array([[1, 2, 0], [0, 0, 0], [3, 1, 1]])) == 0
    B = random_complex(rng, 4, 4)
    blk = np.zeros((n + 4, n + 4), dtype=complex)
    blk[:n, :n] = A
    blk[n:, n:] = B
    assert rel_err(oracle.bbfg(blk, "dd"), p * oracle.bbfg(B, "dd")) < 1e-13
    # multilinearity in a row: perm(row0 = x + y) = perm(row0 = x) + perm(row0 = y)
    x, y = A.copy(), A.copy()
    y[0] = random_complex(rng, 1, n)[0]
    s = A.copy()
    s[0] = x[0] + y[0]
    assert rel_err(oracle.bbfg(s, "dd"), oracle.bbfg(x, "dd") + oracle.bbfg(y, "dd")) < 1e-12


def test_edge_rules():
    # BBFGPermanentCalculator.hpp:90-95 / GlynnPermanentCalculator.hpp:50-53
    assert oracle.bbfg(np.zeros((0, 0))) == 1.0
    assert oracle.bbfg(np.ones((4, 2))) == 0.0
    assert oracle.glynn_recursive(np.ones((4, 2))) == 0.0
    # GlynnPermanentCalculatorSimple.hpp:67-70: non-square -> 0
    assert oracle.glynn_simple(np.ones((2, 3))) == 0.0


# multiplicity special cases of ctests/permanents_speed.cpp:208-231
SPECIAL = [
    ([1, 1, 1], None), ([2, 3, 0], None), ([2, 3, 0, 1, 0], None), ([1, 1, 1, 1, 1, 1, 1, 2], None),
    ([1, 1, 1, 1, 1, 1, 3, 3], None), ([2, 3, 0], [1, 1, 3]), ([1, 1, 3], [2, 3, 0]),
    ([2, 3, 0, 1, 0], [1, 1, 0, 2, 2]), ([1, 1, 0, 2, 2], [2, 3, 0, 1, 0]),
    ([0, 0, 1, 1, 0, 2, 2], [0, 1, 1, 3, 0, 1, 0]), ([1, 1, 1, 1, 1, 1, 1, 2], [2, 1, 1, 1, 1, 1, 1, 1]),
    ([1, 1, 1, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 2],
     [0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 4, 1, 1, 1, 0, 0, 0, 0, 2]),
]


@pytest.mark.parametrize("case", range(len(SPECIAL)))
def test_repeated_equals_expanded(case):
    inp, outp = SPECIAL[case]
    outp = inp if outp is None else outp
    m = len(inp)
    rng = np.random.default_rng(300 + case)
    U = random_complex(rng, m, m, scale=1 / np.sqrt(m))
    E = oracle.adapt_interferometer_glynn_multiplied(U, inp, outp)
    ref = oracle.bbfg(E, "dd") if E.size else 1.0
    # full m x m form with zero multiplicities
    for prec in ("double", "long double", "dd"):
        v = oracle.bbfg_repeated(U, inp, outp, prec)
        assert rel_err(v, ref) < 1e-11, (prec, v, ref)
    v = oracle.glynn_repeated(U, inp, outp, "long double")
    assert rel_err(v, ref) < 1e-11
    # compact form (adaptInterferometer) with the nonzero multiplicities
    Cm = oracle.adapt_interferometer(U, inp, outp)
    if Cm.size:
        ci = np.array([x for x in inp if x > 0])
        co = np.array([x for x in outp if x > 0])
        assert rel_err(oracle.bbfg_repeated(Cm, ci, co, "dd"), ref) < 1e-11


def test_repeated_sum_mismatch_message():
    with pytest.raises(oracle.OracleError, match="Number of input and output states should be equal"):
        oracle.bbfg_repeated(np.eye(2), [1, 1], [2, 1])
    with pytest.raises(oracle.OracleError, match="GlynnPermanentCalculatorRepeated::calculate"):
        oracle.glynn_repeated(np.eye(2), [1, 1], [2, 1])


def test_int_permanent_vs_bruteforce():
    rng = np.random.default_rng(5)
    for n in range(1, 9):
        A = rng.integers(-3, 4, size=(n, n))
        assert oracle.int_permanent(A) == round(oracle.bruteforce(A.astype(float), "quad").real)


def test_golden_haar_consistency(haar_blocks):
    """Long double and double-double restatements agree on the stored blocks
    (the golden values are the extended-precision oracle of the parity tests)."""
    for name in ("c1_n20", "c2a_n28", "c2a_n30"):
        ld = complex(*haar_blocks[name + "_ld"])
        h = haar_blocks[name + "_dd"]
        dd = complex(h[0] + h[1], h[2] + h[3])
        assert rel_err(ld, dd) < 1e-13, name
    # the n=20 Haar block is quoted against the reference GlynnPermanentCalculator (long double)
    gr = complex(*haar_blocks["c1_n20_glynn_ld"])
    assert rel_err(gr, complex(*haar_blocks["c1_n20_ld"])) < 1e-13


def test_golden_haar_unitary_block(haar_blocks):
    """The stored blocks are blocks of unitaries: rows have norm <= 1."""
    for name in ("c1_n20", "c2a_n28", "c2a_n30"):
        A = haar_blocks[name]
        assert np.all(np.linalg.norm(A, axis=1) <= 1 + 1e-12)


def test_adapters_match_product():
    from piquassoboost_amd.sampling import utilities as u

    rng = np.random.default_rng(9)
    U = random_complex(rng, 7, 7)
    inp = [2, 0, 1, 1, 0, 3, 0]
    outp = [0, 1, 2, 0, 3, 0, 1]
    np.testing.assert_array_equal(u.adapt_interferometer(U, inp, outp), oracle.adapt_interferometer(U, inp, outp))
    np.testing.assert_array_equal(
        u.adapt_interferometer_glynn_multiplied(U, inp, outp),
        oracle.adapt_interferometer_glynn_multiplied(U, inp, outp),
    )
