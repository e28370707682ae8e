"""GPU parity of the double-double sweep (PQB_PREC_INF, the engine's form of the
reference's MPFR calculators, lib 1 plain and lib 9 repeated).

The reference's use_inf path evaluates the sum in exact arithmetic and rounds
once (BBFGPermanentCalculator.cpp:100-107, InfinitePrecisionComplex.h), so
the expected value is the correctly rounded exact sum from tests/exact_sum.py
and every comparison below is bit for bit.  Sizes beyond the exact checker use
the oracle's binary128 sweep (pinned to the exact sum in test_exact_sum.py) or
the double-double golden fixtures."""

import numpy as np
import pytest

import oracle
from conftest import random_complex, record_parity, rel_err
from exact_sum import exact_bbfg, exact_repeated, expand
from piquassoboost_amd import _capi as C
from piquassoboost_amd.sampling import Boson_Sampling_Utilities as BSU

pytestmark = pytest.mark.gpu


def inf(A):
    return C.permanent(A, C.ALGO_BBFG, C.PREC_INF)


@pytest.mark.parametrize("n", list(range(1, 13)))
def test_square_equals_correctly_rounded_exact_sum(n):
    rng = np.random.default_rng(4000 + n)
    A = random_complex(rng, n, n)
    got, ref = inf(A), exact_bbfg(A)
    print(f"n={n} got={got!r} exact={ref!r}", flush=True)
    assert got == ref


@pytest.mark.parametrize("shape", [(1, 5), (2, 4), (3, 2), (5, 4), (4, 9), (7, 12), (9, 8), (13, 14)])
def test_rectangular_equals_exact_sum(shape):
    r, c = shape
    rng = np.random.default_rng(4100 + 16 * r + c)
    A = random_complex(rng, r, c)
    assert inf(A) == exact_bbfg(A)


@pytest.mark.parametrize("n", [6, 9, 12])
def test_wide_dynamic_range(n):
    """Entries spread over 2^-20 .. 2^20: the column sums need far more than
    53 bits."""
    rng = np.random.default_rng(4200 + n)
    A = random_complex(rng, n, n) * 2.0 ** rng.integers(-20, 21, size=(n, n))
    got, ref = inf(A), exact_bbfg(A)
    fp64 = C.permanent(A, C.ALGO_BBFG, C.PREC_FP64)
    print(f"n={n} fp64 rel_err={rel_err(fp64, ref):.2e}", flush=True)
    assert got == ref


@pytest.mark.parametrize("n", [8, 10, 12])
def test_cancellation_beyond_fp64(n):
    """Blocks [[1, 1], [1, -1]] (permanent 0) plus 2^-24 noise: the Glynn terms
    are O(1) and cancel to a tiny permanent; the FP64 sweep keeps a few digits,
    the double-double sweep the correctly rounded value."""
    rng = np.random.default_rng(4300 + n)
    H = np.array([[1.0, 1.0], [1.0, -1.0]])
    A = np.kron(np.eye(n // 2), H) + 2.0 ** -24 * random_complex(rng, n, n, scale=1.0)
    got, ref = inf(A), exact_bbfg(A)
    fp64 = C.permanent(A, C.ALGO_BBFG, C.PREC_FP64)
    print(f"n={n} |perm|={abs(ref):.3e} fp64 rel_err={rel_err(fp64, ref):.2e}", flush=True)
    assert got == ref


def test_integer_matrix_takes_the_exact_route():
    rng = np.random.default_rng(4400)
    A = rng.integers(-2, 3, size=(12, 12)).astype(float)
    assert inf(A) == complex(float(oracle.int_permanent(A.astype(np.int64))), 0.0)


REPEATED = [
    ((5, 5), [2, 1, 0, 1, 1], [1, 1, 2, 1, 0]),
    ((4, 6), [3, 2, 1, 0], [1, 1, 1, 1, 1, 1]),
    ((6, 6), [1, 1, 1, 1, 1, 1], [2, 0, 2, 0, 1, 1]),
    ((3, 3), [4, 3, 3], [3, 3, 4]),
    ((7, 7), [2, 0, 1, 3, 0, 2, 1], [1, 2, 1, 1, 2, 1, 1]),
]


@pytest.mark.parametrize("case", range(len(REPEATED)))
def test_repeated_equals_exact_expanded(case):
    (r, c), rows, cols = REPEATED[case]
    rng = np.random.default_rng(4500 + case)
    U = random_complex(rng, r, c)
    got = C.permanent_repeated(U, np.array(rows), np.array(cols), C.ALGO_BBFG, C.PREC_INF)
    ref = exact_repeated(U, rows, cols)
    print(f"case {case}: got={got!r} exact={ref!r}", flush=True)
    assert got == ref


def test_repeated_integer_matrix_is_exact():
    """0/1 and small-integer matrices through the repeated double-double sweep
    (the exact integer route is plain-only): equal to the integer permanent of
    the expanded matrix."""
    rng = np.random.default_rng(4600)
    for k in range(3):
        U = rng.integers(-1, 3, size=(6, 6)).astype(float)
        rows = rng.integers(0, 3, size=6)
        cols = rng.permutation(rows)
        if rows.sum() == 0:
            continue
        got = C.permanent_repeated(U, rows, cols, C.ALGO_BBFG, C.PREC_INF)
        ref = oracle.int_permanent(expand(U, rows, cols).real.astype(np.int64))
        assert got == complex(float(ref), 0.0), (k, got, ref)


@pytest.mark.parametrize("mult", [(12, 12, 12), (9, 11, 10, 8)])
def test_widest_slot_range(mult):
    """33..40 photons (the widest instantiation, up to 10 slots per lane) with few Gray
    steps (three or four rows of large multiplicity): equal to the oracle's binary128
    repeated sweep rounded to double, and the FP64 sweep within 1e-10 of it."""
    rng = np.random.default_rng(4650 + len(mult))
    k = len(mult)
    m = sum(mult)
    U = random_complex(rng, k, m)
    rows = np.array(mult)
    cols = np.ones(m, dtype=np.int64)
    got = C.permanent_repeated(U, rows, cols, C.ALGO_BBFG, C.PREC_INF)
    ref = oracle.bbfg_repeated(U, cols, rows, "quad")
    fp64 = C.permanent_repeated(U, rows, cols, C.ALGO_BBFG, C.PREC_DOUBLE)
    print(f"photons={m}: got={got!r} quad={ref!r} fp64 rel_err={rel_err(fp64, ref):.2e}", flush=True)
    assert got == ref
    assert rel_err(fp64, ref) < 1e-10


def test_batch_of_uneven_tasks():
    rng = np.random.default_rng(4700)
    tasks, refs = [], []
    for k in range(24):
        m = int(rng.integers(3, 8))
        U = random_complex(rng, m, m)
        rows = rng.integers(0, 3, size=m)
        if rows.sum() == 0:
            rows[0] = 1
        cols = rng.permutation(rows)
        tasks.append((U, rows, cols))
        refs.append(exact_repeated(U, rows, cols))
    got = C.permanent_batch(tasks, C.ALGO_BBFG, C.PREC_INF)
    assert got == refs


@pytest.mark.parametrize("n", [14, 17, 20])
def test_haar_blocks_match_the_binary128_oracle(n):
    from piquassoboost_amd.sampling.utilities import haar_unitary

    A = np.ascontiguousarray(haar_unitary(2 * n, 1001)[:n, :n])
    got = inf(A)
    ref = oracle.bbfg(A, "quad")
    print(f"n={n} got={got!r} quad={ref!r}", flush=True)
    assert got == ref


@pytest.mark.parametrize("name", ["c2a_n28", "c2a_n30"])
def test_haar_n28_n30_match_the_double_double_fixture(haar_blocks, name):
    A = haar_blocks[name]
    h = haar_blocks[name + "_dd"]
    ref = complex(h[0] + h[1], h[2] + h[3])
    got = inf(A)
    print(f"{name}: got={got!r} dd fixture={ref!r}", flush=True)
    record_parity(f"double-double sweep: {name[4:]} vs dd fixture (bit-exact)", rel_err(got, ref), 0)
    assert got == ref


def test_lib_codes_route_to_the_double_double_sweep(kat_cases):
    rng = np.random.default_rng(4800)
    A = random_complex(rng, 9, 9)
    assert BSU.GlynnPermanentInf(A).calculate() == exact_bbfg(A)
    for c in kat_cases:
        v = BSU.GlynnRepeatedInfPermanentCalculator(c["matrix"], c["input"], c["output"]).calculate()
        assert rel_err(v, c["value"]) < 1e-12, c["name"]
        assert v == exact_repeated(c["matrix"], c["output"], c["input"]), c["name"]
