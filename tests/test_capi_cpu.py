"""C-ABI boundary checks that need no GPU: the library loads, exports every
symbol include/pqb_perm.h declares, and applies the reference's argument and
edge-case rules before any device work (these paths never launch a kernel)."""

import ctypes
import re

import numpy as np
import pytest

from conftest import ROOT
from piquassoboost_amd import _capi as C
from piquassoboost_amd.sampling import Boson_Sampling_Utilities as BSU
from piquassoboost_amd.sampling.permanent_calculators import permanent_CPU_repeated_double

HEADER = ROOT / "include" / "pqb_perm.h"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pqb_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    L = C.lib()
    names = declared_symbols()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    # and the binding covers every declaration
    assert set(names) <= set(C.SIGNATURES), set(names) - set(C.SIGNATURES)


def test_abi_version_and_device_query():
    assert C.lib().pqb_abi_version() == 1
    assert C.device_count() >= 0  # never fails, 0 on a machine without a GPU


def test_error_text_empty_before_failure():
    assert isinstance(C.lib().pqb_last_error(), bytes)


def _raw_perm(A, algo=C.ALGO_BBFG, prec=C.PREC_DOUBLE):
    M = np.ascontiguousarray(A, dtype=np.complex128).reshape(A.shape[0], -1) if A.size else np.zeros((A.shape[0], A.shape[1]), np.complex128)
    out = (ctypes.c_double * 2)()
    rc = C.lib().pqb_perm(M.ctypes.data_as(C._dp), M.shape[0], M.shape[1], max(M.shape[1], 0), algo, prec, out)
    return rc, complex(out[0], out[1])


def test_edge_rules_without_device():
    # BBFGPermanentCalculator.hpp:90-95, GlynnPermanentCalculator.hpp:50-53
    assert _raw_perm(np.zeros((0, 0))) == (0, 1.0)
    assert _raw_perm(np.zeros((0, 3))) == (0, 1.0)
    assert _raw_perm(np.ones((5, 3))) == (0, 0.0)
    assert _raw_perm(np.ones((5, 3)), C.ALGO_GLYNN) == (0, 0.0)
    # GlynnPermanentCalculatorSimple.hpp:65-70: non-square -> 0
    assert _raw_perm(np.ones((2, 3)), C.ALGO_SIMPLE) == (0, 0.0)


def test_argument_errors():
    out = (ctypes.c_double * 2)()
    A = np.ones((2, 2), dtype=np.complex128)
    L = C.lib()
    assert L.pqb_perm(A.ctypes.data_as(C._dp), -1, 2, 2, 0, 0, out) == C.PQB_ERR_ARG
    assert L.pqb_perm(A.ctypes.data_as(C._dp), 2, 2, 1, 0, 0, out) == C.PQB_ERR_ARG  # ld < cols
    assert L.pqb_perm(A.ctypes.data_as(C._dp), 2, 2, 2, 9, 0, out) == C.PQB_ERR_ARG  # algo
    assert L.pqb_perm(A.ctypes.data_as(C._dp), 2, 2, 2, 0, 9, out) == C.PQB_ERR_ARG  # precision
    assert L.pqb_perm(A.ctypes.data_as(C._dp), 2, 2, 2, 0, C.PREC_INF + 1, out) == C.PQB_ERR_ARG
    assert L.pqb_perm(None, 2, 2, 2, 0, 0, out) == C.PQB_ERR_ARG
    big = np.ones((41, 41), dtype=np.complex128)
    assert L.pqb_perm(big.ctypes.data_as(C._dp), 41, 41, 41, 0, 0, out) == C.PQB_ERR_SIZE
    assert "40 columns" in L.pqb_last_error().decode()


def test_inf_precision_edge_rules_need_no_device():
    """PQB_PREC_INF follows the same shape rules; decided cases never touch the device."""
    assert _raw_perm(np.zeros((0, 0)), C.ALGO_BBFG, C.PREC_INF) == (0, 1.0)
    assert _raw_perm(np.ones((5, 3)), C.ALGO_BBFG, C.PREC_INF) == (0, 0.0)
    z = np.zeros(3, dtype=np.int64)
    assert C.permanent_batch([(np.eye(3), z, z)], C.ALGO_BBFG, C.PREC_INF) == [1.0 + 0.0j]
    with pytest.raises(Exception, match="Number of input and output states should be equal"):
        C.permanent_repeated(np.eye(2), np.array([1, 1]), np.array([2, 1]), C.ALGO_BBFG, C.PREC_INF)
    from piquassoboost_amd.sampling import Boson_Sampling_Utilities_wrapper as W

    assert W.PLAIN_LIBS[W.GlynnInf] == (C.ALGO_BBFG, C.PREC_INF)
    assert W.REPEATED_LIBS[W.GlynnRepInf] == (C.ALGO_BBFG, C.PREC_INF)


def test_exact_mode_rejects_non_integer():
    with pytest.raises(C.EngineError) as ei:
        C.permanent(np.array([[0.5, 1], [1, 1]]), C.ALGO_BBFG, C.PREC_EXACT)
    assert ei.value.code == C.PQB_ERR_EXACT


def test_repeated_state_errors_match_reference():
    # BBFGPermanentCalculatorRepeated.hpp:119-124
    with pytest.raises(Exception, match=r"^BBFGPermanentCalculatorRepeated_Tasks::calculate:  Number of input and "
                                        r"output states should be equal$"):
        BSU.BBFGRepeatedPermanentCalculatorDouble(np.eye(2), np.array([1, 1]), np.array([2, 1])).calculate()
    # GlynnPermanentCalculatorRepeated.hpp:55-60
    with pytest.raises(Exception, match=r"^GlynnPermanentCalculatorRepeated::calculate:  Number of input and "
                                        r"output states should be equal$"):
        BSU.GlynnRepeatedPermanentCalculatorDouble(np.eye(2), np.array([1, 1]), np.array([2, 1])).calculate()
    with pytest.raises(Exception, match="Number of input and output states should be equal"):
        permanent_CPU_repeated_double(np.eye(2), np.array([1, 1]), np.array([0, 1]))


def test_repeated_edge_rules_without_device():
    # zero photons -> 1 (unittests/test_permanent.py:55-78)
    v = BSU.BBFGRepeatedPermanentCalculatorDouble(np.eye(3), np.zeros(3, np.int64), np.zeros(3, np.int64)).calculate()
    assert v == 1.0
    v = BSU.GlynnRepeatedPermanentCalculator(np.eye(3), np.zeros(3, np.int64), np.zeros(3, np.int64)).calculate()
    assert v == 1.0


def test_wrong_lib_codes():
    for lib in (2, 3, 4, 5, 9, 12, -1):
        with pytest.raises(Exception, match="Wrong value set for permanent library."):
            BSU.GlynnPermanentCalculator_wrapper(np.eye(2), lib)
    for lib in (2, 3, 4, 5, 10, 11, 12):
        with pytest.raises(Exception, match="Wrong value set for permanent library."):
            BSU.ChinHuhPermanentCalculator_wrapper(lib, np.eye(2), np.array([1, 1]), np.array([1, 1]))
    for cls in (BSU.GlynnPermanentSingleDFE, BSU.GlynnPermanentDualDFE):
        with pytest.raises(Exception, match="Wrong value set for permanent library."):
            cls(np.eye(2))


def test_state_type_checks():
    with pytest.raises(Exception, match="The input state should be a numpy array"):
        BSU.BBFGRepeatedPermanentCalculatorDouble(np.eye(2), [1, 1], np.array([1, 1]))


def test_state_length_mismatch_raises():
    with pytest.raises(Exception, match="state lengths must match"):
        BSU.BBFGRepeatedPermanentCalculatorDouble(np.eye(3), np.array([1, 1]), np.array([1, 1])).calculate()


def test_plan_is_host_only():
    p = C.plain_plan(30, 30, waves_resident=3072)
    assert p["K"] + p["B"] + p["H"] == 29
    # exactly the resident waves, each a run of ~2.7 groups of 2^10 steps
    assert p["waves"] == 3072 and p["blocks"] * p["block"] == 64 * p["waves"]
    assert 2 ** p["H"] == 8192 and p["K"] == 10
    small = C.plain_plan(3, 3, waves_resident=3072)
    assert small["K"] + small["B"] + small["H"] == 2 and small["waves"] == 2  # one step pair each
    for n in range(1, 41):
        for wr in (1, 7, 1024, 3072, 4096):
            q = C.plain_plan(n, n, waves_resident=wr)
            assert q["K"] + q["B"] + q["H"] == n - 1
            assert 0 <= q["K"] <= 14 and 0 <= q["B"] <= 6
            assert 1 <= q["waves"] <= max(wr, 1)
            assert q["waves"] <= 2 ** (q["H"] + q["K"]) // (2 if q["K"] else 1)
            assert q["blocks"] * q["block"] == 64 * q["waves"]


def test_exact_route_bounds():
    """Which integer matrices the exact 128-bit sweep accepts (host-only check)."""
    import math

    # J_26: prod bound 26^26 is too weak, Bregman's bound (26!) proves 2^25 * 26! < 2^127
    ok, lb = C.exact_bound(np.ones((26, 26)))
    assert ok and abs(lb - (math.log2(math.factorial(26)) + 25)) < 1e-9
    # derangement matrices up to n = 26
    for n in (10, 20, 25, 26):
        ok, _ = C.exact_bound(np.ones((n, n)) - np.eye(n))
        assert ok, n
    # a row of zeros: permanent provably 0
    Z = np.ones((5, 5))
    Z[2] = 0
    assert C.exact_bound(Z)[0]
    # large entries overflow the bound
    assert not C.exact_bound(np.full((20, 20), 1000.0))[0]
    # non-integer or complex input is not eligible
    assert not C.exact_bound(np.array([[0.5, 1.0], [1.0, 1.0]]))[0]
    assert not C.exact_bound(np.array([[1j, 1.0], [1.0, 1.0]]))[0]


def _true_log2_sum(A):
    """log2 |sum_delta (prod delta) prod_j (sum_i delta_i a_ij)| in Python integers."""
    import math

    from exact_sum import exact_sum

    tr, ti, k = exact_sum(np.asarray(A, dtype=np.complex128))
    assert ti == 0
    return (math.log2(abs(tr)) + k + (A.shape[0] - 1)) if tr else float("-inf")


@pytest.mark.parametrize("shape", [(1, 40), (2, 30), (3, 2), (4, 3), (3, 7), (6, 6), (5, 4)])
def test_exact_route_bound_holds_for_every_shape(shape):
    """The exact route's bound must dominate the sweep's sum for every shape it
    accepts: for rows != cols the sum is not a permanent of the matrix, so only
    the column abs-sum bound applies (rows = cols + 1 gives 2^(R-1) perm(A
    without row 0); rows < cols counts maps with odd row multiplicities)."""
    r, c = shape
    rng = np.random.default_rng(7700 + 10 * r + c)
    for A in (np.full(shape, 1000.0), np.full(shape, 50.0), rng.integers(-3, 4, size=shape).astype(float)):
        A = A.copy()
        if r == c + 1:
            A[0] = 0.0  # a zero anchor row does not make the sum zero
        ok, lb = C.exact_bound(A)
        true = _true_log2_sum(A)
        assert true <= lb + 1e-9, (shape, A[0, 0], true, lb)
        if ok:
            assert lb < 126.5


def test_exact_route_rejects_wide_rectangles():
    """1 x 40 of 1000s (sum ~2^398), 2 x 30 of 50s (~2^199): too large for 128 bits;
    rows = cols + 1 with a zero row 0: not provably zero (perm of the other rows)."""
    assert not C.exact_bound(np.full((1, 40), 1000.0))[0]
    assert not C.exact_bound(np.full((2, 30), 50.0))[0]
    Z = np.array([[0.0, 0.0], [1000.0, 7.0], [3.0, 1000.0]])
    ok, lb = C.exact_bound(Z)
    assert ok and np.isfinite(lb)
    assert abs(_true_log2_sum(Z) - (np.log2(1000.0 * 1000.0 + 21.0) + 2)) < 1e-9


def _wrap128(v):
    v &= (1 << 128) - 1
    return v - (1 << 128) if v >> 127 else v


def _grouped_product(c, G):
    """The exact sweep's product of column sums: G sums at a time multiplied in
    int64, each partial product folded into a 128-bit word by one 64 x 64 -> 128
    product of the low word, one 64-bit product of the high word and a negation
    for a negative factor (perm_plain.hip, mul_u128_s64)."""
    M64, M128 = (1 << 64) - 1, (1 << 128) - 1
    pr = None
    for j0 in range(0, len(c), G):
        q = 1
        for x in c[j0:j0 + G]:
            q *= x
        assert -(1 << 63) <= q < (1 << 63), "group product leaves int64"
        if pr is None:
            pr = q & M128
            continue
        m = (-q) & M64 if q < 0 else q
        lo = (pr & M64) * m
        hi = (((pr >> 64) * m) + (lo >> 64)) & M64
        r = ((hi << 64) | (lo & M64)) & M128
        pr = (-r) & M128 if q < 0 else r
    return _wrap128(pr)


@pytest.mark.parametrize("G,cmax", [(13, 28), (6, 1400), (2, (1 << 30) - 1)])
def test_grouped_product_matches_128bit_product(G, cmax):
    """At each threshold the host picks (largest column abs sum 28 -> G = 13,
    1400 -> 6, below 2^30 -> 2), every group's product fits int64 and the folded
    result equals the product modulo 2^128; column sums at +-cmax exactly included."""
    rng = np.random.default_rng(7800 + G)
    assert cmax ** G < (1 << 63)
    for n in (1, G - 1, G, G + 1, 2 * G + 3, 26, 40):
        if n < 1:
            continue
        for trial in range(400):
            if trial % 4 == 0:
                c = [int(cmax) * int(s) for s in rng.choice([-1, 1], size=n)]
            else:
                c = [int(x) for x in rng.integers(-cmax, cmax + 1, size=n)]
            exact = 1
            for x in c:
                exact *= x
            assert _grouped_product(c, G) == _wrap128(exact), (G, n, c)


def test_grouped_product_thresholds_are_tight():
    """The next column abs sum above each threshold would overflow int64 for the
    larger group (29^13 > 2^63) -- the reason the host switches group size there."""
    assert 28 ** 13 < (1 << 63) <= 29 ** 13
    assert 1400 ** 6 < (1 << 63)
    assert ((1 << 30) - 1) ** 2 < (1 << 63)


def test_repeated_plan_host_only():
    """Task planning needs no device: anchor, orientation, digit split."""
    U = np.eye(6, dtype=complex) + 0.1
    # rows (output) multiplicities 1,3,0,2,1,2 ; columns all ones except a zero
    outp = np.array([1, 3, 0, 2, 1, 2])
    inp = np.array([2, 1, 1, 1, 2, 2])
    p = C.repeated_plan(U, outp, inp, 1 << 20)
    assert p["slots"] == 9
    # whole walk inside one chunk
    assert p["high"] == 0 and p["chunks"] == 1
    # S is minimised over the orientation: rows give 1*4*3*2*3 = 72, cols 1*2*2*2*3*3 = 72 -> rows kept
    assert p["walk"] * p["chunks"] == 72
    p1 = C.repeated_plan(U, outp, inp, 1)
    assert p1["low"] == 0 and p1["walk"] == 1 and p1["chunks"] == 72
    # zero photons: no sweep
    z = C.repeated_plan(U, np.zeros(6, np.int64), np.zeros(6, np.int64), 16)
    assert z["chunks"] == 0
    # transposed orientation when the columns repeat more
    # 3 x 2 matrix, rows 1,1,1 (S = 1*2*2 = 4), columns 2,1 (S = 1*3 = 3): transposed
    p2 = C.repeated_plan(np.ones((3, 2)), np.array([1, 1, 1]), np.array([2, 1]), 1 << 20)
    assert p2["walk"] * p2["chunks"] == 3 and p2["slots"] == 3 and p2["rows"] == 2


GRAD_INSTANTIATIONS = {(1, 0), (2, 0), (2, 1), (2, 2), (2, 3), (3, 3), (4, 3), (5, 3), (6, 3), (7, 3), (8, 3),
                       (9, 2), (10, 2), (5, 4), (7, 4)}


def test_grad_plan_host_only():
    """Adjoint sweep plan: the n-1 sign bits split into register (K), lane (L)
    and high (H) bits, one compiled kernel per (columns per lane, K), whole
    blocks of 4 waves per matrix, LDS within the 160 KiB of a CU."""
    for n in range(1, 41):
        for count, wr in ((1, 1024), (8, 2048), (256, 2048), (5000, 512)):
            g = C.grad_plan(count, n, wr)
            assert g["K"] + g["L"] + g["H"] == n - 1
            assert g["M"] == (n + 3) // 4 and g["L"] == min(4, n - 1)
            assert (g["M"], g["K"]) in GRAD_INSTANTIATIONS, (n, g)
            assert g["wpt"] % 4 == 0 and 4 * g["blocks"] == count * g["wpt"]
            assert g["wpt"] <= max(4, 2 ** g["H"] + 3)
            assert g["lds"] == 16 * (n * 4 * g["M"] + 4 * (16 + g["H"]) * 4 * g["M"]) <= 160 * 1024
    # the gradient workload: 256 x n=20 and 8 x n=26 at 2048 resident waves
    assert C.grad_plan(256, 20, 2048)["wpt"] == 32
    assert C.grad_plan(8, 26, 2048)["wpt"] == 1024


def test_grad_argument_errors_without_device():
    p, G = C.permanent_grad(np.zeros((0, 0)))
    assert p == 1.0 and G.shape == (0, 0)
    with pytest.raises(C.EngineError):
        C.permanent_grad(np.ones((3, 4)))
    with pytest.raises(C.EngineError, match="order above"):
        C.permanent_grad(np.ones((41, 41)))
    with pytest.raises(C.EngineError):
        C.grad_plan(0, 5, 1024)


def test_minors_edge_and_shape_errors_without_device():
    assert C.permanent_minors(np.zeros((0, 1))).tolist() == [1.0]
    with pytest.raises(C.EngineError, match="rows x \\(rows \\+ 1\\)"):
        C.permanent_minors(np.ones((3, 3)))
    with pytest.raises(C.EngineError):
        C.permanent_minors(np.ones((40, 41)))  # 41 columns


def test_batch_states_rules_without_device():
    """pqb_perm_batch_states: shapes, the reference's message for unequal photon counts
    (with the pair index), and pairs decided by the edge rules need no device."""
    U = np.eye(4)
    z = np.zeros((3, 4), dtype=np.int64)
    assert C.permanent_batch_states(U, z, z).tolist() == [1.0, 1.0, 1.0]
    with pytest.raises(Exception, match="same count"):
        C.permanent_batch_states(U, z, z[:2])
    with pytest.raises(Exception, match="state lengths"):
        C.permanent_batch_states(U, z[:, :3], z[:, :3])
    bad_in = np.array([[0, 0, 0, 0], [1, 0, 0, 0]])
    bad_out = np.array([[0, 0, 0, 0], [2, 0, 0, 0]])
    with pytest.raises(Exception, match=r"^task 1: BBFGPermanentCalculatorRepeated_Tasks::calculate:  Number of "
                                        r"input and output states should be equal$"):
        C.permanent_batch_states(U, bad_in, bad_out)
    neg = np.array([[-1, 1, 0, 0]])
    with pytest.raises(C.EngineError):
        C.permanent_batch_states(U, neg, neg)
    out = (ctypes.c_double * 2)()
    L = C.lib()
    assert L.pqb_perm_batch_states(None, 4, 4, 4, None, None, 1, 0, 0, out) == C.PQB_ERR_ARG
    assert L.pqb_perm_batch_states(None, 4, 4, 3, None, None, 0, 0, 0, out) == C.PQB_OK  # nothing to do
    assert L.pqb_perm_batch_states(None, -1, 4, 4, None, None, 1, 0, 0, out) == C.PQB_ERR_ARG


def test_batched_front_end_errors_without_device():
    from piquassoboost_amd.sampling.BatchedPermanentCalculator import BatchednPermanentCalculator

    with pytest.raises(Exception, match="interferometer matrix is zero"):
        BatchednPermanentCalculator().calculate()
    b = BatchednPermanentCalculator(np.eye(3))
    assert b.calculate().shape == (1, 0)  # nothing accumulated
    b.reserve_space(1)
    b.add(np.array([1, 1, 0]), np.array([0, 1, 1]), 0)
    with pytest.raises(Exception, match="repeated DFE not implemented yet"):
        b.calculate(lib=2)
    with pytest.raises(Exception, match="Wrong value set for permanent library"):
        b.calculate(lib=5)


def test_batch_reports_the_first_failing_task():
    """Validation of a large batch runs on host threads; the error reported is
    still the one of the lowest task index (no device needed: the state sums
    are checked before the engine starts)."""
    A = np.eye(3, dtype=complex)
    good = (A, np.array([1, 1, 1]), np.array([1, 1, 1]))
    bad = (A, np.array([1, 1, 1]), np.array([1, 1, 2]))
    tasks = [good] * 2000
    for i in (1999, 1500, 700):
        tasks[i] = bad
    with pytest.raises(C.EngineError, match=r"^task 700: .*Number of input and output states should be equal"):
        C.permanent_batch(tasks)


def test_batch_validation_from_concurrent_host_threads():
    """The host worker pool serves one parallel phase at a time; a phase that
    finds it busy starts threads of its own.  Calls from 8 Python threads (ctypes
    releases the GIL), 25 each, every one must report its own first bad task."""
    from concurrent.futures import ThreadPoolExecutor

    A = np.eye(3, dtype=complex)
    good = (A, np.array([1, 1, 1]), np.array([1, 1, 1]))
    bad = (A, np.array([1, 1, 1]), np.array([1, 1, 2]))

    def work(k):
        first = 100 + 37 * k
        tasks = [good] * 3000
        for i in (first, first + 900, 2999):
            tasks[i] = bad
        for _ in range(25):
            try:
                C.permanent_batch(tasks)
            except C.EngineError as e:
                if not str(e).startswith(f"task {first}: "):
                    return str(e)
            else:
                return "no error"
        return None

    with ThreadPoolExecutor(8) as ex:
        assert list(ex.map(work, range(8))) == [None] * 8


def test_row_stride_binding_host_only():
    """Row-strided complex128 arrays go to the C ABI in place with their row
    stride as ld; anything else is copied to a packed array (ld = cols)."""
    P = np.arange(40 * 40, dtype=np.float64).reshape(40, 40) * (1 + 1j)
    M, ld = C.as_rows(P[3:23, 5:25])
    assert ld == 40 and M.ctypes.data == P.ctypes.data + 16 * (3 * 40 + 5)
    M, ld = C.as_rows(P[::2, :10])  # every other row: stride 80
    assert ld == 80
    M, ld = C.as_rows(P[:, ::2])  # column step: copied
    assert ld == 20 and M.flags.c_contiguous
    M, ld = C.as_rows(P.T)  # transposed: copied
    assert ld == 40 and M.flags.c_contiguous and np.array_equal(M, P.T)
    M, ld = C.as_rows(P[::-1])  # negative row stride: copied
    assert ld == 40 and np.array_equal(M, P[::-1])
    M, ld = C.as_rows(P.real[:4, :4])  # float64: converted
    assert ld == 4 and M.dtype == np.complex128
    S = np.zeros((3, 12, 16), dtype=np.complex128)
    B, ld, stride = C._stack_strides(S[:, :10, :11])
    assert (ld, stride) == (16, 192) and B.ctypes.data == S.ctypes.data
    B, ld, stride = C._stack_strides(S[:, :, ::2])
    assert (ld, stride) == (8, 96) and B.flags.c_contiguous


def _read_i64(addr, n):
    import ctypes

    return np.ctypeslib.as_array(ctypes.cast(int(addr), ctypes.POINTER(ctypes.c_int64)), shape=(n,)).copy()


def test_task_array_marshalling():
    """The pqb_task array built for pqb_perm_batch: every field of every task, for
    matrices used in place (C-contiguous complex128, a row slice with ld > cols) or
    converted (float64, a column slice), and states given as int32 arrays or lists."""
    rng = np.random.default_rng(7)
    big = rng.standard_normal((6, 9)) + 1j * rng.standard_normal((6, 9))
    tasks = [
        (np.ascontiguousarray(big[:3, :4]), np.array([1, 2, 1]), np.array([1, 1, 1, 1])),
        (big[1:4, 2:6], np.array([2, 1, 1], dtype=np.int32), [1, 1, 1, 1]),  # rows of `big`: ld = 9
        (rng.standard_normal((2, 2)), [1, 1], np.array([1, 1], dtype=np.int64)),  # float64: copied
        (big[:, ::2][:2, :3], np.array([2, 1]), np.array([1, 1, 1])),  # column stride 2: copied
    ]
    arr, keep = C._task_array(tasks)
    assert arr.shape == (4, 6)
    want_ld = [4, 9, 2, 3]
    for k, (A, r, c) in enumerate(tasks):
        rows, cols = np.shape(A)
        assert arr[k, 1] == rows and arr[k, 2] == cols and arr[k, 3] == want_ld[k]
        assert (_read_i64(arr[k, 4], rows) == np.asarray(r)).all()
        assert (_read_i64(arr[k, 5], cols) == np.asarray(c)).all()
    # task 1 is passed in place: its address is that of big[1, 2]
    assert arr[1, 0] == big.__array_interface__["data"][0] + 16 * (1 * 9 + 2)
    del keep


def test_task_array_errors():
    ok = (np.eye(2, dtype=np.complex128), np.array([1, 1]), np.array([1, 1]))
    with pytest.raises(Exception, match="task 1: state lengths must match"):
        C._task_array([ok, (np.eye(3, dtype=np.complex128), np.array([1, 1]), np.array([1, 1, 1]))])
    with pytest.raises(Exception, match="numpy2PicState_int64: Wrong matrix dimension was given"):
        C._task_array([ok, (np.eye(2, dtype=np.complex128), np.ones((2, 1), dtype=np.int64), np.array([1, 1]))])
    col = np.ones((2, 1), dtype=np.int64)  # every state 2-D: still refused
    with pytest.raises(Exception, match="numpy2PicState_int64: Wrong matrix dimension was given"):
        C._task_array([(np.eye(2, dtype=np.complex128), col, col)] * 2)
