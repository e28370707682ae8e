"""GPU parity of the plain permanent (FP64 sweep and exact integer sweep)
against the CPU oracle and the golden fixtures.  Every call goes through the
C ABI of libpqb_perm.so."""

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, random_complex, record_parity, rel_err
from exact_sum import exact_bbfg
from piquassoboost_amd import _capi as C
from piquassoboost_amd.sampling import Boson_Sampling_Utilities as BSU

pytestmark = pytest.mark.gpu

# The FP64 sweep is compared with the double-double oracle.  Tolerances: 1e-12
# relative error for the n=20 Haar block and 1e-10 for the n=28/30 blocks; the
# small random cases use 1e-12 as well.
TOL_SMALL = 1e-12


def test_device_visible():
    assert C.device_count() >= 1


@pytest.mark.parametrize("n", list(range(1, 17)) + [18, 21, 24])
def test_square_random_vs_oracle(n):
    rng = np.random.default_rng(1000 + n)
    A = random_complex(rng, n, n)
    ref = oracle.bbfg(A, "dd")
    got = C.permanent(A, C.ALGO_BBFG, C.PREC_FP64)
    print(f"n={n} rel_err={rel_err(got, ref):.3e}", flush=True)
    record_parity("plain FP64, random n=1..24 vs oracle dd", rel_err(got, ref), TOL_SMALL)
    assert rel_err(got, ref) < TOL_SMALL


@pytest.mark.parametrize("shape", [(2, 1), (3, 2), (5, 4), (9, 8), (2, 5), (4, 9), (7, 12), (1, 6), (13, 14)])
def test_rectangular_vs_oracle(shape):
    """rows = cols+1 and rows < cols follow the same sum (BBFGPermanentCalculator.hpp:87-235)."""
    r, c = shape
    rng = np.random.default_rng(r * 100 + c)
    A = random_complex(rng, r, c)
    ref = oracle.bbfg(A, "dd")
    got = C.permanent(A, C.ALGO_BBFG, C.PREC_FP64)
    # rows = cols + 1: the sum is perm(A without row 0) and its terms cancel more
    # than in the square case; compare on the scale of the terms
    scale = max(abs(ref), np.prod(np.abs(A).sum(axis=0)) * 1e-3)
    assert abs(got - ref) <= TOL_SMALL * scale, (got, ref)


def test_strided_input_matches_contiguous():
    rng = np.random.default_rng(3)
    big = random_complex(rng, 12, 20)
    A = big[:, 3:15]  # non-contiguous view, converted by the binding
    assert C.permanent(A) == C.permanent(np.ascontiguousarray(A))


def test_kat_plain_all_libs(kat_cases):
    c = [k for k in kat_cases if k["name"] == "no_repetition_6x6"][0]
    for cls in (BSU.GlynnPermanent, BSU.GlynnPermanentInf, BSU.GlynnPermanentDouble, BSU.BBFGPermanentDouble,
                BSU.BBFGPermanentLongDouble, BSU.GlynnPermanentSimpleDouble, BSU.GlynnPermanentSimpleLongDouble):
        v = cls(c["matrix"]).calculate()
        assert isinstance(v, complex)
        assert rel_err(v, c["value"]) < 1e-12, cls.__name__


def test_list_batched_plain():
    rng = np.random.default_rng(8)
    mats = [random_complex(rng, n, n) for n in (3, 5, 8, 2)]
    res = BSU.BBFGPermanentDouble(mats).calculate()
    assert isinstance(res, list) and len(res) == 4
    for m, v in zip(mats, res):
        assert rel_err(v, oracle.bbfg(m, "dd")) < TOL_SMALL


def _list_mix(seed):
    """A list of uneven plain matrices: square n = 1..24 (several of each shape),
    rectangular rows = cols + 1 and rows < cols, decided shapes (empty, rows >=
    cols + 2), integer matrices (exact route) and row slices of a wider array."""
    rng = np.random.default_rng(seed)
    mats = []
    for n in list(range(1, 17)) * 3 + [18, 20, 20, 22, 24]:
        mats.append(random_complex(rng, n, n))
    for r, c in [(5, 4), (9, 8), (2, 5), (4, 9), (7, 12), (5, 4), (0, 3), (6, 3), (1, 1)]:
        mats.append(random_complex(rng, r, c) if r and c else np.zeros((r, c), np.complex128))
    for n in (6, 10, 12):
        mats.append(rng.integers(-2, 3, size=(n, n)).astype(np.complex128))
    wide = random_complex(rng, 14, 30)
    mats += [wide[:, 2:16], wide[1:13, 5:17]]
    order = rng.permutation(len(mats))
    return [mats[i] for i in order]


@pytest.mark.parametrize("lib", [0, 6, 7, 8, 10, 11])
def test_list_surface_equals_per_call_bitwise(lib):
    """GlynnPermanentCalculator_Wrapper.hpp:204-293: a list of matrices in one call
    (pqb_perm_list, one launch per shape on the FP64 route) gives, for every matrix,
    the value of the per-matrix call bit for bit, and the oracle's value within the
    usual bar."""
    from piquassoboost_amd.sampling.Boson_Sampling_Utilities_wrapper import PLAIN_LIBS

    mats = _list_mix(8100 + lib)
    algo, prec = PLAIN_LIBS[lib]
    got = BSU.GlynnPermanentCalculator_wrapper(mats, lib).calculate()
    assert isinstance(got, list) and len(got) == len(mats)
    worst = 0.0
    for k, (m, v) in enumerate(zip(mats, got)):
        one = C.permanent(m, algo, prec)
        assert v == one, (k, m.shape, v, one)
        r, c = m.shape
        if r == 0 or c == 0 or (algo == C.ALGO_SIMPLE and r != c) or r >= c + 2:
            continue
        ref = oracle.bbfg(m, "dd")
        scale = abs(ref) if r == c else max(abs(ref), np.prod(np.abs(m).sum(axis=0)) * 1e-3)
        err = abs(v - ref) / scale
        worst = max(worst, err)
        assert err < TOL_SMALL, (k, m.shape, v, ref)
    print(f"lib {lib}: {len(mats)} matrices, worst rel_err vs oracle {worst:.2e}", flush=True)


def test_list_surface_inf_and_exact():
    """The double-double route (lib 1) in one batched launch, and the exact route of
    integer matrices inside a list: equal to the per-call values, and the
    double-double ones to the correctly rounded exact sum."""
    rng = np.random.default_rng(8200)
    mats = [random_complex(rng, n, n) for n in (3, 7, 9, 9, 11)] + [random_complex(rng, 5, 4)]
    mats.append(rng.integers(-3, 4, size=(8, 8)).astype(np.complex128))
    got = BSU.GlynnPermanentCalculator_wrapper(mats, 1).calculate()
    for m, v in zip(mats, got):
        assert v == C.permanent(m, C.ALGO_BBFG, C.PREC_INF)
        assert v == exact_bbfg(m), (m.shape, v)
    ints = [rng.integers(-2, 3, size=(n, n)).astype(float) for n in (5, 9, 13)]
    vals = C.permanent_list(ints, C.ALGO_BBFG, C.PREC_EXACT)
    for m, v in zip(ints, vals):
        assert v == complex(float(oracle.int_permanent(m.astype(np.int64))), 0.0)


def test_list_surface_errors_name_the_matrix():
    mats = [np.eye(3), np.ones((2, 41)), np.eye(2)]
    with pytest.raises(Exception, match="matrix 1: matrix has more than 40 columns"):
        BSU.GlynnPermanentCalculator_wrapper(mats, 7).calculate()
    with pytest.raises(C.EngineError, match="matrix 1: exact mode needs a real integer-valued matrix"):
        C.permanent_list([np.eye(3), np.eye(3) * 0.5], C.ALGO_BBFG, C.PREC_EXACT)
    assert C.permanent_list([]) == []


def test_list_surface_large_uniform_batch():
    """1000 matrices of n = 12 (one launch, 1000 grid rows) and 64 of n = 20: each
    bit-identical to its own call."""
    rng = np.random.default_rng(8300)
    mats = [random_complex(rng, 12, 12) for _ in range(1000)] + [random_complex(rng, 20, 20) for _ in range(64)]
    got = C.permanent_list(mats, C.ALGO_BBFG, C.PREC_FP64)
    for k in list(range(0, 1000, 37)) + list(range(1000, 1064, 7)):
        assert got[k] == C.permanent(mats[k], C.ALGO_BBFG, C.PREC_FP64), k
    for k in range(0, 1064, 131):
        assert rel_err(got[k], oracle.bbfg(mats[k], "dd")) < TOL_SMALL, k


def test_haar_n20(haar_blocks):
    """n=20 block of a 40-mode Haar unitary, 1e-12 vs the reference
    GlynnPermanentCalculator (long double) restated."""
    A = haar_blocks["c1_n20"]
    ref = complex(*haar_blocks["c1_n20_glynn_ld"])
    for lib_cls in (BSU.GlynnPermanent, BSU.BBFGPermanentDouble):
        got = lib_cls(A).calculate()
        err = rel_err(got, ref)
        print(f"n=20 Haar {lib_cls.__name__}: rel_err={err:.3e}", flush=True)
        record_parity("config 1: n=20 Haar vs Glynn long double", err, 1e-12)
        assert err < 1e-12


@pytest.mark.parametrize("name", ["c2a_n28", "c2a_n30"])
def test_haar_n28_n30(haar_blocks, name):
    """n=28 / n=30 blocks of a 60-mode Haar unitary within 1e-10 of
    the double-double oracle."""
    A = haar_blocks[name]
    h = haar_blocks[name + "_dd"]
    ref = complex(h[0] + h[1], h[2] + h[3])
    got = C.permanent(A, C.ALGO_BBFG, C.PREC_FP64)
    err = rel_err(got, ref)
    print(f"{name}: got={got!r} ref={ref!r} rel_err={err:.3e}", flush=True)
    record_parity(f"config 2a: {name[4:]} Haar vs dd fixture", err, 1e-10)
    assert err < 1e-10


@pytest.mark.parametrize("n", [29, 30])
def test_full_size_invariants(haar_blocks, n):
    """Size-independent properties at the full sizes (n=29 has its own launch
    plan and no fixture): perm(A^T) = perm(A), perm(P A Q) = perm(A) for
    permutations P, Q, perm(c A) = c^n perm(A) for a unit phase c, and
    perm(conj A) = conj perm(A), each within 1e-10 relative.  Every variant
    sums the same terms in another order or another walk."""
    A = np.ascontiguousarray(haar_blocks["c2a_n30"][:n, :n])
    base = C.permanent(A, C.ALGO_BBFG, C.PREC_FP64)
    rng = np.random.default_rng(3000 + n)
    c = np.exp(0.7j)
    cases = {
        "transpose": (A.T, base),
        "permuted": (A[rng.permutation(n)][:, rng.permutation(n)], base),
        "phase": (c * A, c ** n * base),
        "conjugate": (A.conj(), base.conjugate()),
    }
    for name, (M, want) in cases.items():
        got = C.permanent(np.ascontiguousarray(M), C.ALGO_BBFG, C.PREC_FP64)
        err = rel_err(got, want)
        print(f"n={n} {name}: rel_err={err:.3e}", flush=True)
        assert err < 1e-10, name


def test_largest_size_n40_block_diagonal():
    """The largest order the kernels are built for (40 columns, 2^39 Gray steps): a
    block-diagonal of two n=20 Haar blocks with rows and columns permuted has
    perm = perm(B1) perm(B2) (the factors from the double-double oracle).  A lane
    sums ~2^24 FP64 terms here; the bar is 1e-8 (1e-10 is stated up to n=30)."""
    from piquassoboost_amd.sampling.utilities import haar_unitary

    B1 = haar_unitary(40, 2001)[:20, :20]
    B2 = haar_unitary(40, 2002)[:20, :20]
    A = np.zeros((40, 40), dtype=np.complex128)
    A[:20, :20] = B1
    A[20:, 20:] = B2
    rng = np.random.default_rng(2003)
    A = A[rng.permutation(40)][:, rng.permutation(40)]
    ref = oracle.bbfg(B1, "dd") * oracle.bbfg(B2, "dd")
    got = C.permanent(A, C.ALGO_BBFG, C.PREC_FP64)
    err = rel_err(got, ref)
    print(f"n=40: got={got!r} ref={ref!r} rel_err={err:.3e}", flush=True)
    record_parity("plain FP64, n=40 block diagonal", err, 1e-8)
    assert err < 1e-8


@pytest.mark.parametrize("n1,n2", [(16, 15), (16, 16), (17, 16), (18, 18)])
def test_block_diagonal_past_30_columns(n1, n2):
    """The sweep instantiations past n = 30 (31 and 32 columns at three waves per SIMD,
    32 with the single-pair step loop; 33 and 36 at one wave per SIMD): a block-diagonal
    of two Haar blocks with rows and columns permuted has perm = perm(B1) perm(B2), the
    factors from the double-double oracle."""
    from piquassoboost_amd.sampling.utilities import haar_unitary

    n = n1 + n2
    B1 = haar_unitary(2 * n1, 2100 + n)[:n1, :n1]
    B2 = haar_unitary(2 * n2, 2200 + n)[:n2, :n2]
    A = np.zeros((n, n), dtype=np.complex128)
    A[:n1, :n1] = B1
    A[n1:, n1:] = B2
    rng = np.random.default_rng(2300 + n)
    A = A[rng.permutation(n)][:, rng.permutation(n)]
    ref = oracle.bbfg(B1, "dd") * oracle.bbfg(B2, "dd")
    got = C.permanent(np.ascontiguousarray(A), C.ALGO_BBFG, C.PREC_FP64)
    err = rel_err(got, ref)
    print(f"n={n}: rel_err={err:.3e}", flush=True)
    record_parity("plain FP64, n=31..36 block diagonal", err, 1e-10)
    assert err < 1e-10


def test_exact_small_integer_matrices():
    rng = np.random.default_rng(21)
    for n in range(1, 15):
        A = rng.integers(-2, 3, size=(n, n))
        exact = oracle.int_permanent(A)
        v, d = C.permanent_exact(A.astype(float))
        assert v == exact, n
        assert d == float(exact)


def test_fp64_path_bit_exact_on_01_below_2p53():
    """0/1 matrices whose permanents fit in 2^53: the double result equals the
    integer permanent bit for bit (default precision routes integer input to
    the exact sweep)."""
    rng = np.random.default_rng(22)
    for n in (8, 12, 16, 20, 22):
        A = (rng.random((n, n)) < 0.35).astype(float)
        exact = oracle.int_permanent(A.astype(np.int64))
        assert exact < 2 ** 53
        got = BSU.BBFGPermanentDouble(A).calculate()
        assert got.real == float(exact) and got.imag == 0.0, (n, got, exact)


def test_bernoulli26_exact():
    """n=26 Bernoulli(1/2) 0/1 matrices; the exact 128-bit result
    equals the oracle's integer permanent, and the double result is its
    correctly rounded value."""
    d = np.load(GOLDEN / "bernoulli26.npz")
    for k in range(d["mats"].shape[0]):
        A = d["mats"][k].astype(float)
        exact = int(str(d["perm"][k]))
        v, dv = C.permanent_exact(A)
        assert v == exact
        assert dv == float(exact)
        got = BSU.BBFGPermanentDouble(A).calculate()
        record_parity("config 2b: n=26 Bernoulli, integer result vs oracle (bit-exact)", abs(v - exact) / exact, 0)
        assert got.real == float(exact)


def derangements(n):
    """!n by the exact recurrence !n = (n-1) (!(n-1) + !(n-2))."""
    a, b = 1, 0  # !0, !1
    if n == 0:
        return 1
    for k in range(2, n + 1):
        a, b = b, (k - 1) * (a + b)
    return b


def _with_column_abs_sum(rng, n, lo, hi, target):
    """n x n integers in [lo, hi] whose largest column abs sum is exactly `target`."""
    while True:
        A = rng.integers(lo, hi + 1, size=(n, n))
        cs = np.abs(A).sum(axis=0)
        if cs.max() > target:
            continue
        # raise column 0 to the target by growing entries away from zero
        need = target - cs[0]
        for i in range(n):
            if need == 0:
                break
            room = max(abs(lo), abs(hi)) - abs(A[i, 0])
            step = min(room, need)
            A[i, 0] = A[i, 0] + step if A[i, 0] >= 0 else A[i, 0] - step
            need -= step
        if need == 0 and np.abs(A).sum(axis=0).max() == target:
            return A


# (n, entry range, largest column abs sum): the host multiplies 13 column sums at a
# time in int64 up to a column abs sum of 28, 6 at a time up to 1400, else 2
EXACT_GROUPS = [(14, (-2, 2), 28), (14, (-3, 3), 29), (10, (-200, 200), 1400), (10, (-200, 200), 1401),
                (10, (-100, 100), None), (6, (-100000, 100000), None), (4, (-(1 << 28) + 1, (1 << 28) - 1), None)]


@pytest.mark.parametrize("case", range(len(EXACT_GROUPS)))
def test_exact_group_thresholds(case):
    """The exact sweep at each grouped-product variant (G = 13, 6, 2) and at the
    column abs sums where the host switches (28 / 29, 1400 / 1401): equal to the
    oracle's 128-bit Ryser permanent."""
    n, (lo, hi), target = EXACT_GROUPS[case]
    rng = np.random.default_rng(7900 + case)
    A = _with_column_abs_sum(rng, n, lo, hi, target) if target else rng.integers(lo, hi + 1, size=(n, n))
    cmax = int(np.abs(A).sum(axis=0).max())
    ok, lb = C.exact_bound(A.astype(float))
    if not ok:
        pytest.fail(f"case {case}: outside the exact bound (log2 {lb:.1f})")
    exact = oracle.int_permanent(A.astype(np.int64))
    v, d = C.permanent_exact(A.astype(float))
    print(f"n={n} cmax={cmax} group={13 if cmax <= 28 else (6 if cmax <= 1400 else 2)} "
          f"log2 bound={lb:.1f} perm={exact}", flush=True)
    assert v == exact
    assert d == float(exact)
    got = BSU.BBFGPermanentDouble(A.astype(float)).calculate()
    assert got.real == float(exact) and got.imag == 0.0


@pytest.mark.parametrize("shape", [(3, 2), (5, 4), (9, 8), (2, 5), (4, 9), (7, 12), (1, 6), (2, 30)])
def test_exact_rectangular_integer_default_route(shape):
    """Rectangular integer matrices through the default precision (which routes
    provably representable integer sums to the exact sweep): equal to the BBFG
    sum evaluated exactly in Python integers and rounded once.  rows = cols + 1
    cases have a zero row 0 (the sum is then perm of the other rows, not 0)."""
    r, c = shape
    rng = np.random.default_rng(8000 + 10 * r + c)
    A = rng.integers(-3, 4, size=shape).astype(float)
    if r == c + 1:
        A[0] = 0.0
    ok, lb = C.exact_bound(A)
    ref = exact_bbfg(A)
    got = C.permanent(A, C.ALGO_BBFG, C.PREC_DOUBLE)
    print(f"shape={shape} exact route={ok} (log2 bound {lb:.1f}) got={got!r} exact={ref!r}", flush=True)
    assert ok
    assert got == ref
    if r == c + 1:
        assert ref == complex(float(oracle.int_permanent(A[1:].astype(np.int64))), 0.0)


def test_exact_wide_rectangles_fall_back_to_fp64():
    """Sums beyond the 128-bit range (1 x 40 of 1000s, 2 x 30 of 50s) must not take
    the exact route; the FP64 sweep's value is checked against the exact sum."""
    for A in (np.full((1, 40), 1000.0), np.full((2, 30), 50.0)):
        assert not C.exact_bound(A)[0]
        ref = exact_bbfg(A)
        got = C.permanent(A, C.ALGO_BBFG, C.PREC_DOUBLE)
        assert rel_err(got, ref) < 1e-13, (A.shape, got, ref)


def test_exact_derangements():
    for n in (10, 16, 20, 25):
        D = np.ones((n, n)) - np.eye(n)
        v, d = C.permanent_exact(D)
        assert v == derangements(n), n
        assert d == float(derangements(n))


def test_repeated_kat(kat_cases):
    """unittests/test_permanent.py through every repeated calculator class."""
    classes = (BSU.BBFGRepeatedPermanentCalculatorDouble, BSU.BBFGRepeatedPermanentCalculatorLongDouble,
               BSU.GlynnRepeatedPermanentCalculator, BSU.GlynnRepeatedPermanentCalculatorDouble,
               BSU.ChinHuhPermanentCalculator, BSU.GlynnRepeatedInfPermanentCalculator)
    for c in kat_cases:
        for cls in classes:
            v = cls(c["matrix"], c["input"], c["output"]).calculate()
            assert rel_err(v, c["value"]) < 1e-12, (c["name"], cls.__name__, v)


def test_connector_kat(kat_cases):
    from piquassoboost_amd.connector import cpp_permanent_function

    for c in kat_cases:
        assert rel_err(cpp_permanent_function(c["matrix"], c["input"], c["output"]), c["value"]) < 1e-12


def test_repeated_list_batched():
    rng = np.random.default_rng(31)
    U = random_complex(rng, 5, 5, scale=0.4)
    ins = [np.array([1, 1, 0, 1, 0]), np.array([2, 0, 0, 1, 0])]
    outs = [[np.array([0, 1, 1, 1, 0]), np.array([3, 0, 0, 0, 0])], [np.array([1, 0, 0, 0, 2])]]
    res = BSU.BBFGRepeatedPermanentCalculatorDouble(U, ins, outs).calculate()
    assert len(res) == 2 and len(res[0]) == 2 and len(res[1]) == 1
    for i in range(2):
        for j, o in enumerate(outs[i]):
            assert rel_err(res[i][j], oracle.bbfg_repeated(U, ins[i], o, "dd")) < 1e-12


def test_concurrent_host_calls_are_deterministic():
    """Host calls from 8 Python threads (ctypes releases the GIL around them) run on up to
    four engine contexts -- a stream, workspaces and batch graphs each -- at once; every
    entry point gives bit-identical results to the same calls made one after another."""
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(99)
    mats = [random_complex(rng, 14, 14) for _ in range(48)]
    U = random_complex(rng, 12, 12, scale=0.3)
    ins = np.array([[1, 1, 0, 1, 0, 1, 0, 0, 1, 0, 1, 0], [2, 0, 0, 1, 0, 1, 0, 0, 0, 0, 1, 0]])
    outs = np.array([[0, 1, 1, 1, 0, 0, 1, 0, 0, 1, 1, 0], [1, 0, 0, 0, 2, 0, 0, 1, 0, 0, 1, 0]])

    def work(k):
        A = mats[k]
        kind = k % 8
        if kind == 0:
            return C.permanent(A)
        if kind == 1:
            return C.permanent_repeated(A[:5, :5], np.array([1, 2, 1, 1, 2]), np.array([2, 1, 1, 2, 1]))
        if kind == 2:
            return complex(C.permanent_grad(A[:9, :9])[1][2, 3])
        if kind == 3:
            return tuple(C.permanent_list([A[:6, :6], A[:7, :8], A], C.ALGO_BBFG, C.PREC_FP64))
        if kind == 4:
            return tuple(C.permanent_batch_states(U, ins, outs).tolist())
        if kind == 5:
            return C.permanent(A[:10, :10], C.ALGO_BBFG, C.PREC_INF)
        if kind == 6:
            return C.permanent_exact(np.round(A[:11, :11].real * 4))[0]
        return tuple(C.permanent_minors(A[:8, :9]).tolist())

    seq = [work(k) for k in range(len(mats))]
    for _ in range(2):
        with ThreadPoolExecutor(8) as ex:
            par = list(ex.map(work, range(len(mats))))
        assert par == seq
