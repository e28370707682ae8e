"""GPU parity of the repeated-row sweep and of the batched launch against the
restated reference (BBFGPermanentCalculatorRepeated) and the golden fixtures.
All calls go through the C ABI of libpqb_perm.so."""

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, random_complex, record_parity, rel_err
from piquassoboost_amd import _capi as C
from piquassoboost_amd.sampling import Boson_Sampling_Utilities as BSU
from piquassoboost_amd.sampling.permanent_calculators import (
    permanent_CPU_repeated_double,
    permanent_CPU_repeated_long_double,
)
from piquassoboost_amd.sampling.utilities import haar_unitary

pytestmark = pytest.mark.gpu

TOL = 1e-12


def test_30_photons_18_outputs_golden():
    """30 photons, 60 modes, 18 occupied output modes."""
    d = np.load(GOLDEN / "repeated_c3.npz")
    h = d["value_dd"]
    ref = complex(h[0] + h[1], h[2] + h[3])
    got = C.permanent_repeated(d["A"], d["row_mult"], d["col_mult"])
    err = rel_err(got, ref)
    print(f"30 photons / 18 outputs, compact: rel_err={err:.3e}", flush=True)
    record_parity("config 3: 30 photons, compact, vs dd fixture", err, 1e-10)
    assert err < 1e-10
    # the same transition on the full 60-mode unitary with zero multiplicities
    full = BSU.BBFGRepeatedPermanentCalculatorDouble(d["U"], d["input_state"], d["output_state"]).calculate()
    assert rel_err(full, ref) < 1e-10
    # and the expanded 30 x 30 matrix (rows repeated by their multiplicity)
    # through the plain sweep: another kernel and another walk over the same sum
    rows = np.repeat(np.arange(d["A"].shape[0]), d["row_mult"])
    E = np.ascontiguousarray(d["A"][rows])
    plain = C.permanent(E, C.ALGO_BBFG, C.PREC_FP64)  # = the repeated value (no multiplicity factor)
    err = rel_err(plain, ref)
    print(f"30 photons / 18 outputs, expanded 30 x 30 through the plain sweep: rel_err={err:.3e}", flush=True)
    record_parity("config 3: 30 photons, expanded 30x30, vs dd fixture", err, 1e-10)
    assert err < 1e-10


SPECIAL = [
    ([1, 1, 1], None), ([2, 3, 0], None), ([2, 3, 0, 1, 0], None), ([0, 0, 0], None),
    ([1, 1, 1, 1, 1, 1, 1, 2], None), ([1, 1, 1, 1, 1, 1, 3, 3], None), ([2, 3, 0], [1, 1, 3]),
    ([1, 1, 3], [2, 3, 0]), ([2, 3, 0, 1, 0], [1, 1, 0, 2, 2]), ([1, 1, 0, 2, 2], [2, 3, 0, 1, 0]),
    ([0, 0, 1, 1, 0, 2, 2], [0, 1, 1, 3, 0, 1, 0]), ([1, 1, 1, 1, 1, 1, 1, 2], [2, 1, 1, 1, 1, 1, 1, 1]),
    ([1, 1, 1, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 2],
     [0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 4, 1, 1, 1, 0, 0, 0, 0, 2]),
]


@pytest.mark.parametrize("case", range(len(SPECIAL)))
def test_special_cases_vs_oracle(case):
    """Multiplicity special cases of ctests/permanents_speed.cpp:208-231, every
    repeated calculator class, against the restated BBFG repeated (dd) and the
    permanent of the expanded matrix."""
    inp, outp = SPECIAL[case]
    outp = inp if outp is None else outp
    m = len(inp)
    U = haar_unitary(m, 500 + case)
    ref = oracle.bbfg_repeated(U, inp, outp, "dd")
    E = oracle.adapt_interferometer_glynn_multiplied(U, inp, outp)
    if E.size:
        assert rel_err(oracle.bbfg(E, "dd"), ref) < 1e-12
    for cls in (BSU.BBFGRepeatedPermanentCalculatorDouble, BSU.BBFGRepeatedPermanentCalculatorLongDouble,
                BSU.GlynnRepeatedPermanentCalculator, BSU.GlynnRepeatedPermanentCalculatorDouble,
                BSU.ChinHuhPermanentCalculator):
        got = cls(U, np.array(inp), np.array(outp)).calculate()
        assert rel_err(got, ref) < TOL, (cls.__name__, got, ref)
    assert rel_err(permanent_CPU_repeated_double(U, np.array(inp), np.array(outp)), ref) < TOL
    assert rel_err(permanent_CPU_repeated_long_double(U, np.array(inp), np.array(outp)), ref) < TOL


@pytest.mark.parametrize("walk", [1, 2, 5, 16, 64, 1000, 1 << 20])
def test_walk_length_split(walk):
    """Every split of the digits into walked (low) and chunk-index (high)
    digits gives the same permanent: walk=1 puts every digit in the chunk
    index, 2^20 walks all of them inside one chunk."""
    rng = np.random.default_rng(77)
    m = 9
    U = haar_unitary(m, 77)
    outp = np.array([1, 3, 0, 2, 1, 2, 0, 1, 2])
    inp = np.array([2, 1, 1, 0, 3, 1, 2, 1, 1])
    ref = oracle.bbfg_repeated(U, inp, outp, "dd")
    got = C.permanent_batch([(U, outp, inp)], walk_length=walk)[0]
    plan = C.repeated_plan(U, outp, inp, walk)
    print(f"walk={walk} plan={plan} rel_err={rel_err(got, ref):.3e}", flush=True)
    assert rel_err(got, ref) < TOL
    del rng


def test_random_multiplicities_vs_oracle():
    rng = np.random.default_rng(2024)
    for trial in range(25):
        m = int(rng.integers(2, 10))
        n_ph = int(rng.integers(1, 14))
        inp = rng.multinomial(n_ph, np.ones(m) / m)
        outp = rng.multinomial(n_ph, np.ones(m) / m)
        U = random_complex(rng, m, m, scale=1 / np.sqrt(m))
        ref = oracle.bbfg_repeated(U, inp, outp, "dd")
        got = C.permanent_repeated(U, outp, inp)
        scale = max(abs(ref), 1e-300)
        assert abs(got - ref) <= TOL * max(scale, 1e-14), (trial, inp, outp, got, ref)


def test_rectangular_compact_forms():
    """Compact (adaptInterferometer) non-square matrices with multiplicities."""
    rng = np.random.default_rng(5)
    A = random_complex(rng, 3, 5)
    rows = np.array([2, 2, 1])
    cols = np.array([1, 1, 1, 1, 1])
    ref = oracle.bbfg_repeated(A, cols, rows, "dd")
    assert rel_err(C.permanent_repeated(A, rows, cols), ref) < TOL
    B = random_complex(rng, 4, 2)
    assert rel_err(C.permanent_repeated(B, [1, 1, 1, 1], [2, 2]), oracle.bbfg_repeated(B, [2, 2], [1, 1, 1, 1], "dd")) < TOL


def test_batch_equals_per_call_mixed_sizes():
    """Uneven batch (n from 1 to 34 photons, crossing the 32-slot kernel split)
    equals the per-call results and the oracle."""
    rng = np.random.default_rng(99)
    U = haar_unitary(40, 99)
    tasks, refs = [], []
    for n in list(range(1, 14)) + [20, 24, 33, 34]:
        modes_in = rng.choice(40, size=min(n, 40), replace=False)
        inp = np.zeros(40, dtype=np.int64)
        inp[modes_in[:n]] = 1
        outp = rng.multinomial(n, np.ones(40) / 40)
        tasks.append((U, outp, inp))
    res = C.permanent_batch(tasks)
    for k, (M, o, i) in enumerate(tasks):
        single = C.permanent_repeated(M, o, i)
        n = int(i.sum())
        if n <= 22:
            ref = oracle.bbfg_repeated(M, i, o, "dd")
            assert rel_err(res[k], ref) < 1e-11, (k, n)
        assert rel_err(res[k], single) < 1e-12, (k, n, res[k], single)


def test_batch_replay_reports_time():
    rng = np.random.default_rng(3)
    U = haar_unitary(12, 3)
    tasks = [(U, rng.multinomial(8, np.ones(12) / 12), np.array([1] * 8 + [0] * 4)) for _ in range(50)]
    C.permanent_batch(tasks)
    ms = C.batch_replay_ms(3)
    assert ms > 0.0


def test_repeated_list_batched_multi_input():
    rng = np.random.default_rng(41)
    U = random_complex(rng, 4, 4, scale=0.5)
    ins = [[np.array([1, 1, 0, 0]), np.array([0, 1, 1, 0])], [np.array([2, 0, 0, 1])]]
    outs = [np.array([0, 1, 0, 1]), np.array([1, 1, 1, 0])]
    res = BSU.BBFGRepeatedPermanentCalculatorDouble(U, ins, outs).calculate()
    for i in range(2):
        for j, x in enumerate(ins[i]):
            assert rel_err(res[i][j], oracle.bbfg_repeated(U, x, outs[i], "dd")) < TOL
