"""Edge inputs through every host entry point that computes: non-finite entries
propagate as they do in the reference's double sweep (no error, no hang), and a
batch of many tiny tasks (more work items than the persistent grid has waves, most
of them one chunk) agrees task by task with the oracle."""

import numpy as np
import pytest

import oracle
from conftest import random_complex, record_parity, rel_err
from piquassoboost_amd import _capi as C

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_non_finite_entries_propagate(bad):
    """A NaN entry makes every term that touches it NaN, so the BBFG sum is NaN, as in
    BBFGPermanentCalculator.hpp (plain double arithmetic); an infinite entry gives a
    non-finite sum (inf - inf terms)."""
    rng = np.random.default_rng(41)
    A = random_complex(rng, 8, 8)
    A[3, 5] = bad
    for algo, prec in ((C.ALGO_BBFG, C.PREC_FP64), (C.ALGO_GLYNN, C.PREC_FP64), (C.ALGO_BBFG, C.PREC_EXTENDED)):
        v = C.permanent(A, algo, prec)
        print(f"bad={bad} algo={algo} prec={prec}: {v}", flush=True)
        assert not np.isfinite(v.real) or not np.isfinite(v.imag)
    rows = np.array([2, 1, 1, 1, 1, 1, 1], dtype=np.int64)
    B = random_complex(rng, 7, 8)
    B[2, 4] = bad
    v = C.permanent_repeated(B, rows, np.ones(8, dtype=np.int64))
    assert not np.isfinite(v.real) or not np.isfinite(v.imag)
    ok = random_complex(rng, 6, 6)
    one = np.ones(6, dtype=np.int64)
    vals = C.permanent_batch([(ok, one, one), (B, rows, np.ones(8, dtype=np.int64)), (ok, one, one)])
    assert np.isfinite(vals[0]) and np.isfinite(vals[2]) and vals[0] == vals[2]
    assert not np.isfinite(vals[1].real) or not np.isfinite(vals[1].imag)
    # the finite tasks of the batch are unaffected by the NaN one
    assert rel_err(vals[0], oracle.bbfg(ok, "dd")) < 1e-12


def test_many_tiny_tasks():
    """20000 tasks of 1..6 photons (one chunk each for most): more work items than the
    persistent grid has waves.  Every result against the oracle at 1e-12."""
    rng = np.random.default_rng(42)
    tasks, want = [], []
    for k in range(20000):
        n = 1 + k % 6
        A = random_complex(rng, n, n)
        mult = np.ones(n, dtype=np.int64)
        tasks.append((A, mult, mult))
    got = C.permanent_batch(tasks)
    errs = []
    for k in range(0, 20000, 7):  # every 7th task: all six sizes, ~2900 oracle calls
        A = tasks[k][0]
        errs.append(rel_err(got[k], oracle.bbfg(A, "dd")))
    worst = max(errs)
    print(f"20000 tiny tasks, {len(errs)} checked: worst rel_err={worst:.3e}", flush=True)
    record_parity("batch: 20000 tasks of 1..6 photons vs oracle dd", worst, 1e-12)
    assert worst < 1e-12
