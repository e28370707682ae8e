"""4096 uneven repeated-row permanents (n in [8, 26], output
multiplicities in {1, 2, 3}) from one seeded 100-mode Haar unitary, computed
in one batched launch; results checked against the restated reference (dd)
to 1e-10 and against the per-call path."""

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, record_parity, rel_err
from piquassoboost_amd import _capi as C
from piquassoboost_amd.sampling.utilities import adapt_interferometer, uneven_batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def uneven_batch4096():
    U, pairs = uneven_batch(seed=1005, count=4096)
    tasks = []
    for inp, outp in pairs:
        A = adapt_interferometer(U, inp, outp)
        tasks.append((A, outp[outp > 0], inp[inp > 0]))
    res = C.permanent_batch(tasks)
    return U, pairs, tasks, res


def test_uneven_batch_against_oracle(uneven_batch4096):
    U, pairs, tasks, res = uneven_batch4096
    rng = np.random.default_rng(7)
    pick = rng.choice(len(tasks), size=40, replace=False)
    worst = 0.0
    for k in pick:
        A, rows, cols = tasks[k]
        ref = oracle.bbfg_repeated(A, cols, rows, "dd")
        err = rel_err(res[k], ref)
        worst = max(worst, err)
        assert err < 1e-10, (int(k), int(cols.sum()), err)
    print(f"uneven batch, sample of 40: worst rel_err={worst:.3e}", flush=True)


@pytest.fixture(scope="module")
def uneven_fixture():
    """The stored workload (unitary, modes, multiplicities) and the oracle's
    double-double value of every task (tests/golden/make_golden.py)."""
    d = np.load(GOLDEN / "uneven4096.npz")
    U = d["U"]
    m = U.shape[0]
    count = d["in_modes"].shape[0]
    ins = np.zeros((count, m), dtype=np.int64)
    outs = np.zeros((count, m), dtype=np.int64)
    for k in range(count):
        im = d["in_modes"][k]
        im = im[im >= 0]
        ins[k, im] = 1
        om = d["out_modes"][k]
        sel = om >= 0
        outs[k, om[sel]] = d["out_mult"][k][sel]
    v = d["value_dd"]
    ref = (v[:, 0] + v[:, 1]) + 1j * (v[:, 2] + v[:, 3])
    return U, ins, outs, ref


def _worst(got, ref):
    err = np.abs(np.asarray(got) - ref) / np.abs(ref)
    return err, int(np.argmax(err))


def test_uneven_batch_every_result_vs_fixture(uneven_fixture):
    """Every one of the 4096 values of the batched workload, through both batched
    entry points (the task list, pqb_perm_batch, and the unitary with state
    arrays, pqb_perm_batch_states), within 1e-10 of the oracle's double-double
    value (BatchedPermanentCalculator.cpp:118-145 evaluates each pair with the
    reference calculator)."""
    U, ins, outs, ref = uneven_fixture
    tasks = [(adapt_interferometer(U, i, o), o[o > 0], i[i > 0]) for i, o in zip(ins, outs)]
    for name, got in (("pqb_perm_batch", C.permanent_batch(tasks)),
                      ("pqb_perm_batch_states", C.permanent_batch_states(U, ins, outs))):
        err, k = _worst(got, ref)
        photons = int(ins[k].sum())
        print(f"{name}: 4096 results, worst rel_err={err[k]:.3e} (task {k}, {photons} photons), "
              f"median {np.median(err):.2e}", flush=True)
        record_parity(f"config 4: 4096 tasks, {name}, vs dd fixture", err[k], 1e-10)
        assert err[k] < 1e-10, (name, k)


def test_uneven_batch_every_result_per_call_vs_fixture(uneven_fixture):
    """The same 4096 tasks one call each (pqb_perm_repeated), each within 1e-10 of the
    oracle: the per-call path the batch is measured against."""
    U, ins, outs, ref = uneven_fixture
    got = [C.permanent_repeated(adapt_interferometer(U, i, o), o[o > 0], i[i > 0]) for i, o in zip(ins, outs)]
    err, k = _worst(got, ref)
    print(f"pqb_perm_repeated per call: worst rel_err={err[k]:.3e} (task {k})", flush=True)
    record_parity("config 4: 4096 tasks, per call, vs dd fixture", err[k], 1e-10)
    assert err[k] < 1e-10


def test_uneven_batch_full_mode_form_matches_compact(uneven_batch4096):
    """The same transitions given on the full 100-mode unitary with zero
    multiplicities (what BBFGRepeatedPermanentCalculator receives from a
    sampler) give the same values.  The walk length depends on the batch's
    total work, so a 256-task batch is chunked differently from the
    4096-task one: agreement is to the parity bar, not bit for bit."""
    U, pairs, tasks, res = uneven_batch4096
    full = C.permanent_batch([(U, outp, inp) for inp, outp in pairs[:256]])
    for k in range(256):
        assert rel_err(full[k], res[k]) < 1e-10


def test_uneven_batch_states_api_is_bit_identical(uneven_batch4096):
    """pqb_perm_batch_states (the unitary once, (count, modes) occupation arrays) plans
    every pair from the full unitary; dropping the unoccupied modes gives the compact
    matrices of the task list, the same plans and the same batch total, so the 4096
    values are bit-identical; likewise in double-double."""
    U, pairs, tasks, res = uneven_batch4096
    ins = np.stack([p[0] for p in pairs])
    outs = np.stack([p[1] for p in pairs])
    got = C.permanent_batch_states(U, ins, outs)
    assert got.shape == (4096,)
    assert got.tolist() == res
    sub = slice(0, 48)
    dd = C.permanent_batch_states(U, ins[sub], outs[sub], C.ALGO_BBFG, C.PREC_INF)
    assert dd.tolist() == C.permanent_batch(tasks[sub], C.ALGO_BBFG, C.PREC_INF)


def test_uneven_batch_per_call_equals_batch(uneven_batch4096):
    U, pairs, tasks, res = uneven_batch4096
    for k in range(0, len(tasks), 97):
        A, rows, cols = tasks[k]
        assert rel_err(C.permanent_repeated(A, rows, cols), res[k]) < 1e-10


def test_uneven_batch_replay_deterministic(uneven_batch4096):
    """A graph replay recomputes the resident batch; the launch is
    deterministic (fixed item order per task in the reduction)."""
    U, pairs, tasks, res = uneven_batch4096
    res2 = C.permanent_batch(tasks)
    assert res2 == res
    assert C.batch_replay_ms(2) > 0.0


def test_batched_front_end(uneven_batch4096):
    """BatchednPermanentCalculator (one interferometer, accumulated state pairs) on 64 of
    the uneven transitions: both lib codes against the per-call path and the oracle."""
    from piquassoboost_amd.sampling.BatchedPermanentCalculator import BatchednPermanentCalculator

    U, pairs, tasks, res = uneven_batch4096
    b = BatchednPermanentCalculator(U)
    b.reserve_space(64)
    for k in range(64):
        b.add(pairs[k][0], pairs[k][1], k)
    for lib in (0, 1):
        out = b.calculate(lib)
        assert out.shape == (1, 64)
        for k in range(0, 64, 9):
            A, rows, cols = tasks[k]
            assert rel_err(out[0, k], oracle.bbfg_repeated(A, cols, rows, "dd")) < 1e-10, (lib, k)


def test_batches_from_concurrent_host_threads(uneven_batch4096):
    """Batch calls from 4 Python threads at once (ctypes releases the GIL): their
    host phases contend for the worker pool (a phase that finds it busy starts
    its own threads) and their sweeps run on separate engine contexts (own
    stream, buffers and captured graphs) at the same time.  Every call returns
    the values of the same call made alone, bit for bit (the task list and the
    state-array form of one half of the batch plan identically)."""
    from concurrent.futures import ThreadPoolExecutor

    U, pairs, tasks, res = uneven_batch4096
    ins = np.stack([i for i, _ in pairs])
    outs = np.stack([o for _, o in pairs])
    want = [C.permanent_batch(tasks[h::2]) for h in range(2)]

    def work(k):
        h = k // 2
        if k % 2:
            got = C.permanent_batch_states(U, ins[h::2], outs[h::2]).tolist()
        else:
            got = C.permanent_batch(tasks[h::2])
        return got == want[h]

    with ThreadPoolExecutor(4) as ex:
        for rep in range(3):
            assert list(ex.map(work, range(4))) == [True] * 4, rep
