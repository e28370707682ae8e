"""Boson-sampling strategies on the GPU engine: the seeded runs of the
reference's unittests/test_seeding.py (as multisets, see test_sampling_host),
the exact same sample sequence as the sampler driven by the oracle's
permanents, and compute_pmf of a 14-photon step against the oracle."""

import json
from collections import Counter

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from piquassoboost_amd.sampling.BosonSamplingSimulator import BosonSamplingSimulator
from piquassoboost_amd.sampling.simulation_strategies.GeneralizedCliffordsSimulationStrategy import (
    GeneralizedCliffordsBSimulationStrategy,
    GeneralizedCliffordsBUniformLossesSimulationStrategy,
    compute_pmf,
)
from piquassoboost_amd.sampling.utilities import haar_unitary

pytestmark = pytest.mark.gpu


def oracle_batch(tasks):
    return [1.0 + 0j if A.size == 0 else oracle.bbfg_repeated(A, c, r, "dd") for A, r, c in tasks]


@pytest.fixture(scope="module")
def seeding():
    d = json.loads((GOLDEN / "seeding.json").read_text())
    U = np.array(d["U"], dtype=float)
    d["U"] = U[..., 0] + 1j * U[..., 1]
    return d


def run(strategy, d):
    sim = BosonSamplingSimulator(strategy)
    return [tuple(int(x) for x in v) for v in sim.get_classical_simulation_results(d["input_state"], d["shots"])]


def test_lossless_seeded(seeding):
    d = seeding
    gpu = run(GeneralizedCliffordsBSimulationStrategy(d["U"], d["seed"]), d)
    assert Counter(gpu) == Counter(tuple(x) for x in d["lossless"])
    cpu = run(GeneralizedCliffordsBSimulationStrategy(d["U"], d["seed"], permanent_batch=oracle_batch), d)
    assert gpu == cpu


def test_uniform_losses_seeded(seeding):
    d = seeding
    U = d["U"]
    lossy = U @ np.diag([d["uniform_singular_value"]] * 5) @ U @ U.T
    sv = np.linalg.svd(lossy)[1]
    gpu = run(GeneralizedCliffordsBUniformLossesSimulationStrategy(lossy / sv[0], sv[0], d["seed"]), d)
    assert Counter(gpu) == Counter(tuple(x) for x in d["uniform_losses"])


def test_nonuniform_losses_seeded(seeding):
    """unittests/test_seeding.py:88-125 on the GPU: the 2m-mode dilation of the lossy
    interferometer (sampling/calculations.py), its permanents from the engine (lock-step
    column minors), the same samples as the oracle-driven run, and the reference test's
    multiset."""
    from piquassoboost_amd.sampling import calculations

    d = seeding
    U = d["U"]
    lossy = U @ np.diag(d["nonuniform_singular_values"]) @ U @ U.T
    gpu = [tuple(int(x) for x in v) for v in calculations.sample(lossy, d["input_state"], d["shots"], d["seed"])]
    assert Counter(gpu) == Counter(tuple(x) for x in d["nonuniform_losses"])
    cpu = [tuple(int(x) for x in v)
           for v in calculations.sample(lossy, d["input_state"], d["shots"], d["seed"], permanent_batch=oracle_batch)]
    assert gpu == cpu


def test_compute_pmf_large_step():
    """14 photons in 40 modes, collisions in input and sample: the column-minors sweep,
    the batched per-mode permanents and the oracle agree."""
    U = haar_unitary(40, 77)
    rng = np.random.default_rng(78)
    current_input = np.zeros(40, np.int64)
    np.add.at(current_input, rng.choice(40, size=14), 1)
    sample = np.zeros(40, np.int64)
    np.add.at(sample, rng.choice(40, size=13), 1)
    ref = compute_pmf(U, sample, current_input, permanent_batch=oracle_batch)
    for method in ("minors", "batch"):
        got = compute_pmf(U, sample, current_input, method=method)
        assert np.abs(got - ref).max() <= 1e-10 * ref.max(), method


@pytest.mark.parametrize("R", [1, 2, 3, 4, 5, 7, 8, 12, 20, 27])
def test_column_minors_against_oracle(R):
    """pqb_perm_minors on R x (R+1) matrices (every lane and register-bit split up to R=8)."""
    from conftest import random_complex
    from piquassoboost_amd import _capi as C

    rng = np.random.default_rng(500 + R)
    A = random_complex(rng, R, R + 1)
    got = C.permanent_minors(A)
    ref = np.array([oracle.bbfg(np.delete(A, j, 1), "dd") for j in range(R + 1)])
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_column_minors_shapes():
    from piquassoboost_amd import _capi as C

    assert C.permanent_minors(np.zeros((0, 1))).tolist() == [1.0]
    with pytest.raises(C.EngineError):
        C.permanent_minors(np.ones((3, 3)))


@pytest.mark.parametrize("R,count", [(0, 5), (1, 3), (4, 7), (8, 2), (13, 9), (19, 4), (24, 3)])
def test_column_minors_batch(R, count):
    """pqb_perm_minors_batch on a (count, R, R+1) stack equals pqb_perm_minors per matrix
    (to 1e-12; the batch splits the sweep over other wave ranges) and, up to R = 13, the
    oracle's minors."""
    from conftest import random_complex
    from piquassoboost_amd import _capi as C

    rng = np.random.default_rng(700 + R)
    stack = np.stack([random_complex(rng, R, R + 1) for _ in range(count)])
    got = C.permanent_minors_batch(stack)
    assert got.shape == (count, R + 1)
    for b in range(count):
        single = C.permanent_minors(stack[b])
        scale = np.abs(single).max()
        assert np.abs(got[b] - single).max() <= 1e-12 * scale, b
        if 0 < R <= 13:
            ref = np.array([oracle.bbfg(np.delete(stack[b], j, 1), "dd") for j in range(R + 1)])
            assert np.abs(got[b] - ref).max() <= 1e-12 * np.abs(ref).max(), b
    print(f"minors batch R={R} count={count}: ok", flush=True)


def test_lockstep_sampler_equals_sequential_on_the_gpu():
    """20 photons in 60 modes, 4 samples: the lock-step sampler (one batched minors launch
    per photon step) and the per-sample loop (pmf_method="batch") give the same samples."""
    U = haar_unitary(60, 1010)
    inp = np.zeros(60, dtype=np.int64)
    inp[:20] = 1
    a = GeneralizedCliffordsBSimulationStrategy(U, 1010).simulate(inp, 4)
    b = GeneralizedCliffordsBSimulationStrategy(U, 1010, pmf_method="batch").simulate(inp, 4)
    assert [tuple(x) for x in a] == [tuple(x) for x in b]
