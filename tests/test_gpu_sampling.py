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


def test_compute_pmf_large_step():
    """14 photons in 40 modes, collisions in input and sample."""
    U = haar_unitary(40, 77)
    rng = np.random.default_rng(78)
    current_input = np.zeros(40, np.int64)
    np.add.at(current_input, rng.choice(40, size=14), 1)
    sample = np.zeros(40, np.int64)
    np.add.at(sample, rng.choice(40, size=13), 1)
    got = compute_pmf(U, sample, current_input)
    ref = compute_pmf(U, sample, current_input, permanent_batch=oracle_batch)
    assert np.abs(got - ref).max() <= 1e-10 * ref.max()
