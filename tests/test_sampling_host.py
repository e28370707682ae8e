"""Boson-sampling strategies, host side (no device): the std::mt19937
restatement against the C++ standard's known answer and numpy's legacy
Mersenne Twister; compute_pmf against the Laplace identity it relies on; and
the photon-by-photon sampler, run with the oracle's permanents, against the
seeded samples of the reference's unittests/test_seeding.py (fixture
tests/golden/seeding.json).  The reference's samples reach the test through
piquasso, which reorders them; the comparison is therefore of the 20-sample
multisets, which a wrong pmf or random-number use would not reproduce."""

import json
from collections import Counter

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from piquassoboost_amd.sampling.mt19937 import MT19937
from piquassoboost_amd.sampling.simulation_strategies.GeneralizedCliffordsSimulationStrategy import (
    GeneralizedCliffordsBSimulationStrategy,
    GeneralizedCliffordsBUniformLossesSimulationStrategy,
    compute_pmf,
    modes_state_to_particle_state,
)


def oracle_batch(tasks):
    return [1.0 + 0j if A.size == 0 else oracle.bbfg_repeated(A, c, r, "dd") for A, r, c in tasks]


@pytest.fixture(scope="module")
def seeding():
    d = json.loads((GOLDEN / "seeding.json").read_text())
    U = np.array(d["U"], dtype=float)
    d["U"] = U[..., 0] + 1j * U[..., 1]
    return d


def test_mt19937_standard_known_answer():
    g = MT19937()  # default seed 5489
    for _ in range(9999):
        g()
    assert g() == 4123659995  # [rand.predef]: 10000th invocation of a default-constructed mt19937


def test_mt19937_matches_numpy_legacy_seeding():
    bg = np.random.MT19937()
    bg._legacy_seeding(123)
    g = MT19937(123)
    assert [g() for _ in range(1500)] == [int(x) for x in bg.random_raw(1500)]
    g.seed(2**32 + 7)  # seed values are taken modulo 2^32
    h = MT19937(7)
    assert [g() for _ in range(10)] == [h() for _ in range(10)]


def test_particle_state():
    assert modes_state_to_particle_state([2, 1, 1, 0, 1]).tolist() == [0, 0, 1, 2, 4]


def test_compute_pmf_is_the_next_photon_distribution():
    """pmf[m] is proportional to |perm|^2 of the expanded matrix with output row m added
    (the Laplace expansion along the new row that compute_pmf uses)."""
    rng = np.random.default_rng(8)
    m = 7
    Z = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    U, _ = np.linalg.qr(Z)
    current_input = np.array([1, 0, 2, 0, 1, 1, 0])
    sample = np.array([0, 2, 0, 1, 0, 1, 0])  # one photon fewer than the input
    pmf = compute_pmf(U, sample, current_input, permanent_batch=oracle_batch)
    cols = np.repeat(np.arange(m), current_input)
    ref = np.empty(m)
    for mdx in range(m):
        s = sample.copy()
        s[mdx] += 1
        rows = np.repeat(np.arange(m), s)
        ref[mdx] = abs(oracle.ryser(U[np.ix_(rows, cols)], "dd")) ** 2
    ref /= ref.sum()
    assert np.abs(pmf - ref).max() < 1e-13
    assert abs(pmf.sum() - 1.0) < 1e-14


def test_lossless_seeded_samples(seeding):
    d = seeding
    s = GeneralizedCliffordsBSimulationStrategy(d["U"], d["seed"], permanent_batch=oracle_batch)
    got = [tuple(int(x) for x in v) for v in s.simulate(d["input_state"], d["shots"])]
    assert Counter(got) == Counter(tuple(x) for x in d["lossless"])


def test_uniform_losses_seeded_samples(seeding):
    """LossyInterferometer U diag(0.9) U U^T: uniform singular values, so the strategy gets
    the matrix divided by its largest singular value and that value as transmissivity
    (the boosted measurement's uniform-loss branch)."""
    d = seeding
    U = d["U"]
    lossy = U @ np.diag([d["uniform_singular_value"]] * 5) @ U @ U.T
    sv = np.linalg.svd(lossy)[1]
    s = GeneralizedCliffordsBUniformLossesSimulationStrategy(lossy / sv[0], sv[0], d["seed"],
                                                             permanent_batch=oracle_batch)
    got = [tuple(int(x) for x in v) for v in s.simulate(d["input_state"], d["shots"])]
    assert Counter(got) == Counter(tuple(x) for x in d["uniform_losses"])


def test_other_seed_gives_other_samples(seeding):
    """The fixture comparison is sensitive: seed 124 does not reproduce the multiset."""
    d = seeding
    s = GeneralizedCliffordsBSimulationStrategy(d["U"], 124, permanent_batch=oracle_batch)
    got = [tuple(int(x) for x in v) for v in s.simulate(d["input_state"], d["shots"])]
    assert Counter(got) != Counter(tuple(x) for x in d["lossless"])


def oracle_minors_batch(Ms):
    """Column minors of a (count, k-1, k) stack from the oracle (the stand-in for the GPU
    entry the lock-step sampler calls)."""
    Ms = np.asarray(Ms)
    out = np.zeros((Ms.shape[0], Ms.shape[2]), dtype=complex)
    for b, M in enumerate(Ms):
        for j in range(M.shape[1]):
            sub = np.delete(M, j, axis=1)
            out[b, j] = 1.0 if sub.shape[0] == 0 else oracle.bbfg(np.ascontiguousarray(sub), "dd")
    return out


@pytest.fixture
def oracle_minors(monkeypatch):
    from piquassoboost_amd.sampling.simulation_strategies import GeneralizedCliffordsSimulationStrategy as G

    monkeypatch.setattr(G.C, "permanent_minors_batch", oracle_minors_batch)


def test_lockstep_sampler_reproduces_the_sequential_loop(seeding, oracle_minors):
    """Several samples advanced together (one batched minors call per photon step) draw
    the random numbers up front in the sequential loop's order: the samples equal those of
    the sequential oracle-driven loop one by one, and the seeded fixture's multiset."""
    d = seeding
    # one generator is shared by all strategies (as pic::rng_gen): seed, then run, each in turn
    lock = GeneralizedCliffordsBSimulationStrategy(d["U"], d["seed"])  # minors method: lock-step
    a = [tuple(int(x) for x in v) for v in lock.simulate(d["input_state"], d["shots"])]
    seq = GeneralizedCliffordsBSimulationStrategy(d["U"], d["seed"], permanent_batch=oracle_batch)
    b = [tuple(int(x) for x in v) for v in seq.simulate(d["input_state"], d["shots"])]
    assert a == b
    assert Counter(a) == Counter(tuple(x) for x in d["lossless"])


def test_lockstep_uniform_losses_reproduces_the_sequential_loop(seeding, oracle_minors):
    """With losses the photon count differs per sample (drawn first); samples that are
    complete leave the batch while the others continue."""
    d = seeding
    U = d["U"]
    lossy = U @ np.diag([d["uniform_singular_value"]] * 5) @ U @ U.T
    sv = np.linalg.svd(lossy)[1]
    lock = GeneralizedCliffordsBUniformLossesSimulationStrategy(lossy / sv[0], sv[0], d["seed"])
    a = [tuple(int(x) for x in v) for v in lock.simulate(d["input_state"], d["shots"])]
    seq = GeneralizedCliffordsBUniformLossesSimulationStrategy(lossy / sv[0], sv[0], d["seed"],
                                                               permanent_batch=oracle_batch)
    b = [tuple(int(x) for x in v) for v in seq.simulate(d["input_state"], d["shots"])]
    assert a == b
    assert len(set(sum(x) for x in a)) > 1  # the photon counts do differ
    assert Counter(a) == Counter(tuple(x) for x in d["uniform_losses"])


def test_nonuniform_losses_seeded_samples(seeding):
    """unittests/test_seeding.py:88-125: LossyInterferometer U diag(0.9..0.5) U U^T, input
    [2,1,1,0,1], seed 123, 20 shots.  Non-uniform singular values: the 2m-mode unitary
    dilation (piquasso 7.2.1's, restated in sampling/calculations.py), the lossless
    strategy on it with the input padded by empty modes, the samples cut to 5 modes
    (calculations.py:121-145); the 20-sample multiset equals the reference test's."""
    from piquassoboost_amd.sampling import calculations

    d = seeding
    U = d["U"]
    lossy = U @ np.diag(d["nonuniform_singular_values"]) @ U @ U.T
    got = calculations.sample(lossy, d["input_state"], d["shots"], d["seed"], permanent_batch=oracle_batch)
    assert got.shape == (20, 5)
    assert Counter(map(tuple, got.tolist())) == Counter(tuple(x) for x in d["nonuniform_losses"])
    # the reference's Branch frequencies
    res = calculations.particle_number_measurement(lossy, d["input_state"], d["shots"], d["seed"],
                                                   permanent_batch=oracle_batch)
    assert sum(f for _, f in res) == 1
    assert {o: f * 20 for o, f in res} == dict(Counter(tuple(x) for x in d["nonuniform_losses"]))


def test_dilation_is_unitary_on_the_occupied_modes(seeding):
    """The dilation's first m columns are orthonormal and its top-left block is the lossy
    matrix: the marginal on the first m output modes is the lossy process's."""
    from piquassoboost_amd.sampling import calculations

    d = seeding
    U = d["U"]
    lossy = U @ np.diag(d["nonuniform_singular_values"]) @ U @ U.T
    W = calculations.prepare_interferometer_matrix_in_expanded_space(np.linalg.svd(lossy))
    assert W.shape == (10, 10)
    assert np.abs(W[:5, :5] - lossy).max() < 1e-14
    assert np.abs(W[:, :5].conj().T @ W[:, :5] - np.eye(5)).max() < 1e-14


def test_calculations_branches_match_the_strategies(seeding):
    """Lossless and uniform-loss inputs go to the same strategies the existing fixtures
    pin (same samples)."""
    from piquassoboost_amd.sampling import calculations

    d = seeding
    U = d["U"]
    got = calculations.sample(U, d["input_state"], d["shots"], d["seed"], permanent_batch=oracle_batch)
    assert Counter(map(tuple, got.tolist())) == Counter(tuple(x) for x in d["lossless"])
    lossy = U @ np.diag([d["uniform_singular_value"]] * 5) @ U @ U.T
    got = calculations.sample(lossy, d["input_state"], d["shots"], d["seed"], permanent_batch=oracle_batch)
    assert Counter(map(tuple, got.tolist())) == Counter(tuple(x) for x in d["uniform_losses"])
