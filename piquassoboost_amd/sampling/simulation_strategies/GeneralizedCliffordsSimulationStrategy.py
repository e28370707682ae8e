"""Generalized Clifford & Clifford boson sampling (Brod and Oszmaniec 2020) as
the reference runs it, with the permanents of each photon step on the GPU.

Mirrors piquassoboost.sampling.simulation_strategies.GeneralizedCliffordsSimulationStrategy
(the Python classes :33-66) over the reference's C++ strategies:

* ``GeneralizedCliffordsBSimulationStrategy`` -- CGeneralizedCliffordsBSimulationStrategy.cpp:
  per sample, photons enter one at a time in a random order (update_input_by_single_photon,
  samplingHelperFunctions.cpp:370-399); after each, the output distribution over modes for one
  more photon is computed (compute_pmf, :78-342) and an output mode is drawn from it
  (sample_from_pmf, :350-367); fill_r_sample :281-305, simulate :135-275.
* ``GeneralizedCliffordsBUniformLossesSimulationStrategy`` --
  CGeneralizedCliffordsBUniformLossesSimulationStrategy.cpp: the number of surviving photons is
  drawn first from Binomial(n, transmissivity^2) (:285-316), then the same photon loop.

compute_pmf needs, for every occupied input mode, the permanent of the effective matrix
with one photon removed from that mode (the reference calls
BBFGPermanentCalculatorRepeated once per mode).  These are the column minors of the
expanded (k-1) x k matrix, and one column-minors sweep on the GPU (pqb_perm_minors) gives
all of them; the per-mode batched form is kept as method="batch".  The random numbers
come from the shared std::mt19937 restated in
``piquassoboost_amd.sampling.mt19937`` (seeded as the reference does), so seeded runs give
the reference's samples.
"""

from __future__ import annotations

import time

import numpy as np

from ... import _capi as C
from ..mt19937 import MT19937

# one generator shared by every strategy, as pic::rng_gen (samplingHelperFunctions.cpp:56-60)
rng_gen = MT19937()


def seed_random_generator(value: int) -> None:
    rng_gen.seed(int(value))


def modes_state_to_particle_state(mode_state) -> np.ndarray:
    """Mode occupations -> list of the modes of the particles
    (CGeneralizedCliffordsBSimulationStrategy.cpp:322-336)."""
    s = np.asarray(mode_state, dtype=np.int64)
    return np.repeat(np.arange(s.shape[0], dtype=np.int64), s)


def _addends_batched(U, sample, current_input, col_idx, perm_fn):
    """One repeated-row permanent per occupied input mode, as the reference computes them
    (samplingHelperFunctions.cpp:137-172), submitted as one batch."""
    out_rows = np.flatnonzero(sample > 0)
    tasks = []
    for idx in col_idx:
        inp = current_input.copy()
        inp[idx] -= 1
        in_cols = np.flatnonzero(inp > 0)
        A = np.ascontiguousarray(U[np.ix_(out_rows, in_cols)], dtype=np.complex128)
        tasks.append((A, sample[out_rows], inp[in_cols]))
    return np.asarray(perm_fn(tasks), dtype=np.complex128) if tasks else np.zeros(0, np.complex128)


def _expanded(U, sample, current_input):
    """The expanded (k-1) x k matrix of a photon step (rows: sample photons, columns: input
    photons) and the input mode of each column."""
    rows = np.repeat(np.arange(sample.shape[0]), sample)
    cols = np.repeat(np.arange(current_input.shape[0]), current_input)
    return np.ascontiguousarray(U[np.ix_(rows, cols)], dtype=np.complex128), cols


def _addends_from_minors(minors, cols, col_idx):
    """The permanent with one photon removed from input mode idx is the minor without one
    (the first) column of mode idx."""
    first = {int(mode): j for j, mode in reversed(list(enumerate(cols)))}
    return np.array([minors[first[int(idx)]] for idx in col_idx], dtype=np.complex128)


def _addends_minors(U, sample, current_input, col_idx):
    """All addends from one column-minors sweep of the expanded matrix."""
    E, cols = _expanded(U, sample, current_input)
    return _addends_from_minors(C.permanent_minors(E), cols, col_idx)


def _pmf_from_addends(U, sample, current_input, col_idx, addends):
    """pmf[mdx] = |sum_idx current_input[idx] * addend[idx] * U[mdx, idx]|^2, normalised,
    accumulated in the reference's order (samplingHelperFunctions.cpp:174-340)."""
    perms = np.zeros(sample.shape[0], dtype=np.complex128)
    for k, idx in enumerate(col_idx):
        perms = perms + (current_input[idx] * addends[k]) * U[:, idx]
    pmf = perms.real * perms.real + perms.imag * perms.imag
    return pmf / pmf.sum()


def compute_pmf(interferometer_matrix, sample, current_input, permanent_batch=None,
                method: str = "minors") -> np.ndarray:
    """Probability of each output mode receiving the next photon
    (samplingHelperFunctions.cpp:78-342).

    For each occupied input mode idx: permanent_addends[idx] = perm of the effective matrix
    (rows: occupied output modes of `sample`, columns: occupied input modes after removing
    one photon from idx, with their multiplicities).  Then
    pmf[mdx] = |sum_idx current_input[idx] * addend[idx] * U[mdx, idx]|^2, normalised.

    method "minors" (default): all addends from one column-minors sweep on the GPU;
    "batch": one repeated-row permanent per mode in one batched launch, through
    `permanent_batch` when given (a callable taking (matrix, row_mult, col_mult) tasks)."""
    U = np.asarray(interferometer_matrix)
    sample = np.asarray(sample, dtype=np.int64)
    current_input = np.asarray(current_input, dtype=np.int64)
    col_idx = np.flatnonzero(current_input > 0)
    if int(sample.sum()) + 1 != int(current_input.sum()):
        raise ValueError("compute_pmf: the input must hold one photon more than the sample")
    if method == "minors" and permanent_batch is None:
        addends = _addends_minors(U, sample, current_input, col_idx)
    elif method in ("minors", "batch"):
        perm_fn = C.permanent_batch if permanent_batch is None else permanent_batch
        addends = _addends_batched(U, sample, current_input, col_idx, perm_fn)
    else:
        raise ValueError(f"unknown compute_pmf method {method!r}")
    return _pmf_from_addends(U, sample, current_input, col_idx, addends)


def sample_from_pmf(sample: np.ndarray, pmf: np.ndarray) -> None:
    """samplingHelperFunctions.cpp:350-367: first index whose running sum reaches the draw
    (index 0 when rounding leaves the total below it)."""
    _sample_with_draw(sample, pmf, rng_gen.random_double_0_1())


def _sample_with_draw(sample: np.ndarray, pmf: np.ndarray, rand_num: float) -> None:
    sampled_index = 0
    prob_sum = 0.0
    for idx in range(pmf.shape[0]):
        prob_sum = prob_sum + float(pmf[idx])
        if prob_sum >= rand_num:
            sampled_index = idx
            break
    sample[sampled_index] += 1


def update_input_by_single_photon(current_input: np.ndarray, working_input_state: np.ndarray) -> np.ndarray:
    """samplingHelperFunctions.cpp:370-399: move a random remaining particle into the
    input; returns the reduced particle list."""
    if working_input_state.shape[0] == 0:
        raise Exception(
            "Generalized Cliffords B Simulation Strategy::update_input_by_single_photon: "
            "the size of working_input_state is zero"
        )
    rand_index = rng_gen.random_index(working_input_state.shape[0])
    current_input[working_input_state[rand_index]] += 1
    return np.delete(working_input_state, rand_index)


class GeneralizedCliffordsBSimulationStrategy:
    """GeneralizedCliffordsSimulationStrategy.py:33-47 /
    CGeneralizedCliffordsBSimulationStrategy.cpp (lossless)."""

    def __init__(self, interferometer_matrix, seed=None, lib: int = 0, permanent_batch=None,
                 pmf_method: str = "minors"):
        self.interferometer_matrix = np.ascontiguousarray(interferometer_matrix, dtype=np.complex128)
        self.lib = lib
        self._perm = permanent_batch
        self._pmf_method = pmf_method
        seed_random_generator(int(time.time()))  # the C++ constructor seeds with time(NULL)
        if seed is not None:
            self.seed(seed)

    def seed(self, value) -> None:
        seed_random_generator(value)

    def _fill_r_sample(self, sample, current_input, working, photons) -> None:
        while photons > int(sample.sum()):
            working = update_input_by_single_photon(current_input, working)
            pmf = compute_pmf(self.interferometer_matrix, sample, current_input, self._perm, self._pmf_method)
            sample_from_pmf(sample, pmf)

    def _photons(self, input_state) -> int:
        return int(np.sum(input_state))

    def simulate(self, input_state, samples_number: int = 1):
        """Tuple of `samples_number` output states (mode occupations, int64)."""
        input_state = np.asarray(input_state, dtype=np.int64)
        particles = modes_state_to_particle_state(input_state)
        n_samples = max(0, int(samples_number))
        if self._pmf_method == "minors" and self._perm is None and n_samples > 1:
            return self._simulate_lockstep(input_state, particles, n_samples)
        samples = []
        for _ in range(max(0, int(samples_number))):
            sample = np.zeros(input_state.shape[0], dtype=np.int64)
            current_input = np.zeros(input_state.shape[0], dtype=np.int64)
            self._fill_r_sample(sample, current_input, particles.copy(), self._photons(input_state))
            samples.append(sample)
        return tuple(samples)

    def _simulate_lockstep(self, input_state, particles, n_samples: int):
        """The samples advanced together: one column-minors launch per photon step covers
        every sample still filling.  The random numbers depend on the seed alone (the
        photon count, which particle enters at each step, the uniform draw of each step),
        never on the permanents, so they are taken first in the order of the sequential
        loop; the samples are those of the sequential loop."""
        U = self.interferometer_matrix
        m = input_state.shape[0]
        inputs, draws = [], []  # per sample: input state after each step, uniform draw of each step
        for _ in range(n_samples):
            photons = self._photons(input_state)
            working = particles.copy()
            current = np.zeros(m, dtype=np.int64)
            ins, ds = [], []
            for _k in range(photons):
                working = update_input_by_single_photon(current, working)
                ins.append(current.copy())
                ds.append(rng_gen.random_double_0_1())
            inputs.append(ins)
            draws.append(ds)
        samples = [np.zeros(m, dtype=np.int64) for _ in range(n_samples)]
        for k in range(max((len(d) for d in draws), default=0)):
            active = [s for s in range(n_samples) if len(draws[s]) > k]
            expanded = [_expanded(U, samples[s], inputs[s][k]) for s in active]
            minors = C.permanent_minors_batch(np.stack([E for E, _ in expanded]))
            for a, s in enumerate(active):
                current = inputs[s][k]
                col_idx = np.flatnonzero(current > 0)
                addends = _addends_from_minors(minors[a], expanded[a][1], col_idx)
                pmf = _pmf_from_addends(U, samples[s], current, col_idx, addends)
                _sample_with_draw(samples[s], pmf, draws[s][k])
        return tuple(samples)


class GeneralizedCliffordsBUniformLossesSimulationStrategy(GeneralizedCliffordsBSimulationStrategy):
    """GeneralizedCliffordsSimulationStrategy.py:50-64 /
    CGeneralizedCliffordsBUniformLossesSimulationStrategy.cpp: transmissivity t, each photon
    survives with probability t^2."""

    def __init__(self, interferometer_matrix, transmissivity, seed=None, lib: int = 0, permanent_batch=None,
                 pmf_method: str = "minors"):
        self.transmissivity = float(transmissivity)
        super().__init__(interferometer_matrix, seed=seed, lib=lib, permanent_batch=permanent_batch,
                         pmf_method=pmf_method)

    def _photons(self, input_state) -> int:
        """calculate_current_photon_number (:285-296) over the binomial weights (:298-316)."""
        n = int(np.sum(input_state))
        eta = self.transmissivity * self.transmissivity
        from math import comb

        probs = [comb(n, i) * eta ** i * (1.0 - eta) ** (n - i) for i in range(n + 1)]
        rand_num = rng_gen.random_double_0_1()
        photon_number = 0
        weight = probs[0]
        while weight < rand_num and weight < 1.0:
            photon_number += 1
            weight += probs[photon_number]
        return photon_number
