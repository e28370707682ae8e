"""Particle-number measurement through the boosted samplers, lossless and lossy
(piquassoboost/sampling/calculations.py:54-160 without piquasso's state and
program objects: the interferometer, the input occupation numbers, the shot
count and the seed are passed directly).

Three branches, chosen as the reference chooses them (:108-140):

* lossless interferometer: the generalized Clifford & Clifford strategy on it;
* uniform losses (all singular values equal, and the uniform-loss speedup on):
  the strategy with binomial photon survival on the interferometer divided by its
  singular value, which is the transmissivity;
* non-uniform losses: the 2m-mode unitary dilation of the interferometer
  (piquasso's ``_prepare_interferometer_matrix_in_expanded_space``, restated
  below), the lossless strategy on it with the input padded by m empty modes,
  and the samples cut back to the first m modes (:144-145).

Every permanent comes from the HIP engine through the strategies.
"""

from __future__ import annotations

from collections import Counter
from fractions import Fraction

import numpy as np

from .simulation_strategies.GeneralizedCliffordsSimulationStrategy import (
    GeneralizedCliffordsBSimulationStrategy,
    GeneralizedCliffordsBUniformLossesSimulationStrategy,
)


def _is_uniform(values) -> bool:
    """calculations.py:49-50."""
    return bool(np.isclose(np.min(values), np.max(values)))


def calculate_singular_values_matrix_expansion(singular_values) -> np.ndarray:
    """The 2m x 2m middle factor of the dilation: [[diag(s), diag(c)], [diag(c), diag(s)]]
    with c = sqrt(max(1 - s^2, 0)).  Restated from piquasso 7.2.1 (the version
    piquassoboost pins, setup.py:37), sampling utilities, used by the reference at
    calculations.py:24-26 and :125.  Only its first m columns act on a state whose
    last m modes are empty, and those ([diag(s); diag(c)]) have orthonormal columns
    because s^2 + c^2 = 1."""
    s = np.asarray(singular_values, dtype=np.float64)
    c = np.sqrt(np.maximum(1.0 - s * s, 0.0))
    return np.block([[np.diag(s), np.diag(c)], [np.diag(c), np.diag(s)]])


def prepare_interferometer_matrix_in_expanded_space(interferometer_svd) -> np.ndarray:
    """2m-mode dilation of a lossy m-mode interferometer L = V diag(s) W (numpy's SVD
    triple (V, s, W)): diag(V, I) . expansion(s) . diag(W, I), whose top-left m x m block
    is L.  Restated from piquasso 7.2.1 (see calculate_singular_values_matrix_expansion)."""
    v, s, w = interferometer_svd
    m = v.shape[0]
    Z = np.zeros((m, m), dtype=np.complex128)
    I = np.eye(m, dtype=np.complex128)
    ev = np.block([[v, Z], [Z, I]])
    ew = np.block([[w, Z], [Z, I]])
    return ev @ calculate_singular_values_matrix_expansion(s) @ ew


def sample(interferometer, initial_state, shots: int, seed=None, speedup_uniform: bool = True,
           strategy_class=GeneralizedCliffordsBSimulationStrategy, **strategy_kwargs) -> np.ndarray:
    """`shots` samples (rows of mode occupations, the order the strategy draws them)
    of a particle-number measurement after `interferometer` on `initial_state`."""
    L = np.asarray(interferometer, dtype=np.complex128)
    initial_state = np.asarray(initial_state, dtype=np.int64)
    m = initial_state.shape[0]
    svd = np.linalg.svd(L)
    s = svd[1]
    lossy = not bool(np.allclose(s, 1.0))
    if lossy and speedup_uniform and _is_uniform(s):
        strategy = GeneralizedCliffordsBUniformLossesSimulationStrategy(L / s[0], s[0], seed, **strategy_kwargs)
        state = initial_state
    elif lossy:
        strategy = strategy_class(prepare_interferometer_matrix_in_expanded_space(svd), seed, **strategy_kwargs)
        state = np.append(initial_state, np.zeros_like(initial_state))
    else:
        strategy = strategy_class(L, seed, **strategy_kwargs)
        state = initial_state
    out = np.array(strategy.simulate(state, shots), dtype=np.int64).reshape(-1, state.shape[0])
    return out[:, :m]


def particle_number_measurement(interferometer, initial_state, shots: int, seed=None, **kwargs) -> list:
    """The reference's measurement result: one (outcome, frequency) pair per distinct
    outcome, frequencies as Fractions of `shots` (calculations.py:146-150)."""
    samples = sample(interferometer, initial_state, shots, seed, **kwargs)
    counts = Counter(map(tuple, samples.tolist()))
    return [(outcome, Fraction(count, shots)) for outcome, count in counts.items()]
