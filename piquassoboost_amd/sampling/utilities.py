"""Input builders for the permanent calculators.

* ``haar_unitary`` -- seeded Haar-random unitary (complex Gaussian, QR,
  phase fix), the synthetic interferometer every benchmark workload uses;
* ``adapt_interferometer`` / ``adapt_interferometer_glynn_multiplied`` --
  the compact and expanded effective matrices of a boson-sampling transition,
  as BatchedPermanentCalculator.cpp:412-451 and :367-399 build them.
"""

from __future__ import annotations

import numpy as np


def haar_unitary(m: int, seed: int) -> np.ndarray:
    """Haar-random m x m unitary: Z = (N + iN)/sqrt(2), Z = QR, U = Q diag(R_ii/|R_ii|)."""
    rng = np.random.default_rng(seed)
    Z = (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))) / np.sqrt(2.0)
    Q, R = np.linalg.qr(Z)
    d = np.diag(R)
    return np.ascontiguousarray(Q * (d / np.abs(d))[None, :])


def adapt_interferometer(U, input_state, output_state) -> np.ndarray:
    """Rows of occupied output modes x columns of occupied input modes."""
    U = np.asarray(U)
    rows = np.flatnonzero(np.asarray(output_state) > 0)
    cols = np.flatnonzero(np.asarray(input_state) > 0)
    return np.ascontiguousarray(U[np.ix_(rows, cols)], dtype=np.complex128)


def adapt_interferometer_glynn_multiplied(U, input_state, output_state) -> np.ndarray:
    """Row i repeated output_state[i] times, column j repeated input_state[j] times."""
    U = np.asarray(U)
    r = np.repeat(np.arange(U.shape[0]), np.asarray(output_state, dtype=np.int64))
    c = np.repeat(np.arange(U.shape[1]), np.asarray(input_state, dtype=np.int64))
    return np.ascontiguousarray(U[np.ix_(r, c)], dtype=np.complex128)
