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


def uneven_batch(seed: int = 1005, count: int = 4096, modes: int = 100, n_min: int = 8, n_max: int = 26,
                 mult_p=(0.7, 0.2, 0.1)):
    """Batch of boson-sampling transitions of uneven size on one Haar interferometer.

    Per task: n ~ U{n_min..n_max} photons in n distinct input modes (one each);
    output multiplicities drawn iid from {1, 2, 3} with probabilities mult_p,
    the last one truncated so that they sum to n, placed on distinct output
    modes.  Returns (U, [(input_state, output_state), ...]) over `modes` modes.
    """
    U = haar_unitary(modes, seed)
    rng = np.random.default_rng(seed + 1)
    tasks = []
    for _ in range(count):
        n = int(rng.integers(n_min, n_max + 1))
        inp = np.zeros(modes, dtype=np.int64)
        inp[rng.choice(modes, size=n, replace=False)] = 1
        mults = []
        left = n
        while left > 0:
            m = int(rng.choice([1, 2, 3], p=mult_p))
            m = min(m, left)
            mults.append(m)
            left -= m
        outp = np.zeros(modes, dtype=np.int64)
        outp[rng.choice(modes, size=len(mults), replace=False)] = mults
        tasks.append((inp, outp))
    return U, tasks


def gradient_blocks(seed: int = 1006, modes: int = 50, sizes=((20, 256), (26, 8))):
    """Square blocks for the gradient workload: block b (counted over `sizes`
    in order) is the top-left n x n block of its own Haar unitary
    haar_unitary(modes, seed + b).  Returns ({n: array count x n x n},
    {n: [seed of each block's unitary]})."""
    blocks, seeds = {}, {}
    b = 0
    for n, count in sizes:
        B = np.empty((count, n, n), dtype=np.complex128)
        s = []
        for k in range(count):
            B[k] = haar_unitary(modes, seed + b)[:n, :n]
            s.append(seed + b)
            b += 1
        blocks[n] = B
        seeds[n] = s
    return blocks, seeds


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
