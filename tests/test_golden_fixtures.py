"""The full-size fixtures against the workload generators and the mathematics
(CPU only): the 4096-task batch fixture holds exactly the benchmark's
transitions and, on a sample, the oracle's values; the gradient-workload minors
satisfy the Laplace expansion perm(A) = sum_j a_ij perm(minor_ij) along every
stored row and column, with perm(A) computed independently by the oracle."""

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, rel_err
from piquassoboost_amd.sampling.utilities import adapt_interferometer, gradient_blocks, uneven_batch


def _dd(a):
    return (a[..., 0] + a[..., 1]) + 1j * (a[..., 2] + a[..., 3])


@pytest.fixture(scope="module")
def uneven():
    return np.load(GOLDEN / "uneven4096.npz")


def test_uneven_fixture_is_the_benchmark_workload(uneven):
    U, pairs = uneven_batch(seed=1005, count=4096)
    assert np.allclose(uneven["U"], U, rtol=0, atol=1e-13)
    for k in range(0, 4096, 7):
        i, o = pairs[k]
        im = uneven["in_modes"][k]
        assert np.array_equal(np.flatnonzero(i), im[im >= 0])
        om = uneven["out_modes"][k]
        sel = om >= 0
        assert np.array_equal(np.flatnonzero(o), om[sel])
        assert np.array_equal(o[np.flatnonzero(o)], uneven["out_mult"][k][sel])
        assert int(i.sum()) == int(uneven["out_mult"][k].sum())


def test_uneven_fixture_values_sample(uneven):
    """40 tasks recomputed with the oracle in long double (a different accumulation
    type from the stored double-double values) agree to 1e-12."""
    U = uneven["U"]
    ref = _dd(uneven["value_dd"])
    rng = np.random.default_rng(11)
    for k in rng.choice(4096, size=40, replace=False):
        im = uneven["in_modes"][k]
        im = im[im >= 0]
        om = uneven["out_modes"][k]
        sel = om >= 0
        i = np.zeros(U.shape[0], np.int64)
        i[im] = 1
        o = np.zeros(U.shape[0], np.int64)
        o[om[sel]] = uneven["out_mult"][k][sel]
        A = adapt_interferometer(U, i, o)
        v = oracle.bbfg_repeated(A, i[im], o[om[sel]], "long double")
        assert rel_err(v, ref[k]) < 1e-12, k


@pytest.fixture(scope="module")
def minors():
    return np.load(GOLDEN / "grad_minors.npz")


def test_grad_fixture_blocks_are_the_workload(minors):
    blocks, _ = gradient_blocks()
    assert np.array_equal(minors["b26"], blocks[26])
    assert np.array_equal(minors["b20"], blocks[20][:16])


def test_grad_fixture_n26_laplace(minors):
    """Row 0 and column 0 of the n=26 minors: sum_j a_0j M_0j = sum_i a_i0 M_i0 =
    perm(A) for two of the blocks (perm from the oracle's long-double sweep), and the
    two expansions agree with each other on all eight."""
    B = minors["b26"]
    row0, col0 = _dd(minors["b26_row0_dd"]), _dd(minors["b26_col0_dd"])
    assert np.array_equal(row0[:, 0], col0[:, 0])
    for b in range(B.shape[0]):
        lr = np.dot(B[b][0], row0[b])
        lc = np.dot(B[b][:, 0], col0[b])
        assert rel_err(lr, lc) < 1e-11, b
        if b < 2:
            assert rel_err(lr, oracle.bbfg(B[b], "long double")) < 1e-11, b


def test_grad_fixture_n20_laplace(minors):
    """Every row and column expansion of the 16 stored n=20 minor matrices equals
    perm(A) from the oracle."""
    B = minors["b20"]
    M = _dd(minors["b20_minors_dd"])
    for b in range(B.shape[0]):
        p = oracle.bbfg(B[b], "dd")
        for i in range(20):
            assert rel_err(np.dot(B[b][i], M[b][i]), p) < 1e-11, (b, "row", i)
            assert rel_err(np.dot(B[b][:, i], M[b][:, i]), p) < 1e-11, (b, "col", i)
