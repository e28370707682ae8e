"""Clements mesh (host only): decomposition of Haar unitaries into
m(m-1)/2 beam splitters and output phases, and the numpy and torch
reconstructions of U from those phases."""

import numpy as np
import pytest

from piquassoboost_amd.interferometer import Mesh, decompose, unitary, unitary_torch
from piquassoboost_amd.sampling.utilities import haar_unitary


@pytest.mark.parametrize("m", [2, 3, 4, 5, 6, 7, 8, 12, 50])
def test_round_trip(m):
    U = haar_unitary(m, 100 + m)
    M = decompose(U)
    assert M.modes.shape == (m * (m - 1) // 2,) and M.alpha.shape == (m,)
    assert M.modes.min() >= 0 and M.modes.max() <= m - 2
    assert np.abs(unitary(M) - U).max() < 1e-13


def test_rectangular_layout_depth():
    """Each mode pair (k, k+1) carries m/2 or (m-1)/2 beam splitters, the
    balanced rectangular arrangement of the Clements layout."""
    m = 10
    M = decompose(haar_unitary(m, 5))
    counts = np.bincount(M.modes, minlength=m - 1)
    assert set(counts.tolist()) <= {m // 2, (m - 1) // 2, (m + 1) // 2}
    assert counts.sum() == m * (m - 1) // 2


def test_identity_and_permutation_like():
    for U in (np.eye(4, dtype=complex), np.diag(np.exp(1j * np.arange(5)))):
        assert np.abs(unitary(decompose(U)) - U).max() < 1e-13


def test_torch_matches_numpy():
    torch = pytest.importorskip("torch")
    U = haar_unitary(9, 3)
    M = decompose(U)
    W = unitary_torch(M.modes, torch.tensor(M.theta), torch.tensor(M.phi), torch.tensor(M.alpha))
    assert np.abs(W.numpy() - U).max() < 1e-13


def test_torch_phase_gradient_matches_finite_differences():
    """d/dtheta of a real function of U through the torch mesh (CPU, no permanent)."""
    torch = pytest.importorskip("torch")
    U = haar_unitary(6, 8)
    M = decompose(U)
    w = torch.tensor(np.random.default_rng(1).standard_normal((6, 6)))

    def f(th, ph, al):
        W = unitary_torch(M.modes, th, ph, al)
        return (w * W.real).sum() + (w.T * W.imag).sum()

    th = torch.tensor(M.theta, requires_grad=True)
    ph = torch.tensor(M.phi, requires_grad=True)
    al = torch.tensor(M.alpha, requires_grad=True)
    f(th, ph, al).backward()
    h = 1e-6
    for k in range(len(M.theta)):
        for src, g in ((M.theta, th.grad), (M.phi, ph.grad)):
            e = np.zeros_like(src)
            e[k] = h
            args = [torch.tensor(M.theta), torch.tensor(M.phi), torch.tensor(M.alpha)]
            which = 0 if src is M.theta else 1
            args[which] = torch.tensor(src + e)
            fp = f(*args).item()
            args[which] = torch.tensor(src - e)
            fm = f(*args).item()
            assert abs((fp - fm) / (2 * h) - g[k].item()) < 1e-7
