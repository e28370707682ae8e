"""torch.autograd over the adjoint sweep (piquassoboost_amd.autograd).

* Complex-autograd convention: gradients through ``permanent`` equal those
  torch computes itself through a differentiable Ryser permanent (n <= 6,
  CPU, complex128), for holomorphic and non-holomorphic losses and batch
  dimensions.
* Chaining: phase gradients of a Clements mesh (m = 8) through both
  permanents agree; for blocks of the gradient workload (top-left n=20 and
  n=26 blocks of Clements-decomposed 50-mode Haar unitaries, loss
  sum log|perm|^2), the phase gradient agrees with central finite
  differences (h = 1e-6) of the loss evaluated with the forward pass."""

import numpy as np
import pytest

from conftest import record_parity

torch = pytest.importorskip("torch")

from piquassoboost_amd.autograd import permanent  # noqa: E402
from piquassoboost_amd.interferometer import decompose, unitary_torch  # noqa: E402
from piquassoboost_amd.sampling.utilities import gradient_blocks, haar_unitary  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def ryser_torch(A):
    """perm(A) = (-1)^n sum_{S nonempty} (-1)^|S| prod_i sum_{j in S} a_ij (differentiable)."""
    n = A.shape[-1]
    total = torch.zeros(A.shape[:-2], dtype=A.dtype, device=A.device)
    for S in range(1, 1 << n):
        cols = [j for j in range(n) if (S >> j) & 1]
        total = total + (-1) ** len(cols) * torch.prod(A[..., :, cols].sum(dim=-1), dim=-1)
    return (-1) ** n * total


def test_matches_torch_complex_autograd():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((2, 3, 6, 6)) + 1j * rng.standard_normal((2, 3, 6, 6))
    w = torch.tensor(rng.standard_normal((2, 3)) + 1j * rng.standard_normal((2, 3)))

    def losses(p, ww):
        return [torch.log(p.abs() ** 2).sum(), (ww * p).real.sum() + (p * p.conj()).real.sum()]

    for k in range(2):
        A_ref = torch.tensor(X, requires_grad=True)
        losses(ryser_torch(A_ref), w)[k].backward()
        A = torch.tensor(X, device=DEV, requires_grad=True)
        p = permanent(A)
        assert torch.allclose(p.detach().cpu(), ryser_torch(torch.tensor(X)), rtol=1e-12, atol=0)
        losses(p, w.to(DEV))[k].backward()
        g, gr = A.grad.cpu(), A_ref.grad
        assert (g - gr).abs().max() <= 1e-11 * gr.abs().max(), k


def test_no_grad_forward_matches_grad_forward():
    rng = np.random.default_rng(4)
    X = torch.tensor(rng.standard_normal((5, 9, 9)) + 1j * rng.standard_normal((5, 9, 9)), device=DEV)
    with torch.no_grad():
        p0 = permanent(X)
    p1 = permanent(X.clone().requires_grad_(True))
    assert torch.allclose(p0, p1.detach(), rtol=1e-12, atol=0)


def test_rejects_cpu_and_wrong_dtype():
    with pytest.raises(ValueError):
        permanent(torch.zeros((3, 3), dtype=torch.complex128))
    with pytest.raises(TypeError):
        permanent(torch.zeros((3, 3), dtype=torch.complex64, device=DEV))


def _blocks(U, index):
    R = torch.tensor(np.stack([r for r, c in index]), device=U.device)
    Cc = torch.tensor(np.stack([c for r, c in index]), device=U.device)
    return U[R[:, :, None], Cc[:, None, :]]


def test_clements_chain_matches_torch_ryser():
    """m = 8 mesh, three 4 x 4 blocks: d loss / d (theta, phi, alpha) through
    the engine (GPU) and through torch's own Ryser (CPU) agree."""
    M = decompose(haar_unitary(8, 12))
    rng = np.random.default_rng(13)
    index = [(np.sort(rng.choice(8, 4, replace=False)), np.sort(rng.choice(8, 4, replace=False))) for _ in range(3)]
    grads = []
    for dev, perm_fn in ((torch.device("cpu"), ryser_torch), (DEV, permanent)):
        th = torch.tensor(M.theta, device=dev, requires_grad=True)
        ph = torch.tensor(M.phi, device=dev, requires_grad=True)
        al = torch.tensor(M.alpha, device=dev, requires_grad=True)
        U = unitary_torch(M.modes, th, ph, al)
        loss = torch.log(perm_fn(_blocks(U, index)).abs() ** 2).sum()
        loss.backward()
        grads.append(torch.cat([th.grad, ph.grad, al.grad]).cpu())
    assert (grads[0] - grads[1]).abs().max() <= 1e-11 * grads[0].abs().max()


def test_clements_chain_finite_differences():
    """Gradient workload blocks b = 0 (n = 20) and b = 256 (n = 26): the
    50-mode Haar unitaries of seeds 1006 and 1262, Clements-decomposed and
    rebuilt on the GPU; loss = sum log|perm(top-left block)|^2; autograd
    phase gradients against central differences of the loss (h = 1e-6,
    forward pass only)."""
    blocks, seeds = gradient_blocks()
    cases = [(20, seeds[20][0], blocks[20][0]), (26, seeds[26][0], blocks[26][0])]
    meshes = [decompose(haar_unitary(50, s)) for _, s, _ in cases]
    params = [[torch.tensor(getattr(M, k), device=DEV, requires_grad=True) for k in ("theta", "phi", "alpha")]
              for M in meshes]

    def loss_of(ps):
        total = 0.0
        for (n, _, _), M, (t, p, a) in zip(cases, meshes, ps):
            U = unitary_torch(M.modes, t, p, a)
            total = total + torch.log(permanent(U[:n, :n]).abs() ** 2)
        return total

    with torch.no_grad():  # the rebuilt unitaries reproduce the workload's blocks
        for (n, _, B), M, (t, p, a) in zip(cases, meshes, params):
            assert (unitary_torch(M.modes, t, p, a)[:n, :n].cpu() - torch.tensor(B)).abs().max() < 1e-12
    loss = loss_of(params)
    print(f"autograd chain: loss={loss.item():.6f}", flush=True)
    loss.backward()
    rng = np.random.default_rng(55)
    h = 1e-6
    worst = 0.0
    for m, M in enumerate(meshes):
        base = [getattr(M, k) for k in ("theta", "phi", "alpha")]
        for which, count in ((0, 3), (1, 3), (2, 2)):
            for k in rng.choice(len(base[which]), size=count, replace=False):
                vals = []
                for sgn in (1, -1):
                    ps = [[torch.tensor(getattr(Mm, kk), device=DEV) for kk in ("theta", "phi", "alpha")]
                          for Mm in meshes]
                    arr = base[which].copy()
                    arr[k] += sgn * h
                    ps[m][which] = torch.tensor(arr, device=DEV)
                    with torch.no_grad():
                        vals.append(loss_of(ps).item())
                fd = (vals[0] - vals[1]) / (2 * h)
                g = params[m][which].grad[k].item()
                err = abs(fd - g) / max(abs(g), 1.0)
                worst = max(worst, err)
                assert err < 1e-5, (m, which, int(k), fd, g)
    print(f"autograd chain finite differences: worst {worst:.3e}", flush=True)
    record_parity("config 5: Clements phases through autograd vs central differences", worst, 1e-5)
