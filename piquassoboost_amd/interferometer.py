"""Clements mesh of an m-mode interferometer (Clements et al., Optica 3, 1460
(2016), "Optimal design for universal multiport interferometers").

A unitary U is written as

    U = D  T_{k_1}(theta_1, phi_1)  T_{k_2}(theta_2, phi_2) ... T_{k_K}(theta_K, phi_K),

K = m(m-1)/2 beam splitters in the rectangular layout, D = diag(exp(i alpha)),
where T_k acts on modes (k, k+1) as

    [[exp(i phi) cos theta, -sin theta],
     [exp(i phi) sin theta,  cos theta]].

``decompose`` finds the phases of a given unitary with the paper's nulling
sequence (alternately from the right, T^-1 on columns, and from the left, T
on rows), then moves the left-hand factors through the diagonal
(T^-1(theta, phi) D = D' T(theta, phi') with phi' = arg(-d_k / d_{k+1}),
d'_k = -exp(-i phi) d_{k+1}).  ``unitary`` rebuilds U from the phases in
numpy, ``unitary_torch`` in torch so that gradients reach theta, phi, alpha.
The reference has no interferometer model (piquasso, which holds one, is not
part of it); this is the caller a training loop puts in front of the
permanent.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Mesh:
    modes: np.ndarray   # (K,) int: beam splitter k acts on (modes[k], modes[k] + 1)
    theta: np.ndarray   # (K,)
    phi: np.ndarray     # (K,)
    alpha: np.ndarray   # (m,) output phases


def _T(m: int, k: int, theta: float, phi: float) -> np.ndarray:
    T = np.eye(m, dtype=np.complex128)
    c, s = np.cos(theta), np.sin(theta)
    T[k, k] = np.exp(1j * phi) * c
    T[k, k + 1] = -s
    T[k + 1, k] = np.exp(1j * phi) * s
    T[k + 1, k + 1] = c
    return T


def decompose(U) -> Mesh:
    U = np.array(U, dtype=np.complex128)
    m = U.shape[0]
    left, right = [], []
    for i in range(m - 1):
        if i % 2 == 0:
            for j in range(i + 1):  # null (m-1-j, i-j) with T^-1 on columns (i-j, i-j+1)
                k, r = i - j, m - 1 - j
                a, b = U[r, k], U[r, k + 1]
                theta = np.arctan2(abs(a), abs(b))
                phi = np.angle(a) - np.angle(b)
                U = U @ _T(m, k, theta, phi).conj().T
                right.append((k, theta, phi))
        else:
            for j in range(i + 1):  # null (k+1, j) with T on rows (k, k+1)
                k = m - 2 - i + j
                a, b = U[k, j], U[k + 1, j]
                theta = np.arctan2(abs(b), abs(a))
                phi = np.angle(-b) - np.angle(a)
                U = _T(m, k, theta, phi) @ U
                left.append((k, theta, phi))
    d = np.diag(U).copy()
    # U_orig = T_l1^-1 ... T_lK^-1 D T_rR ... T_r1: move the inverses through D
    moved = []
    for k, theta, phi in reversed(left):
        d1, d2 = d[k], d[k + 1]
        phi2 = np.angle(-d1 / d2)
        d[k] = -np.exp(-1j * phi) * d2
        moved.append((k, theta, phi2))
    seq = list(reversed(moved)) + list(reversed(right))
    return Mesh(
        modes=np.array([s[0] for s in seq], dtype=np.int64),
        theta=np.array([s[1] for s in seq]),
        phi=np.array([s[2] for s in seq]),
        alpha=np.angle(d),
    )


def unitary(mesh: Mesh) -> np.ndarray:
    m = mesh.alpha.shape[0]
    U = np.diag(np.exp(1j * mesh.alpha))
    for k, theta, phi in zip(mesh.modes, mesh.theta, mesh.phi):
        c, s, e = np.cos(theta), np.sin(theta), np.exp(1j * phi)
        a, b = U[:, k].copy(), U[:, k + 1].copy()
        U[:, k] = a * e * c + b * e * s
        U[:, k + 1] = -a * s + b * c
    return U


def unitary_torch(modes, theta, phi, alpha):
    """U(theta, phi, alpha) as a differentiable complex128 torch tensor on the
    device of `theta`; `modes` is a sequence of ints."""
    import torch

    one = torch.ones((), dtype=torch.complex128, device=theta.device)
    U = torch.diag(torch.exp(1j * alpha.to(torch.complex128)))
    for idx, k in enumerate(modes):
        k = int(k)
        c = torch.cos(theta[idx]) * one
        s = torch.sin(theta[idx]) * one
        e = torch.exp(1j * phi[idx].to(torch.complex128))
        a, b = U[:, k], U[:, k + 1]
        na = a * e * c + b * e * s
        nb = -a * s + b * c
        U = torch.cat([U[:, :k], na[:, None], nb[:, None], U[:, k + 2:]], dim=1)
    return U
