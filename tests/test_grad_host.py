"""Host-side check of the adjoint sweep's decomposition (no device).

A numpy restatement of the GPU kernel's schedule -- the bit split and wave
ranges come from pqb_grad_plan, the walk is the reflected Gray code over the
register and high bits, delta lanes carry the lane bits, and G is accumulated
by summation by parts exactly as the kernel does (register rows per lane,
high rows per wave at block transitions, lane rows and row 0 at the end) --
is compared with the permanents of the minors from the oracle.  It pins the
plan and the summation-by-parts bookkeeping independently of the GPU."""

import numpy as np
import pytest

import oracle
from conftest import random_complex
from piquassoboost_amd import _capi as C


def _ctz(x):
    return (x & -x).bit_length() - 1


def _leave_one_out(c):
    n = c.shape[1]
    pre = np.ones_like(c)
    suf = np.ones_like(c)
    for j in range(1, n):
        pre[:, j] = pre[:, j - 1] * c[:, j - 1]
    for j in range(n - 2, -1, -1):
        suf[:, j] = suf[:, j + 1] * c[:, j + 1]
    return pre * suf


def adjoint_schedule(A, K, L, H, wpt, drop_high_end=False):
    """perm(A) and G as the kernel computes them (bit b <-> row n-1-b; bits
    [0,K) register, [K,K+L) lane, [K+L,n-1) high).  drop_high_end leaves out
    the delta_h(end) T(end) term of the high rows (a planted mistake)."""
    R = n = A.shape[0]
    nl = 1 << L
    G = np.zeros((n, n), complex)
    perm = 0j
    NB = 1 << H
    for wi in range(wpt):
        g0, g1 = NB * wi // wpt, NB * (wi + 1) // wpt
        if g0 >= g1:
            continue
        tau0 = g0 << K
        gam = tau0 ^ (tau0 >> 1)
        delta = np.ones((nl, R))
        for p in range(K + H):
            bit = p if p < K else p + L
            if (gam >> p) & 1:
                delta[:, R - 1 - bit] = -1
        for d in range(nl):
            for k in range(L):
                if (d >> k) & 1:
                    delta[d, R - 1 - (K + k)] = -1
        s_lane = np.prod(delta, axis=1)
        c = delta @ A
        T = np.zeros((nl, n), complex)
        Z = np.zeros((K, nl, n), complex)
        Zh = np.zeros((H, n), complex)
        acc = np.zeros(nl, complex)
        steps = (g1 - g0) << K
        for it in range(steps):
            tau = tau0 + it
            if it > 0:
                p = _ctz(tau)
                g_new = tau ^ (tau >> 1)
                dnew = 1 - 2 * ((g_new >> p) & 1)
                if p < K:
                    row = R - 1 - p
                    Z[p] -= 2 * dnew * T
                else:
                    h = p - K
                    row = R - 1 - (h + K + L)
                    Zh[h] -= 2 * dnew * T.sum(0)
                c = c + 2 * dnew * A[row]
            s = s_lane * (-1) ** it
            T += s[:, None] * _leave_one_out(c)
            acc += s * np.prod(c, axis=1)
        tau_end = tau0 + steps - 1
        gam = tau_end ^ (tau_end >> 1)
        Ts = T.sum(0)
        G[0] += Ts
        for k in range(L):
            sg = np.array([-1 if (d >> k) & 1 else 1 for d in range(nl)])
            G[R - 1 - (K + k)] += (sg[:, None] * T).sum(0)
        for p in range(K):
            G[R - 1 - p] += Z[p].sum(0) + (1 - 2 * ((gam >> p) & 1)) * Ts
        for h in range(H):
            end = 0.0 if drop_high_end else (1 - 2 * ((gam >> (h + K)) & 1)) * Ts
            G[R - 1 - (h + K + L)] += Zh[h] + end
        perm += acc.sum()
    sc = 2.0 ** -(R - 1)
    return perm * sc, G * sc


def _perm(M):
    """Brute force up to 7 x 7, Ryser (double-double) above: both independent of Glynn."""
    if M.shape[0] == 0:
        return 1.0
    return oracle.bruteforce(M) if M.shape[0] <= 7 else oracle.ryser(M, "dd")


def minors(A):
    n = A.shape[0]
    return np.array([[_perm(np.delete(np.delete(A, i, 0), j, 1)) for j in range(n)] for i in range(n)])


@pytest.mark.parametrize("n", list(range(2, 12)))
@pytest.mark.parametrize("waves_resident", [4, 64])
def test_planned_schedule_matches_minors(n, waves_resident):
    """The kernel's own plan (pqb_grad_plan) for one matrix."""
    plan = C.grad_plan(1, n, waves_resident)
    rng = np.random.default_rng(n)
    A = random_complex(rng, n, n, scale=0.6)
    p, G = adjoint_schedule(A, plan["K"], plan["L"], plan["H"], plan["wpt"])
    ref = minors(A)
    P = _perm(A)
    assert abs(p - P) <= 1e-13 * abs(P)
    assert np.abs(G - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("n", [5, 7, 8])
def test_every_bit_split_matches_minors(n):
    """Any split of the n-1 bits into register, lane and high bits, and wave
    counts that do not divide the number of blocks."""
    rng = np.random.default_rng(10 + n)
    A = random_complex(rng, n, n, scale=0.6)
    ref = minors(A)
    for L in range(0, min(4, n - 1) + 1):
        for K in range(0, min(3, n - 1 - L) + 1):
            H = n - 1 - L - K
            for wpt in (1, 3, 5):
                _, G = adjoint_schedule(A, K, L, H, wpt)
                assert np.abs(G - ref).max() <= 1e-13 * np.abs(ref).max(), (K, L, H, wpt)


def test_planted_mistake_is_caught():
    """Leaving out the high rows' delta(end) T(end) term breaks the agreement."""
    n = 9
    rng = np.random.default_rng(1)
    A = random_complex(rng, n, n)
    ref = minors(A)
    K, L = 1, 2
    H = n - 1 - K - L
    _, G = adjoint_schedule(A, K, L, H, 3)
    _, G_bad = adjoint_schedule(A, K, L, H, 3, drop_high_end=True)
    assert np.abs(G - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.abs(G_bad - ref).max() > 1e-3 * np.abs(ref).max()
