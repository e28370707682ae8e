"""Host model of the plain sweep's schedule (perm_plain.hip, plain_sweep_kernel):
with the kernel's own plan (pqb_plain_plan, host only), every wave's run of
the walk index, its seed at an arbitrary even walk position, the walk and the
group transitions are replayed in numpy for each of the 64 lanes.  Checks that
every sign vector of the Glynn/BBFG sum (BBFGPermanentCalculator.hpp:87-235)
is visited exactly once with the right sign, and that the replayed sum is the
permanent of the oracle.  No device is needed."""

import numpy as np
import pytest

from piquassoboost_amd import _capi as C

import oracle


def replay(A, plan):
    """Sum over all waves and lanes, as the kernel forms it, and the visit count
    of every sign vector (indexed by its row signs, bit i set <=> delta_i = -1)."""
    R, NC = A.shape
    K, B, H, nwaves = plan["K"], plan["B"], plan["H"], plan["waves"]
    W = 1 << K
    M2p, M2m = 2.0 * A, -2.0 * A
    step = 2 if K >= 1 else 1
    units = (1 << (H + K)) // step
    lanes = np.arange(64)
    lane_on = lanes < (1 << B)
    visits = np.zeros(1 << R, dtype=np.int64)
    total = 0j

    def code(g, t):
        """delta bits per lane for group g, walk position t (bit i <=> row i at -1)."""
        gray = t ^ (t >> 1)
        bits = np.zeros(64, dtype=np.int64)
        for b in range(K):
            if (gray >> b) & 1:
                bits |= 1 << (R - 1 - b)
        for b in range(B):
            bits |= ((lanes >> b) & 1) << (R - 1 - K - b)
        for h in range(H):
            if (g >> h) & 1:
                bits |= 1 << (R - 1 - K - B - h)
        return bits

    for wave in range(nwaves):
        tau0 = (units * wave // nwaves) * step
        tau1 = (units * (wave + 1) // nwaves) * step
        if tau0 >= tau1:
            continue
        g, t0 = tau0 >> K, tau0 & (W - 1)
        assert K == 0 or t0 % 2 == 0
        # seed (kernel: +2A row 0, walk rows by the Gray code of t0, lane rows, group rows, halved)
        c = np.tile(M2p[0], (64, 1)).astype(complex)
        gray0 = t0 ^ (t0 >> 1)
        for i in range(R - K, R):
            c += M2m[i] if (gray0 >> (R - 1 - i)) & 1 else M2p[i]
        for b in range(B):
            s = np.where((lanes >> b) & 1, -1.0, 1.0)[:, None]
            c += s * M2p[R - 1 - K - b]
        for h in range(H):
            c += M2m[R - 1 - K - B - h] if (g >> h) & 1 else M2p[R - 1 - K - B - h]
        c *= 0.5
        tr = np.zeros(64, dtype=complex)
        while True:
            gbase = g << K
            t_stop = W if tau1 - gbase >= W else tau1 - gbase
            pm = [np.zeros(64, dtype=complex), np.zeros(64, dtype=complex)]
            t = t0
            pm[0] += np.prod(c, axis=1)
            np.add.at(visits, code(g, t)[lane_on], 1)
            if K >= 1:
                c += M2p[R - 1] if (t0 >> 1) & 1 else M2m[R - 1]
                pm[1] += np.prod(c, axis=1)
                np.add.at(visits, code(g, t0 + 1)[lane_on], 1)
                for t in range(t0 + 2, t_stop, 2):
                    b = (t & -t).bit_length() - 1
                    nb = ((t >> b) ^ (t >> (b + 1))) & 1
                    c += M2m[R - 1 - b] if nb else M2p[R - 1 - b]
                    pm[0] += np.prod(c, axis=1)
                    np.add.at(visits, code(g, t)[lane_on], 1)
                    c += M2p[R - 1] if (t >> 1) & 1 else M2m[R - 1]
                    pm[1] += np.prod(c, axis=1)
                    np.add.at(visits, code(g, t + 1)[lane_on], 1)
            gs = -1.0 if bin(g).count("1") & 1 else 1.0
            tr += gs * (pm[0] - pm[1])
            if gbase + t_stop >= tau1:
                break
            g += 1
            t0 = 0
            if K >= 1:
                c += M2p[R - K]
            diff = (g - 1) ^ g
            for h in range(H):
                if (diff >> h) & 1:
                    c += M2m[R - 1 - K - B - h] if (g >> h) & 1 else M2p[R - 1 - K - B - h]
        ls = np.where(np.array([bin(x).count("1") & 1 for x in lanes]), -1.0, 1.0)
        total += np.sum(np.where(lane_on, ls * tr, 0.0))
    return total / 2.0 ** (R - 1), visits


@pytest.mark.parametrize("n,wr", [(4, 3072), (7, 5), (9, 13), (10, 3072), (11, 96), (12, 24), (12, 7)])
def test_every_sign_vector_once_and_the_sum(n, wr):
    rng = np.random.default_rng(n * 100 + wr)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    plan = C.plain_plan(n, n, waves_resident=wr)
    assert plan["waves"] <= wr
    got, visits = replay(A, plan)
    # delta_0 = +1: exactly the 2^(n-1) vectors with bit 0 clear, each once
    assert visits[0::2].tolist() == [1] * (1 << (n - 1))
    assert not visits[1::2].any()
    ref = oracle.bruteforce(A) if n <= 8 else oracle.ryser(A, "dd")
    assert abs(got - ref) <= 1e-9 * max(1.0, abs(ref))


def test_runs_start_inside_groups():
    """At the n=30 shape the runs do not align with groups (3072 resident waves,
    2^13 groups of 2^10 steps): the seed at an odd pair offset is exercised
    on a small matrix with the same kind of split."""
    plan = C.plain_plan(12, 12, waves_resident=24)
    units = (1 << (plan["H"] + plan["K"])) // 2
    starts = {(units * w // plan["waves"]) * 2 % (1 << plan["K"]) for w in range(plan["waves"])}
    assert any(s != 0 for s in starts)
    A = np.random.default_rng(5).normal(size=(12, 12)) + 0j
    got, visits = replay(A, plan)
    assert visits[0::2].min() == 1 and visits[0::2].max() == 1
    assert abs(got - oracle.ryser(A, "dd")) <= 1e-9 * max(1.0, abs(oracle.ryser(A, "dd")))
