"""Writes the golden fixtures used by the parity tests (run in the build container).

    python tests/golden/make_golden.py [--only NAME]

Inputs are the seeded workloads of the benchmark configurations; because
LAPACK's QR is not bit-stable across builds, the matrices themselves are
stored, not just their seeds.  Expected values come from the CPU oracle
(oracle/pqb_oracle.cpp): the restated reference sweep in long double and in
double-double, and exact 128-bit integer permanents (Ryser) for 0/1 matrices.
The known-answer cases are the data of the reference's unit tests
(unittests/test_permanent.py:47-251).
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from piquassoboost_amd.sampling.utilities import haar_unitary  # noqa: E402


def _c(v):
    return [float(np.real(v)), float(np.imag(v))]


def kat():
    """Known-answer cases of unittests/test_permanent.py (matrix, input, output, value)."""
    cases = []
    cases.append(dict(name="trivial_1x1", matrix=[[4.2]], input=[1], output=[1], value=[4.2, 0.0]))  # :47-52
    U3 = [[0.62270314 + 0.55117657j, -0.0258677 - 0.07171713j, 0.09597446 - 0.54168404j],
          [0.34756795 - 0.29444499j, -0.43514701 + 0.18975153j, 0.71929752 + 0.22304973j],
          [-0.24500645 - 0.20227626j, -0.45222962 - 0.75121057j, 0.06995606 - 0.3540245j]]
    cases.append(dict(name="zero_input", matrix=U3, input=[0, 0, 0], output=[0, 0, 0], value=[1.0, 0.0]))  # :55-78
    U6 = [[0.62113733 - 0.01959968j, -0.15627468 - 0.0772489j, -0.2705819 - 0.18997122j,
           0.26504798 - 0.30838768j, 0.03372169 - 0.11154586j, 0.15278476 + 0.52137824j],
          [-0.1776024 - 0.21000195j, 0.18950753 + 0.20741494j, -0.15537846 + 0.19161071j,
           0.07400899 - 0.37578572j, -0.44458249 - 0.0047501j, -0.62212719 + 0.23055313j],
          [-0.05572001 - 0.20287464j, 0.22359337 + 0.30693557j, -0.13719319 + 0.23245719j,
           0.1102451 + 0.02659467j, 0.81942653 + 0.04327346j, -0.17215559 + 0.15114287j],
          [-0.24319645 - 0.44143551j, -0.50022937 - 0.08513718j, 0.07671116 - 0.05858231j,
           0.0679656 + 0.52109972j, -0.0482276 - 0.12736588j, -0.11768435 + 0.41307881j],
          [-0.29469977 - 0.20027018j, 0.22135149 - 0.02983563j, -0.18587346 - 0.83950064j,
           -0.21606625 - 0.14975436j, 0.11702974 - 0.02297493j, -0.01552763 + 0.01646485j],
          [-0.29741767 + 0.15644426j, -0.61959257 - 0.23497653j, 0.07397837 + 0.05367843j,
           -0.05838964 - 0.57132173j, 0.28736069 - 0.00798998j, -0.13763068 - 0.09058005j]]
    cases.append(dict(name="no_repetition_6x6", matrix=U6, input=[1] * 6, output=[1] * 6,
                      value=[0.022227325527358795, 0.01807052573717885]))  # :81-140
    U2 = (np.array([[1, 1j], [1, -1j]]) / np.sqrt(2)).tolist()
    cases.append(dict(name="asymmetric_2x2", matrix=U2, input=[0, 2], output=[2, 0], value=[-1.0, 0.0]))  # :143-149
    U4 = [[0.50142122 - 0.15131566j, 0.0265964 - 0.67076793j, -0.48706586 - 0.15221542j, 0.1148288 - 0.0381445j],
          [-0.30504953 + 0.02783877j, -0.33442264 + 0.27130083j, -0.52279136 - 0.6644152j, -0.04684954 + 0.06143215j],
          [0.20279518 + 0.57890235j, -0.47584393 - 0.16777288j, 0.08168919 + 0.10673884j, -0.34434342 + 0.48221603j],
          [0.31022502 - 0.39919511j, 0.21927998 + 0.24751057j, -0.06266598 - 0.05334444j, -0.78447544 + 0.11350836j]]
    # _perm(unitary, cols, rows): input = cols, output = rows (:182-187)
    cases.append(dict(name="repeated_4x4", matrix=U4, input=[0, 0, 3, 1], output=[3, 0, 0, 1],
                      value=[0.4302957973670928, 0.3986355418194044]))  # :152-188
    U6b = [[0.00456448 - 0.37857979j, 0.0795763 + 0.10465352j, 0.03365652 + 0.41152925j,
            0.05555206 + 0.18738511j, -0.39720048 + 0.1311295j, 0.40848797 + 0.537455j],
           [0.00901246 + 0.19215191j, 0.4609619 + 0.13507775j, -0.47801241 + 0.19203936j,
            0.23140442 - 0.62025569j, 0.00333535 + 0.14816728j, 0.0783928 + 0.02267907j],
           [0.43945864 - 0.04054749j, 0.00377335 - 0.30622153j, 0.22692424 - 0.43506116j,
            0.08710784 - 0.32486556j, -0.58818325 + 0.08888961j, 0.04507539 - 0.0406845j],
           [-0.29415733 + 0.17675044j, 0.33184779 + 0.39750559j, 0.34950358 - 0.26546825j,
            -0.45096811 - 0.10690984j, -0.12192328 + 0.20376462j, -0.237482 + 0.30640823j],
           [-0.02410776 + 0.30238262j, -0.46322751 + 0.20107186j, -0.09817621 - 0.15398772j,
            -0.29490926 - 0.28320581j, 0.02993443 - 0.3052611j, 0.5788743 + 0.15254606j],
           [0.53706284 + 0.35589712j, -0.26260555 + 0.25708476j, 0.23762169 + 0.20229218j,
            0.16489218 + 0.03524027j, 0.24701479 + 0.48757485j, -0.06188047 + 0.14647274j]]
    cases.append(dict(name="repeated_6x6", matrix=U6b, input=[2, 1, 3, 0, 1, 2], output=[1, 1, 0, 3, 2, 2],
                      value=[0.13160241373727416, 0.36535625772184577]))  # :191-251
    for c in cases:
        m = np.array(c["matrix"], dtype=np.complex128)
        c["matrix"] = [[_c(v) for v in row] for row in m]
    (HERE / "kat.json").write_text(json.dumps(cases, indent=1))
    print("kat.json", len(cases))


def haar_blocks():
    """n=20 block of a 40-mode Haar unitary (seed 1001); n=28, 30 blocks of a
    60-mode Haar unitary (seed 1002), with long double and double-double
    values of the restated reference sweep."""
    out = {}
    U40 = haar_unitary(40, 1001)
    U60 = haar_unitary(60, 1002)
    for name, U, n in (("c1_n20", U40, 20), ("c2a_n28", U60, 28), ("c2a_n30", U60, 30)):
        A = np.ascontiguousarray(U[:n, :n])
        t = time.time()
        ld, _ = oracle.bbfg(A, "long double", full=True)
        t1 = time.time()
        dd, ddp = oracle.bbfg(A, "dd", full=True)
        t2 = time.time()
        out[name] = A
        out[name + "_ld"] = np.array([ld.real, ld.imag])
        out[name + "_dd"] = np.array(ddp)  # (hi_re, lo_re, hi_im, lo_im)
        print(name, ld, dd, f"ld {t1 - t:.1f}s dd {t2 - t1:.1f}s", flush=True)
        if n <= 20:
            gr = oracle.glynn_recursive(A, "long double")
            out[name + "_glynn_ld"] = np.array([gr.real, gr.imag])
            print("  glynn recursive ld", gr, flush=True)
    np.savez_compressed(HERE / "haar_blocks.npz", **out)


def bernoulli():
    """Four 26x26 Bernoulli(1/2) 0/1 matrices (seed 1003) with exact permanents."""
    rng = np.random.default_rng(1003)
    mats = rng.integers(0, 2, size=(4, 26, 26)).astype(np.int64)
    vals = []
    for k in range(4):
        t = time.time()
        v = oracle.int_permanent(mats[k])
        vals.append(str(v))
        print("bernoulli", k, v, f"{time.time() - t:.1f}s", flush=True)
    np.savez_compressed(HERE / "bernoulli26.npz", mats=mats, perm=np.array(vals))


def repeated_c3():
    """30 photons in 60 modes: inputs 0..29, 18 distinct output modes with
    multiplicities 1 + Multinomial(12, 1/18) (seed 1004)."""
    rng = np.random.default_rng(1004)
    U = haar_unitary(60, 1004)
    inp = np.zeros(60, dtype=np.int64)
    inp[:30] = 1
    modes = np.sort(rng.choice(60, size=18, replace=False))
    mult = 1 + rng.multinomial(12, np.full(18, 1.0 / 18))
    outp = np.zeros(60, dtype=np.int64)
    outp[modes] = mult
    A = np.ascontiguousarray(U[np.ix_(modes, np.arange(30))])
    rows = mult.astype(np.int64)
    cols = np.ones(30, dtype=np.int64)
    t = time.time()
    v, vp = oracle.bbfg_repeated(A, cols, rows, "long double", full=True)
    print("c3 ld", v, f"{time.time() - t:.1f}s", flush=True)
    t = time.time()
    d, dp = oracle.bbfg_repeated(A, cols, rows, "dd", full=True)
    print("c3 dd", d, f"{time.time() - t:.1f}s", flush=True)
    np.savez_compressed(HERE / "repeated_c3.npz", U=U, input_state=inp, output_state=outp, A=A,
                        row_mult=rows, col_mult=cols, value_ld=np.array([v.real, v.imag]), value_dd=np.array(dp))


def seeding():
    """Seeded boson-sampling runs of unittests/test_seeding.py (data of the reference's
    tests): the 5-mode interferometer (:30-49), input state [2,1,1,0,1], seed 123, 20 shots;
    expected samples of the lossless run (:63-84), of the lossy one with singular values
    0.9 .. 0.5 (:88-125, run through the 2m-mode dilation) and of the uniformly lossy one
    with singular values 0.9 (:134-155)."""
    U = [[-0.17959207 - 0.29175972j, -0.64550941 - 0.02897055j, -0.023922 - 0.38854671j,
          -0.07908932 + 0.35617293j, -0.39779191 + 0.14902272j],
         [-0.36896208 + 0.19468375j, -0.11545557 + 0.20434514j, 0.25548079 + 0.05220164j,
          -0.51002161 + 0.38442256j, 0.48106678 - 0.25210091j],
         [0.25912844 + 0.16131742j, 0.11886251 + 0.12632645j, 0.69028213 - 0.25734432j,
          0.01276639 + 0.05841739j, 0.03713264 + 0.57364845j],
         [-0.20314019 - 0.18973473j, 0.59146854 + 0.28605532j, -0.11096495 - 0.26870144j,
          -0.47290354 - 0.0489408j, -0.42459838 + 0.01554643j],
         [-0.5021973 - 0.53474291j, 0.24524545 - 0.0741398j, 0.37786104 + 0.10225255j,
          0.46696955 + 0.10636677j, 0.07171789 - 0.09194236j]]
    lossless = [(4, 0, 0, 0, 1), (0, 0, 0, 5, 0), (0, 1, 0, 3, 1), (0, 1, 1, 2, 1), (1, 0, 2, 0, 2),
                (1, 0, 2, 2, 0), (4, 0, 0, 1, 0), (0, 0, 2, 2, 1), (0, 4, 0, 0, 1), (0, 0, 3, 0, 2),
                (2, 0, 3, 0, 0), (1, 2, 1, 0, 1), (0, 0, 2, 1, 2), (0, 2, 1, 1, 1), (0, 1, 1, 1, 2),
                (2, 1, 1, 0, 1), (0, 3, 1, 1, 0), (0, 0, 3, 1, 1), (4, 0, 0, 0, 1), (0, 0, 3, 1, 1)]
    uniform = [(0, 1, 1, 2, 0), (0, 0, 1, 1, 1), (1, 0, 4, 0, 0), (1, 0, 2, 1, 0), (0, 1, 2, 0, 0),
               (0, 0, 5, 0, 0), (1, 0, 1, 0, 2), (0, 1, 1, 1, 0), (0, 0, 2, 3, 0), (0, 0, 1, 2, 0),
               (0, 0, 0, 1, 2), (0, 0, 1, 2, 1), (2, 1, 1, 0, 1), (2, 0, 2, 0, 0), (0, 2, 1, 0, 0),
               (1, 0, 0, 2, 1), (3, 0, 0, 0, 1), (2, 1, 1, 0, 1), (0, 0, 1, 2, 2), (2, 1, 1, 0, 1)]
    nonuniform = [(0, 3, 1, 0, 0), (0, 0, 0, 0, 4), (0, 1, 0, 0, 2), (0, 0, 2, 0, 1), (0, 2, 1, 0, 0),
                  (0, 0, 1, 0, 1), (1, 0, 0, 1, 2), (0, 0, 0, 1, 2), (0, 0, 0, 1, 3), (2, 0, 0, 1, 1),
                  (0, 1, 0, 0, 0), (0, 1, 0, 1, 0), (0, 0, 0, 1, 1), (0, 0, 1, 2, 0), (0, 0, 2, 0, 0),
                  (3, 0, 0, 0, 0), (0, 1, 1, 0, 2), (0, 0, 1, 1, 0), (0, 0, 0, 0, 0), (0, 0, 1, 1, 0)]
    data = dict(U=[[_c(v) for v in row] for row in U], input_state=[2, 1, 1, 0, 1], seed=123, shots=20,
                uniform_singular_value=0.9, lossless=lossless, uniform_losses=uniform,
                nonuniform_singular_values=[0.9, 0.8, 0.7, 0.6, 0.5], nonuniform_losses=nonuniform)
    (HERE / "seeding.json").write_text(json.dumps(data, indent=1))
    print("seeding.json")


def uneven4096():
    """The batched workload in full: 4096 transitions of the seed-1005 100-mode Haar
    unitary (n ~ U{8..26} photons in distinct input modes, output multiplicities
    {1,2,3} with p = .7/.2/.1, as utilities.uneven_batch draws them).  Stored: the
    unitary itself, each task's input modes and (output mode, multiplicity) lists
    (padded with -1 / 0), and the oracle's double-double value of every task
    (restated BBFGPermanentCalculatorRepeated on the compact matrix, as
    BatchedPermanentCalculator.cpp:118-145 evaluates each pair)."""
    from piquassoboost_amd.sampling.utilities import adapt_interferometer, uneven_batch

    U, pairs = uneven_batch(seed=1005, count=4096)
    count = len(pairs)
    in_modes = np.full((count, 26), -1, dtype=np.int8)
    out_modes = np.full((count, 26), -1, dtype=np.int8)
    out_mult = np.zeros((count, 26), dtype=np.int8)
    vals = np.zeros((count, 4), dtype=np.float64)
    t = time.time()
    for k, (i, o) in enumerate(pairs):
        im = np.flatnonzero(i)
        om = np.flatnonzero(o)
        in_modes[k, :im.size] = im
        out_modes[k, :om.size] = om
        out_mult[k, :om.size] = o[om]
        A = adapt_interferometer(U, i, o)
        _, vp = oracle.bbfg_repeated(A, i[im], o[om], "dd", full=True)
        vals[k] = vp
        if k % 512 == 511:
            print("uneven4096", k + 1, f"{time.time() - t:.0f}s", flush=True)
    np.savez_compressed(HERE / "uneven4096.npz", U=U, in_modes=in_modes, out_modes=out_modes, out_mult=out_mult,
                        value_dd=vals)


def grad_minors():
    """Minors of the gradient workload's blocks (block b: top-left n x n of the
    50-mode Haar unitary of seed 1006 + b): for all 8 n=26 blocks the row-0 and
    column-0 minors (d perm / d a_0j and d perm / d a_i0: 51 n=25 permanents per
    block), for the first 16 n=20 blocks every minor (400 n=19 permanents per block),
    each the oracle's double-double restated reference sweep."""
    from piquassoboost_amd.sampling.utilities import gradient_blocks

    blocks, seeds = gradient_blocks()
    out = {}
    t = time.time()
    B26 = blocks[26]
    row0 = np.zeros((B26.shape[0], 26, 4))
    col0 = np.zeros((B26.shape[0], 26, 4))
    for b in range(B26.shape[0]):
        A = B26[b]
        for j in range(26):
            row0[b, j] = oracle.bbfg(np.ascontiguousarray(np.delete(np.delete(A, 0, 0), j, 1)), "dd", full=True)[1]
        col0[b, 0] = row0[b, 0]
        for i in range(1, 26):
            col0[b, i] = oracle.bbfg(np.ascontiguousarray(np.delete(np.delete(A, i, 0), 0, 1)), "dd", full=True)[1]
        print("grad_minors n=26 block", b, f"{time.time() - t:.0f}s", flush=True)
    B20 = blocks[20][:16]
    full20 = np.zeros((B20.shape[0], 20, 20, 4))
    for b in range(B20.shape[0]):
        A = B20[b]
        for i in range(20):
            for j in range(20):
                full20[b, i, j] = oracle.bbfg(np.ascontiguousarray(np.delete(np.delete(A, i, 0), j, 1)), "dd",
                                              full=True)[1]
    print("grad_minors n=20", f"{time.time() - t:.0f}s", flush=True)
    out.update(b26=B26, b26_row0_dd=row0, b26_col0_dd=col0, b20=B20, b20_minors_dd=full20)
    np.savez_compressed(HERE / "grad_minors.npz", **out)


STEPS = {"kat": kat, "haar": haar_blocks, "bernoulli": bernoulli, "c3": repeated_c3, "seeding": seeding,
         "uneven4096": uneven4096, "grad_minors": grad_minors}

if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", choices=sorted(STEPS), action="append")
    a = ap.parse_args()
    for name in a.only or STEPS:
        STEPS[name]()
