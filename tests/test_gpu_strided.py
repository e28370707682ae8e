"""The row-stride contract of the C ABI: every host entry point reads a matrix
through (pointer, rows, cols, ld) with ld >= cols, as pic::matrix carries
`stride` (matrix_base.hpp:48-49) and the reference sweep reads
mtx_data[row_idx * mtx.stride + col_idx] (BBFGPermanentCalculator.hpp:154,161).

Each test calls the entry point through ctypes directly on a sub-block of a
larger row-major buffer (ld = 40, the parent's width, or cols + 3), and
requires the result to be bit-identical to the same call on a packed copy
(ld = cols) and within the usual bar of the oracle on that copy."""

import ctypes

import numpy as np
import pytest

import oracle
from conftest import rel_err
from piquassoboost_amd import _capi as C
from piquassoboost_amd.sampling.utilities import haar_unitary

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
W = 40


@pytest.fixture(scope="module")
def parent():
    U = np.ascontiguousarray(haar_unitary(W, 6001))
    assert U.strides == (16 * W, 16)
    return U


def _ptr(U, r0, c0):
    """Address of element (r0, c0) of the C-contiguous parent as a double*."""
    return ctypes.cast(U.ctypes.data + 16 * (r0 * U.shape[1] + c0), _dp)


def _padded(A, pad=3):
    """A copy of A inside a buffer whose rows are cols + pad wide (pad filled with NaN)."""
    r, c = A.shape
    buf = np.full((r, c + pad), np.nan + 1j * np.nan, dtype=np.complex128)
    buf[:, :c] = A
    return buf


def _perm(ptr, rows, cols, ld, algo=C.ALGO_BBFG, prec=C.PREC_DOUBLE):
    out = (ctypes.c_double * 2)()
    C._check(C.lib().pqb_perm(ptr, rows, cols, ld, algo, prec, out))
    return complex(out[0], out[1])


def _packed(A):
    A = np.ascontiguousarray(A)
    return A, A.ctypes.data_as(_dp)


@pytest.mark.parametrize("r0,c0,rows,cols", [(3, 5, 20, 20), (5, 7, 30, 30), (1, 2, 12, 20), (4, 0, 11, 10),
                                             (0, 9, 1, 31), (7, 3, 26, 26)])
def test_perm_strided_fp64(parent, r0, c0, rows, cols):
    block = parent[r0:r0 + rows, c0:c0 + cols]
    got = _perm(_ptr(parent, r0, c0), rows, cols, W, prec=C.PREC_FP64)
    Ap, pp = _packed(block)
    want = _perm(pp, rows, cols, cols, prec=C.PREC_FP64)
    print(f"pqb_perm {rows}x{cols} ld={W}: {got!r} packed {want!r}", flush=True)
    assert got == want
    if rows <= 26:
        ref = oracle.bbfg(Ap, "dd")
        scale = max(abs(ref), np.prod(np.abs(Ap).sum(axis=0)) * 1e-3) if rows != cols else abs(ref)
        assert abs(got - ref) <= 1e-12 * scale, (got, ref)
    # the Python binding keeps a row slice in place with its stride
    M, ld = C.as_rows(block)
    assert ld == W and M.ctypes.data == parent.ctypes.data + 16 * (r0 * W + c0)
    assert C.permanent(block, C.ALGO_BBFG, C.PREC_FP64) == want


@pytest.mark.parametrize("prec", [C.PREC_DOUBLE, C.PREC_EXTENDED, C.PREC_INF])
def test_perm_strided_precisions(parent, prec):
    block = parent[2:16, 4:18]
    got = _perm(_ptr(parent, 2, 4), 14, 14, W, prec=prec)
    Ap, pp = _packed(block)
    assert got == _perm(pp, 14, 14, 14, prec=prec)
    assert rel_err(got, oracle.bbfg(Ap, "dd")) < 1e-12
    for algo in (C.ALGO_GLYNN, C.ALGO_SIMPLE):
        assert _perm(_ptr(parent, 2, 4), 14, 14, W, algo=algo, prec=prec) == _perm(pp, 14, 14, 14, algo=algo,
                                                                                    prec=prec)


def test_perm_exact_strided():
    """The exact integer route (pqb_perm_exact and the default precision) on a padded buffer."""
    rng = np.random.default_rng(6002)
    A = rng.integers(-2, 3, size=(13, 13)).astype(np.complex128)
    buf = _padded(A, 3)
    exact = oracle.int_permanent(A.real.astype(np.int64))
    w = (ctypes.c_int64 * 2)()
    out = (ctypes.c_double * 2)()
    C._check(C.lib().pqb_perm_exact(buf.ctypes.data_as(_dp), 13, 13, 16, w, out))
    v = (int(w[1]) << 64) | (int(w[0]) & 0xFFFFFFFFFFFFFFFF)
    assert v == exact and out[0] == float(exact)
    assert _perm(buf.ctypes.data_as(_dp), 13, 13, 16) == complex(float(exact), 0.0)
    b = ctypes.c_double(0.0)
    assert C.lib().pqb_exact_bound(buf.ctypes.data_as(_dp), 13, 13, 16, ctypes.byref(b)) == C.PQB_OK


@pytest.mark.parametrize("prec", [C.PREC_DOUBLE, C.PREC_INF])
def test_perm_repeated_strided(parent, prec):
    rows, cols = 7, 9
    r0, c0 = 6, 11
    rm = np.array([2, 1, 0, 3, 1, 1, 2], dtype=np.int64)
    cm = np.array([1, 1, 2, 1, 0, 1, 2, 1, 1], dtype=np.int64)
    out = (ctypes.c_double * 2)()
    C._check(C.lib().pqb_perm_repeated(_ptr(parent, r0, c0), rows, cols, W, rm.ctypes.data_as(_i64p),
                                       cm.ctypes.data_as(_i64p), C.ALGO_BBFG, prec, out))
    got = complex(out[0], out[1])
    Ap, _ = _packed(parent[r0:r0 + rows, c0:c0 + cols])
    assert got == C.permanent_repeated(Ap, rm, cm, C.ALGO_BBFG, prec)
    assert rel_err(got, oracle.bbfg_repeated(Ap, cm, rm, "dd")) < 1e-12


def test_perm_batch_tasks_strided(parent):
    """pqb_task.ld = 40: every task a block of the parent, read in place."""
    rng = np.random.default_rng(6003)
    specs = []
    for k in range(24):
        r = int(rng.integers(3, 13))
        c = r
        r0 = int(rng.integers(0, W - r))
        c0 = int(rng.integers(0, W - c))
        rm = rng.integers(1, 3, size=r).astype(np.int64)
        cm = rng.permutation(rm).astype(np.int64)
        specs.append((r0, c0, r, c, rm, cm))
    arr = np.zeros((len(specs), 6), dtype=np.int64)
    for k, (r0, c0, r, c, rm, cm) in enumerate(specs):
        arr[k] = [parent.ctypes.data + 16 * (r0 * W + c0), r, c, W, rm.ctypes.data, cm.ctypes.data]
    out = np.zeros(2 * len(specs))
    C._check(C.lib().pqb_perm_batch(ctypes.c_void_p(arr.ctypes.data), len(specs), C.ALGO_BBFG, C.PREC_DOUBLE,
                                    out.ctypes.data_as(_dp)))
    got = out[0::2] + 1j * out[1::2]
    packed = [(np.ascontiguousarray(parent[r0:r0 + r, c0:c0 + c]), rm, cm) for r0, c0, r, c, rm, cm in specs]
    want = C.permanent_batch(packed)
    assert list(got) == want
    for k, (Ap, rm, cm) in enumerate(packed):
        assert rel_err(got[k], oracle.bbfg_repeated(Ap, cm, rm, "dd")) < 1e-12, k
    # the Python task list keeps row slices in place (ld = 40 in the task array)
    slices = [(parent[r0:r0 + r, c0:c0 + c], rm, cm) for r0, c0, r, c, rm, cm in specs]
    tarr, _ = C._task_array(slices)
    assert (tarr[:, 3] == W).all()
    assert C.permanent_batch(slices) == want


def test_perm_batch_states_strided(parent):
    """The interferometer as a 30 x 30 block of the 40 x 40 parent (ld = 40)."""
    rng = np.random.default_rng(6004)
    m = 30
    ins, outs = [], []
    for _ in range(40):
        n = int(rng.integers(4, 11))
        i = np.zeros(m, np.int64)
        i[rng.choice(m, n, replace=False)] = 1
        o = np.bincount(rng.choice(m, n), minlength=m).astype(np.int64)
        ins.append(i)
        outs.append(o)
    ins = np.stack(ins)
    outs = np.stack(outs)
    out = np.zeros(2 * len(ins))
    C._check(C.lib().pqb_perm_batch_states(_ptr(parent, 4, 6), m, m, W, ins.ctypes.data_as(_i64p),
                                           outs.ctypes.data_as(_i64p), len(ins), C.ALGO_BBFG, C.PREC_DOUBLE,
                                           out.ctypes.data_as(_dp)))
    got = out[0::2] + 1j * out[1::2]
    Up = np.ascontiguousarray(parent[4:4 + m, 6:6 + m])
    want = C.permanent_batch_states(Up, ins, outs)
    assert np.array_equal(got, want)
    for k in range(len(ins)):
        A = oracle.adapt_interferometer(Up, ins[k], outs[k])
        ref = oracle.bbfg_repeated(A, ins[k][ins[k] > 0], outs[k][outs[k] > 0], "dd")
        assert rel_err(got[k], ref) < 1e-12, k
    assert np.array_equal(C.permanent_batch_states(parent[4:4 + m, 6:6 + m], ins, outs), want)


def test_perm_grad_strided(parent):
    n, r0, c0 = 12, 9, 21
    perm = (ctypes.c_double * 2)()
    G = np.zeros((n, n), dtype=np.complex128)
    C._check(C.lib().pqb_perm_grad(_ptr(parent, r0, c0), n, W, perm, G.ctypes.data_as(_dp)))
    Ap = np.ascontiguousarray(parent[r0:r0 + n, c0:c0 + n])
    p0, G0 = C.permanent_grad(Ap)
    assert complex(perm[0], perm[1]) == p0 and np.array_equal(G, G0)
    ref = np.array([[oracle.bbfg(np.delete(np.delete(Ap, i, 0), j, 1), "dd") for j in range(n)] for i in range(n)])
    assert np.abs(G - ref).max() / np.abs(ref).max() < 1e-12
    # batch form: three blocks of the parent, ld = 40, matrix stride = 40 rows
    count = 3
    stride = n * W + 0
    perms = np.zeros(count, dtype=np.complex128)
    Gb = np.zeros((count, n, n), dtype=np.complex128)
    C._check(C.lib().pqb_perm_grad_batch(_ptr(parent, 0, 2), count, n, W, stride, perms.ctypes.data_as(_dp),
                                         Gb.ctypes.data_as(_dp)))
    blocks = np.stack([parent[b * n:(b + 1) * n, 2:2 + n] for b in range(count)])
    pp, Gp = C.permanent_grad_batch(np.ascontiguousarray(blocks))
    assert np.array_equal(perms, pp) and np.array_equal(Gb, Gp)


def test_perm_minors_strided(parent):
    rows, r0, c0 = 10, 13, 17
    cols = rows + 1
    out = np.zeros(cols, dtype=np.complex128)
    C._check(C.lib().pqb_perm_minors(_ptr(parent, r0, c0), rows, cols, W, out.ctypes.data_as(_dp)))
    Ap = np.ascontiguousarray(parent[r0:r0 + rows, c0:c0 + cols])
    assert np.array_equal(out, C.permanent_minors(Ap))
    for j in range(cols):
        assert rel_err(out[j], oracle.bbfg(np.delete(Ap, j, 1), "dd")) < 1e-12, j
