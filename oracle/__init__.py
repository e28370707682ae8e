"""CPU oracle for the permanent engine -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or as the timed
CPU baseline).  The product path (``piquassoboost_amd``) never imports it.

The numerics live in ``pqb_oracle.cpp`` (a step-by-step restatement of the
reference's CPU calculators plus independent brute-force / Ryser checks); this
module is a thin ctypes wrapper plus numpy restatements of the reference's
matrix adapters.

Precision codes: ``"double"``, ``"long double"``, ``"dd"`` (double-double),
``"quad"`` (binary128).
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "_build" / "libpqb_oracle.so"
_lib = None

PREC = {"double": 0, "long double": 1, "dd": 2, "quad": 3}

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)


class OracleError(Exception):
    pass


def lib():
    """Load (building first if needed) the oracle shared library."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            from . import build as _b

            _b.build()
        L = ctypes.CDLL(str(_LIB_PATH))
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_bbfg.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                               ctypes.c_int64, ctypes.c_int, _dp]
        L.orc_bbfg_partial.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                                       ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, _dp]
        L.orc_bbfg_repeated.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int64, _i64p, _i64p,
                                        ctypes.c_int, ctypes.c_int64, ctypes.c_int, _dp]
        L.orc_glynn_recursive.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int, _dp]
        L.orc_glynn_repeated.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int64, _i64p, _i64p,
                                         ctypes.c_int, _dp]
        L.orc_glynn_simple.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int, _dp]
        L.orc_perm_bruteforce.argtypes = [_dp, ctypes.c_int, ctypes.c_int64, ctypes.c_int, _dp]
        L.orc_perm_ryser.argtypes = [_dp, ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_perm_int128.argtypes = [_i64p, ctypes.c_int, ctypes.c_int64, ctypes.c_int, _i64p]
        L.orc_hardware_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _threads(n):
    return int(n) if n else max(1, os.cpu_count() or 1)


def _mat(A):
    A = np.ascontiguousarray(A, dtype=np.complex128)
    if A.ndim == 1:
        A = A.reshape(-1, 1)
    if A.ndim != 2:
        raise ValueError("matrix must be 1-D or 2-D")
    return A


def _out_to(out, full):
    if full:
        return complex(out[0], out[1]), (out[2], out[3], out[4], out[5])
    return complex(out[0], out[1])


def hardware_threads() -> int:
    return lib().orc_hardware_threads()


def bbfg(A, prec="double", njobs=None, threads=None, full=False):
    """BBFGPermanentCalculator (BBFGPermanentCalculator.hpp:87-235).

    ``njobs`` is the reference's chunk count (4 x hardware threads there)."""
    A = _mat(A)
    t = _threads(threads)
    nj = int(njobs) if njobs else 4 * t
    out = (ctypes.c_double * 6)()
    rc = lib().orc_bbfg(A.ctypes.data_as(_dp), A.shape[0], A.shape[1], A.shape[1], PREC[prec], nj, t, out)
    if rc:
        raise OracleError(lib().orc_last_error().decode())
    return _out_to(out, full)


def bbfg_partial(A, njobs, job_lo, job_hi, prec="double", threads=None):
    """Jobs [job_lo, job_hi) of the reference sweep split into ``njobs`` chunks."""
    A = _mat(A)
    out = (ctypes.c_double * 6)()
    rc = lib().orc_bbfg_partial(A.ctypes.data_as(_dp), A.shape[0], A.shape[1], A.shape[1], PREC[prec],
                                int(njobs), int(job_lo), int(job_hi), _threads(threads), out)
    if rc:
        raise OracleError(lib().orc_last_error().decode())
    return complex(out[0], out[1])


def bbfg_repeated(A, input_state, output_state, prec="double", njobs=None, threads=None, full=False):
    """BBFGPermanentCalculatorRepeated (….hpp:74-307); rows <-> output_state."""
    A = _mat(A)
    ins = np.ascontiguousarray(input_state, dtype=np.int64)
    outs = np.ascontiguousarray(output_state, dtype=np.int64)
    if ins.shape[0] != A.shape[1] or outs.shape[0] != A.shape[0]:
        raise ValueError("state lengths must match the matrix shape")
    t = _threads(threads)
    nj = int(njobs) if njobs else 4 * t
    out = (ctypes.c_double * 6)()
    rc = lib().orc_bbfg_repeated(A.ctypes.data_as(_dp), A.shape[0], A.shape[1], A.shape[1],
                                 outs.ctypes.data_as(_i64p), ins.ctypes.data_as(_i64p), PREC[prec], nj, t, out)
    if rc:
        raise OracleError(lib().orc_last_error().decode())
    return _out_to(out, full)


def glynn_recursive(A, prec="long double", full=False):
    """GlynnPermanentCalculator (GlynnPermanentCalculator.hpp:47-201)."""
    A = _mat(A)
    out = (ctypes.c_double * 6)()
    rc = lib().orc_glynn_recursive(A.ctypes.data_as(_dp), A.shape[0], A.shape[1], A.shape[1], PREC[prec], out)
    if rc:
        raise OracleError(lib().orc_last_error().decode())
    return _out_to(out, full)


def glynn_repeated(A, input_state, output_state, prec="long double", full=False):
    """GlynnPermanentCalculatorRepeated (….hpp:49-258)."""
    A = _mat(A)
    ins = np.ascontiguousarray(input_state, dtype=np.int64)
    outs = np.ascontiguousarray(output_state, dtype=np.int64)
    out = (ctypes.c_double * 6)()
    rc = lib().orc_glynn_repeated(A.ctypes.data_as(_dp), A.shape[0], A.shape[1], A.shape[1],
                                  ins.ctypes.data_as(_i64p), outs.ctypes.data_as(_i64p), PREC[prec], out)
    if rc:
        raise OracleError(lib().orc_last_error().decode())
    return _out_to(out, full)


def glynn_simple(A, prec="double", full=False):
    """GlynnPermanentCalculatorSimple (….hpp:61-168)."""
    A = _mat(A)
    out = (ctypes.c_double * 6)()
    rc = lib().orc_glynn_simple(A.ctypes.data_as(_dp), A.shape[0], A.shape[1], A.shape[1], PREC[prec], out)
    if rc:
        raise OracleError(lib().orc_last_error().decode())
    return _out_to(out, full)


def bruteforce(A, prec="long double", full=False):
    """Sum over all n! permutations (independent of Glynn); n <= 11."""
    A = _mat(A)
    if A.shape[0] != A.shape[1]:
        raise ValueError("square matrix required")
    out = (ctypes.c_double * 6)()
    rc = lib().orc_perm_bruteforce(A.ctypes.data_as(_dp), A.shape[0], A.shape[1], PREC[prec], out)
    if rc:
        raise OracleError("bruteforce: n too large")
    return _out_to(out, full)


def ryser(A, prec="dd", threads=None, full=False):
    """Ryser's formula over column subsets (independent of Glynn)."""
    A = _mat(A)
    if A.shape[0] != A.shape[1]:
        raise ValueError("square matrix required")
    out = (ctypes.c_double * 6)()
    rc = lib().orc_perm_ryser(A.ctypes.data_as(_dp), A.shape[0], A.shape[1], PREC[prec], _threads(threads), out)
    if rc:
        raise OracleError("ryser: n too large")
    return _out_to(out, full)


def int_permanent(A, threads=None) -> int:
    """Exact permanent of an integer matrix (Ryser in 128-bit integers)."""
    Ai = np.ascontiguousarray(A, dtype=np.int64)
    if Ai.ndim != 2 or Ai.shape[0] != Ai.shape[1]:
        raise ValueError("square integer matrix required")
    w = (ctypes.c_int64 * 2)()
    rc = lib().orc_perm_int128(Ai.ctypes.data_as(_i64p), Ai.shape[0], Ai.shape[1], _threads(threads), w)
    if rc:
        raise OracleError(lib().orc_last_error().decode())
    v = (int(w[1]) << 64) | (int(w[0]) & 0xFFFFFFFFFFFFFFFF)
    return v


# ---------------------------------------------------------------- adapters --


def adapt_interferometer(U, input_state, output_state):
    """Compact effective matrix: rows of occupied output modes x columns of
    occupied input modes (BatchedPermanentCalculator.cpp:412-451)."""
    U = np.asarray(U)
    rows = [i for i in range(U.shape[0]) if output_state[i] > 0]
    cols = [j for j in range(U.shape[0]) if input_state[j] > 0]
    M = np.zeros((len(rows), len(cols)), dtype=np.complex128)
    for a, i in enumerate(rows):
        for b, j in enumerate(cols):
            M[a, b] = U[i, j]
    return M


def adapt_interferometer_glynn_multiplied(U, input_state, output_state):
    """Expanded effective matrix: row i repeated output_state[i] times, column
    j repeated input_state[j] times (BatchedPermanentCalculator.cpp:367-399)."""
    U = np.asarray(U)
    n = U.shape[0]
    total = int(sum(input_state))
    M = np.zeros((total, total), dtype=np.complex128)
    r = 0
    for i in range(n):
        for _ in range(int(output_state[i])):
            c = 0
            for j in range(n):
                for _ in range(int(input_state[j])):
                    M[r, c] = U[i, j]
                    c += 1
            r += 1
    return M
