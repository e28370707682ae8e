"""ctypes binding of ``libpqb_perm.so`` (the C ABI declared in include/pqb_perm.h).

The library is the only compute path: when it is missing or no HIP device is
visible, every compute call raises -- there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libpqb_perm.so"

# status codes (include/pqb_perm.h)
PQB_OK = 0
PQB_ERR_ARG = 1
PQB_ERR_SIZE = 2
PQB_ERR_STATE = 3
PQB_ERR_DEVICE = 4
PQB_ERR_EXACT = 5

ALGO_BBFG = 0
ALGO_GLYNN = 1
ALGO_SIMPLE = 2

PREC_DOUBLE = 0
PREC_EXTENDED = 1
PREC_FP64 = 2
PREC_EXACT = 3
PREC_INF = 4  # double-double sweep (the reference's MPFR "Inf" calculators)

MAX_COLS = 40

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)

# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "pqb_last_error": (ctypes.c_char_p, []),
    "pqb_abi_version": (ctypes.c_int, []),
    "pqb_device_count": (ctypes.c_int, []),
    "pqb_init": (ctypes.c_int, [ctypes.c_int]),
    "pqb_perm": (ctypes.c_int, [_dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                ctypes.c_int, _dp]),
    "pqb_perm_list": (ctypes.c_int, [ctypes.c_void_p, _i64p, _i64p, _i64p, ctypes.c_int64, ctypes.c_int,
                                     ctypes.c_int, _dp]),
    "pqb_perm_exact": (ctypes.c_int, [_dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _i64p, _dp]),
    "pqb_exact_bound": (ctypes.c_int, [_dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _dp]),
    "pqb_perm_repeated": (ctypes.c_int, [_dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _i64p, _i64p,
                                         ctypes.c_int, ctypes.c_int, _dp]),
    "pqb_perm_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _dp]),
    "pqb_perm_batch_states": (ctypes.c_int, [_dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _i64p, _i64p,
                                             ctypes.c_int64, ctypes.c_int, ctypes.c_int, _dp]),
    "pqb_perm_batch_tuned": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int64, _dp]),
    "pqb_batch_replay": (ctypes.c_int, [ctypes.c_int, _dp]),
    "pqb_set_walk_lengths": (ctypes.c_int, [_i64p, ctypes.c_int64]),
    "pqb_repeated_plan": (ctypes.c_int, [_dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _i64p, _i64p,
                                         ctypes.c_int64, _i64p]),
    "pqb_perm_device_workspace": (ctypes.c_size_t, [ctypes.c_int64, ctypes.c_int64]),
    "pqb_perm_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "pqb_plain_plan": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _i64p]),
    "pqb_time_perm_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                            ctypes.c_int, _dp]),
    "pqb_perm_grad": (ctypes.c_int, [_dp, ctypes.c_int64, ctypes.c_int64, _dp, _dp]),
    "pqb_perm_grad_batch": (ctypes.c_int, [_dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                           _dp, _dp]),
    "pqb_perm_grad_workspace": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(ctypes.c_size_t)]),
    "pqb_perm_grad_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "pqb_grad_plan": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _i64p]),
    "pqb_perm_minors": (ctypes.c_int, [_dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _dp]),
    "pqb_perm_minors_batch": (ctypes.c_int, [_dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                             ctypes.c_int64, _dp]),
    "pqb_perm_batch_device_workspace": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                                       ctypes.POINTER(ctypes.c_size_t)]),
    "pqb_probe_fp64_fma": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, _dp, _dp]),
    "pqb_clock_stamp": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "pqb_perm_device_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                               ctypes.c_void_p]),
    "pqb_perm_device_noalloc_scratch_bytes": (ctypes.c_size_t, []),
    "pqb_perm_batch_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
}

_lib = None
_lock = threading.Lock()
_C128 = np.dtype(np.complex128)


class EngineError(Exception):
    """Raised for every non-zero status of the C ABI (message from pqb_last_error)."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def _share_torch_hip_runtime() -> None:
    """A process holds one HIP runtime: PyTorch ships its own libamdhip64, and if the engine
    library is loaded first it binds the system one, after which PyTorch finds no GPU.
    When PyTorch is installed it is therefore imported before the engine library, whose
    libamdhip64 dependency then resolves to PyTorch's (same soname)."""
    import importlib.util

    if importlib.util.find_spec("torch") is not None:
        import torch  # noqa: F401


def lib():
    """Load libpqb_perm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not LIB_PATH.exists():
                    raise RuntimeError(
                        f"{LIB_PATH} is missing: build the HIP engine first "
                        "(python -m piquassoboost_amd.build)"
                    )
                _share_torch_hip_runtime()
                L = ctypes.CDLL(str(LIB_PATH))
                for name, (res, args) in SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = L
                _apply_saved_tuning()
    return _lib


def _apply_saved_tuning() -> None:
    """Walk lengths found by the autotuner (piquassoboost_amd/tuning/*.json)."""
    from . import autotune

    table = autotune.load()
    if table:
        set_walk_lengths(table)


def _check(rc: int) -> None:
    if rc != PQB_OK:
        msg = lib().pqb_last_error().decode(errors="replace")
        raise EngineError(rc, msg)


def as_matrix(A) -> np.ndarray:
    """complex128, C-contiguous, 2-D (a 1-D array is a column, as numpy2matrix
    does at numpy_interface.cpp:152-155)."""
    M = np.ascontiguousarray(A, dtype=np.complex128)
    if M.ndim == 1:
        M = M.reshape(-1, 1)
    if M.ndim != 2:
        raise Exception("numpy2matrix: Wrong matrix dimension was given")
    return M


def as_rows(A) -> tuple[np.ndarray, int]:
    """(matrix, ld): a complex128 2-D array whose rows are contiguous and evenly
    spaced (a C-contiguous array, or a block / row slice of one) is passed in
    place with its row stride as ld, as pic::matrix carries `stride`
    (matrix_base.hpp:48-49); anything else is copied to a C-contiguous array
    (ld = cols)."""
    if isinstance(A, np.ndarray) and A.dtype == np.complex128 and A.ndim == 2 and A.size > 0:
        s0, s1 = A.strides
        if s1 == 16 and s0 > 0 and s0 % 16 == 0 and s0 // 16 >= A.shape[1] and A.flags.aligned:
            return A, s0 // 16
    M = as_matrix(A)
    return M, M.shape[1]


def _stack_strides(B) -> tuple[np.ndarray, int, int]:
    """(stack, ld, stride) of a (count, rows, cols) complex128 stack in elements,
    kept in place when rows are contiguous and evenly spaced, else copied."""
    if isinstance(B, np.ndarray) and B.dtype == np.complex128 and B.ndim == 3 and B.size > 0:
        s0, s1, s2 = B.strides
        count, rows, cols = B.shape
        if (s2 == 16 and s1 > 0 and s1 % 16 == 0 and s1 // 16 >= cols and s0 % 16 == 0
                and (count == 1 or s0 // 16 >= rows * (s1 // 16)) and B.flags.aligned):
            return B, s1 // 16, (s0 // 16 if count > 1 else rows * (s1 // 16))
    C = np.ascontiguousarray(B, dtype=np.complex128)
    if C.ndim != 3:
        return C, 0, 0
    return C, C.shape[2], C.shape[1] * C.shape[2]


def as_state(s) -> np.ndarray:
    v = np.ascontiguousarray(s, dtype=np.int64)
    if v.ndim != 1:
        raise Exception("numpy2PicState_int64: Wrong matrix dimension was given")
    return v


def device_count() -> int:
    return int(lib().pqb_device_count())


def permanent(A, algo: int = ALGO_BBFG, precision: int = PREC_DOUBLE) -> complex:
    M, ld = as_rows(A)
    out = (ctypes.c_double * 2)()
    _check(lib().pqb_perm(M.ctypes.data_as(_dp), M.shape[0], M.shape[1], ld, algo, precision, out))
    return complex(out[0], out[1])


def permanent_list(mats, algo: int = ALGO_BBFG, precision: int = PREC_DOUBLE) -> list:
    """Plain permanents of a list of matrices (any shapes) in one call: each value
    equals permanent(m, algo, precision) bit for bit; matrices of one shape share a
    launch (pqb_perm_list)."""
    mats = list(mats)
    n = len(mats)
    if n == 0:
        return []
    keep = [as_rows(m) for m in mats]
    ptrs = np.fromiter((_address(M) if M.size else 0 for M, _ in keep), dtype=np.int64, count=n)
    rows = np.fromiter((M.shape[0] for M, _ in keep), dtype=np.int64, count=n)
    cols = np.fromiter((M.shape[1] for M, _ in keep), dtype=np.int64, count=n)
    lds = np.fromiter((ld for _, ld in keep), dtype=np.int64, count=n)
    out = np.zeros(2 * n, dtype=np.float64)
    _check(lib().pqb_perm_list(ctypes.c_void_p(ptrs.ctypes.data), rows.ctypes.data_as(_i64p),
                               cols.ctypes.data_as(_i64p), lds.ctypes.data_as(_i64p), n, algo, precision,
                               out.ctypes.data_as(_dp)))
    del keep
    return (out[0::2] + 1j * out[1::2]).tolist()


def permanent_exact(A) -> tuple[int, float]:
    """Exact integer permanent (two's complement 128-bit) and its rounded double."""
    M, ld = as_rows(A)
    w = (ctypes.c_int64 * 2)()
    out = (ctypes.c_double * 2)()
    _check(lib().pqb_perm_exact(M.ctypes.data_as(_dp), M.shape[0], M.shape[1], ld, w, out))
    v = (int(w[1]) << 64) | (int(w[0]) & 0xFFFFFFFFFFFFFFFF)
    return v, out[0]


def permanent_repeated(A, row_mult, col_mult, algo: int = ALGO_BBFG, precision: int = PREC_DOUBLE) -> complex:
    M, ld = as_rows(A)
    r = as_state(row_mult)
    c = as_state(col_mult)
    if r.shape[0] != M.shape[0] or c.shape[0] != M.shape[1]:
        raise Exception(
            "state lengths must match the matrix: len(output_state) == rows, len(input_state) == cols"
        )
    out = (ctypes.c_double * 2)()
    _check(lib().pqb_perm_repeated(M.ctypes.data_as(_dp), M.shape[0], M.shape[1], ld,
                                   r.ctypes.data_as(_i64p), c.ctypes.data_as(_i64p), algo, precision, out))
    return complex(out[0], out[1])


def _address(M):
    """Data address of a C-contiguous array: through the buffer protocol (~0.3 us),
    else (read-only or empty arrays) through __array_interface__ (~1 us)."""
    try:
        return ctypes.addressof(ctypes.c_char.from_buffer(M))
    except (TypeError, ValueError):
        return M.__array_interface__["data"][0]


def _task_array(tasks):
    """tasks: sequence of (matrix, row_mult, col_mult).  Returns an (n, 6) int64
    array laid out as n structs pqb_task of include/pqb_perm.h (six 8-byte
    fields: A, rows, cols, ld, row_mult, col_mult; pointers as addresses) and the
    arrays it points into (kept alive by the caller).  Matrices whose rows are
    contiguous complex128 (see as_rows) and C-contiguous int64 states are used
    in place."""
    n = len(tasks)
    mats, lds, shapes, rows, cols = [], [], [], [], []
    for A, r, c in tasks:
        # C-contiguous complex128 matrices (the common case) skip as_rows
        if type(A) is np.ndarray and A.dtype == _C128 and A.ndim == 2 and A.flags.c_contiguous and A.size:
            M, ld = A, A.shape[1]
        else:
            M, ld = as_rows(A)
        mats.append(M)
        lds.append(ld)
        shapes.append(M.shape)
        rows.append(r)
        cols.append(c)
    shapes = np.array(shapes, dtype=np.int64).reshape(n, 2)
    # the occupation numbers of all tasks packed into two int64 arrays (one
    # conversion pass each): a pointer per task is then base + offset
    try:
        rall = np.concatenate(rows, dtype=np.int64, casting="same_kind") if n else np.zeros(0, np.int64)
        call = np.concatenate(cols, dtype=np.int64, casting="same_kind") if n else np.zeros(0, np.int64)
        nr = np.fromiter(map(len, rows), dtype=np.int64, count=n)
        nc = np.fromiter(map(len, cols), dtype=np.int64, count=n)
        if rall.ndim != 1 or call.ndim != 1 or int(nr.sum()) != rall.shape[0] or int(nc.sum()) != call.shape[0]:
            raise ValueError
    except (ValueError, TypeError):
        rows = [as_state(r) for r in rows]  # raises the reference's message for a bad state
        cols = [as_state(c) for c in cols]
        rall = np.concatenate(rows) if n else np.zeros(0, np.int64)
        call = np.concatenate(cols) if n else np.zeros(0, np.int64)
        nr = np.fromiter((v.shape[0] for v in rows), dtype=np.int64, count=n)
        nc = np.fromiter((v.shape[0] for v in cols), dtype=np.int64, count=n)
    bad = np.flatnonzero((nr != shapes[:, 0]) | (nc != shapes[:, 1]))
    if bad.size:
        raise Exception(
            f"task {int(bad[0])}: state lengths must match the matrix: len(output_state) == rows, "
            "len(input_state) == cols"
        )
    arr = np.empty((n, 6), dtype=np.int64)
    arr[:, 0] = np.fromiter((_address(M) for M in mats), dtype=np.int64, count=n)
    arr[:, 1] = nr
    arr[:, 2] = nc
    arr[:, 3] = lds
    arr[:, 4] = rall.__array_interface__["data"][0] + 8 * (np.cumsum(nr) - nr)
    arr[:, 5] = call.__array_interface__["data"][0] + 8 * (np.cumsum(nc) - nc)
    return arr, (mats, rall, call)


def permanent_batch(tasks, algo: int = ALGO_BBFG, precision: int = PREC_DOUBLE, walk_length: int = 0) -> list:
    """Permanents of many repeated-row tasks (matrix, row_mult, col_mult) in one launch."""
    tasks = list(tasks)
    if not tasks:
        return []
    arr, keep = _task_array(tasks)
    out = np.zeros(2 * len(tasks), dtype=np.float64)
    tp = ctypes.c_void_p(arr.ctypes.data)
    if walk_length:
        _check(lib().pqb_perm_batch_tuned(tp, len(tasks), algo, precision, walk_length, out.ctypes.data_as(_dp)))
    else:
        _check(lib().pqb_perm_batch(tp, len(tasks), algo, precision, out.ctypes.data_as(_dp)))
    del keep
    return (out[0::2] + 1j * out[1::2]).tolist()


def permanent_batch_states(U, inputs, outputs, algo: int = ALGO_BBFG, precision: int = PREC_DOUBLE) -> np.ndarray:
    """Permanents of (input state, output state) pairs of one interferometer U in one
    launch: inputs is (count, cols), outputs (count, rows); pair i is the repeated-row
    permanent with row multiplicities outputs[i] and column multiplicities inputs[i]
    (BatchednPermanentCalculator::calculate without per-pair matrices)."""
    M, ld = as_rows(U)
    ins = np.ascontiguousarray(inputs, dtype=np.int64)
    outs = np.ascontiguousarray(outputs, dtype=np.int64)
    if ins.ndim == 1:
        ins = ins.reshape(1, -1)
    if outs.ndim == 1:
        outs = outs.reshape(1, -1)
    if ins.ndim != 2 or outs.ndim != 2 or ins.shape[0] != outs.shape[0]:
        raise Exception("inputs and outputs must be (count, modes) arrays with the same count")
    if ins.shape[1] != M.shape[1] or outs.shape[1] != M.shape[0]:
        raise Exception("state lengths must match the interferometer: outputs per row, inputs per column")
    count = ins.shape[0]
    out = np.zeros(2 * count, dtype=np.float64)
    _check(lib().pqb_perm_batch_states(M.ctypes.data_as(_dp), M.shape[0], M.shape[1], ld,
                                       ins.ctypes.data_as(_i64p), outs.ctypes.data_as(_i64p), count, algo,
                                       precision, out.ctypes.data_as(_dp)))
    return out[0::2] + 1j * out[1::2]


def set_walk_lengths(table) -> None:
    """Walk length per photon count (dict n -> steps, or None to reset)."""
    if not table:
        _check(lib().pqb_set_walk_lengths(None, 0))
        return
    arr = np.zeros(MAX_COLS + 1, dtype=np.int64)
    for n, L in table.items():
        if 0 <= int(n) <= MAX_COLS:
            arr[int(n)] = int(L)
    _check(lib().pqb_set_walk_lengths(arr.ctypes.data_as(_i64p), arr.shape[0]))


def batch_replay_ms(iters: int) -> float:
    """Replay the last batch's graph `iters` times (inputs resident); elapsed ms."""
    ms = ctypes.c_double(0.0)
    _check(lib().pqb_batch_replay(iters, ctypes.byref(ms)))
    return ms.value


def repeated_plan(A, row_mult, col_mult, walk_length: int) -> dict:
    """Host-only decomposition of one repeated task (no device needed)."""
    M, ld = as_rows(A)
    r = as_state(row_mult)
    c = as_state(col_mult)
    info = (ctypes.c_int64 * 6)()
    _check(lib().pqb_repeated_plan(M.ctypes.data_as(_dp), M.shape[0], M.shape[1], ld,
                                   r.ctypes.data_as(_i64p), c.ctypes.data_as(_i64p), walk_length, info))
    return dict(zip(["slots", "rows", "low", "high", "walk", "chunks"], [int(v) for v in info]))


def exact_bound(A) -> tuple[bool, float]:
    """(eligible, log2 bound) of the exact integer route; host only."""
    M, ld = as_rows(A)
    b = ctypes.c_double(0.0)
    rc = lib().pqb_exact_bound(M.ctypes.data_as(_dp), M.shape[0], M.shape[1], ld, ctypes.byref(b))
    if rc == PQB_ERR_EXACT and b.value == 0.0 and "integer-valued" in lib().pqb_last_error().decode():
        return False, float("nan")
    if rc not in (PQB_OK, PQB_ERR_EXACT):
        _check(rc)
    return rc == PQB_OK, b.value


def plain_plan(rows: int, cols: int, waves_resident: int = 0) -> dict:
    p = (ctypes.c_int64 * 6)()
    _check(lib().pqb_plain_plan(rows, cols, waves_resident, p))
    return {"K": p[0], "B": p[1], "H": p[2], "waves": p[3], "block": p[4], "blocks": p[5]}


def device_workspace_bytes(rows: int, cols: int) -> int:
    return int(lib().pqb_perm_device_workspace(rows, cols))


def permanent_device(dA_ptr: int, rows: int, cols: int, d_out_ptr: int, d_work_ptr: int, work_bytes: int,
                     stream: int = 0) -> None:
    """Asynchronous plain permanent on device pointers (see pqb_perm_device)."""
    _check(lib().pqb_perm_device(dA_ptr, rows, cols, d_out_ptr, d_work_ptr, work_bytes, stream or None))


def permanent_device_noalloc(dA_ptr: int, rows: int, cols: int, d_out_ptr: int, stream: int = 0) -> None:
    """Asynchronous plain permanent on device pointers with the library's static device
    scratch buffer as workspace (see pqb_perm_device_noalloc): no allocation, no workspace
    argument; calls that overlap on one GPU race."""
    _check(lib().pqb_perm_device_noalloc(dA_ptr, rows, cols, d_out_ptr, stream or None))


def time_permanent_device(dA_ptr: int, rows: int, cols: int, d_out_ptr: int, iters: int) -> float:
    ms = ctypes.c_double(0.0)
    _check(lib().pqb_time_perm_device(dA_ptr, rows, cols, d_out_ptr, iters, ctypes.byref(ms)))
    return ms.value


# ---- permanent with its gradient (adjoint sweep) ----

def permanent_grad(A) -> tuple[complex, np.ndarray]:
    """(perm(A), G) with G[i, j] = d perm / d a_ij = perm of the (i, j) minor; A square."""
    M, ld = as_rows(A)
    n = M.shape[0]
    if M.shape[1] != n:
        raise EngineError(PQB_ERR_ARG, "the gradient needs a square matrix")
    out = (ctypes.c_double * 2)()
    G = np.zeros((n, n), dtype=np.complex128)
    _check(lib().pqb_perm_grad(M.ctypes.data_as(_dp), n, ld, out, G.ctypes.data_as(_dp)))
    return complex(out[0], out[1]), G


def permanent_grad_batch(As) -> tuple[np.ndarray, np.ndarray]:
    """Batch of equal-size square matrices (count x n x n) -> (perms[count], grads[count, n, n])."""
    B, ld, stride = _stack_strides(As)
    if B.ndim != 3 or B.shape[1] != B.shape[2]:
        raise EngineError(PQB_ERR_ARG, "expected a count x n x n batch of square matrices")
    count, n = B.shape[0], B.shape[1]
    perms = np.zeros(count, dtype=np.complex128)
    G = np.zeros((count, n, n), dtype=np.complex128)
    _check(lib().pqb_perm_grad_batch(B.ctypes.data_as(_dp), count, n, ld, stride, perms.ctypes.data_as(_dp),
                                     G.ctypes.data_as(_dp)))
    return perms, G


def grad_workspace_bytes(count: int, n: int) -> int:
    b = ctypes.c_size_t(0)
    _check(lib().pqb_perm_grad_workspace(count, n, ctypes.byref(b)))
    return int(b.value)


def permanent_grad_device(dA_ptr: int, count: int, n: int, d_perm_ptr: int, d_grad_ptr: int, d_work_ptr: int,
                          work_bytes: int, stream: int = 0) -> None:
    """Asynchronous adjoint sweep on device pointers (see pqb_perm_grad_device)."""
    _check(lib().pqb_perm_grad_device(dA_ptr, count, n, d_perm_ptr, d_grad_ptr, d_work_ptr, work_bytes,
                                      stream or None))


def grad_plan(count: int, n: int, waves_resident: int = 0) -> dict:
    p = (ctypes.c_int64 * 7)()
    _check(lib().pqb_grad_plan(count, n, waves_resident, p))
    return {"M": p[0], "K": p[1], "L": p[2], "H": p[3], "wpt": p[4], "blocks": p[5], "lds": p[6]}


# ---- batch of plain permanents on device pointers ----

def batch_device_workspace_bytes(count: int, rows: int, cols: int) -> int:
    b = ctypes.c_size_t(0)
    _check(lib().pqb_perm_batch_device_workspace(count, rows, cols, ctypes.byref(b)))
    return int(b.value)


def permanent_batch_device(dA_ptr: int, count: int, rows: int, cols: int, d_out_ptr: int, d_work_ptr: int,
                           work_bytes: int, stream: int = 0) -> None:
    """Asynchronous batch of plain permanents on device pointers (see pqb_perm_batch_device)."""
    _check(lib().pqb_perm_batch_device(dA_ptr, count, rows, cols, d_out_ptr, d_work_ptr, work_bytes, stream or None))


def permanent_minors(A) -> np.ndarray:
    """For a rows x (rows+1) matrix: [perm(A without column j) for j in range(rows + 1)]."""
    M, ld = as_rows(A)
    out = np.zeros(M.shape[1], dtype=np.complex128)
    _check(lib().pqb_perm_minors(M.ctypes.data_as(_dp), M.shape[0], M.shape[1], ld,
                                 out.ctypes.data_as(_dp)))
    return out


def permanent_minors_batch(A) -> np.ndarray:
    """For a (count, rows, rows+1) stack: out[b, j] = perm(A[b] without column j), one launch."""
    M, ld, stride = _stack_strides(A)
    if M.ndim != 3:
        raise Exception("permanent_minors_batch: a (count, rows, rows + 1) array is needed")
    count, rows, cols = M.shape
    out = np.zeros((count, cols), dtype=np.complex128)
    if count == 0:
        return out
    _check(lib().pqb_perm_minors_batch(M.ctypes.data_as(_dp), count, rows, cols, ld, stride,
                                       out.ctypes.data_as(_dp)))
    return out


def clock_stamp(d_dst_ptr: int, stream: int = 0) -> None:
    """Asynchronously write the (shader clock, 100 MHz real-time) counters to d_dst (2 x u64)."""
    _check(lib().pqb_clock_stamp(d_dst_ptr, stream or None))


def probe_fp64_fma(waves_per_simd: int, iters: int = 1 << 16) -> tuple[float, float]:
    """(ms, TFLOP/s) of the FP64 FMA issue-rate probe at that occupancy."""
    ms = ctypes.c_double(0.0)
    tf = ctypes.c_double(0.0)
    _check(lib().pqb_probe_fp64_fma(waves_per_simd, iters, ctypes.byref(ms), ctypes.byref(tf)))
    return ms.value, tf.value
