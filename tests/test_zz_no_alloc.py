"""pqb_perm_device_noalloc: the plain permanent on device pointers with its
workspace in the library's static device scratch buffer (no allocation, no
workspace argument).  Checked against pqb_perm_device (the same sweep with a
caller workspace, same launch plan) bit for bit, on two calls in a row -- the
second with a smaller matrix, so stale scratch contents of the first call
would show -- and against the oracle."""

import ctypes

import numpy as np
import pytest

import oracle
from conftest import rel_err
from piquassoboost_amd import _capi as C
from piquassoboost_amd.sampling.utilities import haar_unitary

pytestmark = pytest.mark.gpu


def _noalloc(dA_ptr, rows, cols, d_out_ptr, stream):
    """Straight through the C ABI (include/pqb_perm.h)."""
    C._check(C.lib().pqb_perm_device_noalloc(ctypes.c_void_p(dA_ptr), rows, cols, ctypes.c_void_p(d_out_ptr),
                                             ctypes.c_void_p(stream or None)))


def _with_workspace(torch, dA, rows, cols, stream):
    out = torch.zeros(6, dtype=torch.float64, device="cuda")
    wb = C.device_workspace_bytes(rows, cols)
    work = torch.zeros((wb + 7) // 8, dtype=torch.float64, device="cuda")
    C.permanent_device(dA.data_ptr(), rows, cols, out.data_ptr(), work.data_ptr(), wb, stream)
    return out


def test_noalloc_matches_workspace_entry_two_calls():
    import torch

    stream = torch.cuda.current_stream().cuda_stream
    U = haar_unitary(60, 1002)
    results = []
    for rows, cols in ((26, 26), (14, 14), (20, 21), (9, 8)):
        A = np.ascontiguousarray(U[:rows, :cols])
        dA = torch.from_numpy(A.view(np.float64).copy()).cuda()
        out = torch.full((6,), np.nan, dtype=torch.float64, device="cuda")
        print(f"noalloc {rows}x{cols}", flush=True)
        _noalloc(dA.data_ptr(), rows, cols, out.data_ptr(), stream)
        ref_dev = _with_workspace(torch, dA, rows, cols, stream)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        want = ref_dev.cpu().numpy()
        assert np.array_equal(got, want), (rows, cols, got, want)
        v = complex(got[0], got[1])
        ref = oracle.bbfg(A, "dd")
        scale = abs(ref) if rows == cols else max(abs(ref), np.prod(np.abs(A).sum(axis=0)) * 1e-3)
        assert abs(v - ref) <= 1e-12 * scale, (rows, cols, v, ref)
        results.append(v)
    # the same matrix again after smaller ones: same bits
    A = np.ascontiguousarray(U[:26, :26])
    dA = torch.from_numpy(A.view(np.float64).copy()).cuda()
    out = torch.zeros(6, dtype=torch.float64, device="cuda")
    _noalloc(dA.data_ptr(), 26, 26, out.data_ptr(), stream)
    torch.cuda.synchronize()
    assert complex(*out[:2].cpu().numpy()) == results[0]


def test_noalloc_rejects_sizes_beyond_the_scratch():
    import torch

    dA = torch.zeros(2 * 41 * 41, dtype=torch.float64, device="cuda")
    out = torch.zeros(6, dtype=torch.float64, device="cuda")
    with pytest.raises(C.EngineError, match="static scratch buffer"):
        _noalloc(dA.data_ptr(), 41, 41, out.data_ptr(), 0)
    with pytest.raises(C.EngineError, match="1 <= rows <= cols \\+ 1"):
        _noalloc(dA.data_ptr(), 5, 3, out.data_ptr(), 0)
    assert C.lib().pqb_perm_device_noalloc_scratch_bytes() < 256 * 1024
