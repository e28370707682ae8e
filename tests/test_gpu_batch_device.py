"""Batch of plain permanents on device pointers (pqb_perm_batch_device):
one launch for `count` equal-size matrices, each checked against the oracle
or a closed form; rectangular shapes; more matrices than one grid row holds
(65535); a second call on the same workspace (counters left at zero)."""

import numpy as np
import pytest

import oracle
from conftest import random_complex, rel_err
from piquassoboost_amd import _capi as C

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def run_batch(B):
    count, R, Cc = B.shape
    dA = torch.from_numpy(np.ascontiguousarray(B)).to(DEV)
    out = torch.zeros((count, 6), dtype=torch.float64, device=DEV)
    wb = C.batch_device_workspace_bytes(count, R, Cc)
    work = torch.zeros(wb, dtype=torch.uint8, device=DEV)
    st = torch.cuda.current_stream(DEV).cuda_stream
    res = []
    for _ in range(2):
        C.permanent_batch_device(dA.data_ptr(), count, R, Cc, out.data_ptr(), work.data_ptr(), wb, st)
        torch.cuda.synchronize()
        o = out.cpu().numpy()
        res.append(o[:, 0] + 1j * o[:, 1])
    assert np.array_equal(res[0], res[1])  # deterministic, counters reset by the first call
    return res[0]


@pytest.mark.parametrize("R,Cc,count", [(20, 20, 256), (26, 26, 8), (5, 6, 33), (7, 6, 10), (1, 1, 5), (12, 14, 100)])
def test_against_oracle(R, Cc, count):
    rng = np.random.default_rng(R * 100 + Cc)
    B = np.stack([random_complex(rng, R, Cc) for _ in range(count)])
    print(f"batch device {R}x{Cc} x {count}", flush=True)
    got = run_batch(B)
    for b in rng.choice(count, size=min(count, 6), replace=False):
        assert rel_err(got[b], oracle.bbfg(B[b], "dd")) < 1e-12, (R, Cc, int(b))
        assert rel_err(got[b], C.permanent(B[b])) < 1e-12


def test_more_matrices_than_one_grid_row():
    rng = np.random.default_rng(9)
    count = 70000
    B = rng.standard_normal((count, 3, 3)) + 1j * rng.standard_normal((count, 3, 3))
    got = run_batch(B)
    a = B
    ref = (a[:, 0, 0] * (a[:, 1, 1] * a[:, 2, 2] + a[:, 1, 2] * a[:, 2, 1])
           + a[:, 0, 1] * (a[:, 1, 0] * a[:, 2, 2] + a[:, 1, 2] * a[:, 2, 0])
           + a[:, 0, 2] * (a[:, 1, 0] * a[:, 2, 1] + a[:, 1, 1] * a[:, 2, 0]))
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def test_shape_errors():
    with pytest.raises(C.EngineError):
        C.batch_device_workspace_bytes(4, 8, 6)  # rows >= cols + 2
    with pytest.raises(C.EngineError):
        C.batch_device_workspace_bytes(0, 4, 4)
