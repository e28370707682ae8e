"""The device entry points on several streams at once: each stream with its own
workspace, launches interleaved from the host, results equal to the same calls made one
after another on the default stream, bit for bit.  (The no-allocation entry shares one
static buffer per GPU, so its calls must not overlap; it is not run here.)"""

import numpy as np
import pytest

from conftest import random_complex
from piquassoboost_amd import _capi as C

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def test_plain_and_grad_device_on_four_streams():
    rng = np.random.default_rng(4201)
    sizes = [22, 20, 16, 24]
    mats = [random_complex(rng, n, n) for n in sizes]
    dAs = [torch.from_numpy(A).to(DEV) for A in mats]
    outs = [torch.zeros(6, dtype=torch.float64, device=DEV) for _ in sizes]
    works = [torch.zeros(C.device_workspace_bytes(n, n), dtype=torch.uint8, device=DEV) for n in sizes]
    gcount = 2
    gmats = [np.stack([random_complex(rng, n, n) for _ in range(gcount)]) for n in (12, 13, 14, 15)]
    gA = [torch.from_numpy(B).to(DEV) for B in gmats]
    gp = [torch.zeros(gcount, dtype=torch.complex128, device=DEV) for _ in gmats]
    gG = [torch.zeros_like(a) for a in gA]
    gw = [torch.zeros(C.grad_workspace_bytes(gcount, B.shape[1]), dtype=torch.uint8, device=DEV) for B in gmats]

    def launch(i, st):
        n = sizes[i]
        C.permanent_device(dAs[i].data_ptr(), n, n, outs[i].data_ptr(), works[i].data_ptr(), works[i].numel(), st)
        m = gmats[i].shape[1]
        C.permanent_grad_device(gA[i].data_ptr(), gcount, m, gp[i].data_ptr(), gG[i].data_ptr(), gw[i].data_ptr(),
                                gw[i].numel(), st)

    # one after another on the default stream
    print("streams: sequential reference", flush=True)
    for i in range(4):
        launch(i, torch.cuda.current_stream(DEV).cuda_stream)
    torch.cuda.synchronize()
    ref = [o.clone() for o in outs] + [p.clone() for p in gp] + [g.clone() for g in gG]
    for t in outs + gp + gG:
        t.fill_(float("nan"))
    torch.cuda.synchronize()
    # the same launches on four streams, several rounds interleaved
    print("streams: four streams", flush=True)
    streams = [torch.cuda.Stream(DEV) for _ in range(4)]
    for rnd in range(3):
        for i, s in enumerate(streams):
            launch(i, s.cuda_stream)
        torch.cuda.synchronize()
        got = outs + gp + gG
        for k, (a, b) in enumerate(zip(got, ref)):
            assert torch.equal(a.view(torch.uint8), b.view(torch.uint8)), f"round {rnd}, output {k}"
