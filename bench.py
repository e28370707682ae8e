"""Benchmark: n=30 complex128 Glynn/BBFG permanents per second on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: the 30x30 top-left block of a seeded 60-mode Haar unitary
(complex Gaussian, QR, phase fix; seed 1002), resident in HBM before the
timed region.  One step = one full permanent (2^29 Gray steps) through the
device-pointer entry point of libpqb_perm.so on torch's current stream.

Prints one JSON line with the throughput, the FP64 roofline fraction of the
sweep kernel (algorithmic flops (8n-4) 2^(n-1) per permanent over the mean
launch time measured with HIP events on the launch stream), and a CPU
baseline: the restated reference TBB sweep (double, std::thread, one chunk per
4 x threads as BBFGPermanentCalculator.hpp:120-134) timed on this host.  The same
line carries the other workloads' legs: the 30-photon repeated-row permanent, the
4096-task batch against per-call launches, the gradient workload's forward and
adjoint, the seeded sampler, the double-double and exact sweeps.

Measurement modes that replace the benchmark: --fma-probe (FP64 issue rate at 1..8
waves per SIMD), --walk-sweep (walk lengths of the 30-photon task), --power S
(socket power and clock under each kernel for S seconds, amd-smi), --copy-mb M (a
copy of known traffic for calibrating the HBM counters).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

N = 30
M_MODES = 60
SEED = 1002
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (and matrix) peak
# HBM bytes per launch of the n=30 sweep: FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00 (factors from a
# 256 MiB copy kernel), measured with rocprofv3 --pmc on this workload
# (profiles/plain_sweep_n30_counters.md).
TRAFFIC_BYTES_N30 = 863296


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden")

# Every timed GPU section runs after at least this much of the same work: the clock
# climbs from ~1.8 to its sustained ~2.3 GHz over the first ~60 ms of FP64 load.
SUSTAIN_MS = 100.0


def sustain(fn, ms: float = SUSTAIN_MS) -> None:
    """Run fn (GPU work, synchronized after each call) until `ms` of wall time has passed."""
    import torch

    t0 = time.perf_counter()
    while True:
        fn()
        torch.cuda.synchronize()
        if (time.perf_counter() - t0) * 1e3 >= ms:
            return


def sustain_replays(C, ms: float = SUSTAIN_MS) -> None:
    """Replay the last captured batch launch for about `ms` of GPU time."""
    t = C.batch_replay_ms(3) / 3
    C.batch_replay_ms(int(min(100000, max(1, np.ceil(ms / max(t, 1e-3))))))


def parity_vs_fixture(value: complex, A: np.ndarray, fixture: str, key: str, dd_key: str) -> dict:
    """Relative error of `value` against the stored double-double value of the same
    workload (the oracle's restated reference sweep), and whether the benchmark's
    seeded matrix is bit-identical to the stored one."""
    d = np.load(os.path.join(GOLDEN, fixture))
    h = d[dd_key]
    ref = complex(h[0] + h[1], h[2] + h[3])
    return {"fixture": f"tests/golden/{fixture}:{dd_key}", "rel_err": abs(value - ref) / abs(ref),
            "same_matrix": bool(np.array_equal(d[key], A))}


def flops_plain(n: int) -> float:
    return (8 * n - 4) * 2.0 ** (n - 1)


def host_cpus() -> dict:
    """The host's CPUs as this process sees them: the model, the logical CPU count
    (nproc), the CPUs the process may run on (affinity), the cgroup CPU quota and
    the OMP_NUM_THREADS setting.  The baseline runs one thread per CPU the process
    may use -- the reference sizes its TBB pool by hardware_concurrency()
    (BBFGPermanentCalculator.hpp:120-122) -- capped by the cgroup quota and by
    OMP_NUM_THREADS when either is set: a share of a larger machine gains nothing
    from threads beyond its quota."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(per)
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = affinity
    if quota:
        threads = min(threads, max(1, int(quota)))
    if omp and omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    return {"model": model, "nproc": os.cpu_count(), "affinity": affinity, "cgroup_quota": quota,
            "omp_num_threads": omp, "threads": max(1, threads)}


def cpu_baseline(A: np.ndarray, budget_s: float = 20.0) -> dict:
    """Time the restated reference sweep (double) on the host cores."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle

    cpus = host_cpus()
    threads = cpus["threads"]
    njobs = 4 * threads
    n = A.shape[0]
    # calibrate on a few jobs, then run a bounded number of whole jobs
    t0 = time.perf_counter()
    oracle.bbfg_partial(A, njobs, 0, threads, "double", threads)
    t_cal = time.perf_counter() - t0
    per_round = max(t_cal, 1e-6)  # one job per thread
    rounds = int(max(1, min(4, budget_s / per_round)))
    jobs = min(njobs, rounds * threads)
    t0 = time.perf_counter()
    oracle.bbfg_partial(A, njobs, 0, jobs, "double", threads)
    dt = time.perf_counter() - t0
    frac = jobs / njobs
    perms_per_s = frac / dt
    return {
        "value": perms_per_s,
        "unit": "perms/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{jobs} of {njobs} Gray-code chunks ({frac:.3f} of one n={n} permanent) in {dt:.2f} s, "
                  f"double, {threads} threads",
        "host": cpus,
    }


def batched_leg() -> dict:
    """4096 uneven repeated-row permanents (n in [8, 26], output
    multiplicities 1/2/3 with p = .7/.2/.1) from one 100-mode Haar unitary, as
    one batched launch against 4096 per-call launches on the same GPU."""
    from piquassoboost_amd import _capi as C
    from piquassoboost_amd.sampling.utilities import adapt_interferometer, uneven_batch

    U, pairs = uneven_batch(seed=1005, count=4096)
    tasks = [(adapt_interferometer(U, i, o), o[o > 0], i[i > 0]) for i, o in pairs]
    steps = 0
    flops = 0.0
    for A, r, c in tasks:
        p = C.repeated_plan(A, r, c, 1 << 40)
        s = p["walk"] * p["chunks"]
        steps += s
        flops += s * (8 * p["slots"] - 2)
    C.permanent_batch(tasks)  # warm: buffers sized for the batch, graph capture
    C.permanent_repeated(*tasks[0])
    sustain_replays(C)
    t0 = time.perf_counter()
    C.permanent_batch(tasks)
    t_batch = time.perf_counter() - t0
    replay_ms = C.batch_replay_ms(5) / 5
    # the same pairs as the unitary plus two (count, modes) state arrays
    ins = np.stack([i for i, _ in pairs])
    outs = np.stack([o for _, o in pairs])
    t0 = time.perf_counter()
    C.permanent_batch_states(U, ins, outs)
    t_states = time.perf_counter() - t0
    t0 = time.perf_counter()
    for A, r, c in tasks:
        C.permanent_repeated(A, r, c)
    t_call = time.perf_counter() - t0
    return {
        "workload": "4096 repeated-row permanents, n ~ U[8,26], output multiplicities {1,2,3} p=(.7,.2,.1), "
                    "submatrices of a 100-mode Haar unitary (seed 1005)",
        "tasks": len(tasks),
        "gray_steps": steps,
        "flops": flops,
        "per_call_s": t_call,
        "batched_s": t_batch,
        "batched_states_s": t_states,
        "batched_replay_ms": replay_ms,
        "speedup_batched_vs_per_call": t_call / t_batch,
        "speedup_replay_vs_per_call": t_call / (replay_ms * 1e-3),
        "replay_tflops": flops / (replay_ms * 1e-3) / 1e12,
    }


def thirty_photon_task():
    """(compact matrix, row multiplicities, column multiplicities) of the 30-photon
    transition: inputs in modes 0..29, 18 distinct output modes with multiplicities
    1 + Multinomial(12, 1/18), seeded 60-mode Haar unitary (seed 1004)."""
    from piquassoboost_amd.sampling.utilities import haar_unitary

    rng = np.random.default_rng(1004)
    U = haar_unitary(60, 1004)
    modes = np.sort(rng.choice(60, size=18, replace=False))
    rows = (1 + rng.multinomial(12, np.full(18, 1.0 / 18))).astype(np.int64)
    cols = np.ones(30, dtype=np.int64)
    return np.ascontiguousarray(U[np.ix_(modes, np.arange(30))]), rows, cols


def walk_sweep(walks=(24, 48, 72, 144, 216, 432), reps: int = 30) -> dict:
    """The 30-photon transition at several walk lengths (Gray steps per chunk): the
    walk the digit packing gives for each target, the chunk count, and the best of two
    averages of `reps` replays of the captured launch."""
    from piquassoboost_amd import _capi as C

    A, rows, cols = thirty_photon_task()
    ref = C.permanent_repeated(A, rows, cols)
    out = []
    for L in walks:
        p = C.repeated_plan(A, rows, cols, L)
        v = C.permanent_batch([(A, rows, cols)], walk_length=L)[0]
        sustain_replays(C)
        ms = min(C.batch_replay_ms(reps) / reps, C.batch_replay_ms(reps) / reps)
        out.append({"target": L, "walk": p["walk"], "chunks": p["chunks"], "chunk_digits": p["high"],
                    "replay_ms": ms, "rel_diff_to_default": abs(v - ref) / abs(ref)})
    return {"workload": "30 photons in 60 modes, 18 distinct outputs (Haar seed 1004)", "walks": out}


def repeated_leg(reps: int = 10) -> dict:
    """One repeated-row permanent, 30 photons in 60 modes -- inputs in modes
    0..29, 18 distinct output modes with multiplicities 1 + Multinomial(12, 1/18), seeded
    60-mode Haar unitary (seed 1004); the compact 18 x 30 matrix.  GPU per call (host API,
    wall) and per replay of the captured launch; the restated reference sweep (double,
    std::thread) on the host cores beside it."""
    from piquassoboost_amd import _capi as C

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle

    A, rows, cols = thirty_photon_task()
    plan = C.repeated_plan(A, rows, cols, 1 << 40)
    steps = plan["walk"] * plan["chunks"]
    flops = steps * (2 * 30 + 6 * 29 + 4)
    v = C.permanent_repeated(A, rows, cols)
    t0 = time.perf_counter()
    for _ in range(reps):
        C.permanent_repeated(A, rows, cols)
    call_ms = (time.perf_counter() - t0) / reps * 1e3
    sustain_replays(C)
    replay_ms = C.batch_replay_ms(reps) / reps
    threads = host_cpus()["threads"]
    t0 = time.perf_counter()
    h = oracle.bbfg_repeated(A, cols, rows, "double", threads=threads)
    host_ms = (time.perf_counter() - t0) * 1e3
    return {
        "workload": "30 photons in 60 modes, 18 distinct outputs (multiplicities "
                    + ",".join(str(int(r)) for r in rows) + "), Haar seed 1004",
        "gray_steps": int(steps),
        "flops": float(flops),
        "per_call_ms": call_ms,
        "replay_ms": replay_ms,
        "replay_tflops": flops / (replay_ms * 1e-3) / 1e12,
        "host_reference_ms": host_ms,
        "host_threads": threads,
        "gpu_vs_host": host_ms / call_ms,
        "perm": [v.real, v.imag],
        "parity": parity_vs_fixture(v, A, "repeated_c3.npz", "A", "value_dd"),
        "host_rel_diff": abs(h - v) / abs(v),
    }


def grad_workload():
    """The gradient workload resident on cuda:0: (blocks, stream, forward, sweep), where
    forward() launches the plain sweep of the 264 permanents (pqb_perm_batch_device, one
    launch per size) and sweep() the adjoint sweep (permanents and every d perm / d a_ij,
    one launch per size plus its reduction), both on torch's current stream."""
    import torch
    from piquassoboost_amd import _capi as C
    from piquassoboost_amd.sampling.utilities import gradient_blocks

    blocks, _ = gradient_blocks()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    fbufs, gbufs = [], []
    for n, B in sorted(blocks.items()):
        count = B.shape[0]
        dA = torch.from_numpy(B).to(dev)
        fout = torch.empty((count, 6), dtype=torch.float64, device=dev)
        fwb = C.batch_device_workspace_bytes(count, n, n)
        fwork = torch.zeros(fwb, dtype=torch.uint8, device=dev)
        fbufs.append((n, count, dA, fout, fwork, fwb))
        dp = torch.empty(count, dtype=torch.complex128, device=dev)
        dG = torch.empty_like(dA)
        wb = C.grad_workspace_bytes(count, n)
        work = torch.empty(wb, dtype=torch.uint8, device=dev)
        gbufs.append((n, count, dA, dp, dG, work, wb))

    def forward():
        for n, count, dA, fout, fwork, fwb in fbufs:
            C.permanent_batch_device(dA.data_ptr(), count, n, n, fout.data_ptr(), fwork.data_ptr(), fwb,
                                     stream.cuda_stream)

    def sweep():
        for n, count, dA, dp, dG, work, wb in gbufs:
            C.permanent_grad_device(dA.data_ptr(), count, n, dp.data_ptr(), dG.data_ptr(), work.data_ptr(), wb,
                                    stream.cuda_stream)

    return blocks, stream, forward, sweep


def grad_leg(reps: int = 10) -> dict:
    """Gradient workload: 256 n=20 and 8 n=26 complex128 blocks, block b the
    top-left corner of the 50-mode Haar unitary of seed 1006 + b, inputs
    resident in HBM.  Forward = the plain sweep of
    the 264 permanents (pqb_perm_batch_device, one launch per size).
    Forward+backward = the adjoint sweep, which yields the permanents and every
    d perm / d a_ij (one launch per size plus its reduction).  Both timed with
    HIP events on the launch stream; the same step through torch.autograd
    (loss sum log|perm|^2) by wall clock."""
    import torch
    from piquassoboost_amd import _capi as C
    from piquassoboost_amd.autograd import permanent

    blocks, stream, forward, sweep = grad_workload()
    dev = torch.device("cuda:0")

    def event_ms(fn):
        sustain(fn)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    fwd_ms = event_ms(forward)
    grad_ms = event_ms(sweep)
    # the persistent repeated-row kernel on the same 264 tasks (graph replay)
    tasks = []
    for n, B in sorted(blocks.items()):
        ones = np.ones(n, np.int64)
        tasks += [(B[b], ones, ones) for b in range(B.shape[0])]
    C.permanent_batch(tasks)
    sustain_replays(C)
    rep_ms = C.batch_replay_ms(reps) / reps

    leaves = [torch.from_numpy(B).to(dev).requires_grad_(True) for n, B in sorted(blocks.items())]

    def step():
        loss = sum(torch.log(permanent(A).abs() ** 2).sum() for A in leaves)
        loss.backward()

    sustain(step)
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    autograd_ms = (time.perf_counter() - t0) * 1e3 / reps

    steps = sum(B.shape[0] * 2.0 ** (n - 1) for n, B in blocks.items())
    flops = sum(B.shape[0] * 2.0 ** (n - 1) * (24 * n - 22) for n, B in blocks.items())
    return {
        "workload": "256 n=20 + 8 n=26 complex128 top-left blocks of 50-mode Haar unitaries (seeds 1006+b), "
                    "loss sum log|perm|^2",
        "gray_steps": steps,
        "forward_plain_ms": fwd_ms,
        "forward_repeated_batch_replay_ms": rep_ms,
        "fwd_bwd_adjoint_ms": grad_ms,
        "ratio_fwd_bwd_over_fwd": grad_ms / fwd_ms,
        # against whichever forward path is faster on the 264 blocks
        "ratio_fwd_bwd_over_fastest_fwd": grad_ms / min(fwd_ms, rep_ms),
        "autograd_step_ms": autograd_ms,
        "adjoint_flops": flops,
        "adjoint_tflops": flops / (grad_ms * 1e-3) / 1e12,
    }


def sampling_leg(photons: int = 20, modes: int = 60, shots: int = 8) -> dict:
    """Seeded generalized-Clifford boson sampling (20 photons in modes 0..19 of a 60-mode
    Haar interferometer, seed 1010): time per sample with compute_pmf's permanents from one
    column-minors sweep per photon step (the samples advanced together, one batched launch
    per step), from the per-mode batched permanents (one sample at a time), and from the
    oracle's restated reference sweep on the host (one call per mode, as the reference
    does).  All three draw the same samples.  Each GPU method runs once untimed first (the
    first launches of a process load the kernels)."""
    from piquassoboost_amd.sampling.simulation_strategies.GeneralizedCliffordsSimulationStrategy import (
        GeneralizedCliffordsBSimulationStrategy,
    )
    from piquassoboost_amd.sampling.utilities import haar_unitary

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle

    U = haar_unitary(modes, 1010)
    inp = np.zeros(modes, np.int64)
    inp[:photons] = 1
    threads = host_cpus()["threads"]

    def oracle_batch(tasks):
        return [1.0 + 0j if A.size == 0 else oracle.bbfg_repeated(A, c, r, "double", threads=threads)
                for A, r, c in tasks]

    res = {}
    samples = {}
    for name, kw in (("minors", {"pmf_method": "minors"}), ("batch", {"pmf_method": "batch"}),
                     ("host_reference", {"permanent_batch": oracle_batch, "pmf_method": "batch"})):
        if name != "host_reference":
            GeneralizedCliffordsBSimulationStrategy(U, 1011, **kw).simulate(inp, shots)
        strat = GeneralizedCliffordsBSimulationStrategy(U, 1010, **kw)  # seeds the shared generator
        t0 = time.perf_counter()
        out = strat.simulate(inp, shots)
        res[name] = (time.perf_counter() - t0) / shots
        samples[name] = [tuple(int(x) for x in v) for v in out]
    return {
        "workload": f"{photons} photons in modes 0..{photons - 1} of a {modes}-mode Haar interferometer "
                    f"(seed 1010), {shots} seeded samples",
        "s_per_sample_minors": res["minors"],
        "s_per_sample_batched": res["batch"],
        "s_per_sample_host_reference": res["host_reference"],
        "host_threads": threads,
        "same_samples": samples["minors"] == samples["batch"] == samples["host_reference"],
    }


def exact_leg(n: int = 26, count: int = 4, reps: int = 3) -> dict:
    """Four n=26 Bernoulli(1/2) 0/1 matrices (seed 1003): the exact 128-bit integer sweep
    (the double-precision codes route real integer matrices there), per call from the host
    API, beside the FP64 sweep of the same matrices."""
    from piquassoboost_amd import _capi as C

    rng = np.random.default_rng(1003)
    mats = [np.ascontiguousarray(rng.integers(0, 2, size=(n, n)).astype(np.float64)) for _ in range(count)]
    for M in mats:
        C.permanent_exact(M)
    t0 = time.perf_counter()
    values = []
    for _ in range(reps):
        values = [C.permanent_exact(M)[0] for M in mats]
    exact_ms = (time.perf_counter() - t0) / (reps * count) * 1e3
    t0 = time.perf_counter()
    for _ in range(reps):
        for M in mats:
            C.permanent(M, C.ALGO_BBFG, C.PREC_FP64)
    fp64_ms = (time.perf_counter() - t0) / (reps * count) * 1e3
    return {
        "workload": f"{count} Bernoulli(1/2) 0/1 matrices {n}x{n} (seed 1003), exact integer permanents",
        "exact_ms_per_call": exact_ms,
        "fp64_sweep_ms_per_call": fp64_ms,
        "permanents": [str(v) for v in values],
    }


def inf_leg(fp64_ms_n30: float, reps: int = 5) -> dict:
    """Double-double sweep (PQB_PREC_INF, the exact-arithmetic calculator codes):
    n=20 (40-mode Haar, seed 1001) and n=30 (the headline block) per permanent.
    Device time from replays of the captured launch (sweep + reduction); the
    host figure is the oracle's double-double restatement of the reference
    sweep on a bounded sample of the n=20 permanent."""
    from piquassoboost_amd import _capi as C
    from piquassoboost_amd.sampling.utilities import haar_unitary

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle

    out = {}
    for n, m, seed in ((16, 32, 1001), (20, 40, 1001), (30, M_MODES, SEED)):
        A = np.ascontiguousarray(haar_unitary(m, seed)[:n, :n])
        v = C.permanent(A, C.ALGO_BBFG, C.PREC_INF)
        sustain_replays(C)
        ms = C.batch_replay_ms(reps) / reps
        out[f"n{n}_ms"] = ms
        out[f"n{n}_perm"] = [v.real, v.imag]
    out["n30_vs_fp64_sweep"] = out["n30_ms"] / fp64_ms_n30
    A = np.ascontiguousarray(haar_unitary(40, 1001)[:20, :20])
    threads = host_cpus()["threads"]
    njobs = 64 * threads
    jobs = 4 * threads
    t0 = time.perf_counter()
    oracle.bbfg_partial(A, njobs, 0, jobs, "dd", threads)
    dt = time.perf_counter() - t0
    out["host_dd_n20_ms"] = dt * njobs / jobs * 1e3
    out["host_threads"] = threads
    out["gpu_vs_host_dd_n20"] = out["host_dd_n20_ms"] / out["n20_ms"]
    return out


def _smi_sample():
    """(socket power in W, graphics clock in MHz) of GPU 0 from amd-smi; (None, None)
    where the tool or a field is missing."""
    import subprocess

    try:
        out = subprocess.run(["amd-smi", "metric", "--power", "--clock", "--json"], capture_output=True,
                             text=True, timeout=10).stdout
        d = json.loads(out)
        g = (d["gpu_data"] if isinstance(d, dict) else d)[0]
        return float(g["power"]["socket_power"]["value"]), float(g["clock"]["gfx_0"]["clk"]["value"])
    except Exception:
        return None, None


def power_leg(secs: float) -> dict:
    """Socket power and graphics clock while each workload runs back to back for `secs`
    seconds: the n=30 sweep, the FP64 FMA probe at the sweep's three waves per SIMD, and the
    gradient workload's plain forward and adjoint sweep.  Samples every ~0.5 s after the
    first 2 s (the clock and power settle within ~0.1 s)."""
    import threading

    import torch

    from piquassoboost_amd import _capi as C
    from piquassoboost_amd.sampling.utilities import haar_unitary

    dev = torch.device("cuda:0")
    A = np.ascontiguousarray(haar_unitary(M_MODES, SEED)[:N, :N])
    dA = torch.from_numpy(A.view(np.float64).copy()).to(dev)
    d_out = torch.zeros(6, dtype=torch.float64, device=dev)

    def run(work):
        samples, stop = [], threading.Event()

        def sampler():
            if not stop.wait(2.0):
                while not stop.is_set():
                    samples.append(_smi_sample())
                    stop.wait(0.5)

        th = threading.Thread(target=sampler)
        th.start()
        t0 = time.perf_counter()
        rates = []
        while time.perf_counter() - t0 < secs:
            rates.append(work())  # ctypes releases the GIL inside the engine call
        stop.set()
        th.join()
        pw = [p for p, _ in samples if p is not None]
        mhz = [c for _, c in samples if c is not None]
        return {"socket_power_w": pw, "gfx_clock_mhz": mhz, "rate": rates[-1] if rates else None}

    def sweep():
        ms = C.time_permanent_device(dA.data_ptr(), N, N, d_out.data_ptr(), 100) / 100
        return {"ms_per_perm": ms, "frac_of_peak": flops_plain(N) / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}

    def probe():
        ms, tf = C.probe_fp64_fma(3, 1 << 17)
        return {"tflops": tf, "frac_of_peak": tf / FP64_PEAK_TFLOPS}

    res = {"seconds_each": secs, "n30_sweep": run(sweep), "fp64_fma_probe_3_waves": run(probe)}

    # the gradient workload: its plain forward, then its adjoint sweep
    _, stream, forward, adjoint = grad_workload()

    def timed(fn, reps=10):
        def work():
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                fn()
            e1.record(stream)
            torch.cuda.synchronize()
            return {"ms": e0.elapsed_time(e1) / reps}
        return work

    res["grad_forward_264"] = run(timed(forward))
    res["grad_adjoint_264"] = run(timed(adjoint))
    res["tool"] = "amd-smi metric --power --clock"
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    # The GPU climbs from ~1.8 to its sustained ~2.3 GHz over the first ~60 ms of the
    # n=30 sweep (20 permanents); the default warm-up covers that ramp.
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-per-call leg")
    ap.add_argument("--no-repeated", action="store_true", help="skip the repeated-row (30 photons) leg")
    ap.add_argument("--no-grad", action="store_true", help="skip the forward+backward leg")
    ap.add_argument("--no-sampling", action="store_true", help="skip the boson-sampling leg")
    ap.add_argument("--no-inf", action="store_true", help="skip the double-double (PQB_PREC_INF) leg")
    ap.add_argument("--no-exact", action="store_true", help="skip the exact integer (0/1 matrices) leg")
    ap.add_argument("--fma-probe", action="store_true",
                    help="instead of the benchmark, run the FP64 FMA issue-rate probe at 1, 2, 3, 4 and 8 waves per SIMD "
                         "(the FP64 VALU ceiling of the device; profile it for the clock)")
    ap.add_argument("--walk-sweep", action="store_true",
                    help="instead of the benchmark, time the 30-photon repeated-row permanent at several walk "
                         "lengths (replays of the captured launch)")
    ap.add_argument("--power", type=float, default=0.0,
                    help="instead of the benchmark, run the n=30 sweep, the FP64 FMA probe and the gradient "
                         "workload's forward and adjoint for this many seconds each and report socket power and "
                         "graphics clock (amd-smi) during each")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="instead of the benchmark, run 10 device-to-device copies of this many MiB "
                         "(reference kernel of known traffic for calibrating FETCH_SIZE / WRITE_SIZE)")
    a = ap.parse_args()
    if a.copy_mb:
        import torch

        src = torch.ones(a.copy_mb * (1 << 20) // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 8, "copies": 10}), flush=True)
        return 0
    if a.fma_probe:
        import torch  # noqa: F401  (one HIP runtime: torch's)

        from piquassoboost_amd import _capi as C

        res = {}
        for w in (1, 2, 3, 4, 8):
            ms, tf = C.probe_fp64_fma(w)
            res[f"waves_per_simd_{w}"] = {"ms": ms, "tflops": tf, "frac_of_peak": tf / FP64_PEAK_TFLOPS}
            print(f"fp64 fma probe, {w} waves/SIMD: {ms:.2f} ms, {tf:.1f} TFLOP/s", file=sys.stderr, flush=True)
        print(json.dumps({"fp64_fma_probe": res, "peak": FP64_PEAK_TFLOPS}), flush=True)
        return 0
    if a.power:
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        print(json.dumps(power_leg(a.power)), flush=True)
        return 0
    if a.walk_sweep:
        import torch  # noqa: F401  (one HIP runtime: torch's)

        print(json.dumps(walk_sweep()), flush=True)
        return 0
    if a.gpus != 1:
        print("bench.py measures one GPU: the permanent engine runs on a single MI355X "
              "(multi-GPU sharding is designed, not built); run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from piquassoboost_amd import _capi as C
    from piquassoboost_amd.sampling.utilities import haar_unitary

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 1
    dev = torch.device("cuda:0")
    U = haar_unitary(M_MODES, SEED)
    A = np.ascontiguousarray(U[:N, :N])
    dA = torch.from_numpy(A.view(np.float64).copy()).to(dev)
    d_out = torch.zeros(6, dtype=torch.float64, device=dev)
    wb = C.device_workspace_bytes(N, N)
    d_work = torch.zeros((wb + 7) // 8, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream

    def step():
        C.permanent_device(dA.data_ptr(), N, N, d_out.data_ptr(), d_work.data_ptr(), wb, sp)

    stamps = torch.zeros(4, dtype=torch.int64, device=dev)  # clock counters around the timed steps
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    C.clock_stamp(stamps.data_ptr(), sp)
    ev0.record(stream)
    for _ in range(a.steps):
        step()
    ev1.record(stream)
    C.clock_stamp(stamps.data_ptr() + 16, sp)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = ev0.elapsed_time(ev1)
    ms_per_step = ev_ms / a.steps
    value = a.steps / wall
    c0, r0, c1, r1 = (int(x) for x in stamps.cpu().tolist())
    # average shader clock over the timed steps (s_memtime ticks at it, s_memrealtime at 100 MHz)
    clock_ghz = (c1 - c0) / (r1 - r0) * 0.1 if r1 > r0 else None

    res = d_out.cpu().numpy()
    perm = complex(res[0], res[1])
    F = flops_plain(N)
    achieved = F / (ms_per_step * 1e-3) / 1e12
    line = {
        "metric": "n=30 complex perms/s (% FP64 peak); batched and fwd+bwd vs plain, same GPU",
        "value": value,
        "unit": "perms/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "plain Glynn/BBFG permanent, n=30 complex128 top-left block of a 60-mode Haar unitary "
                        "(seed 1002), one permanent per step",
            "n": N,
            "modes": M_MODES,
            "flops_per_perm": F,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": TRAFFIC_BYTES_N30,
            # the peak is quoted at 2.4 GHz; the GPU holds a lower clock under this load
            "clock_ghz": clock_ghz,
            "frac_at_clock": (achieved / (FP64_PEAK_TFLOPS * clock_ghz / 2.4)) if clock_ghz else None,
        },
        "event_ms_per_perm": ms_per_step,
        "perm": [perm.real, perm.imag],
        "parity": parity_vs_fixture(perm, A, "haar_blocks.npz", "c2a_n30", "c2a_n30_dd"),
    }
    if not a.no_repeated:
        line["repeated"] = repeated_leg()
    if not a.no_batch:
        line["batched"] = batched_leg()
    if not a.no_grad:
        line["grad"] = grad_leg()
    if not a.no_sampling:
        line["sampling"] = sampling_leg()
    if not a.no_inf:
        line["inf"] = inf_leg(ms_per_step)
    if not a.no_exact:
        line["exact"] = exact_leg()
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(A)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
