"""Autotuner for the walk length (Gray steps per chunk) of the batched sweep.

The walk length trades the per-chunk seed (mixed-radix decomposition of the
chunk index, one FMA row update per chunk digit) against parallelism: long
walks amortise the seed, short walks give every resident lane work.  The
tuner groups a batch by photon count, times the graph replay of each group's
sub-batch for a list of candidate walk lengths, and keeps the fastest per
photon count.  `--refine` then starts from the stored table and tunes it on the
whole batch by coordinate descent (each photon count's walk halved or doubled,
kept when the batch's replay gets faster).  The table is stored as JSON next to
this module and applied when the engine library is loaded.

    python -m piquassoboost_amd.autotune [--out PATH]            # per photon count, 4096-task batch
    python -m piquassoboost_amd.autotune --refine [--out PATH]   # refine on the whole batch
"""

from __future__ import annotations

import argparse
import json
from collections import defaultdict
from pathlib import Path

from . import _capi as C

TABLE_PATH = Path(__file__).resolve().parent / "tuning" / "walk_lengths_gfx950.json"
CANDIDATES = (16, 32, 64, 128, 256, 512, 1024, 2048, 4096)


def load(path: Path = TABLE_PATH) -> dict:
    if not Path(path).exists():
        return {}
    data = json.loads(Path(path).read_text())
    return {int(k): int(v) for k, v in data.get("walk_length_by_photons", {}).items()}


def apply(table: dict) -> None:
    C.set_walk_lengths(table or None)


def tune(tasks, candidates=CANDIDATES, reps: int = 3) -> dict:
    """tasks: (matrix, row_mult, col_mult) triples.  Returns {photons: best walk}."""
    groups = defaultdict(list)
    for t in tasks:
        groups[int(C.as_state(t[2]).sum())].append(t)
    best = {}
    timings = {}
    for n, sub in sorted(groups.items()):
        res = {}
        for L in candidates:
            C.permanent_batch(sub, walk_length=L)
            res[L] = C.batch_replay_ms(reps) / reps
        best[n] = min(res, key=res.get)
        timings[n] = res
    return best, timings


def _batch_ms(tasks, table, reps: int = 5) -> float:
    """Replay time of the whole batch under walk-length table `table`: best of three
    averages of `reps` replays after ~50 ms of replays (the clock climbs for the
    first ~60 ms of load)."""
    C.set_walk_lengths(table)
    C.permanent_batch(tasks)
    t = C.batch_replay_ms(2) / 2
    C.batch_replay_ms(int(min(2000, max(1, 50.0 / max(t, 1e-3)))))
    return min(C.batch_replay_ms(reps) / reps for _ in range(3))


def refine(tasks, table: dict, passes: int = 2, gain: float = 0.005):
    """Coordinate descent on the whole batch: for each photon count, halve or double
    its walk length and keep the change when the batch's replay gets faster by more
    than `gain` (relative).  Tuning each photon count's sub-batch alone misses how the
    groups share the persistent grid; this measures the mix itself."""
    table = dict(table)
    best = _batch_ms(tasks, table)
    history = [("start", best)]
    lo, hi = min(CANDIDATES), max(CANDIDATES)
    for _ in range(passes):
        changed = False
        for n in sorted(table):
            for L in (table[n] // 2, table[n] * 2):
                if L < lo or L > hi:
                    continue
                trial = dict(table)
                trial[n] = L
                t = _batch_ms(tasks, trial)
                if t < best * (1.0 - gain):
                    table, best, changed = trial, t, True
                    history.append((f"{n}:{L}", t))
        if not changed:
            break
    C.set_walk_lengths(table)
    return table, best, history


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--out", default=str(TABLE_PATH))
    ap.add_argument("--refine", action="store_true",
                    help="start from the stored table and refine it on the whole batch")
    a = ap.parse_args()
    from .sampling.utilities import adapt_interferometer, uneven_batch

    U, pairs = uneven_batch(seed=1005, count=4096)
    tasks = [(adapt_interferometer(U, i, o), o[o > 0], i[i > 0]) for i, o in pairs]
    if a.refine:
        start = load()
        table, ms, history = refine(tasks, start)
        out = Path(a.out)
        data = json.loads(TABLE_PATH.read_text()) if TABLE_PATH.exists() else {}
        data["walk_length_by_photons"] = {str(k): v for k, v in sorted(table.items())}
        data["refined_on_batch"] = {"replay_ms": round(ms, 4), "steps": [[k, round(t, 4)] for k, t in history]}
        out.parent.mkdir(parents=True, exist_ok=True)
        out.write_text(json.dumps(data, indent=1))
        print(json.dumps({"table": {str(k): v for k, v in sorted(table.items())}, "replay_ms": ms,
                          "history": history}))
        return 0
    C.set_walk_lengths(None)
    best, timings = tune(tasks)
    out = Path(a.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    out.write_text(json.dumps({
        "device": "gfx950",
        "workload": "4096 uneven repeated-row permanents (n ~ U[8,26]), replay ms per photon-count group",
        "walk_length_by_photons": {str(k): v for k, v in best.items()},
        "replay_ms": {str(k): {str(L): round(t, 5) for L, t in v.items()} for k, v in timings.items()},
    }, indent=1))
    print(json.dumps({str(k): v for k, v in best.items()}))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
