"""Autotuner for the walk length (Gray steps per chunk) of the batched sweep.

The walk length trades the per-chunk seed (mixed-radix decomposition of the
chunk index, one FMA row update per chunk digit) against parallelism: long
walks amortise the seed, short walks give every resident lane work.  The
tuner groups a batch by photon count, times the graph replay of each group's
sub-batch for a list of candidate walk lengths, and keeps the fastest per
photon count.  The table is stored as JSON next to this module and applied
when the engine library is loaded.

    python -m piquassoboost_amd.autotune [--out PATH]   # tunes on the 4096-task uneven batch
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


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--out", default=str(TABLE_PATH))
    a = ap.parse_args()
    from .sampling.utilities import adapt_interferometer, uneven_batch

    U, pairs = uneven_batch(seed=1005, count=4096)
    tasks = [(adapt_interferometer(U, i, o), o[o > 0], i[i > 0]) for i, o in pairs]
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
