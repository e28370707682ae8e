"""Shared test configuration.

Markers: ``gpu`` -- needs a visible MI355X (run with ``-m gpu`` on a machine that has one).
Everything unmarked runs on a CPU-only machine: the oracle against the golden
fixtures and closed forms, host-side logic, and that the C-ABI library loads
and exports every symbol of include/pqb_perm.h.
"""

import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


# Worst relative error of each parity check of a run, printed after the tests so that
# the per-result errors are in the run's summary and not only in captured output.
PARITY = {}


def record_parity(name, err, bar):
    """Keep the worst `err` seen for check `name` (with its bar) for the run summary."""
    err = float(err)
    prev = PARITY.get(name)
    if prev is None or err > prev[0]:
        PARITY[name] = (err, float(bar))


def pytest_terminal_summary(terminalreporter, exitstatus, config):
    if not PARITY:
        return
    terminalreporter.write_sep("-", "parity: worst relative error of each check (bar)")
    for name in sorted(PARITY):
        err, bar = PARITY[name]
        terminalreporter.write_line(f"{name}: {err:.2e} (bar {bar:.0e})")


@pytest.fixture(scope="session")
def kat_cases():
    cases = json.loads((GOLDEN / "kat.json").read_text())
    for c in cases:
        m = np.array(c["matrix"], dtype=np.float64)
        c["matrix"] = m[..., 0] + 1j * m[..., 1]
        c["value"] = complex(*c["value"])
        c["input"] = np.array(c["input"], dtype=np.int64)
        c["output"] = np.array(c["output"], dtype=np.int64)
    return cases


@pytest.fixture(scope="session")
def haar_blocks():
    return dict(np.load(GOLDEN / "haar_blocks.npz"))


def rel_err(a, b):
    b = complex(b)
    return abs(complex(a) - b) / max(abs(b), 1e-300)


def random_complex(rng, r, c, scale=None):
    s = scale if scale is not None else 1.0 / np.sqrt(2 * max(r, c))
    return (rng.standard_normal((r, c)) + 1j * rng.standard_normal((r, c))) * s
