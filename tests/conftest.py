import glob
import json
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

GOLDEN_DIR = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (select with -m gpu)")


def load_golden():
    out = []
    for path in sorted(glob.glob(os.path.join(GOLDEN_DIR, "*.json"))):
        with open(path) as f:
            out.append(json.load(f))
    return out


def golden_runs(max_n=None):
    """(fixture, run) pairs, optionally limited to matrices with N <= max_n."""
    pairs = []
    for fx in load_golden():
        if max_n is not None and fx["N"] > max_n:
            continue
        for run in fx["runs"]:
            pairs.append((fx, run))
    return pairs


def run_id(pair):
    fx, run = pair
    extra = "-".join(f"{k}={v}" for k, v in sorted(run["env"].items())
                     if k not in ("CG_MATRIX_FORMAT", "CG_PRECONDITIONER"))
    return "-".join(x for x in (fx["case"], run["env"]["CG_MATRIX_FORMAT"],
                                run["env"]["CG_PRECONDITIONER"], extra) if x)


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    oracle.lib()
    return oracle
