import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


def random_rowptr(rng, nrows, mode):
    """Row-length mixes that exercise every scan branch."""
    if mode == 0:
        lens = rng.integers(0, 20, nrows)
    elif mode == 1:
        lens = rng.choice([0, 1, 2, 3, 50, 81, 130, 400], nrows)
    elif mode == 2:
        lens = rng.geometric(0.05, nrows) - 1
    else:
        lens = np.full(nrows, int(rng.integers(1, 90)))
    lens = np.minimum(lens, 5000)
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)


def random_csr(rng, nrows, ncols, mode):
    rp = random_rowptr(rng, nrows, mode)
    lens = np.diff(rp)
    cols = np.empty(int(rp[-1]), np.int32)
    for r in range(nrows):
        n = int(lens[r])
        if n <= ncols:
            cols[rp[r]:rp[r + 1]] = np.sort(rng.choice(ncols, n, replace=False))
        else:
            cols[rp[r]:rp[r + 1]] = np.sort(rng.integers(0, ncols, n))
    vals = rng.uniform(-1.0, 1.0, int(rp[-1]))
    return rp, cols, vals


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "kat_wpk.json")) as f:
        return json.load(f)
