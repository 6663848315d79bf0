import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
DATA = os.path.join(GOLDEN, "data")

# The drop-in library appends a record to ./gpu_stats.txt per call, like the
# reference; keep test runs from writing into the working tree.
os.environ.setdefault("LSAMD_GPU_STATS", "0")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) GPU")
    config.addinivalue_line("markers", "slow: long CPU check, run with LSAMD_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("LSAMD_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow; set LSAMD_SLOW=1")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)


def load_data(name):
    """(obj_id, t, y, dy|None) of a reference data file."""
    d = np.loadtxt(os.path.join(DATA, name), delimiter=",")
    return (d[:, 0].astype(np.uint32), d[:, 1].copy(), d[:, 2].copy(),
            d[:, 3].copy() if d.shape[1] > 3 else None)


@pytest.fixture(scope="session")
def golden():
    with np.load(os.path.join(GOLDEN, "reference_cpu_fixtures.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}
