"""Test configuration: the `gpu` marker, import paths and shared fixtures.

CPU tests (`-m "not gpu"`) check the oracle against the golden fixtures and
closed forms, the host-side library (generators, cleaning, Matrix Market
reader, exported symbols).  GPU tests (`-m gpu`) call libgraphhip.so through
its C ABI and compare with the oracle on the same inputs.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"
GOLDEN_NAMES = ["karate", "petersen", "grid5", "er60", "plc80"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libgraphhip.so")


def load_golden(name):
    """Fixture arrays plus the symmetric CSR as an SpMtx."""
    from graphkokkos_amd import SpMtx
    z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    n = int(z["n"])
    A = SpMtx(n, n, z["rst"], z["col"])
    return A, {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def gpu():
    from graphkokkos_amd import Context
    ctx = Context(0)
    yield ctx
    ctx.close()
