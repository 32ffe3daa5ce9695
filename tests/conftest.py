import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def load_fixture(name):
    """Fixture npz -> nested dict ('in.lam' -> d['in']['lam'])."""
    z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    out: dict = {}
    for k in z.files:
        parts = k.split(".")
        d = out
        for p in parts[:-1]:
            d = d.setdefault(p, {})
        d[parts[-1]] = z[k]
    return out


def case_from_fixture(fx):
    """The case dict stored under 'in' with scalars unwrapped."""
    c = {}
    for k, v in fx["in"].items():
        c[k] = v.item() if v.ndim == 0 else v
    return c


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    O.load("oracle")
    return O


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
