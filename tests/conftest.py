import json
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "linear-solver_amd"))

GOLDEN_DIR = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(GOLDEN_DIR, "golden.json")) as fp:
        meta = json.load(fp)
    arrs = np.load(os.path.join(GOLDEN_DIR, "golden.npz"), allow_pickle=False)
    return meta, arrs


@pytest.fixture(scope="session")
def golden_deep():
    """The reference solvers at the deep histories 16 < m <= 32
    (tests/golden/make_golden_deep.py)."""
    with open(os.path.join(GOLDEN_DIR, "golden_deep.json")) as fp:
        meta = json.load(fp)
    return meta, np.load(os.path.join(GOLDEN_DIR, "golden_deep.npz"), allow_pickle=False)


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    return O


@pytest.fixture(scope="session")
def gpu():
    """The HIP library, loaded on a GPU host (no CPU fallback exists)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from aar_hip import _lib
    _lib.load()
    import aar_hip
    return aar_hip
