import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "rev-sim_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and librevsim.so")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def box216():
    import revsim

    return revsim.water_box(216, 1.862)


@pytest.fixture(scope="session")
def jittered216(box216):
    rng = np.random.default_rng(1)
    return np.ascontiguousarray(box216.coords + 0.01 * rng.standard_normal(box216.coords.size))


@pytest.fixture(scope="session")
def box3000():
    import revsim

    return revsim.water_box(3000, 4.48)


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def per_atom_rel(a, b):
    """max over atoms of |a_i - b_i| / rms(|b|): relative to the typical force magnitude."""
    a, b = np.asarray(a).reshape(-1, 3), np.asarray(b).reshape(-1, 3)
    rms = np.sqrt(np.mean(np.sum(b * b, axis=1)))
    return float(np.max(np.linalg.norm(a - b, axis=1)) / rms)


@pytest.fixture(scope="session")
def gpu_lib():
    import revsim

    try:
        return revsim.load_library()
    except revsim.LibraryMissing as e:
        pytest.fail(str(e))
