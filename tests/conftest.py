import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")
    config.addinivalue_line("markers", "slow: long CPU or GPU case")


def rel_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def seeded_field(shape, nc, seed, scale=1.0):
    """Complex field from the shared splitmix64 stream (same in oracle and product)."""
    from gp5_amd import synth

    M = int(np.prod(shape)) * nc
    u = synth.uniform(seed, 2 * M)
    a = (u[0::2] + 1j * u[1::2]).reshape(tuple(shape) + (nc,), order="F") * scale
    return np.asfortranarray(a)


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O

    O.build()
    return O
