import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built liblss.so")


@pytest.fixture(scope="session")
def synth1():
    """Config-1 workload (64 ch, 32 units, 2 x 60000 samples), seed 1."""
    from livespikesorter_amd import synth
    return synth.make(1, seed=1)


@pytest.fixture(scope="session")
def small_synth():
    """A reduced workload for fast CPU checks: 32 ch, 8 units, 2 x 10000 samples."""
    from livespikesorter_amd import synth
    return synth.generate(synth.SynthConfig(n_chan=32, n_units=8, n_samples=20000, NT=10000,
                                            nrange=16, nC=4, seed=3))


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))
