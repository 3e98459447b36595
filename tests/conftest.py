import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "isspa-force-cuda_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def adi2():
    """The reference's 44-atom ADI2 deck (timing_tests/ADI2/10000MC), nMC=1000."""
    from isspa_amd import io
    d = os.path.join(GOLDEN, "adi2")
    return io.load_system(os.path.join(d, "ADI2.prmtop"), os.path.join(d, "ADI2.isspatop"),
                          os.path.join(d, "ADI2.rst7"), lbox=200.0, n_mc=1000, seed=8622,
                          name="ADI2")


@pytest.fixture(scope="session")
def cfg0():
    from isspa_amd import synth
    return synth.workload(0)


@pytest.fixture(scope="session")
def cfg1():
    from isspa_amd import synth
    return synth.workload(1)


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
