import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libnrqcd_amd.so")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def oracle():
    from oracle.pyoracle import Oracle, build

    build()
    return Oracle()


@pytest.fixture(scope="session")
def refmods():
    from oracle.pyoracle import RefModules

    if not RefModules.available():
        pytest.skip("oracle/_ref not built (needs /root/reference and amdflang)")
    return RefModules()


def _gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device count only; does not initialise HIP)

        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def nrq():
    """The product library; GPU tests fail loudly when it is missing."""
    import nrqcd_amd

    nrqcd_amd.library()
    return nrqcd_amd


@pytest.fixture()
def lattice(nrq):
    """Factory: (re)initialise the library lattice module for one test."""
    state = {"on": False}

    def init(ext, spacing):
        if state["on"]:
            nrq.lattice_finalize()
        nrq.lattice_init(ext, spacing)
        state["on"] = True

    yield init
    if state["on"]:
        nrq.lattice_finalize()
