import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")
    # The CPU checker is test infrastructure; build it if this tree has not been built.
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"], check=True)


def _gpu_present():
    try:
        import torch  # noqa: F401  (device count only; does not initialise HIP)
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def gpu():
    if not _gpu_present():
        pytest.skip("no GPU")
    lib = os.path.join(ROOT, "spice_amd", "libspice_hip.so")
    assert os.path.exists(lib), "libspice_hip.so missing on a machine with a GPU: build it first"
    return True
