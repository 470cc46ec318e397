import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "dpe-mvs-fast_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

# Load torch's HIP runtime before libdpe_hip.so is loaded by any test, so a
# process that uses both (the autograd wrapper, device-memory tests) shares
# one runtime: libdpe_hip.so then binds to the libamdhip64 already loaded.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C-ABI)")
