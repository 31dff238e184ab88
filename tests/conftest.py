import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and libseiswave.so")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:  # pragma: no cover
        return False


# The oracle's OpenMP threads must not spin-wait when a host runs more
# threads than it has cores (read once, when the oracle library loads).
os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")
