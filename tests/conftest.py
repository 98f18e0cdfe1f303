import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "solfec-1.0_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def gpu_available():
    try:
        import torch  # noqa: F401  (device count only; no GPU initialisation)
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")
