"""Test configuration.

Markers:
  gpu  -- needs an MI355X (runs the HIP library); everything else runs on CPU.

The CPU oracle (oracle/liboracle.so) is the checker; the HIP library is the
thing tested.  Both are built before the tests if missing (the GPU box
receives them prebuilt).
"""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (gfx950)")
    oracle_so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(oracle_so):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=False)
    lib_so = os.path.join(ROOT, "spaghetty_amd", "libspaghetty_amd.so")
    if not os.path.exists(lib_so):
        subprocess.run([sys.executable, "-m", "spaghetty_amd.build"], cwd=ROOT, check=False)


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
