"""GPU: torch still sees the GPU when a shtns_amd configuration (which initializes this
library's HIP runtime) is created before torch is used -- torch bundles its own HIP runtime,
and the package loads it first.  Runs in a fresh interpreter so the order is the one tested."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_config_before_torch(device):
    code = ("import shtns_amd; s = shtns_amd.sht(15); "
            "s.set_grid(32, 64, flags=shtns_amd.SHT_THETA_CONTIGUOUS); "
            "import torch; print(torch.cuda.is_available())")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.strip().endswith("True"), r.stdout
