"""The torch-tensor wrappers (anime4kcpp_amd/hip.py) create their device objects only after
torch's HIP context exists: a net created before torch has touched the device leaves torch
usable (page-locked host memory, device tensors).  Run in a fresh process so that the
library really is the first to touch the device."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CODE = """
import sys
sys.path.insert(0, {root!r})
from anime4kcpp_amd import hip
net = hip.Net("hdn0", 0)
import torch
x = torch.arange(16, dtype=torch.uint8).pin_memory()
y = x.cuda()
torch.cuda.synchronize()
print("ok", int(y.sum()))
"""


def test_net_before_torch(cuda):
    r = subprocess.run([sys.executable, "-u", "-c", CODE.format(root=ROOT)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "ok 120" in r.stdout, r.stdout + r.stderr
