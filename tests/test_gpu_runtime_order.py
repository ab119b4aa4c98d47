"""The device library and PyTorch share one HIP runtime whichever the caller touches first.

A fresh process opens a vr.Device before it uses torch's device memory, then allocates a torch
tensor on the GPU and hands it to the library as a device column."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import sys
sys.path.insert(0, %r)
from vecruntime_amd import abi, vr
dev = vr.Device(0)
import torch
keys = torch.arange(1000, dtype=torch.int64, device="cuda") %% 7
spec = abi.AggSpec([abi.INT64], [], [abi.agg_desc(abi.COUNT_STAR)], key_nullable=[False])
g = vr.GroupedAggregate(dev, spec)
g.update([vr.device_column(abi.INT64, keys)], [], 1000)
assert g.num_groups() == 7, g.num_groups()
g.close()
dev.close()
print("runtime order ok", flush=True)
""" % ROOT


@pytest.mark.gpu
def test_device_before_torch_shares_the_runtime():
    r = subprocess.run([sys.executable, "-c", SCRIPT], capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime order ok" in r.stdout
