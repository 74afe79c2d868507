"""The PyTorch-ROCm front end (mtgp_amd/torch_ext.py), in its own process:
torch imported first, one HIP runtime, numbers written in place into
tensors on a torch stream, against the oracle."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import sys
import numpy as np
import torch
import mtgp_amd.torch_ext as mt
from pyoracle import Oracle
assert len(mt.hip_runtimes()) == 1, mt.hip_runtimes()
o = Oracle()
gs = mt.TensorGenSet([(32, 11213, 0), (64, 23209, 1)])
gs.init_seed([7, 8])
n = 100000
u = torch.empty(n, dtype=torch.int32, device="cuda")
f = torch.empty(n, device="cuda")
d = torch.empty(n // 2, dtype=torch.float64, device="cuda")
s = torch.cuda.Stream()
with torch.cuda.stream(s):
    gs.generate_([(0, mt.KIND_INT, u), (0, mt.KIND_CLOSE_OPEN, f), (1, mt.KIND_CLOSE_OPEN, d, 5)])
p0 = o.params(32, 11213, 0)
w = o.init_state(p0, 7)
e, w = o.generate(p0, w, 0, n)
assert (u.cpu().numpy().view(np.uint32) == e).all()
e, w = o.generate(p0, w, 2, n)
assert (f.cpu().numpy().view(np.uint32) == e.view(np.uint32)).all()
p1 = o.params(64, 23209, 1)
w1 = o.init_state(p1, 8)
_, w1 = o.generate(p1, w1, 0, 5)
e, w1 = o.generate(p1, w1, 2, n // 2)
assert (d.cpu().numpy().view(np.uint64) == e.view(np.uint64)).all()
assert (gs.state(0) == w).all() and (gs.state(1) == w1).all()
try:
    gs.generate_([(0, mt.KIND_CLOSE_OPEN, d)])
    raise SystemExit("dtype mismatch not refused")
except TypeError:
    pass
print("TORCH-OK %.4f" % float(f.mean()))
"""


def test_torch_tensors_in_place():
    env = dict(os.environ)
    env["PYTHONPATH"] = os.pathsep.join([ROOT, os.path.join(ROOT, "oracle"), env.get("PYTHONPATH", "")])
    r = subprocess.run([sys.executable, "-c", SCRIPT], capture_output=True, text=True, timeout=280, env=env, cwd=ROOT)
    assert r.returncode == 0 and "TORCH-OK" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
