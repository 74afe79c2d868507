"""The PyTorch-ROCm front end (mtgp_amd/torch_ext.py), in its own process:
torch imported first, one HIP runtime, numbers written in place into
tensors on a torch stream, and the torch.Generator-style object (fills,
skip by jump-ahead, state), against the oracle."""
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
# the torch.Generator-style object: one generator's stream across fills
g = mt.Generator(32, 11213, 3, seed=99)
p = o.params(32, 11213, 3)
w = o.init_state(p, 99)
x = g.rand(50000)
e, w = o.generate(p, w, 2, 50000)
assert x.dtype == torch.float32 and (x.cpu().numpy().view(np.uint32) == e.view(np.uint32)).all()
r = g.random_(torch.empty(1000, dtype=torch.int32, device="cuda"))
e, w = o.generate(p, w, 0, 1000)
assert (r.cpu().numpy().view(np.uint32) == e).all()
g.skip(1 << 40)
import mtgp_amd as M
P = o.hex_to_poly(M.charpoly_hex(32, 11213, 3), 11213 // 64 + 2)
w = o.jump(p, w, o.poly_to_hex(o.xpow_mod(1 << 40, P)))
y = g.rand(2, 500, kind=mt.KIND_OPEN_OPEN)
e, w = o.generate(p, w, 4, 1000)
assert tuple(y.shape) == (2, 500) and (y.cpu().numpy().reshape(-1).view(np.uint32) == e.view(np.uint32)).all()
assert (g.get_state() == w).all()
# [1,2) into float tensors (mtgp*_genrand_close1_open2 returns float /
# double), and as raw bits into an integer tensor of the same width
q = g.uniform_(torch.empty(777, device="cuda"), kind=mt.KIND_CLOSE1_OPEN2)
e, w = o.generate(p, w, 1, 777)
assert q.dtype == torch.float32 and (q.cpu().numpy().view(np.uint32) == e.view(np.uint32)).all()
assert float(q.min()) >= 1.0 and float(q.max()) < 2.0
qi = g.uniform_(torch.empty(333, dtype=torch.int32, device="cuda"), kind=mt.KIND_CLOSE1_OPEN2)
e, w = o.generate(p, w, 1, 333)
assert (qi.cpu().numpy().view(np.uint32) == e.view(np.uint32)).all()
assert (g.get_state() == w).all()
g64 = mt.Generator(64, 23209, 2, array=[1, 2, 3, 4])
p64 = o.params(64, 23209, 2)
w64 = o.init_by_array(p64, [1, 2, 3, 4])
z = g64.rand(3000)
e, w64 = o.generate(p64, w64, 2, 3000)
assert z.dtype == torch.float64 and (z.cpu().numpy().view(np.uint64) == e.view(np.uint64)).all()
z1 = g64.uniform_(torch.empty(999, dtype=torch.float64, device="cuda"), kind=mt.KIND_CLOSE1_OPEN2)
e, w64 = o.generate(p64, w64, 1, 999)
assert (z1.cpu().numpy().view(np.uint64) == e.view(np.uint64)).all()
print("TORCH-OK %.4f" % float(f.mean()))
"""


def test_torch_tensors_in_place():
    env = dict(os.environ)
    env["PYTHONPATH"] = os.pathsep.join([ROOT, os.path.join(ROOT, "oracle"), env.get("PYTHONPATH", "")])
    r = subprocess.run([sys.executable, "-c", SCRIPT], capture_output=True, text=True, timeout=280, env=env, cwd=ROOT)
    assert r.returncode == 0 and "TORCH-OK" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
