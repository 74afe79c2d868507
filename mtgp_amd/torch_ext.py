"""PyTorch-ROCm front end: MTGP numbers written in place into torch tensors,
on torch's current HIP stream.

The library is the same libmtgp_amd.so.  Import torch before mtgp_amd, so
that the library's dependency libamdhip64.so.7 resolves to the HIP runtime
torch has already loaded (torch's copy carries that SONAME) instead of a
second runtime; importing this module checks that exactly one HIP runtime
is mapped and fails otherwise.

    import torch
    import mtgp_amd.torch_ext as mt
    gs = mt.TensorGenSet([(32, 11213, 0), (64, 23209, 1)])
    gs.init_seed([1, 2])
    x = torch.empty(1 << 20, device="cuda")              # float32
    y = torch.empty(1 << 20, dtype=torch.float64, device="cuda")
    gs.generate_([(0, mt.KIND_CLOSE_OPEN, x), (1, mt.KIND_CLOSE_OPEN, y)])
"""
import ctypes as C
import os

import torch

import mtgp_amd as M  # noqa: E402
from mtgp_amd import (KIND_CLOSE1_OPEN2, KIND_CLOSE_OPEN, KIND_INT, KIND_OPEN_CLOSE,  # noqa: E402,F401
                      KIND_OPEN_OPEN)


def hip_runtimes():
    """Distinct HIP runtime files mapped into this process."""
    found = set()
    with open("/proc/self/maps") as f:
        for line in f:
            path = line.split()[-1] if len(line.split()) >= 6 else ""
            if os.path.basename(path).startswith("libamdhip64.so"):
                found.add(os.path.realpath(path))
    return found


_rt = hip_runtimes()
if len(_rt) > 1:
    raise ImportError("two HIP runtimes are loaded (%s); import torch before mtgp_amd" % ", ".join(sorted(_rt)))

# Output element type per (word size, kind): the bit patterns of uint32 /
# uint64 are exposed as int32 / int64 tensors where torch lacks the unsigned
# type.  The floating-point kinds ([1,2) included) fill float32 / float64
# tensors; [1,2) also fills an integer tensor of the same width with its
# raw bits.
_UINT32 = getattr(torch, "uint32", torch.int32)
_UINT64 = getattr(torch, "uint64", torch.int64)


def _dtypes(bits, kind):
    ints = (_UINT32, torch.int32) if bits == 32 else (_UINT64, torch.int64)
    floats = (torch.float32,) if bits == 32 else (torch.float64,)
    if kind == KIND_INT:
        return ints
    if kind == KIND_CLOSE1_OPEN2:
        return floats + ints
    return floats


class TensorGenSet(M.GenSet):
    """A device generator set whose requests write into torch tensors."""

    def generate_(self, requests, stream=None):
        """requests: (gen, kind, tensor) or (gen, kind, tensor, skip); each
        tensor (contiguous, on the GPU, of the kind's element type) is
        filled completely, in order, on `stream` (default: torch's current
        stream), which is synchronised before returning."""
        reqs = []
        for r in requests:
            gen, kind, t = r[0], r[1], r[2]
            skip = r[3] if len(r) > 3 else 0
            bits = self.specs[gen][0]
            if not t.is_cuda or not t.is_contiguous():
                raise ValueError("tensor must be contiguous and on the GPU")
            if t.dtype not in _dtypes(bits, kind):
                raise TypeError("tensor dtype %s does not hold kind %d of an MTGP%d generator" % (t.dtype, kind, bits))
            reqs.append((gen, kind, t.numel(), skip, t.data_ptr()))
        s = stream if stream is not None else torch.cuda.current_stream()
        arr = M._requests(reqs)
        M._check(M._lib.mtgp_genset_generate(self.h, arr, len(reqs), C.c_void_p(s.cuda_stream)), "generate")


class Generator:
    """One MTGP generator whose stream fills torch tensors on the GPU, in
    the manner of a torch.Generator.  Every fill is one plan: the stream is
    cut into cells started by jump-ahead, so a single generator uses the
    whole device, and consecutive fills continue the stream exactly as the
    reference's genrand calls would.

        g = mt.Generator(32, 11213, 0, seed=1234)
        x = g.rand(1 << 20)                                  # float32, [0, 1)
        g.random_(torch.empty(1 << 20, dtype=torch.int32, device="cuda"))
        g.skip(1 << 40)                                      # by jump-ahead

    MTGP32 generators fill 32-bit tensors (uint32 / int32 bits, float32),
    MTGP64 ones 64-bit tensors (uint64 / int64 bits, float64)."""

    def __init__(self, bits=32, mexp=11213, index=0, seed=0, array=None):
        self.bits = bits
        self._gs = TensorGenSet([(bits, mexp, index)])
        if array is not None:
            self._gs.init_array([[int(k) for k in array]])
        else:
            self._gs.init_seed([int(seed)])

    def random_(self, tensor, stream=None):
        """Raw output words (genrand_uint32 / genrand_uint64) into `tensor`."""
        self._gs.generate_([(0, KIND_INT, tensor)], stream)
        return tensor

    def uniform_(self, tensor, kind=KIND_CLOSE_OPEN, stream=None):
        """Floating-point outputs of `kind` (default [0, 1)) into `tensor`."""
        self._gs.generate_([(0, kind, tensor)], stream)
        return tensor

    def rand(self, *size, kind=KIND_CLOSE_OPEN, device="cuda", stream=None):
        """A new tensor of `size` (float32 for MTGP32, float64 for MTGP64)."""
        dtype = torch.float32 if self.bits == 32 else torch.float64
        return self.uniform_(torch.empty(*size, dtype=dtype, device=device), kind, stream)

    def skip(self, n):
        """Advances the stream by n numbers (stepped over or jumped)."""
        self._gs.generate([(0, KIND_INT, 0, int(n), 0)])

    def get_state(self):
        """The state window (numpy, N words, oldest first)."""
        return self._gs.state(0)

    def set_state(self, window):
        self._gs.set_state(0, window)
