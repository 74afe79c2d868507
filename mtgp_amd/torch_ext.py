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
# type.
_UINT32 = getattr(torch, "uint32", torch.int32)
_UINT64 = getattr(torch, "uint64", torch.int64)


def _dtypes(bits, kind):
    if kind in (KIND_INT, KIND_CLOSE1_OPEN2):
        return (_UINT32, torch.int32) if bits == 32 else (_UINT64, torch.int64)
    return (torch.float32,) if bits == 32 else (torch.float64,)


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
