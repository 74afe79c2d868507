"""ctypes front end of the test oracle.  Test infrastructure only.

Loads ``oracle/libmtgp_oracle.so`` (this tree's C restatement, always
present once built) and, when they were built, the reference library builds
under ``oracle/_ref/`` (compiled from the reference's own sources).  Only
tests, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use this
module; the product package never imports it.

Output kinds (for both word sizes): 0 integer, 1 [1,2), 2 [0,1), 3 (0,1],
4 (0,1), matching genrand_uint*/close1_open2/close_open/open_close/open_open.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

MEXPS = {32: (11213, 23209, 44497), 64: (23209, 44497, 110503)}
KINDS = ("uint", "close1_open2", "close_open", "open_close", "open_open")


class Params(C.Structure):
    _fields_ = [("bits", C.c_int), ("mexp", C.c_int), ("index", C.c_int),
                ("n", C.c_int), ("pos", C.c_int), ("sh1", C.c_int),
                ("sh2", C.c_int), ("mask", C.c_uint64),
                ("tbl", C.c_uint64 * 16), ("tmp", C.c_uint64 * 16),
                ("flt", C.c_uint64 * 16), ("sha1", C.c_char * 41)]


def _dtype(bits, kind):
    if kind in (0, 1):
        return np.uint32 if bits == 32 else np.uint64
    return np.float32 if bits == 32 else np.float64


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """This tree's C restatement (oracle/mtgp_oracle.c)."""

    def __init__(self, path=None):
        path = path or os.path.join(HERE, "libmtgp_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle oracle`")
        L = self.lib = C.CDLL(path, mode=os.RTLD_LOCAL)
        L.orc_params_get.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(Params)]
        for b in (32, 64):
            g = getattr(L, "orc%d_generate" % b)
            g.argtypes = [C.POINTER(Params), C.c_void_p, C.c_int, C.c_uint64, C.c_void_p]
            getattr(L, "orc%d_init_state" % b).argtypes = [C.POINTER(Params), C.c_uint64 if b == 64 else C.c_uint32, C.c_void_p]
            getattr(L, "orc%d_init_by_array" % b).argtypes = [C.POINTER(Params), C.c_void_p, C.c_int, C.c_void_p]
            getattr(L, "orc%d_init_by_str" % b).argtypes = [C.POINTER(Params), C.c_char_p, C.c_void_p]
            getattr(L, "orc%d_jump" % b).argtypes = [C.POINTER(Params), C.c_void_p, C.c_char_p]
        L.orc_poly_deg.argtypes = [C.c_void_p, C.c_long]
        L.orc_poly_deg.restype = C.c_long
        L.orc_poly_mulmod.argtypes = [C.c_void_p] * 3 + [C.c_long, C.c_void_p]
        L.orc_poly_xpow_mod.argtypes = [C.c_void_p, C.c_long, C.c_void_p, C.c_long, C.c_void_p]
        L.orc_bm.argtypes = [C.c_void_p, C.c_long, C.c_void_p, C.c_long]
        L.orc_bm.restype = C.c_long
        L.orc_poly_to_hex.argtypes = [C.c_void_p, C.c_long, C.c_char_p]
        L.orc_hex_to_poly.argtypes = [C.c_char_p, C.c_void_p, C.c_long]

    def params(self, bits, mexp, index):
        p = Params()
        if self.lib.orc_params_get(bits, mexp, index, C.byref(p)):
            raise ValueError("no parameter set %d/%d/%d" % (bits, mexp, index))
        return p

    @staticmethod
    def params_from_record(bits, rec, index=-1):
        """Params from a 73-value record laid out as mtgp*-fast.h:47-58
        (mexp, pos, sh1, sh2, tbl[16], tmp_tbl[16], flt/dbl table[16], mask,
        poly_sha1[20]) -- e.g. one the reference library wrote."""
        rec = [int(v) for v in rec]
        p = Params()
        p.bits, p.mexp, p.index = bits, rec[0], index
        p.n = rec[0] // bits + 1
        p.pos, p.sh1, p.sh2 = rec[1], rec[2], rec[3]
        for j in range(16):
            p.tbl[j], p.tmp[j], p.flt[j] = rec[4 + j], rec[20 + j], rec[36 + j]
        p.mask = rec[52]
        p.sha1 = bytes(rec[53:73]).hex().encode()
        return p

    def _wdtype(self, bits):
        return np.uint32 if bits == 32 else np.uint64

    def init_state(self, p, seed):
        w = np.zeros(p.n, self._wdtype(p.bits))
        getattr(self.lib, "orc%d_init_state" % p.bits)(C.byref(p), seed, _ptr(w))
        return w

    def init_by_array(self, p, key):
        key = np.ascontiguousarray(key, self._wdtype(p.bits))
        w = np.zeros(p.n, self._wdtype(p.bits))
        getattr(self.lib, "orc%d_init_by_array" % p.bits)(C.byref(p), _ptr(key), len(key), _ptr(w))
        return w

    def init_by_str(self, p, s):
        w = np.zeros(p.n, self._wdtype(p.bits))
        getattr(self.lib, "orc%d_init_by_str" % p.bits)(C.byref(p), s, _ptr(w))
        return w

    def generate(self, p, window, kind, count):
        """Returns (outputs, new window); `window` is not modified."""
        w = np.array(window, self._wdtype(p.bits), copy=True)
        out = np.empty(count, _dtype(p.bits, kind))
        rc = getattr(self.lib, "orc%d_generate" % p.bits)(C.byref(p), _ptr(w), kind, count, _ptr(out))
        if rc:
            raise MemoryError
        return out, w

    def jump(self, p, window, hexstr):
        w = np.array(window, self._wdtype(p.bits), copy=True)
        rc = getattr(self.lib, "orc%d_jump" % p.bits)(C.byref(p), _ptr(w), hexstr.encode())
        if rc:
            raise ValueError("bad jump string")
        return w

    # GF(2)[x] helpers; polynomials are numpy uint64 little-endian bit vectors.
    def poly_words(self, degree):
        return degree // 64 + 2

    def hex_to_poly(self, hexstr, words):
        a = np.zeros(words, np.uint64)
        if self.lib.orc_hex_to_poly(hexstr.encode(), _ptr(a), words):
            raise ValueError("bad hex")
        return a

    def poly_to_hex(self, a):
        buf = C.create_string_buffer(len(a) * 16 + 2)
        self.lib.orc_poly_to_hex(_ptr(a), len(a), buf)
        return buf.value.decode()

    def poly_deg(self, a):
        return self.lib.orc_poly_deg(_ptr(a), len(a))

    def mulmod(self, a, b, m):
        out = np.zeros(len(m), np.uint64)
        self.lib.orc_poly_mulmod(_ptr(a), _ptr(b), _ptr(m), len(m), _ptr(out))
        return out

    def xpow_mod(self, e, m):
        """x^e mod m for a Python int e >= 0."""
        ew = max(1, (e.bit_length() + 63) // 64)
        ea = np.array([(e >> (64 * i)) & (2**64 - 1) for i in range(ew)], np.uint64)
        out = np.zeros(len(m), np.uint64)
        self.lib.orc_poly_xpow_mod(_ptr(ea), ew, _ptr(m), len(m), _ptr(out))
        return out

    def bm(self, bits_seq, words):
        s = np.ascontiguousarray(bits_seq, np.uint8)
        out = np.zeros(words, np.uint64)
        L = self.lib.orc_bm(_ptr(s), len(s), _ptr(out), words)
        return out, L


class RefLib:
    """The reference library compiled from its own sources (oracle/_ref)."""

    def __init__(self, variant, bits):
        path = os.path.join(REF_DIR, "libref_%s%d.so" % (variant, bits))
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.bits, self.variant = bits, variant
        L = self.lib = C.CDLL(path, mode=os.RTLD_LOCAL)
        L.rs_new_seed.argtypes = [C.c_int, C.c_int, C.c_uint64]
        L.rs_new_seed.restype = C.c_void_p
        L.rs_new_array.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int]
        L.rs_new_array.restype = C.c_void_p
        L.rs_new_str.argtypes = [C.c_int, C.c_int, C.c_char_p]
        L.rs_new_str.restype = C.c_void_p
        L.rs_free.argtypes = [C.c_void_p]
        L.rs_size.argtypes = [C.c_void_p]
        L.rs_fill.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_void_p]
        L.rs_window.argtypes = [C.c_void_p, C.c_void_p]
        L.rs_idstring.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        if variant in ("fast", "dc", "dccl"):
            L.rs_set_window.argtypes = [C.c_void_p, C.c_void_p]
            L.rs_jump.argtypes = [C.c_void_p, C.c_char_p]
            L.rs_init_state.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_void_p]
            L.rs_params.argtypes = [C.c_int, C.c_int, C.c_void_p]

    @staticmethod
    def available(variant="fast", bits=32):
        return os.path.exists(os.path.join(REF_DIR, "libref_%s%d.so" % (variant, bits)))

    def _w(self):
        return np.uint32 if self.bits == 32 else np.uint64

    def new_seed(self, mexp, index, seed):
        return RefGen(self, self.lib.rs_new_seed(mexp, index, seed))

    def new_array(self, mexp, index, key):
        key = np.ascontiguousarray(key, self._w())
        return RefGen(self, self.lib.rs_new_array(mexp, index, _ptr(key), len(key)), keep=key)

    def new_str(self, mexp, index, s):
        return RefGen(self, self.lib.rs_new_str(mexp, index, s))

    def init_state(self, mexp, index, seed, n):
        out = np.zeros(n, self._w())
        self.lib.rs_init_state(mexp, index, seed, _ptr(out))
        return out

    def params(self, mexp, index):
        out = np.zeros(73, np.uint64)
        if self.lib.rs_params(mexp, index, _ptr(out)):
            raise ValueError
        return out


class RefGen:
    def __init__(self, ref, handle, keep=None):
        if not handle:
            raise MemoryError
        self.ref, self.h, self._keep = ref, handle, keep

    def __del__(self):
        if getattr(self, "h", None):
            self.ref.lib.rs_free(self.h)
            self.h = None

    @property
    def n(self):
        return self.ref.lib.rs_size(self.h)

    def fill(self, kind, count):
        out = np.empty(count, _dtype(self.ref.bits, kind))
        self.ref.lib.rs_fill(self.h, kind, count, _ptr(out))
        return out

    def window(self):
        out = np.zeros(self.n, self.ref._w())
        self.ref.lib.rs_window(self.h, _ptr(out))
        return out

    def set_window(self, w):
        w = np.ascontiguousarray(w, self.ref._w())
        self.ref.lib.rs_set_window(self.h, _ptr(w))

    def jump(self, hexstr):
        self.ref.lib.rs_jump(self.h, hexstr.encode())

    def idstring(self):
        buf = C.create_string_buffer(256)
        self.ref.lib.rs_idstring(self.h, buf, 256)
        return buf.value.decode()
