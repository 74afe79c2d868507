"""MI355X-native MTGP: Python front end over libmtgp_amd.so (ctypes).

The library exports the reference C surface (mtgp32-fast.h, mtgp64-fast.h,
mtgp32/64-fast-jump.h) plus the GPU entry points of include/mtgp_hip.h.
This module mirrors the reference's operator interface in Python:

* ``Generator32`` / ``Generator64`` wrap ``mtgp32_fast_t`` / ``mtgp64_fast_t``
  (init by seed, by array, by string; the genrand family; ``fast_jump``),
  with ``fill()`` producing any number of outputs on the GPU;
* ``charpoly_hex`` / ``jump_hex`` are the NTL-free replacements of
  calc_characteristic / calc_jump;
* ``GenSet`` holds many device-resident generators and runs batched
  requests (``Plan``) as persistent kernel launches.

Device buffers are ``DeviceBuffer`` objects (the library's own hipMalloc),
so no other GPU framework needs to be loaded into the process.  If the
library has not been built, importing this package raises ImportError.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmtgp_amd.so")

KIND_INT, KIND_CLOSE1_OPEN2, KIND_CLOSE_OPEN, KIND_OPEN_CLOSE, KIND_OPEN_OPEN = range(5)
KINDS = {"uint": 0, "close1_open2": 1, "close_open": 2, "open_close": 3, "open_open": 4}
MEXPS = {32: (11213, 23209, 44497), 64: (23209, 44497, 110503)}

if not os.path.exists(LIB_PATH):
    raise ImportError("libmtgp_amd.so is not built: run `python -m mtgp_amd.build`")

_lib = C.CDLL(LIB_PATH, mode=os.RTLD_LOCAL)


class MTGPError(RuntimeError):
    pass


def last_error():
    return _lib.mtgp_last_error().decode()


def _check(rc, what):
    if rc is None or (isinstance(rc, int) and rc < 0):
        raise MTGPError("%s: %s" % (what, last_error()))
    return rc


# --------------------------------------------------------------- structs
class Params32(C.Structure):
    """mtgp32_params_fast_t (mtgp32-fast.h:47-58)."""
    _fields_ = [("mexp", C.c_int), ("pos", C.c_int), ("sh1", C.c_int), ("sh2", C.c_int),
                ("tbl", C.c_uint32 * 16), ("tmp_tbl", C.c_uint32 * 16),
                ("flt_tmp_tbl", C.c_uint32 * 16), ("mask", C.c_uint32),
                ("poly_sha1", C.c_ubyte * 21)]


class Params64(C.Structure):
    """mtgp64_params_fast_t (mtgp64-fast.h:46-57)."""
    _fields_ = [("mexp", C.c_int), ("pos", C.c_int), ("sh1", C.c_int), ("sh2", C.c_int),
                ("tbl", C.c_uint64 * 16), ("tmp_tbl", C.c_uint64 * 16),
                ("dbl_tmp_tbl", C.c_uint64 * 16), ("mask", C.c_uint64),
                ("poly_sha1", C.c_ubyte * 21)]


class Status(C.Structure):
    """Header of mtgp32_status_fast_t / mtgp64_status_fast_t (array follows)."""
    _fields_ = [("idx", C.c_int), ("size", C.c_int), ("large_size", C.c_int),
                ("large_mask", C.c_int)]


class Fast32(C.Structure):
    _fields_ = [("params", Params32), ("status", C.POINTER(Status))]


class Fast64(C.Structure):
    _fields_ = [("params", Params64), ("status", C.POINTER(Status))]


class Request(C.Structure):
    """mtgp_request_t."""
    _fields_ = [("gen", C.c_int32), ("kind", C.c_int32), ("count", C.c_uint64),
                ("skip", C.c_uint64), ("out", C.c_void_p)]


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


vp, i32, u64, sz, cp = C.c_void_p, C.c_int, C.c_uint64, C.c_size_t, C.c_char_p
_sig("mtgp_last_error", C.c_char_p)
_sig("mtgp_set_device", i32, i32)
_sig("mtgp_device_info", i32, C.POINTER(C.c_int), C.POINTER(C.c_size_t))
_sig("mtgp_dev_malloc", vp, sz)
_sig("mtgp_dev_free", i32, vp)
_sig("mtgp_memcpy_htod", i32, vp, vp, sz)
_sig("mtgp_memcpy_dtoh", i32, vp, vp, sz)
_sig("mtgp_device_copy", i32, vp, vp, sz, vp)
_sig("mtgp_device_fill", i32, vp, C.c_uint32, sz, vp)
_sig("mtgp_synchronize", i32)
_sig("mtgp_stream_create", vp)
_sig("mtgp_stream_destroy", i32, vp)
_sig("mtgp_stream_synchronize", i32, vp)
_sig("mtgp_event_create", vp)
_sig("mtgp_event_destroy", i32, vp)
_sig("mtgp_event_record", i32, vp, vp)
_sig("mtgp_event_elapsed_ms", i32, vp, vp, C.POINTER(C.c_float))
for b in (32, 64):
    P, F = (Params32, Fast32) if b == 32 else (Params64, Fast64)
    W = C.c_uint32 if b == 32 else C.c_uint64
    _sig("mtgp%d_init" % b, i32, C.POINTER(F), C.POINTER(P), W)
    _sig("mtgp%d_init_state" % b, None, C.POINTER(W), C.POINTER(P), W)
    _sig("mtgp%d_init_by_array" % b, i32, C.POINTER(F), C.POINTER(P), C.POINTER(W), i32)
    _sig("mtgp%d_init_by_str" % b, i32, C.POINTER(F), C.POINTER(P), cp)
    _sig("mtgp%d_free" % b, None, C.POINTER(F))
    _sig("mtgp%d_fast_jump" % b, None, C.POINTER(F), cp)
    _sig("mtgp%d_fill" % b, i32, C.POINTER(F), i32, vp, sz)
    _sig("mtgp%d_charpoly_hex" % b, i32, C.POINTER(P), cp, sz)
    _sig("mtgp%d_jump_hex" % b, i32, C.POINTER(P), C.POINTER(C.c_uint64), i32, cp, sz)
_sig("mtgp_genset_create", vp, i32, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int))
_sig("mtgp_genset_create_params", vp, i32, C.POINTER(C.c_int), C.POINTER(C.c_void_p))
_sig("mtgp_genset_destroy", None, vp)
_sig("mtgp_genset_count", i32, vp)
_sig("mtgp_genset_state_words", i32, vp, i32)
_sig("mtgp_genset_init_seed", i32, vp, C.POINTER(C.c_uint64))
_sig("mtgp_genset_init_array", i32, vp, C.POINTER(C.c_uint64), C.POINTER(C.c_int), C.POINTER(C.c_int))
_sig("mtgp_genset_init_seed_device", i32, vp, C.POINTER(C.c_uint64), vp)
_sig("mtgp_genset_init_array_device", i32, vp, C.POINTER(C.c_uint64), C.POINTER(C.c_int), C.POINTER(C.c_int), vp)
_sig("mtgp_genset_set_state", i32, vp, i32, vp)
_sig("mtgp_genset_get_state", i32, vp, i32, vp)
_sig("mtgp_genset_jump", i32, vp, i32, cp)
_sig("mtgp_genset_substreams", i32, vp, vp, cp)
_sig("mtgp_plan_create", vp, vp, C.POINTER(Request), i32, u64)
_sig("mtgp_plan_launch", i32, vp, vp)
_sig("mtgp_plan_commit", i32, vp)
_sig("mtgp_plan_rebase", i32, vp, vp)
_sig("mtgp_plan_launch_info", i32, vp, i32, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
     C.POINTER(C.c_int), C.POINTER(C.c_int), cp, sz)
_sig("mtgp_plan_destroy", None, vp)
_sig("mtgp_plan_configure", i32, vp, i32, i32)
_sig("mtgp_plan_set_overlap", i32, vp, i32)
_sig("mtgp_tuning_set", i32, C.c_char_p, C.c_double)
_sig("mtgp_tuning_get", i32, C.c_char_p, C.POINTER(C.c_double))
_sig("mtgp_plan_launch_part", i32, vp, i32, vp)
_sig("mtgp_plan_stats", i32, vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double))
_sig("mtgp_plan_jump_stats", i32, vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int))
_sig("mtgp_plan_capture", i32, vp, vp)
_sig("mtgp_plan_launch_graph", i32, vp, vp)
_sig("mtgp_plan_autotune", i32, vp, vp, i32, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_float))
_sig("mtgp_genset_generate", i32, vp, C.POINTER(Request), i32, vp)


class ServerConfig(C.Structure):
    _fields_ = [("arena_bytes", C.c_size_t), ("cell", C.c_uint64), ("cached_cells", C.c_int),
                ("streams", C.c_int), ("max_inflight", C.c_int), ("max_batch", C.c_int), ("flags", C.c_int)]


class Completion(C.Structure):
    _fields_ = [("ticket", C.c_int64), ("out", C.c_void_p), ("bytes", C.c_uint64), ("submit_us", C.c_double),
                ("launch_us", C.c_double), ("done_us", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("batches", C.c_int64),
                ("launches", C.c_int64), ("numbers", C.c_double), ("plan_us", C.c_double),
                ("max_batch_requests", C.c_int64)]


SERVER_PLAIN = 1
SERVER_FUSED_JUMPS = 2
_dp, _i32p, _u64p = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_uint64)
_sig("mtgp_server_create", vp, vp, C.POINTER(ServerConfig))
_sig("mtgp_server_destroy", i32, vp)
_sig("mtgp_server_submit", C.c_int64, vp, i32, i32, u64, u64)
_sig("mtgp_server_poll", i32, vp, C.POINTER(Completion), i32, i32)
_sig("mtgp_server_release", i32, vp, C.c_int64)
_sig("mtgp_server_drain", i32, vp)
_sig("mtgp_server_stats", i32, vp, C.POINTER(ServerStats))
_sig("mtgp_server_replay", i32, vp, i32, _dp, _i32p, _i32p, _u64p, _u64p, _dp, _dp, _u64p, _dp)
_sig("mtgp_digest64", C.c_uint64, vp, sz)

lib = _lib


def params(bits, mexp, index):
    """The reference parameter record mtgp{bits}_params_fast_{mexp}[index]."""
    if mexp not in MEXPS.get(bits, ()) or not 0 <= index < 128:
        raise ValueError("no parameter set mtgp%d-%d #%d" % (bits, mexp, index))
    P = Params32 if bits == 32 else Params64
    table = (P * 128).in_dll(_lib, "mtgp%d_params_fast_%d" % (bits, mexp))
    return table[index]


def dc_params(bits, index):
    """The dynamic-creator record mtgp{bits}dc_params_fast_11213[index]
    (200 sets; cuda-sample/mtgp{bits}dc-param-11213.c)."""
    if bits not in (32, 64) or not 0 <= index < 200:
        raise ValueError("no dynamic-creator set mtgp%d-11213 #%d" % (bits, index))
    P = Params32 if bits == 32 else Params64
    return (P * 200).in_dll(_lib, "mtgp%ddc_params_fast_11213" % bits)[index]


def _wdtype(bits):
    return np.uint32 if bits == 32 else np.uint64


def out_dtype(bits, kind):
    """Element type of `kind`: unsigned words for genrand_uint*, float32 /
    float64 for the four floating-point kinds (close1_open2 included, as
    mtgp*_genrand_close1_open2 returns a float / double)."""
    if kind == KIND_INT:
        return np.uint32 if bits == 32 else np.uint64
    return np.float32 if bits == 32 else np.float64


# ------------------------------------------------------------ device memory
def device_info():
    cus, hbm = C.c_int(), C.c_size_t()
    _check(_lib.mtgp_device_info(C.byref(cus), C.byref(hbm)), "mtgp_device_info")
    return cus.value, hbm.value


def tuning_set(name, value):
    """Sets a planner tuning knob (mtgp_tuning_set) for plans created later."""
    _check(_lib.mtgp_tuning_set(name.encode(), float(value)), "mtgp_tuning_set")


def tuning_get(name):
    v = C.c_double()
    _check(_lib.mtgp_tuning_get(name.encode(), C.byref(v)), "mtgp_tuning_get")
    return v.value


class tuning:
    """Context manager: `with tuning(two_size_rho=1e9): ...` sets knobs and
    restores their previous values on exit."""

    def __init__(self, **knobs):
        self.knobs, self.saved = knobs, {}

    def __enter__(self):
        for k, v in self.knobs.items():
            self.saved[k] = tuning_get(k)
            tuning_set(k, v)
        return self

    def __exit__(self, *exc):
        for k, v in self.saved.items():
            tuning_set(k, v)
        return False


class DeviceBuffer:
    """A device allocation made by the library (hipMalloc)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = _lib.mtgp_dev_malloc(max(1, self.nbytes))
        if not self.ptr:
            raise MTGPError("mtgp_dev_malloc(%d): %s" % (nbytes, last_error()))

    def free(self):
        if self.ptr:
            _lib.mtgp_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def address(self, offset=0):
        return self.ptr + int(offset)

    def to_host(self, dtype, count=None, offset=0):
        dtype = np.dtype(dtype)
        count = self.nbytes // dtype.itemsize if count is None else int(count)
        a = np.empty(count, dtype)
        if count:
            _check(_lib.mtgp_memcpy_dtoh(a.ctypes.data, self.ptr + offset, a.nbytes), "memcpy dtoh")
        return a

    def from_host(self, a, offset=0):
        a = np.ascontiguousarray(a)
        _check(_lib.mtgp_memcpy_htod(self.ptr + offset, a.ctypes.data, a.nbytes), "memcpy htod")


def synchronize():
    _check(_lib.mtgp_synchronize(), "mtgp_synchronize")


def memcpy_dtoh(host, device_address):
    """Copies host.nbytes bytes from a device address into a host array."""
    _check(_lib.mtgp_memcpy_dtoh(host.ctypes.data, device_address, host.nbytes), "memcpy dtoh")
    return host


def device_fill(dst, value, nbytes, stream=None):
    """Fills nbytes of a DeviceBuffer / device address with a 32-bit value."""
    d = dst.ptr if isinstance(dst, DeviceBuffer) else dst
    _check(_lib.mtgp_device_fill(d, int(value) & 0xFFFFFFFF, int(nbytes), stream.handle if stream else None),
           "device_fill")


def device_copy(dst, src, nbytes, stream=None):
    """dst <- src (DeviceBuffers or device addresses) by a plain copy kernel."""
    d = dst.ptr if isinstance(dst, DeviceBuffer) else dst
    s = src.ptr if isinstance(src, DeviceBuffer) else src
    _check(_lib.mtgp_device_copy(d, s, int(nbytes), stream.handle if stream else None), "device_copy")


class Stream:
    def __init__(self):
        self.handle = _lib.mtgp_stream_create()
        if not self.handle:
            raise MTGPError(last_error())

    def synchronize(self):
        _check(_lib.mtgp_stream_synchronize(self.handle), "stream sync")

    def __del__(self):
        if getattr(self, "handle", None):
            _lib.mtgp_stream_destroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        self.handle = _lib.mtgp_event_create()
        if not self.handle:
            raise MTGPError(last_error())

    def record(self, stream=None):
        _check(_lib.mtgp_event_record(self.handle, stream.handle if stream else None), "event record")

    def elapsed_ms(self, end):
        ms = C.c_float()
        _check(_lib.mtgp_event_elapsed_ms(self.handle, end.handle, C.byref(ms)), "elapsed")
        return ms.value

    def __del__(self):
        if getattr(self, "handle", None):
            _lib.mtgp_event_destroy(self.handle)
            self.handle = None


# ------------------------------------------------------ host-side generators
class _Generator:
    bits = 32

    def __init__(self, mexp=None, index=None, seed=None, array=None, string=None, record=None):
        """A table set (mexp, index), or any parameter record (`record`)."""
        self.p = record if record is not None else params(self.bits, mexp, index)
        F = Fast32 if self.bits == 32 else Fast64
        self.g = F()
        W = C.c_uint32 if self.bits == 32 else C.c_uint64
        if array is not None:
            key = (W * len(array))(*[int(v) for v in array])
            rc = getattr(_lib, "mtgp%d_init_by_array" % self.bits)(C.byref(self.g), C.byref(self.p), key, len(array))
        elif string is not None:
            rc = getattr(_lib, "mtgp%d_init_by_str" % self.bits)(C.byref(self.g), C.byref(self.p), string)
        else:
            rc = getattr(_lib, "mtgp%d_init" % self.bits)(C.byref(self.g), C.byref(self.p), int(seed or 0))
        if rc:
            raise MemoryError("mtgp%d_init failed" % self.bits)

    def __del__(self):
        g = getattr(self, "g", None)
        if g is not None and g.status:
            getattr(_lib, "mtgp%d_free" % self.bits)(C.byref(g))

    @property
    def n(self):
        return self.g.status.contents.size

    def window(self):
        """Canonical state window (oldest word first)."""
        st = self.g.status.contents
        W = C.c_uint32 if self.bits == 32 else C.c_uint64
        arr = (W * st.large_size).from_address(C.addressof(st) + C.sizeof(Status))
        ring = np.frombuffer(arr, dtype=_wdtype(self.bits)).copy()
        idx = (st.idx - st.size + 1 + np.arange(st.size) + st.large_size) & st.large_mask
        return ring[idx]

    def fill(self, kind, n, out=None):
        """n numbers of `kind` on the GPU.  out: None (returns a numpy array),
        a numpy array (filled in place), or a DeviceBuffer."""
        kind = KINDS.get(kind, kind)
        dt = out_dtype(self.bits, kind)
        if out is None:
            out = np.empty(n, dt)
        if isinstance(out, DeviceBuffer) or hasattr(out, "ptr"):
            ptr = out.ptr
        else:
            # a float kind may also be written as its raw bits (an unsigned
            # array of the same width)
            ok = out.dtype == dt or out.dtype == _wdtype(self.bits)
            if not ok or out.size < n or not out.flags.c_contiguous:
                raise ValueError("output array must be contiguous %s of >= %d" % (dt, n))
            ptr = out.ctypes.data
        _check(getattr(_lib, "mtgp%d_fill" % self.bits)(C.byref(self.g), kind, ptr, n), "fill")
        return out

    def fast_jump(self, hexstr):
        """mtgp*_fast_jump on the GPU (mtgp32-fast-jump.c:51-79)."""
        getattr(_lib, "mtgp%d_fast_jump" % self.bits)(C.byref(self.g), hexstr.encode())
        err = last_error()
        if err:
            raise MTGPError("fast_jump: " + err)


class Generator32(_Generator):
    bits = 32


class Generator64(_Generator):
    bits = 64


def charpoly_hex(bits, mexp=None, index=None, record=None):
    """Characteristic polynomial of a table set or of a record, as hex."""
    p = record if record is not None else params(bits, mexp, index)
    buf = C.create_string_buffer(p.mexp // 4 + 8)
    _check(getattr(_lib, "mtgp%d_charpoly_hex" % bits)(C.byref(p), buf, len(buf)), "charpoly")
    return buf.value.decode()


def jump_hex(bits, mexp=None, index=None, steps=0, record=None):
    """x^steps mod P as a jump string (steps: Python int >= 0), for a table
    set or a record."""
    p = record if record is not None else params(bits, mexp, index)
    words = max(1, (int(steps).bit_length() + 63) // 64)
    e = (C.c_uint64 * words)(*[(int(steps) >> (64 * i)) & (2**64 - 1) for i in range(words)])
    buf = C.create_string_buffer(p.mexp // 4 + 8)
    _check(getattr(_lib, "mtgp%d_jump_hex" % bits)(C.byref(p), e, words, buf, len(buf)), "jump_hex")
    return buf.value.decode()


# ------------------------------------------------------------ device sets
class GenSet:
    """Device-resident generators (mtgp_genset_t)."""

    def __init__(self, specs):
        """specs: list of (bits, mexp, index) table sets, or of parameter
        records (Params32 / Params64 instances, e.g. dc_params(32, i))."""
        n = len(specs)
        if all(isinstance(s, (Params32, Params64)) for s in specs):
            self._recs = list(specs)
            self.specs = [(32 if isinstance(s, Params32) else 64, s.mexp, -1) for s in specs]
            b = (C.c_int * n)(*[s[0] for s in self.specs])
            ptrs = (C.c_void_p * n)(*[C.cast(C.pointer(r), C.c_void_p) for r in self._recs])
            self.h = _lib.mtgp_genset_create_params(n, b, ptrs)
        else:
            self.specs = [tuple(int(v) for v in s) for s in specs]
            b = (C.c_int * n)(*[s[0] for s in self.specs])
            m = (C.c_int * n)(*[s[1] for s in self.specs])
            i = (C.c_int * n)(*[s[2] for s in self.specs])
            self.h = _lib.mtgp_genset_create(n, b, m, i)
        if not self.h:
            raise MTGPError("genset_create: " + last_error())

    def __del__(self):
        if getattr(self, "h", None):
            _lib.mtgp_genset_destroy(self.h)
            self.h = None

    def __len__(self):
        return len(self.specs)

    def init_seed(self, seeds, device=False, stream=None):
        """mtgp*_init for every generator (device=True: on the GPU)."""
        s = (C.c_uint64 * len(self))(*[int(v) & (2**64 - 1) for v in seeds])
        if device:
            _check(_lib.mtgp_genset_init_seed_device(self.h, s, stream.handle if stream else None), "init_seed")
        else:
            _check(_lib.mtgp_genset_init_seed(self.h, s), "init_seed")

    def init_array(self, keys, device=False, stream=None):
        """mtgp*_init_by_array for every generator (device=True: on the GPU)."""
        flat, offs, lens = [], [], []
        for k in keys:
            offs.append(len(flat))
            lens.append(len(k))
            flat.extend(int(v) & (2**64 - 1) for v in k)
        n = len(self)
        args = ((C.c_uint64 * max(1, len(flat)))(*flat), (C.c_int * n)(*offs), (C.c_int * n)(*lens))
        if device:
            _check(_lib.mtgp_genset_init_array_device(self.h, *args, stream.handle if stream else None),
                   "init_array")
        else:
            _check(_lib.mtgp_genset_init_array(self.h, *args), "init_array")

    def state(self, g):
        n = _lib.mtgp_genset_state_words(self.h, g)
        w = np.empty(n, _wdtype(self.specs[g][0]))
        _check(_lib.mtgp_genset_get_state(self.h, g, w.ctypes.data), "get_state")
        return w

    def set_state(self, g, window):
        w = np.ascontiguousarray(window, _wdtype(self.specs[g][0]))
        _check(_lib.mtgp_genset_set_state(self.h, g, w.ctypes.data), "set_state")

    def jump(self, g, hexstr):
        _check(_lib.mtgp_genset_jump(self.h, g, hexstr.encode()), "jump")

    def substreams(self, base_window, hexstr):
        w = np.ascontiguousarray(base_window, _wdtype(self.specs[0][0]))
        _check(_lib.mtgp_genset_substreams(self.h, w.ctypes.data, hexstr.encode()), "substreams")

    def plan(self, requests, segment=0):
        return Plan(self, requests, segment)

    def generate(self, requests, stream=None):
        reqs = _requests(requests)
        _check(_lib.mtgp_genset_generate(self.h, reqs, len(requests), stream.handle if stream else None),
               "generate")


def _requests(requests):
    arr = (Request * max(1, len(requests)))()
    for k, r in enumerate(requests):
        gen, kind, count, skip, out = r
        arr[k].gen, arr[k].kind, arr[k].count, arr[k].skip = gen, KINDS.get(kind, kind), count, skip
        arr[k].out = out.ptr if isinstance(out, DeviceBuffer) else out
    return arr


class Plan:
    """A batch of requests compiled to device work items (mtgp_plan_t).
    requests: list of (gen, kind, count, skip, out) with out a DeviceBuffer
    or a device address."""

    def __init__(self, genset, requests, segment=0):
        self.genset = genset
        self._reqs = _requests(requests)
        self.h = _lib.mtgp_plan_create(genset.h, self._reqs, len(requests), int(segment))
        if not self.h:
            raise MTGPError("plan_create: " + last_error())

    def launch(self, stream=None):
        _check(_lib.mtgp_plan_launch(self.h, stream.handle if stream else None), "plan_launch")

    def launch_part(self, part, stream=None):
        _check(_lib.mtgp_plan_launch_part(self.h, part, stream.handle if stream else None), "plan_launch_part")

    def commit(self):
        _check(_lib.mtgp_plan_commit(self.h), "plan_commit")

    def rebase(self, stream=None):
        """Points the plan at the generators' current states: the next launch
        writes the next numbers of every stream into the same outputs."""
        _check(_lib.mtgp_plan_rebase(self.h, stream.handle if stream else None), "plan_rebase")

    def configure(self, layout=-1, blocks_per_cu=-1):
        """Step layout (0 quad, 1 lane-strided, 2 two-wave cells, 3 four-wave
        cells; launches that do not admit 2 or 3 take the next narrower one)
        and resident workgroups per CU; -1 leaves a knob unchanged."""
        _check(_lib.mtgp_plan_configure(self.h, layout, blocks_per_cu), "plan_configure")

    def set_overlap(self, on=True):
        """Two-phase launch order on/off; returns True if the plan now
        launches in two phases."""
        r = _lib.mtgp_plan_set_overlap(self.h, 1 if on else 0)
        _check(0 if r >= 0 else -1, "plan_set_overlap")
        return r == 1

    def capture(self, stream):
        """Records the plan's launches in a hipGraph (stream: a Stream)."""
        _check(_lib.mtgp_plan_capture(self.h, stream.handle), "plan_capture")

    def launch_graph(self, stream=None):
        _check(_lib.mtgp_plan_launch_graph(self.h, stream.handle if stream else None), "plan_launch_graph")

    def autotune(self, stream=None, reps=2):
        """Times the generation launch per (layout, workgroups per CU) and
        keeps the fastest; returns (layout, blocks_per_cu, ms)."""
        lay, bpc, ms = C.c_int(), C.c_int(), C.c_float()
        _check(_lib.mtgp_plan_autotune(self.h, stream.handle if stream else None, reps, C.byref(lay), C.byref(bpc),
                                       C.byref(ms)), "plan_autotune")
        return lay.value, bpc.value, ms.value

    def launches(self):
        """The generation launches in serial order: dicts of bits, class,
        layout, grid, items and the kernel's template name."""
        n = _check(_lib.mtgp_plan_launch_info(self.h, -1, None, None, None, None, None, None, 0), "plan_launch_info")
        out = []
        for k in range(n):
            v = [C.c_int() for _ in range(5)]
            name = C.create_string_buffer(128)
            _check(_lib.mtgp_plan_launch_info(self.h, k, *[C.byref(x) for x in v], name, len(name)),
                   "plan_launch_info")
            out.append(dict(zip(("bits", "cls", "layout", "grid", "items"), [x.value for x in v]),
                            kernel=name.value.decode()))
        return out

    def stats(self):
        it, ln, ms = C.c_int(), C.c_int(), C.c_double()
        _lib.mtgp_plan_stats(self.h, C.byref(it), C.byref(ln), C.byref(ms))
        j, mj, mg = C.c_int(), C.c_int(), C.c_int()
        _lib.mtgp_plan_jump_stats(self.h, C.byref(j), C.byref(mj), C.byref(mg))
        return {"items": it.value, "launches": ln.value, "plan_ms": ms.value, "jumps": j.value,
                "matrix_jumps": mj.value, "matrix_groups": mg.value}

    def __del__(self):
        if getattr(self, "h", None):
            _lib.mtgp_plan_destroy(self.h)
            self.h = None


# ---------------------------------------------------------------- serving
def digest64(a):
    """The 64-bit digest the server's replay computes, of a host array."""
    a = np.ascontiguousarray(a)
    return int(_lib.mtgp_digest64(a.ctypes.data, a.nbytes))


class Server:
    """Continuous-batching request server over a GenSet (mtgp_server_t)."""

    def __init__(self, genset, arena_bytes=0, cell=0, cached_cells=0, streams=0, max_inflight=0, max_batch=0,
                 plain=False, fused_jumps=False):
        flags = (SERVER_PLAIN if plain else 0) | (SERVER_FUSED_JUMPS if fused_jumps else 0)
        cfg = ServerConfig(arena_bytes, cell, cached_cells, streams, max_inflight, max_batch, flags)
        self.genset = genset
        self.h = _lib.mtgp_server_create(genset.h, C.byref(cfg))
        if not self.h:
            raise MTGPError("server_create: " + last_error())

    def submit(self, gen, kind, count, skip=0):
        t = _lib.mtgp_server_submit(self.h, gen, KINDS.get(kind, kind), count, skip)
        if t < 0:
            raise MTGPError("server_submit: " + last_error())
        return t

    def poll(self, max_n=256, wait_ms=0):
        arr = (Completion * max_n)()
        k = _lib.mtgp_server_poll(self.h, arr, max_n, wait_ms)
        if k < 0:
            raise MTGPError("server_poll: " + last_error())
        return [dict(ticket=c.ticket, out=c.out, bytes=c.bytes, submit_us=c.submit_us, launch_us=c.launch_us,
                     done_us=c.done_us) for c in arr[:k]]

    def release(self, ticket):
        _check(_lib.mtgp_server_release(self.h, ticket), "server_release")

    def drain(self):
        _check(_lib.mtgp_server_drain(self.h), "server_drain")

    def stats(self):
        st = ServerStats()
        _check(_lib.mtgp_server_stats(self.h, C.byref(st)), "server_stats")
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def replay(self, t_us, gen, kind, count, skip, digests=False, queue=None):
        """Submits request i at t_us[i] microseconds; returns (latency_us,
        digests or None, wall_us).  `queue`: optional float64 array that
        receives each request's launch - submission delay."""
        n = len(t_us)
        t = np.ascontiguousarray(t_us, np.float64)
        g = np.ascontiguousarray(gen, np.int32)
        k = np.ascontiguousarray(kind, np.int32)
        c = np.ascontiguousarray(count, np.uint64)
        sk = np.ascontiguousarray(skip, np.uint64)
        lat = np.zeros(n, np.float64)
        dg = np.zeros(n, np.uint64) if digests else None
        wall = C.c_double(0)
        _check(_lib.mtgp_server_replay(self.h, n, t.ctypes.data_as(_dp), g.ctypes.data_as(_i32p),
                                       k.ctypes.data_as(_i32p), c.ctypes.data_as(_u64p), sk.ctypes.data_as(_u64p),
                                       lat.ctypes.data_as(_dp),
                                       queue.ctypes.data_as(_dp) if queue is not None else None,
                                       dg.ctypes.data_as(_u64p) if digests else None,
                                       C.byref(wall)), "server_replay")
        return lat, dg, wall.value

    def close(self):
        if getattr(self, "h", None):
            h, self.h = self.h, None
            _check(_lib.mtgp_server_destroy(h), "server_destroy")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

