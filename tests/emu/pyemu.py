"""ctypes front end of the kernel emulation library (test infrastructure).

Items point at host numpy arrays, which the emulated kernels read and write
like device memory."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ITEM = np.dtype([("out", "<u8"), ("count", "<u8"), ("state_in", "<u8"), ("state_out", "<u8"),
                 ("poly", "<u8"), ("poly_words", "<i4"), ("param", "<i4"), ("kind", "<i4"),
                 ("parts", "<i4"), ("part_bytes", "<u8"), ("segs", "<u8"), ("nsegs", "<i4"), ("pad0", "<i4")])
assert ITEM.itemsize == 80
INIT_ITEM = np.dtype([("state", "<u8"), ("seed", "<u8"), ("hidden", "<u8"), ("n", "<i4"), ("bits", "<i4"),
                      ("key_len", "<i4"), ("mode", "<i4"), ("pad", "<u8")])
assert INIT_ITEM.itemsize == 48
SEGMENT = np.dtype([("out", "<u8"), ("count", "<u8"), ("kind", "<i4"), ("pad0", "<i4"), ("pad1", "<u8")])
assert SEGMENT.itemsize == 32


def load():
    subprocess.check_call(["make", "-s", "-C", HERE])
    lib = C.CDLL(os.path.join(HERE, "libmtgp_emu.so"), mode=os.RTLD_LOCAL)
    lib.emu_run_items.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int]
    lib.emu_run_jumps.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int]
    lib.emu_run_init.argtypes = [C.c_void_p, C.c_int]
    lib.emu_run_items_p.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    lib.emu_run_jumps_p.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    lib.emu_run_mjumps.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int]
    return lib


DEVPARAM = np.dtype([("bits", "<i4"), ("mexp", "<i4"), ("n", "<i4"), ("pos", "<i4"), ("sh1", "<i4"), ("sh2", "<i4"),
                     ("mask_lo", "<u4"), ("mask_hi", "<u4"), ("tbl", "<u4", (16,)), ("tmp", "<u4", (16,)),
                     ("flt", "<u4", (16,))])
assert DEVPARAM.itemsize == 224


def devparam(bits, rec):
    """The kernels' parameter record from a 73-value reference record
    (64-bit tables keep the high halves, whose low halves are zero)."""
    rec = [int(v) for v in rec]
    d = np.zeros(1, DEVPARAM)
    d["bits"], d["mexp"], d["n"] = bits, rec[0], rec[0] // bits + 1
    d["pos"], d["sh1"], d["sh2"] = rec[1], rec[2], rec[3]
    shift = 32 if bits == 64 else 0
    d["tbl"][0] = [v >> shift for v in rec[4:20]]
    d["tmp"][0] = [v >> shift for v in rec[20:36]]
    d["flt"][0] = [v >> shift for v in rec[36:52]]
    d["mask_lo"], d["mask_hi"] = rec[52] & 0xFFFFFFFF, rec[52] >> 32
    return d


JUMP_ITEM = np.dtype([("state_in", "<u8"), ("masks", "<u8"), ("states_out", "<u8"), ("nb", "<i4"),
                      ("c0", "<i4"), ("c1", "<i4"), ("param", "<i4"), ("state_bytes", "<i4"), ("in_parts", "<i4"),
                      ("in_part_bytes", "<u8"), ("pad", "<u8")])
assert JUMP_ITEM.itemsize == 64
BATCH, CHUNK = 4, 1024


def jump_masks(coeff_lists):
    """Per-residue masks of up to BATCH jumps (lists of set coefficient
    indices): word c*64+rho holds bit 16b+q for coefficient c*1024+64q+rho."""
    top = max([max(c) for c in coeff_lists if len(c)] + [0])
    nchunks = top // CHUNK + 1
    m = np.zeros(nchunks * 64, np.uint64)
    for b, cs in enumerate(coeff_lists):
        for k in cs:
            c, r = divmod(int(k), CHUNK)
            q, rho = divmod(r, 64)
            m[c * 64 + rho] |= np.uint64(1 << (16 * b + q))
    return m, nchunks


def set_id(bits, mexp, index):
    t = {(32, 11213): 0, (32, 23209): 1, (32, 44497): 2, (64, 23209): 3, (64, 44497): 4, (64, 110503): 5}
    return t[(bits, mexp)] * 128 + index


def kernel_class(bits, n):
    lim = (351, 726, 1391) if bits == 32 else (363, 696, 1727)
    return next(i for i, v in enumerate(lim) if n <= v)


def run(lib, bits, items, layout, grid=None):
    arr = np.zeros(len(items), ITEM)
    n_max = 0
    for k, it in enumerate(items):
        for f, v in it.items():
            arr[k][f] = v
    cls = max(kernel_class(bits, it["n"]) for it in items) if items else 0
    rc = lib.emu_run_items(bits, cls, layout, arr.ctypes.data, len(items), grid or len(items))
    assert rc == 0, rc


MJUMP_GROUP = np.dtype([("state_in", "<u8"), ("in_part_bytes", "<u8"), ("planes", "<u8"), ("coef", "<u8"),
                        ("out", "<u8"), ("out_stride", "<u8"), ("nj", "<i4"), ("njp", "<i4"), ("kpad", "<i4"),
                        ("pw", "<i4"), ("param", "<i4"), ("in_parts", "<i4"), ("n", "<i4"), ("pad", "<i4")])
assert MJUMP_GROUP.itemsize == 80
MJUMP_ITEM = np.dtype([("group", "<i4"), ("r0", "<i4"), ("j0", "<i4"), ("pad", "<i4")])
MJ_WORDS = 8


def mj_words(rt):
    return MJ_WORDS // rt


def mj_kpad(mexp):
    return (mexp + 63) // 64 * 64


def mj_pw(kpad, n):
    pw = kpad // 32 + (n + MJ_WORDS + 31) // 32 + 3
    return (pw + 1) & ~1
