"""mtgp32_fast_jump_noalloc (device memory from the library's static array)
against mtgp32_fast_jump (per-call hipMalloc), through the C ABI.

Two calls in a row, the second on a smaller state and with a shorter jump
string, so anything the first call left in the shared array would show;
each result must equal the allocating entry point's bit for bit, and the
oracle's.  Error paths leave the state unchanged.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(mtgp):
    f = mtgp._lib.mtgp32_fast_jump_noalloc
    f.restype = C.c_int
    f.argtypes = [C.POINTER(mtgp.Fast32), C.c_char_p]
    return f


def _pair(mtgp, mexp, index, seed):
    return mtgp.Generator32(mexp, index, seed=seed), mtgp.Generator32(mexp, index, seed=seed)


def test_noalloc_jump_matches_fast_jump(mtgp, oracle):
    noalloc = _noalloc(mtgp)
    cases = [
        # larger state and a full-length polynomial first ...
        (23209, 5, 99, 3**162),
        # ... then a smaller state and a shorter string
        (11213, 0, 1234, 10**6),
    ]
    for mexp, index, seed, steps in cases:
        print("noalloc jump: MTGP32-%d #%d seed %d, %d steps" % (mexp, index, seed, steps), flush=True)
        a, b = _pair(mtgp, mexp, index, seed)
        hexs = mtgp.jump_hex(32, mexp, index, steps)
        if steps == 10**6:
            hexs = hexs[:400]  # a shorter (still valid) jump polynomial
        a.fast_jump(hexs)
        assert noalloc(C.byref(b.g), hexs.encode()) == 0, mtgp.last_error()
        assert mtgp.last_error() == ""
        wa, wb = a.window(), b.window()
        bad = np.nonzero(wa != wb)[0]
        assert bad.size == 0, "MTGP32-%d: %d words differ, first %d" % (mexp, bad.size, bad[0])
        p = oracle.params(32, mexp, index)
        w = oracle.jump(p, oracle.init_state(p, seed), hexs)
        assert (wb == w).all(), "MTGP32-%d: differs from the oracle" % mexp
        # the jumped generators continue identically
        assert (a.fill(0, 5000) == b.fill(0, 5000)).all()


def test_noalloc_jump_errors(mtgp):
    noalloc = _noalloc(mtgp)
    g = mtgp.Generator32(11213, 3, seed=7)
    before = g.window()
    assert noalloc(C.byref(g.g), b"12g4") == -1
    assert "non-hex" in mtgp.last_error()
    assert (g.window() == before).all()
    assert noalloc(C.byref(g.g), b"1" * 16385) == -1
    assert "limit" in mtgp.last_error()
    assert (g.window() == before).all()
    # the empty string is the zero polynomial, as for mtgp32_fast_jump
    h = mtgp.Generator32(11213, 3, seed=7)
    h.fast_jump("")
    assert noalloc(C.byref(g.g), b"") == 0
    assert (g.window() == h.window()).all()
