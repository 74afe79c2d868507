"""Pins the oracle (oracle/mtgp_oracle.c) before anything is checked
against it: golden vectors written by the reference library itself, the
reference build (when present), the reference's shipped jump data, and
what the mathematics fixes (SHA-1 of the characteristic polynomial,
x^(4^k * 3^162) relations, jump == stepping)."""
import hashlib
import json
import os

import numpy as np
import pytest

from pyoracle import RefLib

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MEXPS = {32: (11213, 23209, 44497), 64: (23209, 44497, 110503)}


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def gold_main():
    return np.load(os.path.join(GOLD, "main_outputs.npz"))


@pytest.fixture(scope="module")
def jdata():
    return np.load(os.path.join(GOLD, "reference_jump_data.npz"))


def _record(p):
    """The 73-value record layout of the golden parameter digests."""
    r = [p.mexp, p.pos, p.sh1, p.sh2] + list(p.tbl) + list(p.tmp) + list(p.flt) + [p.mask]
    r += list(bytes.fromhex(p.sha1.decode()))
    return np.array(r, np.uint64)


def test_param_expansion_matches_reference_tables(oracle):
    gold = json.load(open(os.path.join(GOLD, "params_digests.json")))
    for bits, mexps in MEXPS.items():
        for mexp in mexps:
            digests = gold["digests"]["%d/%d" % (bits, mexp)]
            for i in range(128):
                assert _sha(_record(oracle.params(bits, mexp, i))) == digests[i], (bits, mexp, i)
    for key, rec in gold["records"].items():
        bits, mexp, i = map(int, key.split("/"))
        assert _record(oracle.params(bits, mexp, i)).tolist() == rec


@pytest.mark.parametrize("bits,mexp", [(b, m) for b in MEXPS for m in MEXPS[b]])
@pytest.mark.parametrize("idx", [0, 127])
def test_main_sequences(oracle, gold_main, bits, mexp, idx):
    """The reference test programs' sequences (mtgp32-fast.c:436-514)."""
    key = "%d_%d_%d" % (bits, mexp, idx)
    p = oracle.params(bits, mexp, idx)
    out, _ = oracle.generate(p, oracle.init_state(p, 1), 0, 1000)
    assert (out == gold_main[key + "_init"]).all()
    out, _ = oracle.generate(p, oracle.init_by_array(p, [1, 2, 3, 4]), 0, 1000)
    assert (out == gold_main[key + "_array"]).all()
    w = oracle.init_by_str(p, b"\x01\x02\x03\x04")
    out, w = oracle.generate(p, w, 0, 1000)
    assert (out == gold_main[key + "_str"]).all()
    for kind in (1, 2, 3, 4):
        out, w = oracle.generate(p, w, kind, 1000)
        g = gold_main[key + "_str_k%d" % kind]
        assert (out.view(g.dtype) == g).all() if kind == 1 else (out.view(np.uint8) == g.view(np.uint8)).all()
    out, _ = oracle.generate(p, oracle.init_by_str(p, b"MTGP\xff\x80\x7f!"), 0, 256)
    assert (out == gold_main[key + "_str_signed"]).all()


def test_planted_mistake_is_caught(oracle, gold_main):
    """A parameter off by one (pos + 1) must fail the golden comparison."""
    p = oracle.params(32, 11213, 0)
    p.pos += 1
    out, _ = oracle.generate(p, oracle.init_state(p, 1), 0, 1000)
    assert not (out == gold_main["32_11213_0_init"]).all()


def test_float_family_identities(oracle):
    """close_open = [1,2) - 1, open_close = 2 - [1,2), open_open sets the
    lowest mantissa bit: exact identities on the same stream positions."""
    p = oracle.params(32, 11213, 9)
    w = oracle.init_state(p, 5)
    f12, _ = oracle.generate(p, w, 1, 5000)
    f12 = f12.view(np.float32)
    c0, _ = oracle.generate(p, w, 2, 5000)
    oc, _ = oracle.generate(p, w, 3, 5000)
    oo, _ = oracle.generate(p, w, 4, 5000)
    assert (c0 == f12 - np.float32(1)).all() and (c0 >= 0).all() and (c0 < 1).all()
    assert (oc == np.float32(2) - f12).all() and (oc > 0).all() and (oc <= 1).all()
    assert ((oo + np.float32(1)).view(np.uint32) == (f12.view(np.uint32) | 1)).all()
    u, _ = oracle.generate(p, w, 0, 5000)
    assert (f12.view(np.uint32) == ((u >> 9) | 0x3F800000)).all()


def test_reference_jump_data(oracle, jdata):
    """x^(3^162) mod CHAR32_11213 is the shipped jump string; the jump table
    rows are x^(4^k 3^162) mod P, i.e. row k+1 = row k ^ 4 mod P."""
    char, js, table = str(jdata["char32_11213"]), str(jdata["jump3_162"]), jdata["jump_table"]
    words = 11213 // 64 + 2
    P = oracle.hex_to_poly(char, words)
    assert oracle.poly_deg(P) == 11213
    J = oracle.xpow_mod(3**162, P)
    assert oracle.poly_to_hex(J) == js
    rows = [np.frombuffer(np.ascontiguousarray(r).tobytes() + b"\0" * (8 * words - 4 * 351), np.uint64)
            for r in table]
    assert (rows[0] == J).all()
    for k in range(5):
        sq = oracle.mulmod(rows[k], rows[k], P)
        assert (oracle.mulmod(sq, sq, P) == rows[k + 1]).all(), k


def test_charpoly_sha1_and_bm(oracle, jdata):
    """Berlekamp-Massey on 2*MEXP output LSBs (mtgp32-calc-poly.cpp:36-45)
    gives CHAR32_11213, and the SHA-1 of its coefficient string (degree 0
    first) is the table's poly_sha1 -- for several sets."""
    for bits, mexp, idx in [(32, 11213, 0), (32, 11213, 77), (64, 23209, 3)]:
        p = oracle.params(bits, mexp, idx)
        out, _ = oracle.generate(p, oracle.init_state(p, 0), 0, 2 * mexp)
        words = mexp // 64 + 2
        C, L = oracle.bm((out & 1).astype(np.uint8), words)
        assert L == mexp
        s = "".join(str((int(C[i // 64]) >> (i % 64)) & 1) for i in range(L + 1))
        assert hashlib.sha1(s.encode()).hexdigest() == p.sha1.decode()
        if (bits, mexp, idx) == (32, 11213, 0):
            assert oracle.poly_to_hex(C) == str(jdata["char32_11213"])


def test_jump_equals_stepping(oracle):
    """test-jump32.cpp:218-264: jumping by x^s mod P equals stepping s, for
    the reference's step list (checked against golden stepping outputs)."""
    gold = np.load(os.path.join(GOLD, "jump_steps.npz"))
    for bits, mexp in [(32, 11213), (64, 23209)]:
        p = oracle.params(bits, mexp, 0)
        words = mexp // 64 + 2
        out, _ = oracle.generate(p, oracle.init_state(p, 0), 0, 2 * mexp)
        P, _ = oracle.bm((out & 1).astype(np.uint8), words)
        _, w = oracle.generate(p, oracle.init_state(p, 0), 0, 3)
        n = p.n
        for s in (1, 2, n + 1, 32 * n - 1, 32 * n + 1, 3003, 200004, 10000005):
            wj = oracle.jump(p, w, oracle.poly_to_hex(oracle.xpow_mod(s, P)))
            got, _ = oracle.generate(p, wj, 0, 100)
            assert (got == gold["%d_%d_%d" % (bits, mexp, s)]).all(), s


def test_plain_stream_digest(oracle):
    """MTGP32-11213 set 0, seed 1234: 2^24 uint32 then 2^24 close1_open2."""
    st = json.load(open(os.path.join(GOLD, "streams.json")))["plain"]
    p = oracle.params(32, 11213, 0)
    a, w = oracle.generate(p, oracle.init_state(p, 1234), 0, 1 << 24)
    b, w = oracle.generate(p, w, 1, 1 << 24)
    assert _sha(a) == st["uint32"] and _sha(b) == st["close1_open2"]
    assert w.tolist() == st["window"]


def test_substreams(oracle, jdata):
    """Substream i = the seed-777 state jumped i times by 3^162 (the
    reference applies mtgp32_fast_jump i times).  The oracle composes
    J^i mod P into one jump: outputs are identical; the window agrees except
    in the 19 low bits of word 0, which the recursion masks off."""
    sub = np.load(os.path.join(GOLD, "substreams.npz"))
    p = oracle.params(32, 11213, 0)
    words = 11213 // 64 + 2
    P = oracle.hex_to_poly(str(jdata["char32_11213"]), words)
    J = oracle.hex_to_poly(str(jdata["jump3_162"]), words)
    base = oracle.init_state(p, 777)
    assert (base == sub["base"]).all()
    for i in sub["indices"]:
        i = int(i)
        c = np.zeros(words, np.uint64)
        c[0] = 1
        for bit in reversed(range(max(1, i.bit_length()))):
            c = oracle.mulmod(c, c, P)
            if i >> bit & 1:
                c = oracle.mulmod(c, J, P)
        w = oracle.jump(p, base, oracle.poly_to_hex(c)) if i else base
        out, _ = oracle.generate(p, w, 0, 1 << 20)
        assert (out[:1024] == sub["head_%d" % i]).all(), i
        assert _sha(out) == str(sub["sha_%d" % i]), i
        ref_w = sub["window_%d" % i]
        assert (w[1:] == ref_w[1:]).all() and (w[0] & p.mask) == (ref_w[0] & p.mask)


@pytest.mark.skipif(not RefLib.available("fast", 32), reason="reference build (oracle/_ref) absent")
def test_oracle_against_reference_build(oracle):
    """Random seeds, keys and strings through both, every kind, 32 and 64."""
    rng = np.random.default_rng(1)
    for bits in (32, 64):
        ref, refref = RefLib("fast", bits), RefLib("ref", bits)
        for mexp in MEXPS[bits]:
            for _ in range(3):
                idx = int(rng.integers(0, 128))
                p = oracle.params(bits, mexp, idx)
                seed = int(rng.integers(0, 2**32 if bits == 32 else 2**63))
                g = ref.new_seed(mexp, idx, seed)
                w = oracle.init_state(p, seed)
                for kind in range(5):
                    n = int(rng.integers(1, 3000))
                    a, w = oracle.generate(p, w, kind, n)
                    assert (a.view(np.uint8) == g.fill(kind, n).view(np.uint8)).all()
                assert (g.window() == w).all()
                key = rng.integers(0, 2**32, int(rng.integers(1, 40)), dtype=np.uint64)
                if bits == 32:
                    key = key.astype(np.uint32)
                assert (ref.new_array(mexp, idx, key).window() == oracle.init_by_array(p, key)).all()
                # the empty key (init_by_array with length 0)
                assert (ref.new_array(mexp, idx, key[:0]).window() == oracle.init_by_array(p, key[:0])).all()
                s = bytes(rng.integers(1, 256, int(rng.integers(1, 50)), dtype=np.uint8))
                assert (ref.new_str(mexp, idx, s).window() == oracle.init_by_str(p, s)).all()
                # the independent "ref" formulation (mtgp32-ref.c) agrees too
                a, _ = oracle.generate(p, oracle.init_by_str(p, s), 0, 500)
                assert (refref.new_str(mexp, idx, s).fill(0, 500) == a).all()
