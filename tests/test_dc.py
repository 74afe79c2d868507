"""Dynamic-creator parameter sets (cuda-sample/mtgp32dc-param-11213.c,
mtgp64dc-param-11213.c; 200 MEXP-11213 sets each), on the host: the
library's exported tables against digests the reference build wrote, the
oracle (driven by the reference's own records) against the reference's
outputs, and the library's init family through the oracle."""
import hashlib
import json
import os

import numpy as np
import pytest

from pyoracle import Oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SETS = {32: (0, 16, 199), 64: (0, 172, 199)}


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def gold():
    return json.load(open(os.path.join(GOLD, "dc_params_digests.json"))), np.load(os.path.join(GOLD, "dc_outputs.npz"))


def _record(p):
    r = [p.mexp, p.pos, p.sh1, p.sh2] + list(p.tbl) + list(p.tmp_tbl)
    r += list(p.flt_tmp_tbl if hasattr(p, "flt_tmp_tbl") else p.dbl_tmp_tbl) + [p.mask]
    r += list(bytes(p.poly_sha1)[:20])
    return np.array(r, np.uint64)


@pytest.mark.parametrize("bits", [32, 64])
def test_exported_dc_tables_match_reference(mtgp, gold, bits):
    digests, _ = gold
    for i in range(200):
        assert _sha(_record(mtgp.dc_params(bits, i))) == digests["digests"][str(bits)][i], (bits, i)


@pytest.mark.parametrize("bits", [32, 64])
def test_oracle_on_dc_sets(oracle, gold, bits):
    """The oracle on the reference's own records (full records are in the
    fixture for sets 0 and 199) reproduces the reference's outputs."""
    digests, arrays = gold
    for idx in (0, 199):
        p = Oracle.params_from_record(bits, digests["records"]["%d/%d" % (bits, idx)])
        key = "%d_%d" % (bits, idx)
        w = oracle.init_state(p, 1)
        out, w = oracle.generate(p, w, 0, 1000)
        assert (out == arrays[key + "_init"]).all()
        for kind in (1, 2, 3, 4):
            out, w = oracle.generate(p, w, kind, 1000)
            assert (out.view(np.uint8) == arrays[key + "_init_k%d" % kind].view(np.uint8)).all(), kind
        out, _ = oracle.generate(p, oracle.init_by_array(p, [1, 2, 3, 4]), 0, 1000)
        assert (out == arrays[key + "_array"]).all()


@pytest.mark.parametrize("bits", [32, 64])
def test_library_init_on_dc_sets(mtgp, oracle, gold, bits):
    """mtgp*_init / init_by_array with DC records (including the pos = 3
    sets), continued by the oracle, give the reference's sequences."""
    _, arrays = gold
    G = mtgp.Generator32 if bits == 32 else mtgp.Generator64
    for idx in SETS[bits]:
        rec = mtgp.dc_params(bits, idx)
        p = Oracle.params_from_record(bits, _record(rec))
        key = "%d_%d" % (bits, idx)
        out, _ = oracle.generate(p, G(record=rec, seed=1).window(), 0, 1000)
        assert (out == arrays[key + "_init"]).all(), idx
        out, _ = oracle.generate(p, G(record=rec, array=[1, 2, 3, 4]).window(), 0, 1000)
        assert (out == arrays[key + "_array"]).all(), idx


@pytest.mark.parametrize("bits", [32, 64])
def test_dc_jump_polynomials(mtgp, oracle, gold, bits):
    """The library's charpoly and x^s mod P for DC records: jumping the
    oracle by them equals the reference's stepping."""
    _, arrays = gold
    n = 11213 // bits + 1
    for idx in SETS[bits][::2]:
        rec = mtgp.dc_params(bits, idx)
        p = Oracle.params_from_record(bits, _record(rec))
        G = mtgp.Generator32 if bits == 32 else mtgp.Generator64
        w0 = G(record=rec, seed=0).window()
        _, w = oracle.generate(p, w0, 0, 3)
        for s in (1, n + 1, 3003, 200004):
            wj = oracle.jump(p, w, mtgp.jump_hex(bits, steps=s, record=rec))
            out, _ = oracle.generate(p, wj, 0, 100)
            assert (out == arrays["%d_%d_step_%d" % (bits, idx, s)]).all(), (idx, s)
