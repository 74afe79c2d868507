"""Dynamic-creator sets on the GPU, against the reference's outputs
(tests/golden/dc_outputs.npz) and the oracle: MTGP32-DC on the regular
kernels, MTGP64-DC (N - pos < 256) on the 128-word-step kernels."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def arrays():
    return np.load(os.path.join(GOLD, "dc_outputs.npz"))


@pytest.mark.parametrize("idx", [0, 16, 199])
def test_dc32_fill_and_jump(mtgp, arrays, idx):
    rec = mtgp.dc_params(32, idx)
    key = "32_%d" % idx
    g = mtgp.Generator32(record=rec, seed=1)
    assert (g.fill(0, 1000) == arrays[key + "_init"]).all()
    for kind in (1, 2, 3, 4):
        got = g.fill(kind, 1000)
        assert (got.view(np.uint8) == arrays[key + "_init_k%d" % kind].view(np.uint8)).all(), kind
    g = mtgp.Generator32(record=rec, array=[1, 2, 3, 4])
    assert (g.fill(0, 1000) == arrays[key + "_array"]).all()
    for s in (1, 352, 3003, 200004):
        g = mtgp.Generator32(record=rec, seed=0)
        g.fill(0, 3)
        g.fast_jump(mtgp.jump_hex(32, steps=s, record=rec))
        assert (g.fill(0, 100) == arrays["%s_step_%d" % (key, s)]).all(), s


def test_dc32_genset_plan(mtgp, oracle, arrays):
    """A generator set built from DC records, each stream segmented across
    the GPU by jump-ahead (jump polynomials from the records' own
    characteristic polynomials): the reference's head and the oracle's
    full stream."""
    from pyoracle import Oracle
    from test_dc import _record
    idxs = (0, 16, 199)
    recs = [mtgp.dc_params(32, i) for i in idxs]
    gs = mtgp.GenSet(recs)
    gs.init_seed([1, 1, 1])
    n = 300000
    bufs = [mtgp.DeviceBuffer(4 * n) for _ in recs]
    plan = gs.plan([(g, 0, n, 0, bufs[g]) for g in range(3)], segment=16384)
    assert plan.stats()["items"] > 30
    plan.launch()
    mtgp.synchronize()
    for g, idx in enumerate(idxs):
        got = bufs[g].to_host(np.uint32, n)
        assert (got[:1000] == arrays["32_%d_init" % idx]).all(), idx
        p = Oracle.params_from_record(32, _record(recs[g]))
        want, _ = oracle.generate(p, oracle.init_state(p, 1), 0, n)
        assert (got == want).all(), idx


@pytest.mark.parametrize("idx", [0, 172, 199])
def test_dc64_fill_and_jump(mtgp, arrays, idx):
    """MTGP64-DC sets (N = 176, N - pos down to 130) run on the 128-word-step
    kernels: fills of every kind and device jumps against the reference."""
    rec = mtgp.dc_params(64, idx)
    key = "64_%d" % idx
    g = mtgp.Generator64(record=rec, seed=1)
    assert (g.fill(0, 1000) == arrays[key + "_init"]).all()
    for kind in (1, 2, 3, 4):
        got = g.fill(kind, 1000)
        assert (got.view(np.uint8) == arrays[key + "_init_k%d" % kind].view(np.uint8)).all(), kind
    g = mtgp.Generator64(record=rec, array=[1, 2, 3, 4])
    assert (g.fill(0, 1000) == arrays[key + "_array"]).all()
    for s in (1, 177, 3003, 200004):
        g = mtgp.Generator64(record=rec, seed=0)
        g.fill(0, 3)
        g.fast_jump(mtgp.jump_hex(64, steps=s, record=rec))
        assert (g.fill(0, 100) == arrays["%s_step_%d" % (key, s)]).all(), s


def test_dc64_genset_plan(mtgp, oracle):
    """A set of MTGP64-DC generators, each stream segmented across the GPU
    (batched jumps on the 128-word-step jump kernel), against the oracle."""
    from pyoracle import Oracle
    from test_dc import _record
    idxs = (0, 172, 199)
    recs = [mtgp.dc_params(64, i) for i in idxs]
    gs = mtgp.GenSet(recs)
    gs.init_seed([3, 4, 5])
    n = 200000
    bufs = [mtgp.DeviceBuffer(8 * n) for _ in recs]
    plan = gs.plan([(g, 2 if g == 1 else 0, n, 0, bufs[g]) for g in range(3)], segment=8192)
    assert plan.stats()["items"] > 60
    plan.launch()
    mtgp.synchronize()
    plan.commit()
    for g, idx in enumerate(idxs):
        p = Oracle.params_from_record(64, _record(recs[g]))
        want, w = oracle.generate(p, oracle.init_state(p, 3 + g), 2 if g == 1 else 0, n)
        got = bufs[g].to_host(want.dtype, n)
        assert (got.view(np.uint64) == want.view(np.uint64)).all(), idx
        assert (gs.state(g) == w).all(), idx


def test_mtgp64_fast_jump_reference_string(mtgp):
    """The drop-in mtgp64_fast_jump on the GPU with the reference's shipped
    64-bit jump string (openCL-sample/mtgp64-jump-string.h, x^(3^162) mod
    the characteristic polynomial of set 0 of the OpenCL sample's
    MTGP64-DC-11213 table, a 128-word-step set): the window and the
    outputs after it against the reference mtgp64_fast_jump's."""
    from test_gf2 import params64_of_record
    d = np.load(os.path.join(GOLD, "dc64cl_jump.npz"))
    p = params64_of_record(mtgp, d["record"])
    g = mtgp.Generator64(record=p, seed=4321)
    g.fast_jump(str(d["jump3_162"]))
    assert (g.window() == d["window"]).all()
    assert (g.fill(0, 1000) == d["uint64"]).all()
    assert (g.fill(2, 1000).view(np.uint64) == d["close_open"].view(np.uint64)).all()
