"""GPU parity at the benchmark sizes, through the C ABI.

The expected values are SHA-256 digests of every output array and final
window, written by the reference library itself (tests/golden/streams.json,
substreams.npz; see tests/golden/make_golden.py), so the whole multi-GB
output of each workload is compared bit for bit without an oracle run at
full size.  Samples are also checked against the oracle directly.
"""
import hashlib
import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _sha(a):
    return hashlib.sha256(memoryview(np.ascontiguousarray(a)).cast("B")).hexdigest()


def _digests(buf, parts, workers=8):
    """SHA-256 of each (dtype, count, byte offset) slice of a device buffer."""
    def one(p):
        dt, n, off = p
        return _sha(buf.to_host(dt, n, off))
    with ThreadPoolExecutor(workers) as ex:
        return list(ex.map(one, parts))


@pytest.fixture(scope="module")
def streams():
    return json.load(open(os.path.join(GOLD, "streams.json")))


def test_batched32_workload(mtgp, streams):
    """All 128 MTGP32-11213 sets, splitmix64 seeds, L_g uint32 then L_g
    close_open floats each (the benchmark workload), in one plan: every
    output and every final window against the reference's digests; then
    the plan rebased onto the committed states makes the next batch of
    every stream, against the reference's digests of that batch."""
    gold = streams["batched32"]
    seeds, lengths = gold["seeds"], gold["lengths"]
    G = len(seeds)
    total = 8 * sum(lengths)
    buf = mtgp.DeviceBuffer(total)
    gs = mtgp.GenSet([(32, 11213, g) for g in range(G)])
    gs.init_seed(seeds)
    reqs, parts, off = [], [], 0
    for g in range(G):
        L = lengths[g]
        reqs.append((g, 0, L, 0, buf.address(off)))
        parts.append((np.uint32, L, off))
        off += 4 * L
        reqs.append((g, 2, L, 0, buf.address(off)))
        parts.append((np.float32, L, off))
        off += 4 * L
    plan = gs.plan(reqs)
    assert plan.stats()["items"] > G  # the streams really are segmented
    plan.launch()
    mtgp.synchronize()
    plan.commit()
    got = _digests(buf, parts)
    bad = [g for g in range(G) if got[2 * g] != gold["uint32"][g] or got[2 * g + 1] != gold["close_open"][g]]
    assert not bad, "generators with wrong output: %s" % bad[:10]
    bad = [g for g in range(G) if _sha(gs.state(g)) != gold["window"][g]]
    assert not bad, "generators with wrong final window: %s" % bad[:10]
    nxt = streams["batched32_next"]
    plan.rebase()
    plan.launch()
    mtgp.synchronize()
    plan.commit()
    got = _digests(buf, parts)
    bad = [g for g in range(G) if got[2 * g] != nxt["uint32"][g] or got[2 * g + 1] != nxt["close_open"][g]]
    assert not bad, "second batch: generators with wrong output: %s" % bad[:10]
    bad = [g for g in range(G) if _sha(gs.state(g)) != nxt["window"][g]]
    assert not bad, "second batch: generators with wrong final window: %s" % bad[:10]


@pytest.mark.parametrize("mexp", [23209, 44497])
def test_batched64_workload(mtgp, streams, mexp):
    """MTGP64 sets 0..63 of one MEXP, init_by_array with 4-word splitmix64
    keys, 2^24 uint64 then 2^24 close_open doubles each (16 GiB), in one
    plan, against the reference's digests."""
    gold = streams["batched64"]
    n = 1 << 24
    G = 64
    buf = mtgp.DeviceBuffer(G * 16 * n)
    gs = mtgp.GenSet([(64, mexp, i) for i in range(G)])
    gs.init_array([[int(k) for k in gold["keys"]["%d/%d" % (mexp, i)]] for i in range(G)])
    reqs, parts, off = [], [], 0
    for i in range(G):
        reqs.append((i, 0, n, 0, buf.address(off)))
        parts.append((np.uint64, n, off))
        off += 8 * n
        reqs.append((i, 2, n, 0, buf.address(off)))
        parts.append((np.float64, n, off))
        off += 8 * n
    gs.generate(reqs)
    got = _digests(buf, parts)
    for i in range(G):
        key = "%d/%d" % (mexp, i)
        assert got[2 * i] == gold["uint64"][key], key
        assert got[2 * i + 1] == gold["close_open"][key], key
        assert _sha(gs.state(i)) == gold["window"][key], key


def _substreams(mtgp, base, hexstr, count, keep=()):
    """4096 generators = substreams i of `base` by powers of `hexstr`,
    each generating `count` uint32 into one buffer.  Returns the buffer and
    the start windows of the substreams in `keep` (all with keep=None)."""
    S = 4096
    gs = mtgp.GenSet([(32, 11213, 0)] * S)
    gs.substreams(base, hexstr)
    starts = {i: gs.state(i) for i in (range(S) if keep is None else keep)}
    buf = mtgp.DeviceBuffer(S * 4 * count)
    gs.generate([(i, 0, count, 0, buf.address(4 * count * i)) for i in range(S)])
    return buf, starts


def test_substreams_3_162(mtgp, oracle):
    """Substream i = set 0, seed 777, jumped i times by 3^162 steps (the
    reference's jump string), for i < 4096, 2^20 uint32 each (16 GiB).
    Every substream's outputs and start window against digests of the
    reference's own repeated mtgp32_fast_jump; sampled substreams in full;
    for a spread of i, the first 1024 outputs of substream i+1 against the
    oracle's jump of substream i's start window."""
    sub = np.load(os.path.join(GOLD, "substreams.npz"))
    jd = np.load(os.path.join(GOLD, "reference_jump_data.npz"))
    js = str(jd["jump3_162"])
    count = 1 << 20
    p = oracle.params(32, 11213, 0)
    base = oracle.init_state(p, 777)
    rng = np.random.default_rng(5)
    chain = sorted(set(rng.integers(0, 4095, 24).tolist()) | {0, 4094})
    buf, starts = _substreams(mtgp, base, js, count, keep=None)
    out16 = sub["all_out_sha16"]
    win16 = sub["all_window_sha16"]

    def digest16(i):
        out = hashlib.sha256(memoryview(buf.to_host(np.uint32, count, 4 * count * i)).cast("B")).digest()[:16]
        w = starts[i].copy()
        w[0] &= np.uint32(p.mask)
        return out, hashlib.sha256(w.tobytes()).digest()[:16]
    with ThreadPoolExecutor(8) as ex:
        got = list(ex.map(digest16, range(4096)))
    bad = [i for i in range(4096) if got[i][0] != bytes(out16[i])]
    assert not bad, "substreams with wrong output: %d, first %s" % (len(bad), bad[:10])
    bad = [i for i in range(4096) if got[i][1] != bytes(win16[i])]
    assert not bad, "substreams with wrong start window: %d, first %s" % (len(bad), bad[:10])
    for i in sub["indices"]:
        i = int(i)
        got = buf.to_host(np.uint32, count, 4 * count * i)
        assert (got[:1024] == sub["head_%d" % i]).all(), i
        assert _sha(got) == str(sub["sha_%d" % i]), i
        w = starts[i]
        ref = sub["window_%d" % i]
        assert (w[1:] == ref[1:]).all() and (w[0] & p.mask) == (ref[0] & p.mask), i
    # chain check on a spread of indices: start(i+1) == jump(start(i))
    for i in chain:
        wn = oracle.jump(p, starts[i], js)
        want, _ = oracle.generate(p, wn, 0, 1024)
        got = buf.to_host(np.uint32, 1024, 4 * count * (i + 1))
        assert (got == want).all(), i


@pytest.mark.parametrize("mfma_min", [None, 0])
def test_substreams_2_64(mtgp, oracle, mfma_min):
    """Substream i = set 0, seed 777, jumped by i * 2^64 steps (a host-made
    polynomial), 2^20 uint32 each: samples against the oracle's direct
    x^(i 2^64) mod P jump -- start states from the matrix cores (default
    for 4096 substreams) and from the radix-4 levels of batched jumps
    (matrix cores off)."""
    print("substreams 2^64, jump_mfma_min %s" % mfma_min, flush=True)
    count = 1 << 20
    p = oracle.params(32, 11213, 0)
    base = oracle.init_state(p, 777)
    js = mtgp.jump_hex(32, 11213, 0, 1 << 64)
    words = 11213 // 64 + 2
    P = oracle.hex_to_poly(mtgp.charpoly_hex(32, 11213, 0), words)
    assert oracle.poly_to_hex(oracle.xpow_mod(1 << 64, P)) == js
    if mfma_min is None:
        buf, _ = _substreams(mtgp, base, js, count)
    else:
        with mtgp.tuning(jump_mfma_min=mfma_min):
            buf, _ = _substreams(mtgp, base, js, count)
    for i in (0, 1, 2, 3, 100, 2047, 4095):
        w = oracle.jump(p, base, oracle.poly_to_hex(oracle.xpow_mod(i << 64, P))) if i else base
        want, _ = oracle.generate(p, w, 0, 1 << 16)
        got = buf.to_host(np.uint32, 1 << 16, 4 * count * i)
        assert (got == want).all(), i
