"""Randomised plans against the oracle: mixed word sizes and state classes,
table and dynamic-creator sets (records), every output kind, short and
long skips (stepped over or jumped), zero-count
requests, automatic and explicit cell grids, every step layout, the
two-size grid forced on small batches, the two-phase launch order, and a
re-launch before the commit.  Every output number and every final window is
compared bit for bit."""
import numpy as np
import pytest
from pyoracle import Oracle
from test_dc import _record

pytestmark = pytest.mark.gpu

SETS = [(32, 11213), (32, 23209), (32, 44497), (64, 23209), (64, 44497), (64, 110503)]


def _bits(a):
    return a.view(np.uint32 if a.dtype.itemsize == 4 else np.uint64)


@pytest.mark.parametrize("seed", range(20))
def test_random_plan(mtgp, oracle, seed):
    rng = np.random.default_rng(1000 + seed)
    G = int(rng.integers(2, 8))
    specs = []
    for _ in range(G):
        bits, mexp = SETS[int(rng.choice(len(SETS), p=[0.3, 0.15, 0.1, 0.25, 0.12, 0.08]))]
        specs.append((bits, mexp, int(rng.integers(0, 128))))
    seeds = [int(v) for v in rng.integers(1, 2**32, G)]
    # About a third of the cases build the set from parameter records, half
    # of them dynamic-creator sets (MTGP64-DC runs on the 128-word-step
    # class); a separate generator keeps the other draws as they were.
    rrng = np.random.default_rng(5000 + seed)
    recs = None
    if rrng.random() < 0.35:
        recs = []
        for g, (bits, mexp, idx) in enumerate(specs):
            if rrng.random() < 0.5:
                bits = int(rrng.choice([32, 64]))
                recs.append(mtgp.dc_params(bits, int(rrng.integers(0, 200))))
                specs[g] = (bits, 11213, -1)
            else:
                recs.append(mtgp.params(bits, mexp, idx))
    gs = mtgp.GenSet(recs if recs is not None else specs)
    gs.init_seed(seeds)
    st = []
    for g, (bits, mexp, idx) in enumerate(specs):
        if idx < 0:
            p = Oracle.params_from_record(bits, _record(recs[g]))
        else:
            p = oracle.params(bits, mexp, idx)
        st.append([p, oracle.init_state(p, seeds[g] if bits == 64 else seeds[g] & 0xFFFFFFFF), None])
    reqs, expect, offs = [], [], []
    off = 0
    total = 0
    for k in range(int(rng.integers(G, 4 * G + 1))):
        g = int(rng.integers(0, G))
        bits, mexp, idx = specs[g]
        p, w, P = st[g]
        kind = int(rng.integers(0, 5))
        r = rng.random()
        count = 0 if r < 0.08 else int(rng.integers(1, 300)) if r < 0.3 else \
            int(rng.integers(300, 60000)) if r < 0.8 else int(rng.integers(60000, 700000))
        r = rng.random()
        skip = 0 if r < 0.65 else int(rng.integers(1, 5000)) if r < 0.82 else \
            int(rng.integers(1 << 16, 1 << 21)) if r < 0.95 else int(rng.integers(1 << 40, 1 << 50))
        if total + count + (skip if skip < 1 << 30 else 0) > 6_000_000:
            continue
        total += count + (skip if skip < 1 << 30 else 0)
        if skip >= 1 << 30:
            if P is None:
                hexp = mtgp.charpoly_hex(bits, record=recs[g]) if recs is not None else \
                    mtgp.charpoly_hex(bits, mexp, idx)
                P = st[g][2] = oracle.hex_to_poly(hexp, mexp // 64 + 2)
            w = oracle.jump(p, w, oracle.poly_to_hex(oracle.xpow_mod(skip, P)))
        elif skip:
            _, w = oracle.generate(p, w, 0, skip)
        want, w = oracle.generate(p, w, kind, count)
        st[g][1] = w
        size = np.dtype(mtgp.out_dtype(bits, kind)).itemsize
        off = (off + size - 1) // size * size + size * int(rng.integers(0, 3))  # any element alignment
        reqs.append([g, kind, count, skip, off])
        expect.append(want)
        offs.append(off)
        off += max(1, count) * size
    buf = mtgp.DeviceBuffer(off + 64)
    for rq in reqs:
        rq[4] = buf.address(rq[4])
    reqs = [tuple(rq) for rq in reqs]
    segment = 0 if rng.random() < 0.5 else int(rng.choice([4096, 8192, 65536, 300000]))
    force_two_size = rng.random() < 0.4
    with mtgp.tuning(two_size_rho=1e9 if force_two_size else mtgp.tuning_get("two_size_rho")):
        plan = gs.plan(reqs, segment=segment)
    overlap = plan.set_overlap(rng.random() < 0.5)
    if rng.random() < 0.3:
        plan.configure(layout=int(rng.integers(0, 4)))
    what = "seed %d: %d gens %s (records %s), %d requests, segment %d, two-size %s, two-phase %s, items %d" % (
        seed, G, specs, recs is not None, len(reqs), segment, force_two_size, overlap, plan.stats()["items"])
    buf.from_host(np.full(off + 64, 0xA5, np.uint8))
    plan.launch()
    mtgp.synchronize()
    plan.launch()  # the same numbers again, from the same states
    mtgp.synchronize()
    plan.commit()
    raw = buf.to_host(np.uint8, off + 64)
    for k, (rq, want) in enumerate(zip(reqs, expect)):
        g, kind, count = rq[0], rq[1], rq[2]
        if not count:
            continue
        dt = mtgp.out_dtype(specs[g][0], kind)
        got = raw[offs[k]:offs[k] + count * np.dtype(dt).itemsize].view(dt)
        bad = np.nonzero(_bits(got) != _bits(want))[0]
        assert bad.size == 0, "%s; request %d (gen %d kind %d count %d skip %d): %d mismatches from %d" % (
            what, k, g, kind, count, rq[3], bad.size, bad[0])
    # bytes outside every request are untouched
    mask = np.ones(off + 64, bool)
    for k, rq in enumerate(reqs):
        mask[offs[k]:offs[k] + rq[2] * np.dtype(mtgp.out_dtype(specs[rq[0]][0], rq[1])).itemsize] = False
    assert (raw[mask] == 0xA5).all(), what + ": write outside a request"
    for g in range(G):
        assert (gs.state(g) == st[g][1]).all(), "%s: state of generator %d" % (what, g)
