"""The request server (mtgp_server_*): continuous batching must not change
any result.  Each request's output is compared between the batched server
(requests coalesced and overlapping on the device) and the plain path (one
launch per request, in arrival order), and against the oracle."""
import os

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu


def _fresh_set(mtgp, specs, seeds):
    gs = mtgp.GenSet(specs)
    gs.init_seed(seeds)
    return gs


def _oracle_digests(mtgp, oracle, specs, seeds, tr, which):
    """Oracle outputs of the requests `which` (each generator's stream in
    submission order) and the oracle's final windows."""
    states = {}
    for g, (bits, mexp, idx) in enumerate(specs):
        p = oracle.params(bits, mexp, idx)
        words = mexp // 64 + 2
        P = oracle.hex_to_poly(mtgp.charpoly_hex(bits, mexp, idx), words)
        states[g] = [p, oracle.init_state(p, seeds[g]), P]
    out = {}
    for i in range(len(tr["gen"])):
        g = int(tr["gen"][i])
        p, w, P = states[g]
        skip, count, kind = int(tr["skip"][i]), int(tr["count"][i]), int(tr["kind"][i])
        if skip >= 1 << 20:
            w = oracle.jump(p, w, oracle.poly_to_hex(oracle.xpow_mod(skip, P)))
        elif skip:
            _, w = oracle.generate(p, w, 0, skip)
        if i in which:
            want, w = oracle.generate(p, w, kind, count)
            out[i] = mtgp.digest64(want)
        else:
            _, w = oracle.generate(p, w, kind, count)
        states[g][1] = w
    return out, {g: states[g][1] for g in states}


@pytest.mark.parametrize("cell,planners", [(1 << 12, 0), (1 << 17, 0), (1 << 12, 3)])
def test_server_same_alone_and_in_mix(mtgp, oracle, cell, planners):
    """A Poisson trace over mixed MTGP32-11213 / MTGP32-23209 / MTGP64-23209
    generators, request sizes 2^6..2^16, every kind, 10% skips: identical
    digests batched and plain, oracle digests for every request, and the
    final states; batches planned on the scheduler thread or on planner
    threads (MTGP_SERVER_PLANNERS)."""
    specs, seeds, tr = bench.serving_workload(n_requests=400, rate=20000.0, n_gen=12, log2_mean=11.0,
                                              log2_sd=2.0, log2_min=6, log2_max=16, seed=3)
    res = {}
    for plain in (False, True):
        gs = _fresh_set(mtgp, specs, seeds)
        os.environ["MTGP_SERVER_PLANNERS"] = str(planners)
        try:
            srv = mtgp.Server(gs, arena_bytes=64 << 20, cell=cell, cached_cells=64, plain=plain)
        finally:
            del os.environ["MTGP_SERVER_PLANNERS"]
        lat, dg, wall = srv.replay(tr["t_us"], tr["gen"], tr["kind"], tr["count"], tr["skip"], digests=True)
        st = srv.stats()
        srv.close()
        res[plain] = (dg, [gs.state(g) for g in range(len(specs))], st, lat)
    assert (res[False][0] == res[True][0]).all(), np.nonzero(res[False][0] != res[True][0])[0][:10]
    assert res[True][2]["batches"] == len(tr["gen"])
    assert res[False][2]["batches"] < len(tr["gen"])  # requests were coalesced
    want, final = _oracle_digests(mtgp, oracle, specs, seeds, tr, set(range(len(tr["gen"]))))
    bad = [i for i in want if want[i] != int(res[False][0][i])]
    assert not bad, "requests differing from the oracle: %s" % bad[:10]
    for g in final:
        assert (res[False][1][g] == final[g]).all() and (res[True][1][g] == final[g]).all(), g
    assert (res[False][3] > 0).all()


def _oracle_gen_digests(mtgp, oracle, spec, seed, reqs):
    """Oracle digests of every request of one generator, `reqs` its
    (trace index, kind, count, skip) in submission order; skips of 2^20 or
    more are applied as x^skip mod P jumps."""
    bits, mexp, idx = spec
    p = oracle.params(bits, mexp, idx)
    w = oracle.init_state(p, seed)
    P = None
    out = {}
    for i, kind, count, skip in reqs:
        if skip >= 1 << 20:
            if P is None:
                P = oracle.hex_to_poly(mtgp.charpoly_hex(bits, mexp, idx), mexp // 64 + 2)
            w = oracle.jump(p, w, oracle.poly_to_hex(oracle.xpow_mod(skip, P)))
        elif skip:
            _, w = oracle.generate(p, w, 0, skip)
        want, w = oracle.generate(p, w, kind, count)
        out[i] = mtgp.digest64(want)
    return out, w


def test_server_full_serving_mix(mtgp, oracle):
    """The benchmark's serving configuration at full size: 12,000 Poisson
    requests at 20,000/s over 256 generators (128 MTGP32-11213, 64
    MTGP32-23209, 64 MTGP64-23209), 2^10..2^26 numbers each, 10% with skips
    up to 2^30, a 4 GiB output budget, through the batched server.  Every
    request of a sample of generators -- the ones holding the largest
    requests, skips stepped over inside a cell and skips jumped -- and their
    final windows against the oracle."""
    from concurrent.futures import ThreadPoolExecutor
    specs, seeds, tr = bench.serving_workload(n_requests=12000, rate=20000.0)
    gs = _fresh_set(mtgp, specs, seeds)
    srv = mtgp.Server(gs, arena_bytes=4 << 30)
    print("full serving mix: replay with digests", flush=True)
    lat, dg, wall = srv.replay(tr["t_us"], tr["gen"], tr["kind"], tr["count"], tr["skip"], digests=True)
    srv.close()
    gen, count, skip = tr["gen"], tr["count"].astype(np.int64), tr["skip"].astype(np.int64)
    largest = [int(gen[i]) for i in np.argsort(-count)[:6]]
    stepped = [int(g) for g in dict.fromkeys(gen[(skip > 0) & (skip < 1 << 17)])][:4]
    jumped = [int(g) for g in dict.fromkeys(gen[skip >= 1 << 20])][:4]
    sample = list(dict.fromkeys(largest + stepped + jumped))
    assert count.max() == 1 << 26 and len(stepped) == 4 and len(jumped) == 4
    per = {g: [] for g in sample}
    for i in range(len(gen)):
        if int(gen[i]) in per:
            per[int(gen[i])].append((i, int(tr["kind"][i]), int(count[i]), int(skip[i])))
    print("full serving mix: oracle for generators %s" % sample, flush=True)
    with ThreadPoolExecutor(8) as ex:
        res = dict(zip(sample, ex.map(lambda g: _oracle_gen_digests(mtgp, oracle, specs[g], seeds[g], per[g]),
                                      sample)))
    n_checked = 0
    for g in sample:
        want, w = res[g]
        bad = [i for i in want if want[i] != int(dg[i])]
        assert not bad, "generator %d: requests differing from the oracle: %s" % (g, bad[:10])
        assert (gs.state(g) == w).all(), "generator %d: final window" % g
        n_checked += len(want)
    assert n_checked > 100


def test_server_budget_admission(mtgp, oracle):
    """An output budget far below the offered load: requests wait for
    admission until earlier outputs are released, and every result is still
    exact (digests against the plain path and the oracle on a sample)."""
    specs, seeds, tr = bench.serving_workload(n_requests=200, rate=1e6, n_gen=6, log2_mean=14.0, log2_sd=1.0,
                                              log2_min=12, log2_max=16, seed=9)
    digs = []
    for plain in (False, True):
        gs = _fresh_set(mtgp, specs, seeds)
        srv = mtgp.Server(gs, arena_bytes=1 << 20, cell=1 << 14, cached_cells=16, plain=plain)
        _, dg, _ = srv.replay(tr["t_us"], tr["gen"], tr["kind"], tr["count"], tr["skip"], digests=True)
        srv.close()
        digs.append(dg)
    assert (digs[0] == digs[1]).all()
    sample = set(range(0, 200, 7))
    want, _ = _oracle_digests(mtgp, oracle, specs, seeds, tr, sample)
    assert all(want[i] == int(digs[0][i]) for i in sample)


def test_server_submit_poll_release(mtgp, oracle):
    """The submit / poll / release interface directly: outputs read from
    the arena before release match the oracle; errors for bad input."""
    gs = mtgp.GenSet([(32, 11213, 0), (64, 23209, 1)])
    gs.init_seed([1, 2])
    srv = mtgp.Server(gs, arena_bytes=8 << 20, cell=1 << 12, cached_cells=8)
    t0 = srv.submit(0, 0, 5000)
    t1 = srv.submit(1, 2, 3000, skip=10)
    t2 = srv.submit(0, 1, 7, skip=100000)
    with pytest.raises(mtgp.MTGPError):
        srv.submit(5, 0, 10)
    with pytest.raises(mtgp.MTGPError):
        srv.submit(0, 0, 1 << 40)
    srv.drain()
    got = {}
    while len(got) < 3:
        for c in srv.poll(wait_ms=100):
            got[c["ticket"]] = c
    p0, p1 = oracle.params(32, 11213, 0), oracle.params(64, 23209, 1)
    w0, w1 = oracle.init_state(p0, 1), oracle.init_state(p1, 2)
    e0, w0 = oracle.generate(p0, w0, 0, 5000)
    _, w1 = oracle.generate(p1, w1, 0, 10)
    e1, w1 = oracle.generate(p1, w1, 2, 3000)
    _, w0 = oracle.generate(p0, w0, 0, 100000)
    e2, w0 = oracle.generate(p0, w0, 1, 7)
    for t, e in ((t0, e0), (t1, e1), (t2, e2)):
        c = got[t]
        assert c["bytes"] == e.nbytes and c["done_us"] >= c["launch_us"] >= c["submit_us"]
        host = np.empty(e.size, e.dtype)
        mtgp.memcpy_dtoh(host, c["out"])
        assert (host.view(np.uint8) == e.view(np.uint8)).all(), t
        srv.release(t)
    with pytest.raises(mtgp.MTGPError):
        srv.release(t0)
    srv.close()
    assert (gs.state(0) == w0).all() and (gs.state(1) == w1).all()
