"""Repeated drop-in fills (mtgp32_fill / mtgp64_fill) and genset generate
calls: one reused plan per device (per set), relaunched without planning
for calls of the same shape, and jump polynomials from the process-wide
cache of uniform grids.  Streams continue bit-exactly across calls of
changing sizes, generators of one parameter set share a cache, other sets
and word sizes do not get its polynomials, host and device destinations
agree, and other requests, jumps and knob changes between calls of one
shape leave every stream exact -- all against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(got, want, what):
    g = got.view(np.uint32 if got.dtype.itemsize == 4 else np.uint64)
    w = want.view(g.dtype)
    bad = np.nonzero(g != w)[0]
    assert bad.size == 0, "%s: %d mismatches, first at %d" % (what, bad.size, bad[0])


@pytest.mark.parametrize("bits,mexp,index", [(32, 11213, 3), (32, 23209, 2), (64, 23209, 1)])
def test_fills_of_changing_sizes_continue_the_stream(mtgp, oracle, bits, mexp, index):
    print("fills MTGP%d-%d set %d" % (bits, mexp, index), flush=True)
    Gen = mtgp.Generator32 if bits == 32 else mtgp.Generator64
    gen = Gen(mexp, index, seed=900 + index)
    p = oracle.params(bits, mexp, index)
    w = oracle.init_state(p, 900 + index)
    for k, (kind, n, dev) in enumerate([(0, 1 << 22, False), (2, 1 << 22, True), (0, 1 << 21, False),
                                        (1, 1 << 22, True), (0, (1 << 22) + 333, False), (2, 1 << 22, False)]):
        if dev:
            buf = mtgp.DeviceBuffer(n * bits // 8)
            gen.fill(kind, n, buf)
            got = buf.to_host(np.uint32 if bits == 32 else np.uint64, n)
        else:
            got = gen.fill(kind, n)
        want, w = oracle.generate(p, w, kind, n)
        _same(got, want, "fill %d (kind %d, n %d, device %s)" % (k, kind, n, dev))
    assert (gen.window() == w).all()


def test_generators_of_one_set_share_the_cache(mtgp, oracle):
    """Two generators of MTGP32-11213 set 7 (different seeds) and one of set
    8, fills of one size in turn: each stream against its own oracle."""
    print("fills sharing a jump cache", flush=True)
    specs = [(7, 11), (7, 22), (8, 33)]
    gens = [mtgp.Generator32(11213, i, seed=s) for i, s in specs]
    ref = []
    for i, s in specs:
        p = oracle.params(32, 11213, i)
        ref.append([p, oracle.init_state(p, s)])
    n = 3 << 20
    for rnd in range(2):
        for k, g in enumerate(gens):
            got = g.fill(0, n)
            want, ref[k][1] = oracle.generate(ref[k][0], ref[k][1], 0, n)
            _same(got, want, "round %d generator %d" % (rnd, k))


@pytest.mark.parametrize("bits,mexp", [(32, 11213), (64, 23209)])
def test_repeated_fills_of_one_shape(mtgp, oracle, bits, mexp):
    """Fills of one shape in a row launch the plan built for the first (no
    planning); a changed planner knob, another generator of the set and a
    different kind in between each still continue every stream exactly."""
    print("repeated fills MTGP%d-%d" % (bits, mexp), flush=True)
    Gen = mtgp.Generator32 if bits == 32 else mtgp.Generator64
    gens = [Gen(mexp, 4, seed=71), Gen(mexp, 4, seed=72)]
    p = oracle.params(bits, mexp, 4)
    ws = [oracle.init_state(p, 71), oracle.init_state(p, 72)]
    n = 1 << 21
    buf = mtgp.DeviceBuffer(n * bits // 8)
    wd = np.uint32 if bits == 32 else np.uint64
    steps = [(0, 0), (0, 0), (1, 0), (0, 0), ("knob", None), (0, 0), (0, 2), (0, 0)]
    old = mtgp.tuning_get("jump_waves_per_cu")
    try:
        for k, (g, kind) in enumerate(steps):
            if g == "knob":
                mtgp.tuning_set("jump_waves_per_cu", 12 if old != 12 else 8)
                continue
            gens[g].fill(kind, n, buf)
            want, ws[g] = oracle.generate(p, ws[g], kind, n)
            _same(buf.to_host(wd, n), want, "fill %d (generator %d, kind %d)" % (k, g, kind))
    finally:
        mtgp.tuning_set("jump_waves_per_cu", old)
    for g in range(2):
        assert (gens[g].window() == ws[g]).all()


def test_repeated_generate_calls_of_one_shape(mtgp, oracle):
    """mtgp_genset_generate with the same requests in a row re-points its
    built plan at the generators' current windows; other requests, a jump
    and a changed planner knob in between each still give every stream
    exactly."""
    print("repeated genset generate calls", flush=True)
    specs = [(32, 11213, 2), (64, 23209, 3), (32, 23209, 4)]
    seeds = [5, 6, 7]
    gs = mtgp.GenSet(specs)
    gs.init_seed(seeds)
    ref = []
    for (bits, mexp, idx), sd in zip(specs, seeds):
        p = oracle.params(bits, mexp, idx)
        ref.append([p, oracle.init_state(p, sd), bits])
    counts = [3 << 19, 1 << 20, (1 << 20) + 77]
    bufs = [mtgp.DeviceBuffer(c * b // 8) for c, (b, _, _) in zip(counts, specs)]
    shape = [(g, (0, 2, 1)[g], counts[g], 0, bufs[g].address(0)) for g in range(3)]

    def run(reqs):
        gs.generate(reqs)
        for g, kind, count, skip, _ in reqs:
            p, w, bits = ref[g]
            if skip:
                _, w = oracle.generate(p, w, 0, skip)
            e, w = oracle.generate(p, w, kind, count)
            ref[g][1] = w
            got = bufs[g].to_host(np.uint32 if bits == 32 else np.uint64, count)
            _same(got, e, "generator %d" % g)

    old = mtgp.tuning_get("jump_waves_per_cu")
    try:
        run(shape)
        run(shape)
        run([shape[1], (0, 0, 5000, 12345, bufs[0].address(0))])
        run(shape)
        js = mtgp.jump_hex(64, 23209, 3, 10 ** 9 + 7)
        gs.jump(1, js)
        ref[1][1] = oracle.jump(ref[1][0], ref[1][1], js)
        run(shape)
        mtgp.tuning_set("jump_waves_per_cu", 12 if old != 12 else 8)
        run(shape)
        run(shape)
    finally:
        mtgp.tuning_set("jump_waves_per_cu", old)
    for g in range(3):
        assert (gs.state(g) == ref[g][1]).all(), "window of generator %d" % g
