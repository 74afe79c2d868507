"""Repeated drop-in fills (mtgp32_fill / mtgp64_fill): one reused plan per
device and jump polynomials from the process-wide cache of uniform grids.
Streams continue bit-exactly across calls of changing sizes, generators of
one parameter set share a cache, other sets and word sizes do not get its
polynomials, and host and device destinations agree -- all against the
oracle."""
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
