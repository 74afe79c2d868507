"""Plans and generator states: rebasing a committed plan continues every
stream (batch after batch, into the same outputs, also by graph replay),
and a commit that would rewind a generator is refused.  Every output and
final window against the oracle, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    return a.view(np.uint32 if a.dtype.itemsize == 4 else np.uint64)


def _batch(mtgp, oracle, specs, seeds):
    """A mixed request list (kinds, short and jumped skips) and the oracle's
    per-generator states."""
    gs = mtgp.GenSet(specs)
    gs.init_seed(seeds)
    st = []
    for (bits, mexp, idx), s in zip(specs, seeds):
        p = oracle.params(bits, mexp, idx)
        st.append([p, oracle.init_state(p, s if bits == 64 else s & 0xFFFFFFFF)])
    rng = np.random.default_rng(17)
    shape = []
    for k in range(3 * len(specs)):
        g = k % len(specs)
        kind = int(rng.integers(0, 5))
        count = int(rng.integers(1, 120000))
        skip = [0, 0, 3000, 200000][int(rng.integers(0, 4))]
        shape.append((g, kind, count, skip))
    return gs, st, shape


def _expect(mtgp, oracle, specs, st, shape):
    want = []
    for g, kind, count, skip in shape:
        p, w = st[g]
        if skip:
            _, w = oracle.generate(p, w, 0, skip)
        out, w = oracle.generate(p, w, kind, count)
        st[g][1] = w
        want.append(out)
    return want


def _layout(mtgp, specs, shape):
    off, offs = 0, []
    for g, kind, count, skip in shape:
        size = np.dtype(mtgp.out_dtype(specs[g][0], kind)).itemsize
        off = (off + 7) // 8 * 8
        offs.append(off)
        off += count * size
    return offs, off


def _check(mtgp, buf, specs, shape, offs, want, what):
    for k, ((g, kind, count, skip), o) in enumerate(zip(shape, offs)):
        dt = mtgp.out_dtype(specs[g][0], kind)
        got = buf.to_host(dt, count, o)
        bad = np.nonzero(_bits(got) != _bits(want[k]))[0]
        assert bad.size == 0, "%s: request %d (gen %d kind %d): %d mismatches from %d" % (
            what, k, g, kind, bad.size, bad[0])


@pytest.mark.parametrize("segment,graph", [(0, False), (16384, False), (16384, True)])
def test_rebase_continues_streams(mtgp, oracle, segment, graph):
    specs = [(32, 11213, 3), (32, 23209, 7), (64, 23209, 1), (32, 11213, 90), (64, 44497, 2)]
    seeds = [11, 22, 33, 44, 55]
    gs, st, shape = _batch(mtgp, oracle, specs, seeds)
    offs, nbytes = _layout(mtgp, specs, shape)
    buf = mtgp.DeviceBuffer(nbytes)
    plan = gs.plan([(g, kind, count, skip, buf.address(o)) for (g, kind, count, skip), o in zip(shape, offs)],
                   segment=segment)
    stream = mtgp.Stream()
    if graph:
        plan.capture(stream)
    for batch in range(3):
        what = "segment %d graph %s batch %d" % (segment, graph, batch)
        print(what, flush=True)
        want = _expect(mtgp, oracle, specs, st, shape)
        if batch:
            plan.rebase(stream)
        if graph:
            plan.launch_graph(stream)
        else:
            plan.launch(stream)
        plan.commit()
        stream.synchronize()
        _check(mtgp, buf, specs, shape, offs, want, what)
        for g in range(len(specs)):
            assert (gs.state(g) == st[g][1]).all(), "%s: window of generator %d" % (what, g)


def test_commit_refuses_stale_plans(mtgp, oracle):
    specs = [(32, 11213, 0), (64, 23209, 5)]
    seeds = [1234, 99]
    gs, st, _ = _batch(mtgp, oracle, specs, seeds)
    n = 50000
    a = mtgp.DeviceBuffer(4 * n)
    b = mtgp.DeviceBuffer(8 * n)
    reqs = [(0, 0, n, 0, a), (1, 2, n, 0, b)]
    # never launched: refused, nothing moves
    p0 = gs.plan(reqs)
    with pytest.raises(mtgp.MTGPError, match="never launched"):
        p0.commit()
    w0 = [gs.state(g) for g in range(2)]
    # two plans from the same states: the second commit would rewind
    p1 = gs.plan(reqs, segment=8192)
    p1.launch()
    p0.launch()
    mtgp.synchronize()
    p0.commit()
    with pytest.raises(mtgp.MTGPError, match="state changed"):
        p1.commit()
    want = _expect(mtgp, oracle, specs, st, [(0, 0, n, 0), (1, 2, n, 0)])
    for g in range(2):
        assert (gs.state(g) == st[g][1]).all(), "state after the refused commit, generator %d" % g
    # rebased, the refused plan continues from the committed state
    p1.rebase()
    p1.launch()
    mtgp.synchronize()
    p1.commit()
    want = _expect(mtgp, oracle, specs, st, [(0, 0, n, 0), (1, 2, n, 0)])
    assert (a.to_host(np.uint32, n) == want[0]).all()
    assert (_bits(b.to_host(np.float64, n)) == _bits(want[1])).all()
    # set_state after building: refused, and the uploaded window stays
    p2 = gs.plan(reqs)
    gs.set_state(0, w0[0])
    p2.launch()
    mtgp.synchronize()
    with pytest.raises(mtgp.MTGPError, match="state changed"):
        p2.commit()
    assert (gs.state(0) == w0[0]).all()
    # a second commit of one launch does nothing
    p3 = gs.plan([(1, 0, 100, 0, b)])
    p3.launch()
    mtgp.synchronize()
    p3.commit()
    s1 = gs.state(1)
    p3.commit()
    assert (gs.state(1) == s1).all()
