"""GPU parity: the HIP kernels (through the C ABI) against the oracle.

Every comparison is bit-exact: integers and the raw bits of every float or
double output.  Sizes are chosen so the oracle finishes in seconds.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SETS = [(32, 11213), (32, 23209), (32, 44497), (64, 23209), (64, 44497), (64, 110503)]


def bits_of(a):
    return a.view(np.uint32 if a.dtype.itemsize == 4 else np.uint64)


def assert_same(got, want, what):
    g, w = bits_of(np.asarray(got)), bits_of(np.asarray(want))
    assert g.shape == w.shape, what
    bad = np.nonzero(g != w)[0]
    assert bad.size == 0, "%s: %d mismatches, first at %d: %x != %x" % (
        what, bad.size, bad[0], int(g[bad[0]]), int(w[bad[0]]))


@pytest.mark.parametrize("bits,mexp", SETS)
@pytest.mark.parametrize("index", [0, 127])
def test_fill_sequence_all_kinds(mtgp, oracle, bits, mexp, index):
    """Short fills of every kind, continuing one stream, with lengths that
    hit the head (unaligned), full-step and tail paths."""
    gen = (mtgp.Generator32 if bits == 32 else mtgp.Generator64)(mexp, index, seed=4321 + index)
    p = oracle.params(bits, mexp, index)
    w = oracle.init_state(p, 4321 + index)
    assert_same(gen.window(), w, "init")
    for kind, n in [(0, 1), (1, 3), (2, 255), (3, 256), (4, 257), (0, 1000), (2, 5000), (1, 0), (0, 777)]:
        got = gen.fill(kind, n)
        want, w = oracle.generate(p, w, kind, n)
        assert_same(got, want, "kind %d n %d" % (kind, n))
        assert_same(gen.window(), w, "window after kind %d n %d" % (kind, n))


def test_close1_open2_is_floating_point(mtgp, oracle):
    """[1,2) comes back as float32 / float64 (mtgp*_genrand_close1_open2
    returns a float / double), or as raw bits into an unsigned array."""
    for bits, mexp in ((32, 11213), (64, 23209)):
        gen = (mtgp.Generator32 if bits == 32 else mtgp.Generator64)(mexp, 5, seed=77)
        p = oracle.params(bits, mexp, 5)
        w = oracle.init_state(p, 77)
        got = gen.fill(mtgp.KIND_CLOSE1_OPEN2, 3000)
        want, w = oracle.generate(p, w, 1, 3000)
        assert got.dtype == (np.float32 if bits == 32 else np.float64)
        assert_same(got, want, "close1_open2 MTGP%d" % bits)
        assert got.min() >= 1.0 and got.max() < 2.0
        raw = np.empty(1000, np.uint32 if bits == 32 else np.uint64)
        gen.fill(mtgp.KIND_CLOSE1_OPEN2, 1000, out=raw)
        want, w = oracle.generate(p, w, 1, 1000)
        assert_same(raw, want, "close1_open2 raw bits MTGP%d" % bits)


def test_config1_single_generator(mtgp, oracle):
    """MTGP32-11213 set 0, mtgp32_init(1234): 2^24 uint32 then 2^24
    close1_open2 continuing the stream, bit-exact, with the final state."""
    n = 1 << 24
    gen = mtgp.Generator32(11213, 0, seed=1234)
    p = oracle.params(32, 11213, 0)
    w = oracle.init_state(p, 1234)
    got = gen.fill(0, n)
    want, w = oracle.generate(p, w, 0, n)
    assert_same(got, want, "uint32")
    got = gen.fill(1, n)
    want, w = oracle.generate(p, w, 1, n)
    assert_same(got, want, "close1_open2")
    assert_same(gen.window(), w, "final window")


@pytest.mark.parametrize("offset", [0, 4, 8, 12])
def test_unaligned_device_output(mtgp, oracle, offset):
    n = 10000
    buf = mtgp.DeviceBuffer(n * 4 + 64)
    gen = mtgp.Generator32(11213, 5, seed=99)
    gen.fill(0, n, out=_View(buf, offset))
    got = buf.to_host(np.uint32, n, offset)
    p = oracle.params(32, 11213, 5)
    want, _ = oracle.generate(p, oracle.init_state(p, 99), 0, n)
    assert_same(got, want, "offset %d" % offset)


@pytest.mark.parametrize("offset", [0, 8])
def test_unaligned_device_output_64(mtgp, oracle, offset):
    n = 3001
    buf = mtgp.DeviceBuffer(n * 8 + 64)
    gen = mtgp.Generator64(23209, 7, seed=5)
    gen.fill(2, n, out=_View(buf, offset))
    got = buf.to_host(np.float64, n, offset)
    p = oracle.params(64, 23209, 7)
    want, _ = oracle.generate(p, oracle.init_state(p, 5), 2, n)
    assert_same(got, want, "offset %d" % offset)


class _View:
    """A DeviceBuffer seen from a byte offset (for fill(out=...))."""

    def __init__(self, buf, offset):
        self.ptr = buf.ptr + offset
        self.buf = buf


def test_fast_jump_matches_oracle(mtgp, oracle):
    hexs = mtgp.jump_hex(32, 11213, 0, 3**162)
    gen = mtgp.Generator32(11213, 0, seed=777)
    p = oracle.params(32, 11213, 0)
    w = oracle.init_state(p, 777)
    gen.fast_jump(hexs)
    w = oracle.jump(p, w, hexs)
    assert_same(gen.window(), w, "jumped window")
    got = gen.fill(0, 5000)
    want, _ = oracle.generate(p, w, 0, 5000)
    assert_same(got, want, "after jump")


@pytest.mark.parametrize("bits,mexp", SETS)
def test_jump_equals_stepping(mtgp, oracle, bits, mexp):
    """test-jump32.cpp:218-264 pattern: jumping by x^s mod P equals
    stepping s numbers, for s around the window and step sizes."""
    n = mexp // bits + 1
    for s in (1, 2, n + 1, 32 * n - 1, 3003, 200004):
        gen = (mtgp.Generator32 if bits == 32 else mtgp.Generator64)(mexp, 1, seed=0)
        gen.fill(0, 3)
        ref = gen.window()
        p = oracle.params(bits, mexp, 1)
        _, stepped = oracle.generate(p, ref, 0, s)
        gen.fast_jump(mtgp.jump_hex(bits, mexp, 1, s))
        a = gen.fill(0, 100)
        b, _ = oracle.generate(p, stepped, 0, 100)
        assert_same(a, b, "s=%d" % s)


@pytest.mark.parametrize("layout", [0, 1, 2])
def test_genset_segmented_requests(mtgp, oracle, layout):
    """Several generators, requests with skips, split into short segments so
    that most segments start from a device jump; every step layout (layout
    2, two-wave cells, applies to the MTGP32-23209 launch)."""
    specs = [(32, 11213, 3), (32, 23209, 9), (64, 23209, 4), (32, 11213, 100), (64, 44497, 2)]
    gs = mtgp.GenSet(specs)
    seeds = [11, 22, 33, 44, 55]
    gs.init_seed(seeds)
    plan_reqs, expect, bufs = [], [], []
    windows = {}
    for g, (bits, mexp, idx) in enumerate(specs):
        p = oracle.params(bits, mexp, idx)
        windows[g] = (p, oracle.init_state(p, seeds[g]))
    requests = [(0, 0, 70000, 0), (0, 2, 30001, 1234), (1, 1, 50000, 7), (2, 0, 40000, 0),
                (3, 4, 123457, 99999), (2, 3, 1, 3), (4, 2, 20000, 0), (4, 0, 5, 5)]
    for g, kind, count, skip in requests:
        bits = specs[g][0]
        dt = mtgp.out_dtype(bits, kind)
        buf = mtgp.DeviceBuffer(count * np.dtype(dt).itemsize)
        bufs.append((buf, dt, count))
        plan_reqs.append((g, kind, count, skip, buf))
        p, w = windows[g]
        if skip:
            _, w = oracle.generate(p, w, 0, skip)
        want, w = oracle.generate(p, w, kind, count)
        windows[g] = (p, w)
        expect.append(want)
    plan = gs.plan(plan_reqs, segment=8192)
    assert plan.stats()["items"] > 40
    plan.configure(layout=layout)
    plan.launch()
    mtgp.synchronize()
    plan.commit()
    for k, (buf, dt, count) in enumerate(bufs):
        assert_same(buf.to_host(dt, count), expect[k], "request %d" % k)
    for g in range(len(specs)):
        assert_same(gs.state(g), windows[g][1], "state %d" % g)


@pytest.mark.parametrize("layout", [0, 1, 2, 3])
@pytest.mark.parametrize("bits,mexp", [(32, 11213), (32, 44497), (64, 23209), (64, 44497)])
def test_no_write_past_requested_outputs(mtgp, oracle, layout, bits, mexp):
    """Every request of a plan writes exactly its own range: guard words
    around each output stay untouched (tails of every length mod 256 and
    odd output offsets)."""
    eb = bits // 8
    counts = [1, 2, 3, 5, 63, 64, 65, 255, 256, 257, 511, 1000, 4099]
    gs = mtgp.GenSet([(bits, mexp, 3)] * len(counts))
    gs.init_seed(list(range(len(counts))))
    guard = 64
    layout_words = [(guard + c + guard) for c in counts]
    buf = mtgp.DeviceBuffer(sum(layout_words) * eb + 64)
    pattern = np.full(sum(layout_words) * eb + 64, 0xA5, np.uint8)
    buf.from_host(pattern)
    reqs, offs, off = [], [], 0
    for g, c in enumerate(counts):
        start = off + guard * eb + (eb if g % 2 else 0)
        reqs.append((g, g % 5, c, 0, buf.address(start)))
        offs.append(start)
        off += layout_words[g] * eb
    plan = gs.plan(reqs, segment=1 << 20)
    plan.configure(layout=layout)
    plan.launch()
    mtgp.synchronize()
    host = buf.to_host(np.uint8)
    mask = np.ones(host.size, bool)
    for g, c in enumerate(counts):
        mask[offs[g]:offs[g] + c * eb] = False
        p = oracle.params(bits, mexp, 3)
        want, _ = oracle.generate(p, oracle.init_state(p, g), g % 5, c)
        got = host[offs[g]:offs[g] + c * eb].view(want.dtype)
        assert_same(got, want, "request %d" % g)
    assert (host[mask] == 0xA5).all(), "bytes outside the requested ranges were written"


@pytest.mark.parametrize("jump_mode", ["0", "1"])
def test_genset_chains_and_gaps(mtgp, oracle, monkeypatch, jump_mode):
    """Many short requests per generator (chains of segments inside one
    work item), short skips stepped over, long skips reached by a jump, a
    2^40 skip, zero-count requests and a trailing skip, on a 4096-number
    cell grid, against the oracle; both start-state jump modes."""
    monkeypatch.setenv("MTGP_JUMP_MODE", jump_mode)
    specs = [(32, 11213, 1), (64, 23209, 2), (32, 44497, 3)]
    gs = mtgp.GenSet(specs)
    seeds = [5, 6, 7]
    gs.init_seed(seeds)
    rng = np.random.default_rng(int(jump_mode) + 11)
    state = {}
    for g, (bits, mexp, idx) in enumerate(specs):
        p = oracle.params(bits, mexp, idx)
        words = mexp // 64 + 2
        P = oracle.hex_to_poly(mtgp.charpoly_hex(bits, mexp, idx), words)
        state[g] = [p, oracle.init_state(p, seeds[g]), P]
    reqs, expect, bufs = [], [], []
    for k in range(90):
        g = k % 3
        p, w, P = state[g]
        kind = int(rng.integers(0, 5))
        count = int(rng.integers(0, 3000)) if k % 17 else 0
        r = rng.random()
        skip = 0 if r < 0.7 else int(rng.integers(1, 500)) if r < 0.85 else int(rng.integers(70000, 200000)) \
            if r < 0.97 else 1 << 40
        if skip >= 1 << 30:
            w = oracle.jump(p, w, oracle.poly_to_hex(oracle.xpow_mod(skip, P)))
        elif skip:
            _, w = oracle.generate(p, w, 0, skip)
        want, w = oracle.generate(p, w, kind, count)
        state[g][1] = w
        dt = mtgp.out_dtype(specs[g][0], kind)
        buf = mtgp.DeviceBuffer(max(1, count) * np.dtype(dt).itemsize)
        bufs.append((buf, dt, count))
        reqs.append((g, kind, count, skip, buf))
        expect.append(want)
    for g in range(3):  # trailing skip
        p, w, P = state[g]
        _, state[g][1] = oracle.generate(p, w, 0, 33)
        reqs.append((g, 0, 0, 33, 0))
    plan = gs.plan(reqs, segment=4096)
    plan.launch()
    mtgp.synchronize()
    plan.commit()
    for k, (buf, dt, count) in enumerate(bufs):
        if count:
            assert_same(buf.to_host(dt, count), expect[k], "request %d" % k)
    for g in range(3):
        assert_same(gs.state(g), state[g][1], "state %d" % g)


def test_device_fill_and_copy(mtgp):
    """The plain fill and copy kernels (counter and ceiling references, and
    the graph-safe way to clear memory): exact ranges, nothing beyond."""
    n = (1 << 20) + 12
    buf = mtgp.DeviceBuffer(4 * n + 64)
    buf.from_host(np.zeros(n + 16, np.uint32))
    mtgp.device_fill(buf, 0xDEADBEEF, 4 * (n - 3))
    mtgp.synchronize()
    h = buf.to_host(np.uint32, n + 16)
    assert (h[:n - 3] == 0xDEADBEEF).all() and (h[n - 3:] == 0).all()
    src = mtgp.DeviceBuffer(4096)
    src.from_host(np.arange(1024, dtype=np.uint32))
    mtgp.device_copy(buf, src, 4096)
    mtgp.synchronize()
    h = buf.to_host(np.uint32, 1100)
    assert (h[:1024] == np.arange(1024)).all() and (h[1024:1100] == 0xDEADBEEF).all()


def test_invalid_requests_are_refused(mtgp):
    """Planning and jump errors are reported (mtgp_last_error) and change
    nothing: out-of-range generator or kind, NULL or misaligned outputs,
    non-hex jump strings, an empty request list."""
    gs = mtgp.GenSet([(32, 11213, 0), (64, 23209, 0)])
    gs.init_seed([1, 2])
    w0, w1 = gs.state(0), gs.state(1)
    buf = mtgp.DeviceBuffer(1 << 16)
    bad = [[(2, 0, 10, 0, buf)], [(0, 5, 10, 0, buf)], [(-1, 0, 10, 0, buf)], [(0, 0, 10, 0, 0)],
           [(1, 0, 10, 0, buf.address(4))]]
    for reqs in bad:
        with pytest.raises(mtgp.MTGPError):
            gs.generate(reqs)
    with pytest.raises(mtgp.MTGPError):
        gs.jump(0, "12xz")
    gs.generate([])
    assert (gs.state(0) == w0).all() and (gs.state(1) == w1).all()
    g = mtgp.Generator32(11213, 0, seed=1)
    w = g.window().copy()
    with pytest.raises(mtgp.MTGPError):
        g.fast_jump("not hex")
    assert (g.window() == w).all()
    with pytest.raises(mtgp.MTGPError):
        g.fill(9, 10)
