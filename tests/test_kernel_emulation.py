"""The library's kernels.hip, compiled as host C++ and run on host threads
(tests/emu), checked bit-exactly against the oracle.  This exercises the
kernel logic -- step layouts, heads and tails, LDS relocation, the device
jump, the self-resetting work queue -- without a GPU."""
import numpy as np
import pytest

import pyemu

SETS = [(32, 11213), (32, 23209), (32, 44497), (64, 23209), (64, 44497), (64, 110503)]


@pytest.fixture(scope="module")
def emu():
    return pyemu.load()


def _bits(a):
    return a.view(np.uint32 if a.dtype.itemsize == 4 else np.uint64)


def _run(emu, bits, items, layout, grid=None):
    arr = np.zeros(len(items), pyemu.ITEM)
    for k, it in enumerate(items):
        for f, v in it.items():
            if f != "n":
                arr[k][f] = v
    cls = max(pyemu.kernel_class(bits, it["n"]) for it in items)
    rc = emu.emu_run_items(bits, cls, layout, arr.ctypes.data, len(items), grid or len(items))
    assert rc == 0, "emulated launch failed (%d)" % rc


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("bits,mexp", SETS)
def test_items_all_kinds(emu, oracle, bits, mexp, layout):
    """One launch, one item per kind with different counts (covering the
    head, whole steps, LDS relocation and the tail), unaligned outputs,
    fewer workgroups than items so that the queue hands out work."""
    p = oracle.params(bits, mexp, 5)
    wd = np.uint32 if bits == 32 else np.uint64
    eb = bits // 8
    w0 = oracle.init_state(p, 2024)
    counts = [(0, 1), (1, 259), (2, 1030), (3, 0), (4, 777), (0, 2049)]
    items, keep, want = [], [], []
    for k, (kind, n) in enumerate(counts):
        raw = np.zeros((n + 16) * eb + 16, np.uint8)
        off = (4 * k) % 16 if bits == 32 else (8 * k) % 16
        sin, sout = w0.copy(), np.zeros_like(w0)
        base = raw.ctypes.data + (-raw.ctypes.data) % 16 + off
        items.append(dict(out=base, count=n, state_in=sin.ctypes.data, state_out=sout.ctypes.data,
                          param=pyemu.set_id(bits, mexp, 5), kind=kind, n=p.n))
        keep.append((raw, base, sin, sout, n))
        e, w = oracle.generate(p, w0, kind, n)
        want.append((e, w))
    _run(emu, bits, items, layout, grid=2)
    for (raw, base, sin, sout, n), (e, w) in zip(keep, want):
        start = base - raw.ctypes.data
        got = raw[start:start + n * eb].view(e.dtype)
        assert (_bits(got) == _bits(e)).all()
        assert (raw[start + n * eb:] == 0).all(), "write past the requested outputs"
        assert (sout == w).all()


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("bits,mexp", SETS[:4])
def test_item_jump(emu, oracle, bits, mexp, layout):
    """Item with a jump polynomial: the window is replaced by its jump
    before generating; compared with the oracle's jump."""
    p = oracle.params(bits, mexp, 9)
    w0 = oracle.init_state(p, 31337)
    words = (mexp + 31) // 32
    rng = np.random.default_rng(mexp)
    poly = rng.integers(0, 2**32, words, dtype=np.uint64).astype(np.uint32)
    poly[-1] &= (1 << (mexp % 32)) - 1 if mexp % 32 else 0xFFFFFFFF
    # hex nibbles low first (polytostring format), from the bit vector
    bitsv = np.unpackbits(poly.view(np.uint8), bitorder="little")
    nib = [int(bitsv[4 * i]) | int(bitsv[4 * i + 1]) << 1 | int(bitsv[4 * i + 2]) << 2 | int(bitsv[4 * i + 3]) << 3
           for i in range(len(bitsv) // 4)]
    hexs = "".join("0123456789abcdef"[v] for v in nib).rstrip("0")
    wj = oracle.jump(p, w0, hexs)
    n = 600
    eb = bits // 8
    out = np.zeros(n + 8, np.uint32 if bits == 32 else np.uint64)
    sin, sout = w0.copy(), np.zeros_like(w0)
    _run(emu, bits, [dict(out=out.ctypes.data, count=n, state_in=sin.ctypes.data, state_out=sout.ctypes.data,
                          poly=poly.ctypes.data, poly_words=words, param=pyemu.set_id(bits, mexp, 9), kind=0,
                          n=p.n)], layout)
    e, w = oracle.generate(p, wj, 0, n)
    assert (out[:n] == e).all()
    assert (out[n:] == 0).all()
    assert (sout == w).all()


def _hex_of(coeffs, top):
    bits = np.zeros(top + 1, np.uint8)
    bits[list(coeffs)] = 1
    bits = np.concatenate([bits, np.zeros((-len(bits)) % 4, np.uint8)])
    nib = bits[0::4] | bits[1::4] << 1 | bits[2::4] << 2 | bits[3::4] << 3
    return "".join("0123456789abcdef"[v] for v in nib).rstrip("0")


@pytest.mark.parametrize("bits,mexp", SETS[:5])
def test_batched_jumps(emu, oracle, bits, mexp):
    """The batched jump kernel: several items of 1..4 jumps from one base
    window, each compared with the oracle's literal jump -- including a
    polynomial of degree above mexp (the reference applies every
    coefficient of the string)."""
    p = oracle.params(bits, mexp, 11)
    wd = np.uint32 if bits == 32 else np.uint64
    base = oracle.init_state(p, 99)
    rng = np.random.default_rng(mexp + bits)
    jobs = []
    for nb in (1, 3, 4):
        cl = []
        for b in range(nb):
            top = mexp - 1 if b != 2 else mexp + 700
            cl.append(sorted(set(rng.integers(0, top + 1, size=top // 2).tolist()) | {top}))
        jobs.append(cl)
    # each job runs as `parts` items over disjoint chunk ranges; their
    # partial windows XOR to the jump (the generation kernel's Item.parts)
    items, keep = [], []
    for k, cl in enumerate(jobs):
        masks, nchunks = pyemu.jump_masks(cl)
        parts = min(k + 1, nchunks)
        outs = np.zeros((parts, len(cl), p.n), wd)
        keep.append((masks, outs))
        for q in range(parts):
            items.append(dict(state_in=base.ctypes.data, masks=masks.ctypes.data,
                              states_out=outs[q].ctypes.data, nb=len(cl), c0=nchunks * q // parts,
                              c1=nchunks * (q + 1) // parts, param=pyemu.set_id(bits, mexp, 11),
                              state_bytes=p.n * (bits // 8)))
    arr = np.zeros(len(items), pyemu.JUMP_ITEM)
    for k, it in enumerate(items):
        for f, v in it.items():
            arr[k][f] = v
    rc = emu.emu_run_jumps(bits, pyemu.kernel_class(bits, p.n), arr.ctypes.data, len(items), 2)
    assert rc == 0
    for (masks, outs), cl in zip(keep, jobs):
        got = np.bitwise_xor.reduce(outs, axis=0)
        for b, cs in enumerate(cl):
            want = oracle.jump(p, base, _hex_of(cs, cs[-1]))
            assert (got[b] == want).all(), "jump %d of %d" % (b, len(cl))


@pytest.mark.parametrize("bits,mexp", [(32, 11213), (64, 23209)])
def test_batched_jump_from_split_base(emu, oracle, bits, mexp):
    """JumpItem.in_parts: the base window is the XOR of partial windows (a
    split jump's result), as the substream doubling chains them."""
    p = oracle.params(bits, mexp, 6)
    w = oracle.init_state(p, 4242)
    rng = np.random.default_rng(7)
    wd = np.uint32 if bits == 32 else np.uint64
    parts = rng.integers(0, np.iinfo(wd).max, (3, p.n), dtype=wd)
    parts[2] = w ^ parts[0] ^ parts[1]
    cs = sorted(set(rng.integers(0, mexp, size=mexp // 2).tolist()) | {mexp - 1})
    masks, nchunks = pyemu.jump_masks([cs])
    out = np.zeros(p.n, wd)
    arr = np.zeros(1, pyemu.JUMP_ITEM)
    for f, v in dict(state_in=parts.ctypes.data, masks=masks.ctypes.data, states_out=out.ctypes.data, nb=1, c0=0,
                     c1=nchunks, param=pyemu.set_id(bits, mexp, 6), state_bytes=p.n * (bits // 8), in_parts=3,
                     in_part_bytes=p.n * (bits // 8)).items():
        arr[0][f] = v
    assert emu.emu_run_jumps(bits, pyemu.kernel_class(bits, p.n), arr.ctypes.data, 1, 1) == 0
    assert (out == oracle.jump(p, w, _hex_of(cs, cs[-1]))).all()


@pytest.mark.parametrize("bits,mexp", [(32, 11213), (64, 23209)])
def test_item_start_window_from_parts(emu, oracle, bits, mexp):
    """Item.parts: the start window is the XOR of partial windows."""
    p = oracle.params(bits, mexp, 4)
    w = oracle.init_state(p, 8)
    rng = np.random.default_rng(3)
    wd = np.uint32 if bits == 32 else np.uint64
    parts = rng.integers(0, np.iinfo(wd).max, (3, p.n), dtype=wd)
    parts[2] = w ^ parts[0] ^ parts[1]
    n = 777
    out = np.zeros(n, wd)
    sout = np.zeros_like(w)
    _run(emu, bits, [dict(out=out.ctypes.data, count=n, state_in=parts.ctypes.data, state_out=sout.ctypes.data,
                          param=pyemu.set_id(bits, mexp, 4), kind=0, parts=3, part_bytes=p.n * bits // 8,
                          n=p.n)], 1)
    e, we = oracle.generate(p, w, 0, n)
    assert (out == e).all() and (sout == we).all()


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("bits,mexp", [(32, 11213), (32, 44497), (64, 23209), (64, 110503)])
def test_item_segment_chain(emu, oracle, bits, mexp, layout):
    """One item running a chain of segments: consecutive pieces of the
    stream with different kinds, unaligned outputs and numbers stepped over
    without output (out == 0), continuing one window from segment to
    segment, then the final window."""
    p = oracle.params(bits, mexp, 9)
    eb = bits // 8
    w0 = oracle.init_state(p, 77)
    plan = [(0, 3), (2, 300), (None, 1000), (1, 257), (None, 5), (4, 2100), (3, 1), (0, 0), (2, 513)]
    segs = np.zeros(len(plan), pyemu.SEGMENT)
    keep, want, w = [], [], w0
    for k, (kind, n) in enumerate(plan):
        if kind is None:
            segs[k] = (0, n, 0, 0, 0)
            _, w = oracle.generate(p, w, 0, n)
            continue
        raw = np.zeros((n + 40) * eb + 256, np.uint8)
        base = raw.ctypes.data + (-raw.ctypes.data) % 128 + eb * (5 * k % 16)
        segs[k] = (base, n, kind, 0, 0)
        e, w = oracle.generate(p, w, kind, n)
        keep.append((raw, base, n, e))
    sin, sout = w0.copy(), np.zeros_like(w0)
    item = dict(state_in=sin.ctypes.data, state_out=sout.ctypes.data, param=pyemu.set_id(bits, mexp, 9),
                segs=segs.ctypes.data, nsegs=len(plan), n=p.n)
    _run(emu, bits, [item], layout)
    for raw, base, n, e in keep:
        start = base - raw.ctypes.data
        got = raw[start:start + n * eb].view(e.dtype)
        assert (_bits(got) == _bits(e)).all()
        assert (raw[start + n * eb:] == 0).all() and (raw[:start] == 0).all()
    assert (sout == w).all()


@pytest.mark.parametrize("bits,mexp", SETS)
def test_init_kernel(emu, oracle, bits, mexp):
    """init_kernels.hip (device init_state / init_by_array), emulated, for
    key lengths below, at and above the state size, against the oracle."""
    p = oracle.params(bits, mexp, 3)
    n = p.n
    tbl = list(p.tbl)
    hidden = (tbl[4] ^ (tbl[8] << 16)) & ((1 << bits) - 1)
    wd = np.uint32 if bits == 32 else np.uint64
    rng = np.random.default_rng(mexp)
    cases = [("seed", 12345), ("seed", 2**31 + 7), ("array", 1), ("array", n - 1), ("array", n + 5)]
    items = np.zeros(len(cases), pyemu.INIT_ITEM)
    outs, keys, want = [], [], []
    for k, (mode, v) in enumerate(cases):
        out = np.zeros(n, wd)
        outs.append(out)
        if mode == "seed":
            items[k] = (out.ctypes.data, v, hidden, n, bits, 0, 0, 0)
            want.append(oracle.init_state(p, v))
        else:
            key = rng.integers(0, 2**63, v, dtype=np.uint64)
            keys.append(key)
            items[k] = (out.ctypes.data, key.ctypes.data, hidden, n, bits, v, 1, 0)
            want.append(oracle.init_by_array(p, key.astype(wd)))
    assert emu.emu_run_init(items.ctypes.data, len(cases)) == 0
    for k, (out, w) in enumerate(zip(outs, want)):
        assert (out == w).all(), cases[k]


@pytest.mark.parametrize("idx", [0, 172, 199])
def test_narrow_class_items_and_jumps(emu, oracle, idx):
    """The 128-word-step kernels (class 3: MTGP64 dynamic-creator sets, N =
    176, N - pos down to 130) on the reference's own records: a chain of
    segments of every kind with unaligned outputs and stepped-over gaps, and
    a batched jump with parts, against the oracle."""
    import json
    import os
    from pyoracle import Oracle
    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "dc_params_digests.json")))
    # records for 0 and 199 are in the fixture; 172 (pos 3) comes from the
    # library's table, itself pinned to the reference digests (test_dc).
    if "64/%d" % idx in gold["records"]:
        rec = gold["records"]["64/%d" % idx]
    else:
        import mtgp_amd
        from test_dc import _record
        rec = _record(mtgp_amd.dc_params(64, idx))
    p = Oracle.params_from_record(64, rec)
    dp = pyemu.devparam(64, rec)
    assert dp["n"][0] == 176 and dp["n"][0] - dp["pos"][0] >= 128
    w0 = oracle.init_state(p, 11)
    plan = [(0, 5), (2, 200), (None, 333), (1, 129), (4, 1000), (3, 1)]
    segs = np.zeros(len(plan), pyemu.SEGMENT)
    keep, w = [], w0
    for k, (kind, n) in enumerate(plan):
        if kind is None:
            segs[k] = (0, n, 0, 0, 0)
            _, w = oracle.generate(p, w, 0, n)
            continue
        raw = np.zeros((n + 40) * 8 + 256, np.uint8)
        base = raw.ctypes.data + (-raw.ctypes.data) % 128 + 8 * (3 * k % 16)
        segs[k] = (base, n, kind, 0, 0)
        e, w = oracle.generate(p, w, kind, n)
        keep.append((raw, base, n, e))
    sin, sout = w0.copy(), np.zeros_like(w0)
    items = np.zeros(1, pyemu.ITEM)
    items[0]["state_in"], items[0]["state_out"] = sin.ctypes.data, sout.ctypes.data
    items[0]["segs"], items[0]["nsegs"], items[0]["param"] = segs.ctypes.data, len(plan), 0
    assert emu.emu_run_items_p(64, 3, 1, items.ctypes.data, 1, 1, dp.ctypes.data) == 0
    for raw, base, n, e in keep:
        start = base - raw.ctypes.data
        assert (raw[start:start + n * 8].view(np.uint64) == e.view(np.uint64)).all()
        assert (raw[start + n * 8:] == 0).all() and (raw[:start] == 0).all()
    assert (sout == w).all()
    # batched jump, two jumps split in two parts
    words = 11213 // 64 + 2
    out, _ = oracle.generate(p, oracle.init_state(p, 0), 0, 2 * 11213)
    P, _ = oracle.bm((out & 1).astype(np.uint8), words)
    polys = [oracle.xpow_mod(s, P) for s in (1000, 77777)]
    coeffs = [[k for k in range(64 * len(q)) if int(q[k // 64]) >> (k % 64) & 1] for q in polys]
    masks, nch = pyemu.jump_masks(coeffs)
    jout = np.zeros((2, 2, 176), np.uint64)
    jitems = np.zeros(2, pyemu.JUMP_ITEM)
    for q in range(2):
        jitems[q] = (w0.ctypes.data, masks.ctypes.data, jout[q].ctypes.data, 2, nch * q // 2, nch * (q + 1) // 2,
                     0, 176 * 8, 0, 0, 0)
    assert emu.emu_run_jumps_p(64, 3, jitems.ctypes.data, 2, 2, dp.ctypes.data) == 0
    for b, q in enumerate(polys):
        want = oracle.jump(p, w0, oracle.poly_to_hex(q))
        got = jout[0, b] ^ jout[1, b]
        assert (got[1:] == want[1:]).all() and (got[0] & p.mask) == (want[0] & p.mask)


@pytest.mark.parametrize("bits,mexp,layout", [(32, 11213, 2), (32, 23209, 2), (32, 44497, 2), (64, 23209, 2),
                                              (64, 44497, 2), (64, 110503, 2), (32, 23209, 3), (32, 44497, 3),
                                              (64, 44497, 3), (64, 110503, 3)])
def test_multi_wave_cells(emu, oracle, bits, mexp, layout):
    """Layout 2 (two waves per cell; 512-word steps, 256-word steps of two
    words per lane for MEXP 11213 / MTGP64-23209) and layout 3 (four waves,
    1024-word steps): every kind with counts around the step size,
    unaligned outputs, a start window from partial windows, a segment chain
    with numbers stepped over, and the queue with fewer workgroups than
    items."""
    p = oracle.params(bits, mexp, 3)
    step = {2: 256 if p.n <= 363 else 512, 3: 512 if p.n <= 726 else 1024}[layout]
    assert p.n - p.pos >= step
    eb = bits // 8
    wd = np.uint32 if bits == 32 else np.uint64
    w0 = oracle.init_state(p, 99)
    counts = [(0, 1), (1, 511), (2, 513), (3, 1030), (4, 3000), (0, 0), (2, 4099)]
    items, keep, want = [], [], []
    for k, (kind, n) in enumerate(counts):
        raw = np.zeros((n + 16) * eb + 128, np.uint8)
        base = raw.ctypes.data + (-raw.ctypes.data) % 128 + eb * (7 * k % 16)
        sin, sout = w0.copy(), np.zeros_like(w0)
        items.append(dict(out=base, count=n, state_in=sin.ctypes.data, state_out=sout.ctypes.data,
                          param=pyemu.set_id(bits, mexp, 3), kind=kind, n=p.n))
        keep.append((raw, base, sin, sout, n))
        want.append(oracle.generate(p, w0, kind, n))
    # start window as the XOR of three partial windows
    rng = np.random.default_rng(mexp)
    parts = rng.integers(0, np.iinfo(wd).max, (3, p.n), dtype=wd)
    parts[2] = w0 ^ parts[0] ^ parts[1]
    plan = [(0, 5), (None, 700), (2, 1500), (None, 3), (1, 513)]
    segs = np.zeros(len(plan), pyemu.SEGMENT)
    chain, w = [], w0
    for k, (kind, n) in enumerate(plan):
        if kind is None:
            segs[k] = (0, n, 0, 0, 0)
            _, w = oracle.generate(p, w, 0, n)
            continue
        raw = np.zeros((n + 40) * eb + 256, np.uint8)
        base = raw.ctypes.data + (-raw.ctypes.data) % 128 + eb * (5 * k % 16)
        segs[k] = (base, n, kind, 0, 0)
        e, w = oracle.generate(p, w, kind, n)
        chain.append((raw, base, n, e))
    cout = np.zeros_like(w0)
    items.append(dict(state_in=parts.ctypes.data, state_out=cout.ctypes.data, param=pyemu.set_id(bits, mexp, 3),
                      segs=segs.ctypes.data, nsegs=len(plan), parts=3, part_bytes=p.n * eb, n=p.n))
    _run(emu, bits, items, layout, grid=3)
    for (raw, base, sin, sout, n), (e, we) in zip(keep, want):
        start = base - raw.ctypes.data
        got = raw[start:start + n * eb].view(e.dtype)
        assert (_bits(got) == _bits(e)).all()
        assert (raw[start + n * eb:] == 0).all() and (raw[:start] == 0).all(), "write outside the outputs"
        assert (sout == we).all()
    for raw, base, n, e in chain:
        start = base - raw.ctypes.data
        got = raw[start:start + n * eb].view(e.dtype)
        assert (_bits(got) == _bits(e)).all()
        assert (raw[start + n * eb:] == 0).all() and (raw[:start] == 0).all()
    assert (cout == w).all()


def _coef_words(cs, wc):
    w = np.zeros(wc, np.uint32)
    for k in cs:
        w[k // 32] |= np.uint32(1 << (k % 32))
    return w


def _mjump(emu, p, mexp, idx, base, jobs, blocks, words, in_parts=1, part_bytes=0):
    """Runs the matrix-core jump of the coefficient lists `jobs` from `base`
    (in_parts windows) for the row blocks `blocks` ((r0, row tiles)) and the
    window words starting at each of `words`; returns the jumped windows
    (words not covered stay zero)."""
    nj = len(jobs)
    kpad = pyemu.mj_kpad(mexp)
    wc, njp = kpad // 32, (nj + 31) // 32 * 32
    coef = np.zeros((wc, njp), np.uint32)
    for b, cs in enumerate(jobs):
        coef[:, b] = _coef_words(cs, wc)
    pw = pyemu.mj_pw(kpad, p.n)
    planes = np.zeros(32 * pw, np.uint32)
    out = np.zeros((nj, p.n), np.uint32)
    g = np.zeros(1, pyemu.MJUMP_GROUP)
    for f, v in dict(state_in=base.ctypes.data, in_part_bytes=part_bytes, planes=planes.ctypes.data,
                     coef=coef.ctypes.data, out=out.ctypes.data, out_stride=p.n * 4, nj=nj, njp=njp, kpad=kpad,
                     pw=pw, param=pyemu.set_id(32, mexp, idx), in_parts=in_parts, n=p.n).items():
        g[0][f] = v
    for r0, rt in blocks:
        items = np.zeros(len(words), pyemu.MJUMP_ITEM)
        items["r0"] = r0
        items["j0"] = words
        assert emu.emu_run_mjumps(g.ctypes.data, 1, items.ctypes.data, len(items), rt) == 0
    return out, planes


def _covered(words, rt, n):
    return [(j0, min(j0 + pyemu.mj_words(rt), n)) for j0 in words]


@pytest.mark.parametrize("mexp,idx,nj,blocks,words,parts", [
    (11213, 5, 40, [(0, 2)], [0, 28, 29, 348], 1),
    (23209, 1, 9, [(0, 1)], [0, 722], 2),
    (11213, 90, 70, [(0, 2), (64, 1)], [0, 101], 1),
])
def test_matrix_core_jumps(emu, oracle, mexp, idx, nj, blocks, words, parts):
    """The matrix-core jump (bit planes, fp4 products, parities): every
    jump of a group at the window words its items cover, against the
    oracle's literal jump -- identity and one-step polynomials, dense
    random ones, row blocks of 1 and 2 tiles, windows starting inside a
    plane word, the window's last words, and a base given as partial
    windows."""
    p = oracle.params(32, mexp, idx)
    base = oracle.init_state(p, 77 + idx)
    rng = np.random.default_rng(mexp + nj)
    jobs = [[0], [1], [mexp - 1]]
    while len(jobs) < nj:
        jobs.append(sorted(set(rng.integers(0, mexp, size=mexp // 2).tolist())))
    jobs = jobs[:nj]
    src = base
    if parts > 1:
        src = rng.integers(0, 2**32, (parts, p.n), dtype=np.uint32)
        src[-1] = base ^ np.bitwise_xor.reduce(src[:-1], axis=0)
    out, _ = _mjump(emu, p, mexp, idx, src, jobs, blocks, words, in_parts=parts, part_bytes=p.n * 4)
    for b, cs in enumerate(jobs):
        want = oracle.jump(p, base, _hex_of(cs, max(cs)))
        rt = next(t for r0, t in blocks if r0 <= b < r0 + 32 * t)
        covered = np.zeros(p.n, bool)
        for j0, j1 in _covered(words, rt, p.n):
            assert (out[b, j0:j1] == want[j0:j1]).all(), "jump %d words %d..%d" % (b, j0, j1)
            covered[j0:j1] = True
        assert not out[b, ~covered].any(), "jump %d: words outside its items written" % b
