"""Start-state jumps on the matrix cores: plans whose MTGP32 generators
have many jumped cells (jump_mfma_min) give every number and final window
bit-exactly as the oracle -- every MTGP32 state class, one and two row tiles
per wave, both plan phases, rebased plans, graph replays and one long
stream cut into ~1000 cells -- and the same bytes as the batched jump
kernel."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(mtgp, oracle, specs, counts, seeds):
    """One request per generator (kinds cycling), its oracle output and the
    oracle's window after it."""
    gs = mtgp.GenSet(specs)
    gs.init_seed(seeds)
    reqs, want, bufs, ends = [], [], [], []
    for g, ((bits, mexp, idx), count) in enumerate(zip(specs, counts)):
        p = oracle.params(bits, mexp, idx)
        w = oracle.init_state(p, seeds[g] if bits == 64 else seeds[g] & 0xFFFFFFFF)
        kind = (0, 2, 1, 3)[g % 4]
        e, w = oracle.generate(p, w, kind, count)
        buf = mtgp.DeviceBuffer(e.nbytes)
        reqs.append((g, kind, count, 0, buf))
        want.append(e)
        bufs.append(buf)
        ends.append(w)
    return gs, reqs, want, bufs, ends


def _check(bufs, want, what):
    for k, (b, e) in enumerate(zip(bufs, want)):
        got = b.to_host(e.dtype, e.size)
        bad = np.nonzero(got.view(np.uint8) != e.view(np.uint8))[0]
        assert bad.size == 0, "%s: request %d, %d bytes differ from byte %d" % (what, k, bad.size, bad[0])


# Cells of 16384 numbers: 81 cells (two row tiles, then one), 30 (one), 64
# (two), 100 (two and two) jumped cells per generator; an MTGP64 generator
# keeps the batched jump kernel.
SPECS = [(32, 11213, 3), (32, 23209, 7), (32, 44497, 1), (64, 23209, 2), (32, 11213, 90)]
COUNTS = [81 * 16384, 30 * 16384 + 5, 64 * 16384 - 7, 40 * 16384, 100 * 16384]
SEEDS = [101, 202, 303, 404, 505]


@pytest.mark.parametrize("mfma_min", [1, 24])
def test_matrix_core_jumps_match_oracle(mtgp, oracle, mfma_min):
    print("matrix-core jumps, jump_mfma_min %d" % mfma_min, flush=True)
    gs, reqs, want, bufs, ends = _case(mtgp, oracle, SPECS, COUNTS, SEEDS)
    with mtgp.tuning(jump_mfma_min=mfma_min):
        plan = gs.plan(reqs, segment=16384)
    st = plan.stats()
    assert st["matrix_groups"] == 4 and st["matrix_jumps"] == 80 + 30 + 63 + 99, st
    plan.launch()
    mtgp.synchronize()
    _check(bufs, want, "jump_mfma_min %d" % mfma_min)
    plan.commit()
    for g, w in enumerate(ends):
        assert (gs.state(g) == w).all(), "window of generator %d" % g


def test_matrix_core_jumps_equal_batched_kernel(mtgp, oracle):
    """The same plan with the matrix cores off: identical bytes."""
    print("matrix-core jumps against the batched jump kernel", flush=True)
    out = []
    for mfma_min in (1, 0):
        gs, reqs, want, bufs, _ = _case(mtgp, oracle, SPECS[:3], [m // 2 for m in COUNTS[:3]], SEEDS[:3])
        with mtgp.tuning(jump_mfma_min=mfma_min):
            plan = gs.plan(reqs, segment=8192)
        assert (plan.stats()["matrix_jumps"] > 0) == (mfma_min > 0)
        plan.launch()
        mtgp.synchronize()
        out.append([b.to_host(np.uint8, e.nbytes) for b, e in zip(bufs, want)])
        _check(bufs, want, "jump_mfma_min %d" % mfma_min)
    for a, b in zip(*out):
        assert (a == b).all()


def test_matrix_core_jumps_rebase_and_graph(mtgp, oracle):
    """A plan with matrix-core groups rebased twice onto its committed
    states (the groups read the new base windows), eagerly and replayed
    from a captured graph."""
    print("matrix-core jumps: rebase and graph", flush=True)
    specs, counts, seeds = SPECS[:3], [40 * 8192, 33 * 8192, 25 * 8192], SEEDS[:3]
    gs = mtgp.GenSet(specs)
    gs.init_seed(seeds)
    st = []
    for (bits, mexp, idx), s in zip(specs, seeds):
        p = oracle.params(bits, mexp, idx)
        st.append([p, oracle.init_state(p, s)])
    bufs = [mtgp.DeviceBuffer(c * 4) for c in counts]
    with mtgp.tuning(jump_mfma_min=1):
        plan = gs.plan([(g, 0, c, 0, b) for g, (c, b) in enumerate(zip(counts, bufs))], segment=8192)
    assert plan.stats()["matrix_groups"] == 3
    stream = mtgp.Stream()
    for graph in (False, True):
        if graph:
            plan.capture(stream)
        for batch in range(2):
            want = []
            for g, c in enumerate(counts):
                e, st[g][1] = oracle.generate(st[g][0], st[g][1], 0, c)
                want.append(e)
            if graph:
                plan.launch_graph(stream)
            else:
                plan.launch(stream)
            plan.commit()
            stream.synchronize()
            _check(bufs, want, "graph %s batch %d" % (graph, batch))
            plan.rebase(stream)


def test_matrix_core_jumps_two_phase(mtgp, oracle):
    """The automatic two-size grid with matrix-core groups (each holding a
    generator's jumps of both phases, run with phase 0), launched in two
    phases (side stream) and in serial order."""
    print("matrix-core jumps: two-phase launch", flush=True)
    specs = [(32, 11213, 0), (32, 23209, 1), (64, 23209, 2), (32, 11213, 3)]
    gs, reqs, want, bufs, _ = _case(mtgp, oracle, specs, [600000] * 4, [10, 20, 30, 40])
    with mtgp.tuning(two_size_rho=1e9, jump_mfma_min=1):
        plan = gs.plan(reqs)
    assert plan.stats()["matrix_groups"] >= 3
    assert plan.set_overlap(True)
    stream = mtgp.Stream()
    plan.launch(stream)
    stream.synchronize()
    _check(bufs, want, "two-phase")
    for b, e in zip(bufs, want):
        b.from_host(np.full(e.nbytes, 0x5A, np.uint8))
    plan.launch_part(0, stream)
    plan.launch_part(1, stream)
    stream.synchronize()
    _check(bufs, want, "serial")


def test_matrix_core_jumps_one_long_stream(mtgp, oracle):
    """One MTGP32-11213 stream of 2^24 numbers in cells of 16384: 1023
    jumps from one base window (16 row blocks, the matrix cores by default)
    against the oracle."""
    print("matrix-core jumps: one stream, 1023 jumps", flush=True)
    gs, reqs, want, bufs, ends = _case(mtgp, oracle, [(32, 11213, 5)], [1 << 24], [4242])
    plan = gs.plan(reqs, segment=16384)
    assert plan.stats()["matrix_jumps"] == 1023
    plan.launch()
    mtgp.synchronize()
    _check(bufs, want, "one stream")
    plan.commit()
    assert (gs.state(0) == ends[0]).all()
