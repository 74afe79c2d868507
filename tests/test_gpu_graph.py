"""Plans captured in a hipGraph and replayed, and the autotuner."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plan_case(mtgp, oracle, count=60000):
    specs = [(32, 11213, 0), (32, 23209, 1), (64, 23209, 2), (32, 11213, 3)]
    gs = mtgp.GenSet(specs)
    seeds = [10, 20, 30, 40]
    gs.init_seed(seeds)
    reqs, want, bufs = [], [], []
    for g, (bits, mexp, idx) in enumerate(specs):
        p = oracle.params(bits, mexp, idx)
        w = oracle.init_state(p, seeds[g])
        for kind, skip in ((0, 0), (2, 777), (1, 0)):
            _, w = oracle.generate(p, w, 0, skip) if skip else (None, w)
            e, w = oracle.generate(p, w, kind, count)
            buf = mtgp.DeviceBuffer(e.nbytes)
            reqs.append((g, kind, count, skip, buf))
            want.append(e)
            bufs.append(buf)
    return gs, reqs, want, bufs


def _check(bufs, want, what):
    for k, (b, e) in enumerate(zip(bufs, want)):
        got = b.to_host(e.dtype, e.size)
        assert (got.view(np.uint8) == e.view(np.uint8)).all(), "%s: request %d" % (what, k)


def test_graph_replays_match_eager(mtgp, oracle):
    """A captured plan (jump launch + generation launch) replayed three
    times, with eager work and synchronisations between replays and the
    outputs scribbled over before each replay; every replay must equal the
    oracle (and so the eager launch)."""
    gs, reqs, want, bufs = _plan_case(mtgp, oracle)
    plan = gs.plan(reqs, segment=8192)
    assert plan.stats()["items"] > 20
    stream = mtgp.Stream()
    plan.launch(stream)
    stream.synchronize()
    _check(bufs, want, "eager")
    plan.capture(stream)
    junk = np.full(max(e.nbytes for e in want), 0x5A, np.uint8)
    other = mtgp.Generator32(11213, 9, seed=3)
    for rep in range(3):
        for b, e in zip(bufs, want):
            b.from_host(junk[:e.nbytes])
        other.fill(0, 50000)  # eager work between replays
        mtgp.synchronize()
        plan.launch_graph(stream)
        stream.synchronize()
        _check(bufs, want, "replay %d" % rep)


def test_autotune_keeps_results_exact(mtgp, oracle):
    """The autotuner tries every (layout, workgroups per CU) on the plan's
    own buffers and keeps the fastest; the outputs after it are exact."""
    gs, reqs, want, bufs = _plan_case(mtgp, oracle, count=200000)
    plan = gs.plan(reqs, segment=16384)
    layout, bpc, ms = plan.autotune(reps=1)
    assert layout in (0, 1, 2, 3) and 1 <= bpc <= 32 and ms > 0
    plan.launch()
    mtgp.synchronize()
    _check(bufs, want, "after autotune layout %d bpc %d" % (layout, bpc))


def test_two_phase_launch_serial_and_graph(mtgp, oracle):
    """A plan on the automatic two-size grid launches in two phases (the
    large cells' jumps, then their generation beside the small cells'
    jumps and generation on a side stream).  Its outputs equal the oracle
    eagerly, in serial order (launch_part) and replayed from a graph
    captured across the two streams."""
    gs, reqs, want, bufs = _plan_case(mtgp, oracle, count=200000)
    # automatic two-size grid, kept whatever the jump cost of this small case
    with mtgp.tuning(two_size_rho=1e9):
        plan = gs.plan(reqs)
    assert plan.set_overlap(True)
    st = plan.stats()
    # three (word size, class) groups: 3 jump + 3 generation launches
    # merged; more when the small cells form a second phase
    assert st["launches"] > 6, st
    junk = np.full(max(e.nbytes for e in want), 0x5A, np.uint8)

    def scribble():
        for b, e in zip(bufs, want):
            b.from_host(junk[:e.nbytes])
        mtgp.synchronize()

    stream = mtgp.Stream()
    scribble()
    plan.launch(stream)
    stream.synchronize()
    _check(bufs, want, "two-phase eager")
    scribble()
    plan.launch_part(0, stream)
    plan.launch_part(1, stream)
    stream.synchronize()
    _check(bufs, want, "serial")
    plan.capture(stream)
    other = mtgp.Generator32(11213, 9, seed=3)
    for rep in range(3):
        scribble()
        other.fill(0, 50000)
        mtgp.synchronize()
        plan.launch_graph(stream)
        stream.synchronize()
        _check(bufs, want, "two-phase replay %d" % rep)
    # captured before any eager launch: the side stream is made outside
    # the capture
    with mtgp.tuning(two_size_rho=1e9):
        fresh = gs.plan(reqs)
    assert fresh.set_overlap(True)
    fresh.capture(stream)
    scribble()
    fresh.launch_graph(stream)
    stream.synchronize()
    _check(bufs, want, "two-phase graph captured first")


def test_tuning_knobs_and_jump_cost_grid(mtgp, oracle):
    """Tuning knobs round-trip and reject bad input; a plan whose jumps are
    dear next to its cells takes the jump-cost rule's grid (the grid that
    two_size_rho = 0 forces, not the two-size grid that a huge
    two_size_rho forces), and both grids give the oracle's numbers.  (Cell
    counts depend on the default workgroups per CU, which an earlier
    autotune in this process may have changed, so the grids are compared
    with each other, not with fixed numbers.)"""
    assert mtgp.tuning_get("two_size_rho") > 0
    with pytest.raises(mtgp.MTGPError):
        mtgp.tuning_set("no_such_knob", 1)
    with pytest.raises(mtgp.MTGPError):
        mtgp.tuning_set("segment_split", 1.5)
    # 24 M numbers: large enough that the grids differ (cells of at least
    # 4096 numbers)
    gs, reqs, want, bufs = _plan_case(mtgp, oracle, count=2000000)
    with mtgp.tuning(two_size_rho=1e9):
        two = gs.plan(reqs)
    with mtgp.tuning(two_size_rho=0):
        rule = gs.plan(reqs)
    cheap = gs.plan(reqs)
    assert cheap.stats()["items"] == rule.stats()["items"] != two.stats()["items"]
    for plan in (two, cheap):
        for b, e in zip(bufs, want):
            b.from_host(np.full(e.nbytes, 0x5A, np.uint8))
        plan.launch()
        mtgp.synchronize()
        _check(bufs, want, "grid with %d items" % plan.stats()["items"])
