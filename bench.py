#!/usr/bin/env python3
"""Benchmark: batched MTGP32-11213 generation on one MI355X.

Workload (the batched configuration the metric is quoted on):
  * all 128 parameter sets of the MTGP32 MEXP=11213 table, one generator
    each, seeded by mtgp32_init with seed_g = low 32 bits of the (g+1)-th
    output of splitmix64 started at 42;
  * generator g emits L_g uint32 (genrand_uint32) followed by L_g floats in
    [0,1) (genrand_close_open), L_g = floor(2^U_g), U_g = 20 + 6 u_g with
    u_g the (g+1)-th output of splitmix64 started at 7, scaled to [0,1);
  * each stream is split into segments; one batched jump-ahead launch
    computes the segment start states, then ONE persistent generation launch
    writes every output.

A step is one launch of that plan: the jump launch plus the generation
launch (inputs -- states, parameter tables, work list -- resident in HBM).  value = output bytes of all steps / time of the
timed steps.  Bit-exactness is checked on samples of every step's output
against the CPU oracle, and the CPU baseline runs the reference library
(compiled from its own sources) on the host cores.
"""
import argparse
import ctypes as C
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
METRIC = "generated GB/s (% HBM peak), bit-exact; p50/p99 latency batched vs plain"
MASK64 = (1 << 64) - 1


def splitmix64(state, count):
    out = []
    for _ in range(count):
        state = (state + 0x9E3779B97F4A7C15) & MASK64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        out.append(z ^ (z >> 31))
    return out


def batched_workload(n_gen=128):
    seeds = [v & 0xFFFFFFFF for v in splitmix64(42, n_gen)]
    us = [(v >> 11) * 2.0**-53 for v in splitmix64(7, n_gen)]
    lengths = [int(2.0 ** (20 + 6 * u)) for u in us]
    return seeds, lengths


def serving_workload(n_requests=6000, rate=10000.0, n_gen=256, log2_mean=16.0, log2_sd=2.5, log2_min=10,
                     log2_max=26, jump_frac=0.1, seed=1):
    """The serving configuration: Poisson arrivals at `rate` per second over
    n_gen device generators -- half MTGP32-11213, a quarter MTGP32-23209, a
    quarter MTGP64-23209 (sets 0, 1, ... of each table), seeded with
    splitmix64 started at 2024 -- each request asking for uint32, float in
    [1,2) or double in [0,1) with equal probability (doubles from the MTGP64
    generators, the others from the MTGP32 ones, uniformly), 2^x numbers
    with x ~ Normal(log2_mean, log2_sd) clipped to [log2_min, log2_max],
    and with probability jump_frac a skip of 2^u numbers, u ~ U[10, 30],
    before its output.  Returns (specs, seeds, trace)."""
    n32a, n32b = n_gen // 2, n_gen // 4
    n64 = n_gen - n32a - n32b
    specs = ([(32, 11213, i % 128) for i in range(n32a)] + [(32, 23209, i % 128) for i in range(n32b)]
             + [(64, 23209, i % 128) for i in range(n64)])
    seeds = splitmix64(2024, n_gen)
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1e6 / rate, n_requests))
    dtype = rng.integers(0, 3, n_requests)
    gen = np.where(dtype < 2, rng.integers(0, n32a + n32b, n_requests), n32a + n32b + rng.integers(0, n64, n_requests))
    kind = np.array([0, 1, 2])[dtype]
    x = np.clip(rng.normal(log2_mean, log2_sd, n_requests), log2_min, log2_max)
    count = np.round(2.0 ** x).astype(np.uint64)
    skip = np.where(rng.random(n_requests) < jump_frac, np.floor(2.0 ** rng.uniform(10, 30, n_requests)), 0)
    trace = {"t_us": t, "gen": gen.astype(np.int32), "kind": kind.astype(np.int32), "count": count,
             "skip": skip.astype(np.uint64)}
    return specs, seeds, trace


def run_serving(M, n_requests, rate, **server_args):
    """Replays one serving trace through the batched server and through the
    plain path (one launch per request, arrival order), from the same
    initial states; returns latency percentiles, throughput and whether the
    two produced identical outputs (digest of every request, in a second
    replay of each)."""
    specs, seeds, tr = serving_workload(n_requests=n_requests, rate=rate)
    out = {}
    digests = {}
    for mode in ("batched", "plain"):
        for check in (False, True):
            gs = M.GenSet(specs)
            gs.init_seed(seeds)
            srv = M.Server(gs, plain=(mode == "plain"), **server_args)
            queue = np.zeros(len(tr["t_us"]))
            lat, dg, wall = srv.replay(tr["t_us"], tr["gen"], tr["kind"], tr["count"], tr["skip"], digests=check,
                                       queue=queue)
            st = srv.stats()
            srv.close()
            if check:
                digests[mode] = dg
                continue
            nbytes = float(sum(int(c) * (8 if k == 2 else 4) for c, k in zip(tr["count"], tr["kind"])))
            out[mode] = {"p50_ms": round(float(np.percentile(lat, 50)) / 1e3, 4),
                         "p99_ms": round(float(np.percentile(lat, 99)) / 1e3, 4),
                         "mean_ms": round(float(lat.mean()) / 1e3, 4),
                         "queue_p50_ms": round(float(np.percentile(queue, 50)) / 1e3, 4),
                         "queue_p99_ms": round(float(np.percentile(queue, 99)) / 1e3, 4),
                         "gbs": round(nbytes / (wall * 1e-6) / 1e9, 3), "wall_s": round(wall * 1e-6, 4),
                         "batches": int(st["batches"]), "max_batch_requests": int(st["max_batch_requests"]),
                         "plan_us_per_batch": round(st["plan_us"] / max(1, st["batches"]), 1)}
    out["bit_exact_equal"] = bool((digests["batched"] == digests["plain"]).all())
    out["requests"] = n_requests
    out["rate_per_s"] = rate
    out["offered_gbs"] = round(float(sum(int(c) * (8 if k == 2 else 4) for c, k in zip(tr["count"], tr["kind"])))
                               / (tr["t_us"][-1] * 1e-6) / 1e9, 3)
    return out


def run_serving_sweep(M, rates, seconds=0.5):
    """The batched server alone on the serving mix at increasing arrival
    rates (`seconds` of arrivals each): offered and served GB/s, p50 / p99
    latency and host planning per batch, up to saturation."""
    out = []
    for rate in rates:
        n = max(200, int(rate * seconds))
        specs, seeds, tr = serving_workload(n_requests=n, rate=rate)
        gs = M.GenSet(specs)
        gs.init_seed(seeds)
        srv = M.Server(gs)
        lat, _, wall = srv.replay(tr["t_us"], tr["gen"], tr["kind"], tr["count"], tr["skip"])
        st = srv.stats()
        srv.close()
        nbytes = float(sum(int(c) * (8 if k == 2 else 4) for c, k in zip(tr["count"], tr["kind"])))
        out.append({"rate_per_s": rate, "requests": n,
                    "offered_gbs": round(nbytes / (tr["t_us"][-1] * 1e-6) / 1e9, 2),
                    "served_gbs": round(nbytes / (wall * 1e-6) / 1e9, 2),
                    "p50_ms": round(float(np.percentile(lat, 50)) / 1e3, 3),
                    "p99_ms": round(float(np.percentile(lat, 99)) / 1e3, 3),
                    "batches": int(st["batches"]),
                    "plan_us_per_batch": round(st["plan_us"] / max(1, st["batches"]), 1)})
    return out


def run_gpu(args):
    import mtgp_amd as M
    seeds, lengths = batched_workload()
    G = len(seeds)
    cus, hbm = M.device_info()
    total_numbers = 2 * sum(lengths)
    out_bytes = total_numbers * 4
    buf = M.DeviceBuffer(out_bytes)
    gs = M.GenSet([(32, 11213, g) for g in range(G)])
    gs.init_seed(seeds)
    reqs, off = [], 0
    for g in range(G):
        reqs.append((g, M.KIND_INT, lengths[g], 0, buf.address(off)))
        off += lengths[g] * 4
        reqs.append((g, M.KIND_CLOSE_OPEN, lengths[g], 0, buf.address(off)))
        off += lengths[g] * 4
    t0 = time.time()
    plan = gs.plan(reqs, segment=args.segment)
    stats = plan.stats()
    stream = M.Stream()
    for _ in range(args.warmup):
        plan.launch(stream)
    stream.synchronize()
    M.synchronize()
    # Whole job, one step = plan.launch: the jump launch for the cells'
    # start states, then the persistent generation launch (or, with the
    # plan's two-phase order switched on, the large cells' jumps and
    # generation beside the small cells' on a side stream).
    evw = [M.Event() for _ in range(args.steps + 1)]
    evw[0].record(stream)
    for k in range(args.steps):
        plan.launch(stream)
        evw[k + 1].record(stream)
    stream.synchronize()
    M.synchronize()
    per = [evw[k].elapsed_ms(evw[k + 1]) for k in range(args.steps)]
    total_ms = evw[0].elapsed_ms(evw[-1])
    # Kernels (roofline): the same plan's work in serial order -- one jump
    # launch, then one generation launch over every cell -- so that each
    # kernel's duration is bracketed by events on the stream it runs on.
    ev = [M.Event() for _ in range(2 * args.steps + 1)]
    ev[0].record(stream)
    for k in range(args.steps):
        plan.launch_part(0, stream)
        ev[2 * k + 1].record(stream)
        plan.launch_part(1, stream)
        ev[2 * k + 2].record(stream)
    stream.synchronize()
    M.synchronize()
    jump_ms = [ev[2 * k].elapsed_ms(ev[2 * k + 1]) for k in range(args.steps)]
    gen_ms = [ev[2 * k + 1].elapsed_ms(ev[2 * k + 2]) for k in range(args.steps)]
    serial_ms = ev[0].elapsed_ms(ev[-1])
    # Bit-exact spot check of the last step's output against the oracle.
    exact = spot_check(M, buf, seeds, lengths) if not args.no_check else None
    kernels = plan.launches()
    sustained = None
    if args.cycles > 0:
        sustained = run_sustained(M, gs, plan, buf, reqs, seeds, lengths, stream, args.cycles, not args.no_check)
    return dict(G=G, cus=cus, hbm=hbm, out_bytes=out_bytes, numbers=total_numbers, per_ms=per,
                jump_ms=jump_ms, gen_ms=gen_ms, total_ms=total_ms, serial_ms=serial_ms, stats=stats,
                plan_s=time.time() - t0, exact=exact, kernels=kernels, sustained=sustained)


def run_sustained(M, gs, plan_a, buf_a, reqs_a, seeds, lengths, stream, cycles, check):
    """Fresh batches back to back: two plans of the workload's shape over two
    output buffers (double buffering).  Cycle k rebases plan k % 2 onto the
    states cycle k-1 committed, launches it and commits, all enqueued on one
    stream while earlier cycles run (no synchronisation between cycles), so
    every cycle writes the NEXT L_g uint32 + L_g floats of every stream.
    The first cycle continues from the workload's initial states."""
    out_bytes = sum(8 * L for L in lengths)
    buf_b = M.DeviceBuffer(out_bytes)
    delta = buf_b.address(0) - buf_a.address(0)
    reqs_b = [(g, k, n, sk, out + delta) for g, k, n, sk, out in reqs_a]
    t = time.perf_counter()
    plan_b = gs.plan(reqs_b)
    fresh_ms = (time.perf_counter() - t) * 1e3
    plans = [plan_a, plan_b]
    stream.synchronize()
    ev0, ev1 = M.Event(), M.Event()
    host = []
    # `warm` untimed cycles first (fresh batches too): the GPU idled while
    # the second plan was built, and its clock settles over the first
    # launches after that
    warm = 4
    for k in range(warm + cycles):
        if k == warm:
            ev0.record(stream)
        p = plans[k % 2]
        t = time.perf_counter()
        if k:
            p.rebase(stream)
        p.launch(stream)
        p.commit()
        if k >= warm:
            host.append((time.perf_counter() - t) * 1e3)
    ev1.record(stream)
    stream.synchronize()
    ms = ev0.elapsed_ms(ev1)
    # host work of one cycle alone (the loop above also waits, two cycles
    # behind, for the GPU to free a plan's staging copy)
    work = []
    for k in range(4):
        p = plans[(warm + cycles + k) % 2]
        stream.synchronize()
        t = time.perf_counter()
        p.rebase(stream)
        p.launch(stream)
        p.commit()
        work.append((time.perf_counter() - t) * 1e3)
    stream.synchronize()
    cycles_run = warm + cycles + 4
    exact = None
    if check:
        # the last cycle's batch (the cycles-th batch of every stream) in
        # the buffer it went to, against the oracle jumped over the earlier
        # batches
        exact = spot_check(M, buf_a if (cycles_run - 1) % 2 == 0 else buf_b, seeds, lengths, batch=cycles_run - 1)
    return {"cycles": cycles, "gbs": round(out_bytes * cycles / (ms * 1e-3) / 1e9, 2),
            "ms_per_cycle": round(ms / cycles, 4),
            "host_ms_per_cycle": round(float(np.mean(host[1:])) if cycles > 1 else host[0], 4),
            "host_work_ms_per_cycle": round(float(np.median(work)), 4),
            "fresh_plan_ms": round(fresh_ms, 1), "bit_exact": exact,
            "how": "cycle k: plan k%2 rebased onto the states cycle k-1 committed (its work items re-pointed, "
                   "no polynomial work), launched, committed; enqueued back to back on one stream over two "
                   "output buffers, timed after 4 such cycles untimed; host_ms_per_cycle includes waiting for "
                   "the GPU (the host runs at most two "
                   "cycles ahead), host_work_ms_per_cycle is one cycle's host calls with the GPU idle; "
                   "fresh_plan_ms is a second mtgp_plan_create of the same shape; bit_exact checks the last "
                   "batch written against the oracle jumped over the earlier batches"}


def _time_plan(M, plan, stream, reps=5, warmup=2):
    """(jump_ms, generate_ms) of a plan: mean over `reps` launches after
    `warmup` ones, HIP events on the stream the kernels run on."""
    for _ in range(warmup):
        plan.launch(stream)
    stream.synchronize()
    ev = [M.Event() for _ in range(3)]
    j = g = 0.0
    for _ in range(reps):
        ev[0].record(stream)
        plan.launch_part(0, stream)
        ev[1].record(stream)
        plan.launch_part(1, stream)
        ev[2].record(stream)
        stream.synchronize()
        j += ev[0].elapsed_ms(ev[1])
        g += ev[1].elapsed_ms(ev[2])
    return j / reps, g / reps


def run_configs(M):
    """The other configurations, timed like the main one (start-state jump
    launch + generation launch per step, inputs resident): 1 (one stream),
    3 (MTGP64 batches per MEXP) and 4 (4096 substreams), plus batches of the
    larger-state MTGP32 classes and of MTGP64 dynamic-creator sets (the
    128-word-step kernels)."""
    stream = M.Stream()
    out = {}

    def line(nbytes, jump_ms, gen_ms, items, **extra):
        d = {"output_bytes": nbytes, "jump_ms": round(jump_ms, 4), "generate_ms": round(gen_ms, 4),
             "ms": round(jump_ms + gen_ms, 4), "gbs": round(nbytes / ((jump_ms + gen_ms) * 1e-3) / 1e9, 2),
             "generate_gbs": round(nbytes / (gen_ms * 1e-3) / 1e9, 2), "work_items": items}
        d.update(extra)
        return d
    # 1: set 0 of MTGP32-11213, mtgp32_init(1234), 2^24 uint32 then 2^24 [1,2) floats
    n = 1 << 24
    gs = M.GenSet([(32, 11213, 0)])
    gs.init_seed([1234])
    buf = M.DeviceBuffer(8 * n)
    plan = gs.plan([(0, M.KIND_INT, n, 0, buf.address(0)), (0, M.KIND_CLOSE1_OPEN2, n, 0, buf.address(4 * n))])
    out["1"] = line(8 * n, *_time_plan(M, plan, stream), plan.stats()["items"],
                    workload="MTGP32-11213 set 0, seed 1234: 2^24 uint32 then 2^24 float [1,2), one stream")
    del plan, buf, gs
    # 3: MTGP64 sets 0..63 per MEXP, init_by_array with 4-word splitmix64 keys,
    # 2^24 uint64 then 2^24 double [0,1) each
    for mexp in (23209, 44497):
        G = 64
        gs = M.GenSet([(64, mexp, i) for i in range(G)])
        gs.init_array([splitmix64(1000 * mexp + i, 4) for i in range(G)])
        buf = M.DeviceBuffer(G * 16 * n)
        reqs = []
        for i in range(G):
            reqs.append((i, M.KIND_INT, n, 0, buf.address(16 * n * i)))
            reqs.append((i, M.KIND_CLOSE_OPEN, n, 0, buf.address(16 * n * i + 8 * n)))
        plan = gs.plan(reqs)
        out["3-%d" % mexp] = line(G * 16 * n, *_time_plan(M, plan, stream), plan.stats()["items"],
                                  workload="MTGP64-%d sets 0..63: 2^24 uint64 then 2^24 double [0,1) each" % mexp)
        del plan, buf, gs
    # The larger-state MTGP32 classes: sets 0..63 of 23209 and 44497, seeds
    # 1..64, 2^24 uint32 then 2^24 float [0,1) each
    for mexp in (23209, 44497):
        G = 64
        gs = M.GenSet([(32, mexp, i) for i in range(G)])
        gs.init_seed(list(range(1, G + 1)))
        buf = M.DeviceBuffer(G * 8 * n)
        reqs = []
        for i in range(G):
            reqs.append((i, M.KIND_INT, n, 0, buf.address(8 * n * i)))
            reqs.append((i, M.KIND_CLOSE_OPEN, n, 0, buf.address(8 * n * i + 4 * n)))
        plan = gs.plan(reqs)
        out["32-%d" % mexp] = line(G * 8 * n, *_time_plan(M, plan, stream), plan.stats()["items"],
                                   workload="MTGP32-%d sets 0..63: 2^24 uint32 then 2^24 float [0,1) each" % mexp)
        del plan, buf, gs
    # MTGP64 dynamic-creator sets 0..63 (N = 176: the 128-word-step class),
    # seeds 1..64, 2^24 uint64 then 2^24 double [0,1) each
    G = 64
    gs = M.GenSet([M.dc_params(64, i) for i in range(G)])
    gs.init_seed(list(range(1, G + 1)))
    buf = M.DeviceBuffer(G * 16 * n)
    reqs = []
    for i in range(G):
        reqs.append((i, M.KIND_INT, n, 0, buf.address(16 * n * i)))
        reqs.append((i, M.KIND_CLOSE_OPEN, n, 0, buf.address(16 * n * i + 8 * n)))
    plan = gs.plan(reqs)
    out["dc64"] = line(G * 16 * n, *_time_plan(M, plan, stream), plan.stats()["items"],
                       workload="MTGP64-DC-11213 sets 0..63 (128-word steps): 2^24 uint64 then 2^24 double "
                                "[0,1) each")
    del plan, buf, gs
    # 4: 4096 substreams i * 3^162 of set 0, seed 777, 2^20 uint32 each
    S, count = 4096, 1 << 20
    base = M.GenSet([(32, 11213, 0)])
    base.init_seed([777])
    w0 = base.state(0)
    gs = M.GenSet([(32, 11213, 0)] * S)
    js = M.jump_hex(32, 11213, 0, 3 ** 162)
    buf = M.DeviceBuffer(S * 4 * count)
    M.synchronize()
    t = time.perf_counter()
    gs.substreams(w0, js)  # first call: level polynomials and the work list built, then run
    M.synchronize()
    first_ms = (time.perf_counter() - t) * 1e3
    plan = gs.plan([(i, M.KIND_INT, count, 0, buf.address(4 * count * i)) for i in range(S)])
    setups = []
    for _ in range(3):
        M.synchronize()
        t = time.perf_counter()
        gs.substreams(w0, js)
        M.synchronize()
        setups.append((time.perf_counter() - t) * 1e3)
    setup_ms = float(np.median(setups))
    # the plan was built on the same slots: its cells start from the windows
    # substreams just wrote
    jump_ms, gen_ms = _time_plan(M, plan, stream)
    d = line(S * 4 * count, jump_ms, gen_ms, plan.stats()["items"], substream_setup_ms=round(setup_ms, 3),
             substream_setup_first_ms=round(first_ms, 3),
             workload="4096 substreams i*3^162 of MTGP32-11213 set 0 (seed 777), 2^20 uint32 each, end to end: "
                      "substream start states made on the device as one matrix-core group (the 4095 states "
                      "J^(i 3^162) w0 as fp4 products of the coefficient rows with the base window's bit planes; "
                      "host time included, wall clock, median of 3 calls after the first, which also builds the "
                      "coefficient rows: substream_setup_first_ms) + generation")
    d["ms"] = round(setup_ms + jump_ms + gen_ms, 4)
    d["gbs"] = round(S * 4 * count / (d["ms"] * 1e-3) / 1e9, 2)
    out["4"] = d
    return out


def jump_latency(M, reps=20):
    """Latency of one 3^162-step jump of MTGP32-11213 set 0 (the jump the
    reference's test-jump32.cpp times): the drop-in mtgp32_fast_jump (host
    state in and out, GPU jump), the library's x^(3^162) mod P (the NTL
    calc_jump replacement), and the reference library's mtgp32_fast_jump on
    one host core."""
    js = M.jump_hex(32, 11213, 0, 3 ** 162)
    g = M.Generator32(11213, 0, seed=1234)
    g.fast_jump(js)  # warm-up
    t = time.perf_counter()
    for _ in range(reps):
        g.fast_jump(js)
    drop_in = (time.perf_counter() - t) * 1e3 / reps
    t = time.perf_counter()
    for _ in range(reps):
        M.jump_hex(32, 11213, 0, 3 ** 162 + 1)
    poly = (time.perf_counter() - t) * 1e3 / reps
    out = {"fast_jump_ms": round(drop_in, 4), "jump_poly_ms": round(poly, 4)}
    noalloc = getattr(M._lib, "mtgp32_fast_jump_noalloc", None)
    if noalloc is not None:
        noalloc.restype = C.c_int
        noalloc.argtypes = [C.POINTER(M.Fast32), C.c_char_p]
        noalloc(C.byref(g.g), js.encode())
        t = time.perf_counter()
        for _ in range(reps):
            noalloc(C.byref(g.g), js.encode())
        out["fast_jump_noalloc_ms"] = round((time.perf_counter() - t) * 1e3 / reps, 4)
    # mtgp32_fill, the drop-in's bulk call (window up, a plan over the GPU,
    # window back, synchronous): per-call time for a host array of 2^10 and
    # 2^20 numbers and a device buffer of 2^20 and 2^24, after one call of
    # the shape (which plans; the calls timed relaunch that plan)
    fills = {}
    for n, dev in ((1 << 10, False), (1 << 20, False), (1 << 20, True), (1 << 24, True)):
        dst = M.DeviceBuffer(4 * n) if dev else np.empty(n, np.uint32)
        g.fill(0, n, dst)
        k = 5 if n >= (1 << 24) else reps
        t = time.perf_counter()
        for _ in range(k):
            g.fill(0, n, dst)
        fills["%s_2^%d_ms" % ("device" if dev else "host", n.bit_length() - 1)] = round(
            (time.perf_counter() - t) * 1e3 / k, 4)
    out["fill_ms"] = fills
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import RefLib
    if RefLib.available("fast", 32):
        r = RefLib("fast", 32).new_seed(11213, 0, 1234)
        r.jump(js)
        t = time.perf_counter()
        for _ in range(reps):
            r.jump(js)
        out["reference_cpu_fast_jump_ms"] = round((time.perf_counter() - t) * 1e3 / reps, 4)
        r.fill(0, 1 << 20)
        t = time.perf_counter()
        for _ in range(5):
            r.fill(0, 1 << 20)
        out["reference_cpu_genrand_2^20_ms"] = round((time.perf_counter() - t) * 1e3 / 5, 4)
    out["what"] = ("one jump by 3^162 steps of MTGP32-11213 set 0 (2803-character string): fast_jump = the "
                   "drop-in on a host mtgp32_fast_t (window up, the jump split over 11 waves by chunk range and "
                   "their partial windows XORed by a second launch, window back, synchronous); "
                   "jump_poly = x^(3^162) mod P on the host; reference = the reference library on one core; "
                   "fill_ms = one mtgp32_fill of 2^k uint32 into a host array or a device buffer (window up, "
                   "launches of the plan built by the first call of that shape, window back, synchronous); "
                   "reference_cpu_genrand = the same count of "
                   "mtgp32_genrand_uint32 on one core")
    return out


def spot_check(M, buf, seeds, lengths, n=1 << 16, batch=0):
    """Heads and tails of a few generators' outputs in `buf` against the
    oracle; `batch` > 0: the batch-th batch of each stream (the oracle jumps
    over the earlier ones)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle
    o = Oracle()
    off = 0
    ok = True
    for g in range(len(seeds)):
        L = lengths[g]
        if g in (0, 1, len(seeds) - 1) or L == max(lengths):
            p = o.params(32, 11213, g)
            w = o.init_state(p, seeds[g])
            if batch:
                P = o.hex_to_poly(M.charpoly_hex(32, 11213, g), 11213 // 64 + 2)
                w = o.jump(p, w, o.poly_to_hex(o.xpow_mod(2 * L * batch, P)))
            want, w = o.generate(p, w, 0, min(n, L))
            got = buf.to_host(np.uint32, min(n, L), off)
            ok &= bool((got == want).all())
            # the tail of the float part (the last segment of the stream)
            _, w = o.generate(p, w, 0, L - min(n, L))
            _, w = o.generate(p, w, 2, L - min(n, L))
            want, _ = o.generate(p, w, 2, min(n, L))
            got = buf.to_host(np.float32, min(n, L), off + 4 * L + 4 * (L - min(n, L)))
            ok &= bool((got.view(np.uint32) == want.view(np.uint32)).all())
        off += 8 * L
    return ok


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_run(threads, seconds, chunk):
    """Numbers generated by `threads` host threads in `seconds`: thread t
    runs generator t of the workload (MTGP32-11213 set t, its seed) through
    the reference library (oracle/_ref, compiled from the reference's own
    sources; the C restatement in oracle/ when that build is absent),
    alternating chunks of uint32 and close_open floats into fresh arrays."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle, RefLib
    seeds, _ = batched_workload()
    kind = "reference" if RefLib.available("fast", 32) else "port"
    done = [0] * threads
    if kind == "reference":
        ref = RefLib("fast", 32)
        gens = [ref.new_seed(11213, g, seeds[g]) for g in range(threads)]

        def step(g, k):
            gens[g].fill(0 if k % 2 == 0 else 2, chunk)
    else:
        o = Oracle()
        ps = [o.params(32, 11213, g) for g in range(threads)]
        ws = [o.init_state(ps[g], seeds[g]) for g in range(threads)]

        def step(g, k):
            _, ws[g] = o.generate(ps[g], ws[g], 0 if k % 2 == 0 else 2, chunk)
    deadline = time.time() + seconds

    def work(g):
        k = 0
        while time.time() < deadline:
            step(g, k)
            done[g] += chunk
            k += 1
    th = [threading.Thread(target=work, args=(g,)) for g in range(threads)]
    t = time.time()
    for x in th:
        x.start()
    for x in th:
        x.join()
    return kind, sum(done), time.time() - t


def cpu_baseline(seconds=8.0, chunk=1 << 22):
    """The reference CPU path on the host cores, on a time-bounded sample:
    one core for `seconds`, then one thread per core (at most 16) for
    `seconds`.  `value` is the all-core rate."""
    cores = max(1, min(16, os.cpu_count() or 1))
    _, n1, t1 = _cpu_run(1, seconds, chunk)
    kind, n, dt = _cpu_run(cores, seconds, chunk)
    return {"value": round(4 * n / dt / 1e9, 3), "unit": "GB/s", "cores": cores, "kind": kind,
            "single_core_gbs": round(4 * n1 / t1 / 1e9, 3), "nproc": os.cpu_count(), "cpu": _cpu_model(),
            "sample": "MTGP32-11213 sets 0..%d of the workload with their seeds, one host thread each, "
                      "alternating 2^22 uint32 / 2^22 close_open chunks: %.3g numbers in %.1f s wall on %d "
                      "threads (single core: %.3g numbers in %.1f s)" % (cores - 1, n, dt, cores, n1, t1)}


def _threads_for(threads, seconds, make, step):
    """Bytes produced by `threads` host threads in `seconds`: thread t runs
    step(state) -> bytes on its own state make(t) until the deadline."""
    states = [make(t) for t in range(threads)]
    done = [0] * threads
    deadline = time.time() + seconds

    def work(t):
        while time.time() < deadline:
            done[t] += step(states[t])
    th = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
    t0 = time.time()
    for x in th:
        x.start()
    for x in th:
        x.join()
    return sum(done), time.time() - t0


def cpu_config_baselines(seconds=3.0):
    """The reference library (oracle/_ref, built from the reference's own
    sources) on the host cores, in the shape of each other configuration:
    bounded samples of `seconds` on one core and on one thread per core (at
    most 16); GB/s of output bytes.  None when the reference build is
    absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import RefLib
    if not (RefLib.available("fast", 32) and RefLib.available("fast", 64)):
        return None
    cores = max(1, min(16, os.cpu_count() or 1))
    out = {"cores": cores, "kind": "reference", "seconds_per_sample": seconds}
    r32, r64 = RefLib("fast", 32), RefLib("fast", 64)
    # 1: the one stream (a single generator: one core)
    n = 1 << 24
    g = r32.new_seed(11213, 0, 1234)
    t = time.perf_counter()
    g.fill(0, n)
    g.fill(1, n)
    dt = time.perf_counter() - t
    out["1"] = {"gbs": round(8 * n / dt / 1e9, 3), "cores": 1,
                "sample": "the whole workload: 2^24 uint32 + 2^24 close1_open2 of set 0, seed 1234, one core"}
    chunk = 1 << 22

    def alt(bits, kind2):
        def step(st):
            st[0].fill(0 if st[1] % 2 == 0 else kind2, chunk)
            st[1] += 1
            return chunk * bits // 8
        return step
    # 3: MTGP64 sets with init_by_array keys, uint64 / close_open double chunks
    for mexp in (23209, 44497):
        def make(t, mexp=mexp):
            return [r64.new_array(mexp, t, splitmix64(1000 * mexp + t, 4)), 0]
        b1, t1 = _threads_for(1, seconds, make, alt(64, 2))
        b, tt = _threads_for(cores, seconds, make, alt(64, 2))
        out["3-%d" % mexp] = {"gbs": round(b / tt / 1e9, 3), "single_core_gbs": round(b1 / t1 / 1e9, 3),
                              "sample": "sets 0..%d, one thread each, alternating 2^22 uint64 / 2^22 close_open "
                                        "double chunks" % (cores - 1)}
    # 4: substreams: one 3^162 jump (mtgp32_fast_jump) then 2^20 uint32 each
    js = None
    try:
        import mtgp_amd as M
        js = M.jump_hex(32, 11213, 0, 3 ** 162)
    except Exception:
        pass
    if js:
        def make4(t):
            return r32.new_seed(11213, 0, 777 + t)

        def step4(h):
            h.jump(js)
            h.fill(0, 1 << 20)
            return 4 << 20
        b1, t1 = _threads_for(1, seconds, make4, step4)
        b, tt = _threads_for(cores, seconds, make4, step4)
        out["4"] = {"gbs": round(b / tt / 1e9, 3), "single_core_gbs": round(b1 / t1 / 1e9, 3),
                    "sample": "per substream one mtgp32_fast_jump by 3^162 steps then 2^20 uint32 (the reference's "
                              "substream recipe), one generator per thread"}
    return out


def measured_traffic(path=os.path.join(ROOT, "profiles", "mtgp32_items_batched11213_pmc.txt")):
    """HBM bytes per generation launch from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE and WRITE_SIZE passes, each scaled
    by its copy-kernel calibration factor); None when the summary is absent."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1].split()[0])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    # the generation launch's clock settles over its first ~8 launches (the
    # board's power management), so the default warm-up covers them
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--segment", type=int, default=0, help="numbers per segment (0 = auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--only-serving", action="store_true",
                    help="run only the serving comparison and print its JSON (for profiling it alone)")
    ap.add_argument("--serving-rate", type=float, default=20000.0, help="serving requests per second")
    ap.add_argument("--serving-requests", type=int, default=12000)
    ap.add_argument("--serving-sweep", type=str, default="",
                    help="comma-separated arrival rates: print only the batched server's sweep over them")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--configs", action="store_true",
                    help="time the other configurations (1, 3, 4) and print their JSON instead")
    ap.add_argument("--no-configs", action="store_true", help="leave the other configurations out of the line")
    ap.add_argument("--cycles", type=int, default=10,
                    help="sustained fresh batches: rebase -> launch -> commit cycles (0 = none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every generator of the workload fits one MI355X and the "
              "path has no cross-GPU step; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.configs:
        import mtgp_amd as M
        print(json.dumps({"configs": run_configs(M)}), flush=True)
        return
    if args.serving_sweep:
        import mtgp_amd as M
        rates = [float(x) for x in args.serving_sweep.split(",")]
        print(json.dumps({"serving_sweep": run_serving_sweep(M, rates)}), flush=True)
        return
    if args.only_serving:
        import mtgp_amd as M
        print(json.dumps({"serving": run_serving(M, args.serving_requests, args.serving_rate)}), flush=True)
        return
    r = run_gpu(args)
    gbs = r["out_bytes"] * args.steps / (r["total_ms"] * 1e-3) / 1e9
    # Dominant kernel: the generation launch (serial region: one launch over
    # every cell).  Algorithmic bytes are the output bytes (4 per uint32 /
    # float); start windows and tables, under 0.01%, are left out.
    gen_avg = sum(r["gen_ms"]) / len(r["gen_ms"])
    alg_bytes = r["out_bytes"]
    kern_gbs = alg_bytes / (gen_avg * 1e-3) / 1e9
    # the plan's generation launches (one per word size and state class;
    # this workload has one)
    kernel = max(r["kernels"], key=lambda k: k["items"])["kernel"] if r["kernels"] else None
    traffic = measured_traffic()
    line = {
        "metric": METRIC,
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["total_ms"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "bit_exact": r["exact"],
        "numbers_per_s": round(r["numbers"] * args.steps / (r["total_ms"] * 1e-3), 1),
        "pct_hbm_peak": round(100 * gbs / HBM_PEAK_GBS, 2),
        "config": {"workload": "batched MTGP32-11213: 128 parameter sets, uneven lengths 2^20..2^26 "
                               "uint32 + close_open float32 each; per step one jump launch (cell start states) + "
                               "one persistent generation launch",
                   "generators": r["G"], "output_bytes": r["out_bytes"], "numbers": r["numbers"],
                   "work_items": r["stats"]["items"], "launches_per_step": r["stats"]["launches"],
                   "plan_ms": round(r["stats"]["plan_ms"], 1), "compute_units": r["cus"],
                   "serial_ms_per_step": round(r["serial_ms"] / args.steps, 4),
                   "jump_ms": round(sum(r["jump_ms"]) / len(r["jump_ms"]), 4),
                   "generate_ms": round(gen_avg, 4)},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": round(kern_gbs, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(kern_gbs / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes": alg_bytes, "traffic": traffic,
                     "traffic_source": ("rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this workload, calibrated "
                                        "against a plain copy kernel: profiles/mtgp32_items_batched11213_pmc.txt")
                     if traffic else None},
    }
    if r["sustained"]:
        line["sustained"] = r["sustained"]
    if not args.no_serving:
        import mtgp_amd as M
        sv = run_serving(M, args.serving_requests, args.serving_rate)
        sv["workload"] = ("Poisson arrivals over 256 device generators (128 MTGP32-11213, 64 MTGP32-23209, "
                          "64 MTGP64-23209); uint32 / float [1,2) / double [0,1) equally likely; 2^x numbers, "
                          "x ~ N(16, 2.5) clipped to [10, 26]; 10% skip 2^U numbers first, U ~ U[10, 30]; "
                          "4 GiB output budget")
        # batched against plain at a rate the plain path sustains
        sv["sustainable"] = run_serving(M, 3000, 5000.0)
        # the batched server alone up to saturation
        sv["sweep"] = run_serving_sweep(M, [10000.0, 20000.0, 40000.0, 80000.0, 160000.0, 320000.0])
        line["serving"] = sv
    import mtgp_amd as M
    line["jump_latency"] = jump_latency(M)
    if not args.no_configs:
        line["configs"] = run_configs(M)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
        if not args.no_configs:
            line["cpu_baseline"]["configs"] = cpu_config_baselines()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
