#!/usr/bin/env python3
"""Benchmark: batched MTGP32-11213 generation on one MI355X.

Workload (the batched configuration the metric is quoted on):
  * all 128 parameter sets of the MTGP32 MEXP=11213 table, one generator
    each, seeded by mtgp32_init with seed_g = low 32 bits of the (g+1)-th
    output of splitmix64 started at 42;
  * generator g emits L_g uint32 (genrand_uint32) followed by L_g floats in
    [0,1) (genrand_close_open), L_g = floor(2^U_g), U_g = 20 + 6 u_g with
    u_g the (g+1)-th output of splitmix64 started at 7, scaled to [0,1);
  * all of it in ONE persistent kernel launch per step (the plan splits each
    stream into segments whose start states are reached by device jump).

A step is one launch of that plan (inputs -- states, parameter tables, work
list -- resident in HBM).  value = output bytes of all steps / time of the
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
    ev = [M.Event() for _ in range(args.steps + 1)]
    ev[0].record(stream)
    for k in range(args.steps):
        plan.launch(stream)
        ev[k + 1].record(stream)
    stream.synchronize()
    M.synchronize()
    per = [ev[k].elapsed_ms(ev[k + 1]) for k in range(args.steps)]
    total_ms = ev[0].elapsed_ms(ev[-1])
    # Bit-exact spot check of the last step's output against the oracle.
    exact = spot_check(M, buf, seeds, lengths) if not args.no_check else None
    return dict(G=G, cus=cus, hbm=hbm, out_bytes=out_bytes, numbers=total_numbers, per_ms=per,
                total_ms=total_ms, stats=stats, plan_s=time.time() - t0, exact=exact)


def spot_check(M, buf, seeds, lengths, n=1 << 16):
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


def cpu_baseline(seconds_target=15.0):
    """The reference library (oracle/_ref, built from the reference sources)
    on the host cores: one thread per generator of the same workload, each
    producing 2^24 uint32 then 2^24 close_open floats."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle, RefLib
    seeds, _ = batched_workload()
    threads = max(1, min(16, os.cpu_count() or 1))
    per = 1 << 24
    kind = "reference" if RefLib.available("fast", 32) else "port"
    if kind == "reference":
        ref = RefLib("fast", 32)
        gens = [ref.new_seed(11213, g, seeds[g]) for g in range(threads)]

        def work(g):
            gens[g].fill(0, per)
            gens[g].fill(2, per)
    else:
        o = Oracle()
        ps = [o.params(32, 11213, g) for g in range(threads)]
        ws = [o.init_state(ps[g], seeds[g]) for g in range(threads)]

        def work(g):
            _, w = o.generate(ps[g], ws[g], 0, per)
            o.generate(ps[g], w, 2, per)
    th = [threading.Thread(target=work, args=(g,)) for g in range(threads)]
    t = time.time()
    for x in th:
        x.start()
    for x in th:
        x.join()
    dt = time.time() - t
    nbytes = threads * 2 * per * 4
    return {"value": round(nbytes / dt / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": kind,
            "sample": "%d generators (MTGP32-11213 sets 0..%d, same seeds) x (2^24 uint32 + 2^24 close_open), "
                      "one host thread each, %.1f s wall" % (threads, threads - 1, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--segment", type=int, default=0, help="numbers per segment (0 = auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every generator of the workload fits one MI355X and the "
              "path has no cross-GPU step; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    r = run_gpu(args)
    avg_ms = sum(r["per_ms"]) / len(r["per_ms"])
    gbs = r["out_bytes"] * args.steps / (r["total_ms"] * 1e-3) / 1e9
    kern_gbs = r["out_bytes"] / (avg_ms * 1e-3) / 1e9
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
                               "uint32 + close_open float32 each, one persistent launch",
                   "generators": r["G"], "output_bytes": r["out_bytes"], "numbers": r["numbers"],
                   "work_items": r["stats"]["items"], "launches_per_step": r["stats"]["launches"],
                   "plan_ms": round(r["stats"]["plan_ms"], 1), "compute_units": r["cus"]},
        "roofline": {"bound": "hbm", "achieved": round(kern_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(kern_gbs / HBM_PEAK_GBS, 4), "traffic": None},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
