#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the reference itself.

Everything here comes from the reference library compiled from its own
sources (oracle/_ref, built by `make -C oracle ref`) or from data files the
reference ships (openCL-sample/mtgp32-jump-string.h, mtgp32-jump-table.cpp,
cuda-sample/mtgp32-char11213.h).  No output of this repository's code is
used.  Run in the container that has the reference tree:

    make -C oracle ref && python3 tests/golden/make_golden.py

Files written (all small, loaded with numpy allow_pickle=False or json):
  params_digests.json   SHA-256 of every full parameter record (768 sets)
                        and two full records per table
  main_outputs.npz      the reference's own test-program sequences
                        (mtgp*-fast.c MAIN: init(1), init_by_array(1,2,3,4),
                        init_by_str("\\1\\2\\3\\4") + the float family)
  jump_steps.npz        test-jump32.cpp pattern: outputs after stepping s
  reference_jump_data.npz  CHAR32_11213, JUMP3_162, mtgp32_jump_table[6][351]
  streams.json          SHA-256 digests of the benchmark-shaped streams
  substreams.npz        substreams i*3^162 of set 0, seed 777
  dc_params_digests.json  the same for the dynamic-creator tables (200+200)
  dc_outputs.npz        test-program sequences and stepped outputs of
                        dynamic-creator sets (reference built with
                        cuda-sample/mtgp*dc-param-11213.c)
  dc64cl_jump.npz       the reference's 64-bit jump data (openCL-sample/:
                        characteristic, 3^162 jump string, jump table) for
                        set 0 of its MTGP64-DC table, that set's record and
                        the reference mtgp64_fast_jump by the string

Steps can be selected by name: make_golden.py dc_golden
"""
import hashlib
import json
import os
import re
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)
REF = "/root/reference"

from pyoracle import RefLib  # noqa: E402

MEXPS = {32: (11213, 23209, 44497), 64: (23209, 44497, 110503)}
MASK64 = (1 << 64) - 1


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def splitmix64(state, count):
    out = []
    for _ in range(count):
        state = (state + 0x9E3779B97F4A7C15) & MASK64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        out.append(z ^ (z >> 31))
    return out


def params_digests():
    out = {"digests": {}, "records": {}}
    for bits, mexps in MEXPS.items():
        ref = RefLib("fast", bits)
        for mexp in mexps:
            ds = []
            for i in range(128):
                rec = ref.params(mexp, i)
                ds.append(sha(rec))
                if i in (0, 127):
                    out["records"]["%d/%d/%d" % (bits, mexp, i)] = [int(v) for v in rec]
            out["digests"]["%d/%d" % (bits, mexp)] = ds
    with open(os.path.join(HERE, "params_digests.json"), "w") as f:
        json.dump(out, f)


def main_outputs():
    arrays = {}
    for bits, mexps in MEXPS.items():
        ref = RefLib("fast", bits)
        for mexp in mexps:
            for idx in (0, 127):
                key = "%d_%d_%d" % (bits, mexp, idx)
                g = ref.new_seed(mexp, idx, 1)
                arrays[key + "_init"] = g.fill(0, 1000)
                g = ref.new_array(mexp, idx, [1, 2, 3, 4])
                arrays[key + "_array"] = g.fill(0, 1000)
                g = ref.new_str(mexp, idx, b"\x01\x02\x03\x04")
                arrays[key + "_str"] = g.fill(0, 1000)
                for kind in (1, 2, 3, 4):
                    arrays[key + "_str_k%d" % kind] = g.fill(kind, 1000)
                # a string with bytes >= 0x80 (init_by_str adds them as signed char)
                g = ref.new_str(mexp, idx, b"MTGP\xff\x80\x7f!")
                arrays[key + "_str_signed"] = g.fill(0, 256)
    np.savez_compressed(os.path.join(HERE, "main_outputs.npz"), **arrays)


def jump_steps():
    """After mtgp_init(seed 0) and three draws, step s numbers and record the
    next 100 (the comparison test-jump32.cpp:218-264 makes after jumping)."""
    arrays = {}
    for bits, mexps in MEXPS.items():
        ref = RefLib("fast", bits)
        for mexp in mexps:
            if mexp == 110503:
                continue
            n = mexp // bits + 1
            for s in (1, 2, n + 1, 32 * n - 1, 32 * n + 1, 3003, 200004, 10000005):
                g = ref.new_seed(mexp, 0, 0)
                g.fill(0, 3)
                g.fill(0, s)
                arrays["%d_%d_%d" % (bits, mexp, s)] = g.fill(0, 100)
    np.savez_compressed(os.path.join(HERE, "jump_steps.npz"), **arrays)


def reference_jump_data():
    char = re.search(r'"([0-9a-f]+)"', open(REF + "/cuda-sample/mtgp32-char11213.h").read()).group(1)
    js = re.search(r'"([0-9a-f]+)"', open(REF + "/openCL-sample/mtgp32-jump-string.h").read()).group(1)
    body = open(REF + "/openCL-sample/mtgp32-jump-table.cpp").read()
    body = body[body.index("mtgp32_jump_table[] = {"):]
    vals = [int(v, 16) for v in re.findall(r"0x([0-9a-fA-F]{8})", body)]
    assert len(vals) == 6 * 351, len(vals)
    np.savez_compressed(os.path.join(HERE, "reference_jump_data.npz"),
                        char32_11213=np.array(char), jump3_162=np.array(js),
                        jump_table=np.array(vals, np.uint32).reshape(6, 351))


def _batched32(g, seed, L):
    """Two consecutive batches of generator g: L uint32 then L close_open
    floats, twice (the second continues the stream: the next batch of a
    rebased plan)."""
    ref = RefLib("fast", 32)
    gen = ref.new_seed(11213, g, seed)
    out = []
    for _ in range(2):
        a = gen.fill(0, L)
        b = gen.fill(2, L)
        out += [sha(a), sha(b), sha(gen.window())]
    return (g,) + tuple(out)


def _batched64(mexp, i, key):
    ref = RefLib("fast", 64)
    gen = ref.new_array(mexp, i, key)
    a = gen.fill(0, 1 << 24)
    b = gen.fill(2, 1 << 24)
    return mexp, i, sha(a), sha(b), sha(gen.window())


def streams():
    import bench
    out = {}
    # plain MTGP32-11213 set 0, seed 1234: 2^24 uint32 then 2^24 close1_open2
    ref = RefLib("fast", 32)
    g = ref.new_seed(11213, 0, 1234)
    a = g.fill(0, 1 << 24)
    b = g.fill(1, 1 << 24)
    out["plain"] = {"uint32": sha(a), "close1_open2": sha(b), "window": g.window().tolist(),
                    "uint32_head": a[:16].tolist(), "close1_open2_head": b[:16].tolist()}
    seeds, lengths = bench.batched_workload()
    with ProcessPoolExecutor(8) as ex:
        res = list(ex.map(_batched32, range(128), seeds, lengths))
    out["batched32"] = {"seeds": seeds, "lengths": lengths,
                        "uint32": [r[1] for r in res], "close_open": [r[2] for r in res],
                        "window": [r[3] for r in res]}
    out["batched32_next"] = {"uint32": [r[4] for r in res], "close_open": [r[5] for r in res],
                             "window": [r[6] for r in res]}
    jobs = []
    for mexp in (23209, 44497):
        for i in range(64):
            key = splitmix64((mexp << 8) | i, 4)
            jobs.append((mexp, i, key))
    with ProcessPoolExecutor(8) as ex:
        res = list(ex.map(_batched64, *zip(*jobs)))
    out["batched64"] = {"keys": {"%d/%d" % (m, i): [str(k) for k in key] for m, i, key in jobs},
                        "uint64": {"%d/%d" % (r[0], r[1]): r[2] for r in res},
                        "close_open": {"%d/%d" % (r[0], r[1]): r[3] for r in res},
                        "window": {"%d/%d" % (r[0], r[1]): r[4] for r in res}}
    with open(os.path.join(HERE, "streams.json"), "w") as f:
        json.dump(out, f)


def substreams():
    """Substream i of set 0, seed 777 = mtgp32_fast_jump(JUMP3_162) applied
    i times (as the reference's jump samples check it)."""
    js = re.search(r'"([0-9a-f]+)"', open(REF + "/openCL-sample/mtgp32-jump-string.h").read()).group(1)
    ref = RefLib("fast", 32)
    sample = [0, 1, 2, 3, 7, 64, 1000, 4095]
    arrays = {"indices": np.array(sample)}
    g = ref.new_seed(11213, 0, 777)
    base = g.window()
    cur = 0
    for i in sample:
        while cur < i:
            g.jump(js)
            cur += 1
        h = ref.new_seed(11213, 0, 777)
        h.set_window(g.window())
        arrays["window_%d" % i] = h.window()
        out = h.fill(0, 1 << 20)
        arrays["head_%d" % i] = out[:1024]
        arrays["sha_%d" % i] = np.array(sha(out))
    arrays["base"] = base
    # Every substream: the first 16 bytes of the SHA-256 of its 2^20
    # uint32 outputs and of its start window (word 0 under the mask: its
    # other bits are outside the state and depend on how the jump
    # polynomial is composed).
    g = ref.new_seed(11213, 0, 777)
    windows = []
    for i in range(4096):
        if i:
            g.jump(js)
        windows.append(g.window())
    with ProcessPoolExecutor(8) as ex:
        res = list(ex.map(_substream_digests, windows))
    arrays["all_out_sha16"] = np.array([r[0] for r in res], np.uint8)
    arrays["all_window_sha16"] = np.array([r[1] for r in res], np.uint8)
    np.savez_compressed(os.path.join(HERE, "substreams.npz"), **arrays)


def _substream_digests(w):
    ref = RefLib("fast", 32)
    h = ref.new_seed(11213, 0, 777)
    h.set_window(w)
    out = h.fill(0, 1 << 20)
    m = np.array(w, np.uint32)
    m[0] &= np.uint32(int(ref.params(11213, 0)[52]))  # the record's mask
    return (list(hashlib.sha256(out.tobytes()).digest()[:16]), list(hashlib.sha256(m.tobytes()).digest()[:16]))


DC_SETS = {32: (0, 16, 199), 64: (0, 172, 199)}  # 16 and 172 have pos = 3


def dc_golden():
    digests = {"digests": {}, "records": {}}
    arrays = {}
    for bits in (32, 64):
        ref = RefLib("dc", bits)
        ds = []
        for i in range(200):
            rec = ref.params(11213, i)
            ds.append(sha(rec))
            if i in (0, 199):
                digests["records"]["%d/%d" % (bits, i)] = [int(v) for v in rec]
        digests["digests"]["%d" % bits] = ds
        n = 11213 // bits + 1
        for idx in DC_SETS[bits]:
            key = "%d_%d" % (bits, idx)
            g = ref.new_seed(11213, idx, 1)
            arrays[key + "_init"] = g.fill(0, 1000)
            for kind in (1, 2, 3, 4):
                arrays[key + "_init_k%d" % kind] = g.fill(kind, 1000)
            arrays[key + "_array"] = ref.new_array(11213, idx, [1, 2, 3, 4]).fill(0, 1000)
            for s in (1, n + 1, 3003, 200004):
                g = ref.new_seed(11213, idx, 0)
                g.fill(0, 3)
                g.fill(0, s)
                arrays[key + "_step_%d" % s] = g.fill(0, 100)
    with open(os.path.join(HERE, "dc_params_digests.json"), "w") as f:
        json.dump(digests, f)
    np.savez_compressed(os.path.join(HERE, "dc_outputs.npz"), **arrays)


def dc64cl_jump():
    """The reference's 64-bit jump data (openCL-sample/): the characteristic
    polynomial of set 0 of the OpenCL sample's MTGP64-DC-11213 table
    (characteristic.64.11213.txt), x^(3^162) mod it (mtgp64-jump-string.h),
    the jump table rows x^(4^k 3^162) (mtgp64-jump-table.cpp), that set's
    record from the reference build, and the reference mtgp64_fast_jump of a
    seeded state by the shipped string with outputs after it."""
    lines = open(REF + "/openCL-sample/characteristic.64.11213.txt").read().split("\n")
    assert lines[0].startswith('# "0e4c448cbacfa0f8ec70c51d24b9ffcee74843f0"')
    char = lines[1].strip()
    js = re.search(r'"([0-9a-f]+)"', open(REF + "/openCL-sample/mtgp64-jump-string.h").read()).group(1)
    body = open(REF + "/openCL-sample/mtgp64-jump-table.cpp").read()
    body = body[body.index("mtgp64_jump_table[MTGP64_JMP_TBL_SIZE * MTGP64_JTS] = {"):]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    vals = [int(v, 16) for v in re.findall(r"0x([0-9a-fA-F]{8})", body)]
    jts = 11213 // 32 + 1
    assert len(vals) == 6 * jts, len(vals)
    ref = RefLib("dccl", 64)
    rec = ref.params(11213, 0)
    g = ref.new_seed(11213, 0, 4321)
    g.jump(js)
    arrays = dict(char64=np.array(char), jump3_162=np.array(js),
                  jump_table=np.array(vals, np.uint32).reshape(6, jts), record=np.array(rec, np.uint64),
                  window=g.window(), uint64=g.fill(0, 1000), close_open=g.fill(2, 1000))
    np.savez_compressed(os.path.join(HERE, "dc64cl_jump.npz"), **arrays)


STEPS = (params_digests, main_outputs, jump_steps, reference_jump_data, substreams, streams, dc_golden, dc64cl_jump)

if __name__ == "__main__":
    wanted = sys.argv[1:]
    for step in STEPS:
        if wanted and step.__name__ not in wanted:
            continue
        print("writing", step.__name__, flush=True)
        step()
