"""The C-ABI library on the CPU: every function include/*.h declares is
exported, the parameter-table symbols hold the reference's records, and the
drop-in headers compile as C and reproduce the reference test programs'
sequences through the library's host functions and the inline genrand
family (no GPU involved: these are host functions)."""
import ctypes
import hashlib
import json
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
LIB = os.path.join(ROOT, "mtgp_amd", "libmtgp_amd.so")
INC = os.path.join(ROOT, "include")


def _strip_bodies(text):
    """Drops comments, preprocessor lines and every function body, keeping
    the top-level declarations (extern "C" blocks are opened, not dropped)."""
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = "\n".join(ln for ln in text.splitlines() if not ln.lstrip().startswith("#"))
    text = text.replace('extern "C" {', "")
    out, depth, i = [], 0, 0
    while i < len(text):
        c = text[i]
        if c == "{":
            # struct/enum bodies stay (their fields are not calls); function
            # bodies go.
            head = text[max(0, i - 200):i]
            is_func = head.rstrip().endswith(")")
            if depth == 0 and is_func:
                j, d = i, 0
                while True:
                    if text[j] == "{":
                        d += 1
                    elif text[j] == "}":
                        d -= 1
                        if d == 0:
                            break
                    j += 1
                out.append(";")
                i = j + 1
                continue
            depth += 1
        elif c == "}":
            depth = max(0, depth - 1)
        out.append(c)
        i += 1
    return "".join(out)


def declared_functions():
    """Names of the non-inline functions the public headers declare."""
    names = set()
    for h in sorted(os.listdir(INC)):
        text = _strip_bodies(open(os.path.join(INC, h)).read())
        for decl in text.split(";"):
            decl = " ".join(decl.split())
            if "inline" in decl or "typedef" in decl or "(" not in decl:
                continue
            m = re.search(r"(\w+)\s*\(", decl)
            if m:
                names.add(m.group(1))
    return names


def test_library_exports_declared_symbols():
    out = subprocess.check_output(["nm", "-D", "--defined-only", LIB], text=True)
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    want = declared_functions()
    assert len(want) > 50
    missing = sorted(want - exported)
    assert not missing, missing
    for t in ("mtgp32_params_fast_11213", "mtgp32_params_fast_23209", "mtgp32_params_fast_44497",
              "mtgp64_params_fast_23209", "mtgp64_params_fast_44497", "mtgp64_params_fast_110503"):
        assert t in exported, t


def test_parameter_tables_match_reference(mtgp):
    gold = json.load(open(os.path.join(GOLD, "params_digests.json")))
    for bits, mexps in mtgp.MEXPS.items():
        for mexp in mexps:
            for i in range(128):
                p = mtgp.params(bits, mexp, i)
                tbl = list(p.tbl)
                tmp = list(p.tmp_tbl)
                flt = list(p.flt_tmp_tbl if bits == 32 else p.dbl_tmp_tbl)
                rec = [p.mexp, p.pos, p.sh1, p.sh2] + tbl + tmp + flt + [p.mask] + list(bytes(p.poly_sha1)[:20])
                d = hashlib.sha256(np.array(rec, np.uint64).tobytes()).hexdigest()
                assert d == gold["digests"]["%d/%d" % (bits, mexp)][i], (bits, mexp, i)


C_PROGRAM = r"""
#include <stdio.h>
#include <string.h>
#include "mtgp32-fast.h"
#include "mtgp64-fast.h"
#include "mtgp32-fast-jump.h"
#include "mtgp64-fast-jump.h"

static void dump32(mtgp32_fast_t *g, FILE *f, int kind, int n) {
    for (int i = 0; i < n; i++) {
        uint32_t u; float x;
        switch (kind) {
        case 0: u = mtgp32_genrand_uint32(g); break;
        case 1: x = mtgp32_genrand_close1_open2(g); memcpy(&u, &x, 4); break;
        case 2: x = mtgp32_genrand_close_open(g); memcpy(&u, &x, 4); break;
        case 3: x = mtgp32_genrand_open_close(g); memcpy(&u, &x, 4); break;
        default: x = mtgp32_genrand_open_open(g); memcpy(&u, &x, 4); break;
        }
        fwrite(&u, 4, 1, f);
    }
}

static void dump64(mtgp64_fast_t *g, FILE *f, int kind, int n) {
    for (int i = 0; i < n; i++) {
        uint64_t u; double x;
        switch (kind) {
        case 0: u = mtgp64_genrand_uint64(g); break;
        case 1: x = mtgp64_genrand_close1_open2(g); memcpy(&u, &x, 8); break;
        case 2: x = mtgp64_genrand_close_open(g); memcpy(&u, &x, 8); break;
        case 3: x = mtgp64_genrand_open_close(g); memcpy(&u, &x, 8); break;
        default: x = mtgp64_genrand_open_open(g); memcpy(&u, &x, 8); break;
        }
        fwrite(&u, 8, 1, f);
    }
}

int main(int argc, char **argv) {
    FILE *f = fopen(argv[1], "wb");
    mtgp32_params_fast_t *t32[3] = {mtgp32_params_fast_11213, mtgp32_params_fast_23209, mtgp32_params_fast_44497};
    mtgp64_params_fast_t *t64[3] = {mtgp64_params_fast_23209, mtgp64_params_fast_44497, mtgp64_params_fast_110503};
    int idx[2] = {0, 127};
    uint32_t ar32[4] = {1, 2, 3, 4};
    uint64_t ar64[4] = {1, 2, 3, 4};
    char s1[] = "\001\002\003\004";
    char s2[] = "MTGP\377\200\177!";
    for (int t = 0; t < 3; t++) for (int k = 0; k < 2; k++) {
        mtgp32_fast_t g;
        if (mtgp32_init(&g, &t32[t][idx[k]], 1)) return 1;
        dump32(&g, f, 0, 1000); mtgp32_free(&g);
        if (mtgp32_init_by_array(&g, &t32[t][idx[k]], ar32, 4)) return 1;
        dump32(&g, f, 0, 1000); mtgp32_free(&g);
        if (mtgp32_init_by_str(&g, &t32[t][idx[k]], s1)) return 1;
        dump32(&g, f, 0, 1000);
        for (int kind = 1; kind < 5; kind++) dump32(&g, f, kind, 1000);
        mtgp32_free(&g);
        if (mtgp32_init_by_str(&g, &t32[t][idx[k]], s2)) return 1;
        dump32(&g, f, 0, 256);
        mtgp32_print_idstring(&g, f);
        mtgp32_free(&g);
    }
    for (int t = 0; t < 3; t++) for (int k = 0; k < 2; k++) {
        mtgp64_fast_t g;
        if (mtgp64_init(&g, &t64[t][idx[k]], 1)) return 1;
        dump64(&g, f, 0, 1000); mtgp64_free(&g);
        if (mtgp64_init_by_array(&g, &t64[t][idx[k]], ar64, 4)) return 1;
        dump64(&g, f, 0, 1000); mtgp64_free(&g);
        if (mtgp64_init_by_str(&g, &t64[t][idx[k]], s1)) return 1;
        dump64(&g, f, 0, 1000);
        for (int kind = 1; kind < 5; kind++) dump64(&g, f, kind, 1000);
        mtgp64_free(&g);
        if (mtgp64_init_by_str(&g, &t64[t][idx[k]], s2)) return 1;
        dump64(&g, f, 0, 256);
        mtgp64_print_idstring(&g, f);
        mtgp64_free(&g);
    }
    /* mtgp32_copy keeps the stream: the copy continues identically */
    {
        mtgp32_fast_t a, b;
        mtgp32_init(&a, &t32[0][5], 42);
        mtgp32_init(&b, &t32[0][6], 0);
        for (int i = 0; i < 777; i++) mtgp32_genrand_uint32(&a);
        mtgp32_copy(&b, &a);
        for (int i = 0; i < 1000; i++) if (mtgp32_genrand_uint32(&a) != mtgp32_genrand_uint32(&b)) return 2;
        mtgp32_free(&a); mtgp32_free(&b);
    }
    fclose(f);
    return 0;
}
"""


def test_dropin_headers_reproduce_reference_programs(tmp_path):
    src = tmp_path / "prog.c"
    exe = tmp_path / "prog"
    src.write_text(C_PROGRAM)
    subprocess.check_call(["gcc", "-std=c99", "-O2", "-Wall", "-Werror", "-I", INC, str(src), "-o", str(exe),
                           LIB, "-Wl,-rpath," + os.path.dirname(LIB)])
    out = tmp_path / "out.bin"
    subprocess.check_call([str(exe), str(out)])
    data = out.read_bytes()
    gold = np.load(os.path.join(GOLD, "main_outputs.npz"))
    gp = json.load(open(os.path.join(GOLD, "params_digests.json")))
    pos = 0
    for bits, mexps in ((32, (11213, 23209, 44497)), (64, (23209, 44497, 110503))):
        w = bits // 8
        for mexp in mexps:
            for idx in (0, 127):
                key = "%d_%d_%d" % (bits, mexp, idx)
                for name, n in [("_init", 1000), ("_array", 1000), ("_str", 1000), ("_str_k1", 1000),
                                ("_str_k2", 1000), ("_str_k3", 1000), ("_str_k4", 1000), ("_str_signed", 256)]:
                    chunk = data[pos:pos + n * w]
                    pos += n * w
                    assert chunk == np.ascontiguousarray(gold[key + name]).tobytes(), key + name
                line_end = data.index(b"\n", pos) + 1
                line = data[pos:line_end].decode()
                pos = line_end
                sha = bytes(gp["records"]["%d/%d/%d" % (bits, mexp, idx)][53:73]).hex()
                assert line == "mtgp%d:%d:%s\n" % (bits, mexp, sha)
    assert pos == len(data)


def test_gpu_entry_points_fail_loudly_without_gpu(mtgp):
    """No CPU fallback: without a visible GPU every entry point that computes
    numbers in bulk (and the device jump) returns an error and computes
    nothing; the host init family still works."""
    try:
        mtgp.device_info()
    except mtgp.MTGPError:
        pass
    else:
        pytest.skip("a GPU is visible")
    g = mtgp.Generator32(11213, 0, seed=1)
    w0 = g.window().copy()
    out = np.zeros(1000, np.uint32)
    with pytest.raises(mtgp.MTGPError):
        g.fill(0, 1000, out=out)
    assert (out == 0).all() and (g.window() == w0).all()
    with pytest.raises(mtgp.MTGPError):
        g.fast_jump(mtgp.jump_hex(32, 11213, 0, 12345))
    assert (g.window() == w0).all()
    f = mtgp._lib.mtgp32_fast_jump_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.POINTER(mtgp.Fast32), ctypes.c_char_p]
    assert f(ctypes.byref(g.g), mtgp.jump_hex(32, 11213, 0, 12345).encode()) == -1
    assert mtgp.last_error() and (g.window() == w0).all()
    with pytest.raises(mtgp.MTGPError):
        mtgp.GenSet([(32, 11213, 0)])
    with pytest.raises(mtgp.MTGPError):
        mtgp.DeviceBuffer(1024)
