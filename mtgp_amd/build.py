"""Builds libmtgp_amd.so (gfx950) in-tree: python mtgp_amd/build.py

Host C++ (parameter tables, init family, GF(2)[x]) is compiled with g++,
the runtime and kernels with hipcc for gfx950, and everything is linked
into mtgp_amd/libmtgp_amd.so.  Objects go to mtgp_amd/_build/.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmtgp_amd.so")
ARCH = os.environ.get("MTGP_OFFLOAD_ARCH", "gfx950")
ROCM_INC = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "include")

HOST_SOURCES = ["host.cpp", "gf2x.cpp"]
HIP_HOST_SOURCES = ["runtime.cpp", "serve.cpp"]
KERNEL_SOURCES = ["kernels.hip", "init_kernels.hip", "noalloc.hip"]

# Packed FP32 instructions are kept out of all device code.
NO_PK_FP32 = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    lines = [ln for ln in r.stdout.splitlines() if "packed-fp32-ops" not in ln]
    if r.returncode != 0:
        sys.stderr.write("\n".join(lines) + "\n")
        raise RuntimeError("build step failed: %s" % " ".join(cmd))
    if verbose and lines:
        print("\n".join(lines))


def build(verbose=False, jobs=4):
    """Compiles every source and links libmtgp_amd.so; returns its path."""
    os.makedirs(OUT, exist_ok=True)
    hipcc = _hipcc()
    common = ["-O3", "-fPIC", "-std=c++17", "-I", INC, "-I", SRC, "-Wall"]
    cmds, objs = [], []
    for s in HOST_SOURCES:
        o = os.path.join(OUT, s + ".o")
        cmds.append(["g++"] + common + ["-c", os.path.join(SRC, s), "-o", o])
        objs.append(o)
    for s in HIP_HOST_SOURCES:
        if not os.path.exists(os.path.join(SRC, s)):
            continue
        # Host code that only calls the HIP runtime API: plain g++.
        o = os.path.join(OUT, s + ".o")
        cmds.append(["g++"] + common + ["-D__HIP_PLATFORM_AMD__", "-I", ROCM_INC, "-c",
                                        os.path.join(SRC, s), "-o", o])
        objs.append(o)
    for s in KERNEL_SOURCES:
        o = os.path.join(OUT, s + ".o")
        cmds.append([hipcc] + common + ["-std=c++20", "--offload-arch=" + ARCH] + NO_PK_FP32 +
                    ["-c", os.path.join(SRC, s), "-o", o])
        objs.append(o)
    procs = []
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        procs = [ex.submit(_run, c, verbose) for c in cmds]
        for p in procs:
            p.result()
    # Device assembly of every kernel source, scanned for the SCC select
    # hazard.
    sys.path.insert(0, HERE)
    import isa_check
    for src in KERNEL_SOURCES:
        asm = os.path.join(OUT, src.replace(".hip", "") + "-gfx950.s")
        _run([hipcc] + common + ["-std=c++20", "--offload-arch=" + ARCH, "--cuda-device-only", "-S"] + NO_PK_FP32 +
             [os.path.join(SRC, src), "-o", asm], verbose)
        bad = isa_check.scan(open(asm).read())
        if bad:
            raise RuntimeError("%s: s_cselect after a carry/borrow write of SCC at %s" %
                               (src, ", ".join("line %d" % b[0] for b in bad)))
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
