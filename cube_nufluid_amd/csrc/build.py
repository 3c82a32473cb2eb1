"""Build libcubehip.so (gfx950) from the HIP sources in this directory.

Every translation unit is compiled with hipcc for gfx950 only.  Files whose
arithmetic must reproduce the reference bit for bit (codecs, drift, kicks,
CIC weights) are compiled with -ffp-contract=off so that no multiply-add is
fused; the FFT keeps the compiler's default contraction.  Packed FP32
instructions are disabled for all device code.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
OUT = os.path.join(PKG, "libcubehip.so")
OBJDIR = os.path.join(PKG, "_build")

SOURCES = {
    "cube_api.hip": ["-ffp-contract=off"],
    "drift.hip": ["-ffp-contract=off"],
    "pm.hip": ["-ffp-contract=off"],
    "pp.hip": ["-ffp-contract=off"],
    "scan.hip": [],
    "fft.hip": [],
    "ic.hip": ["-ffp-contract=off"],
    "pk.hip": [],
    "probe.hip": [],
    "serve.hip": ["-ffp-contract=off"],
    "fluid.hip": ["-ffp-contract=off"],
    "halo.hip": ["-ffp-contract=off"],
}

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-fast-math", "-Wno-unused-result",
    "-I" + INCLUDE, "-I" + HERE,
]


def hipcc():
    for p in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(p) or p == "hipcc":
            return p
    raise RuntimeError("hipcc not found")


def compile_one(src, extra):
    obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
    cmd = [hipcc(), "-c", os.path.join(HERE, src), "-o", obj] + COMMON + extra
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=True):
    os.makedirs(OBJDIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        futs = {ex.submit(compile_one, s, e): s for s, e in SOURCES.items()}
        objs = []
        for f in concurrent.futures.as_completed(futs):
            objs.append(f.result())
            if verbose:
                print("  compiled", futs[f], flush=True)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + sorted(objs)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, OUT)
    if verbose:
        print("  linked", OUT, flush=True)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
