"""Build the native library ``ml_prec_amd/libmlp.so`` for gfx950 with hipcc.

Every source is compiled with ``-ffp-contract=off`` so that floating-point
expressions are evaluated as written (a rounded multiply, then a rounded add),
which keeps FP64 results identical to the reference's CPU loops.
"""
from __future__ import annotations

import os
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
CPP = os.path.join(HERE, "cpp")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmlp.so")
DEMO = os.path.join(BUILD, "dropin_demo")
ARCH = os.environ.get("MLP_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-I" + os.path.join(HERE, "..", "include")]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]

SOURCES = [
    (os.path.join(CSRC, "hierarchy.cc"), False),
    (os.path.join(CSRC, "plan.cc"), False),
    (os.path.join(CSRC, "assembly.cc"), False),
    (os.path.join(CPP, "ml_prec.cc"), False),
    (os.path.join(CSRC, "capi.hip"), True),
    (os.path.join(CSRC, "devbuild.hip"), True),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build(verbose: bool = False) -> str:
    """Compile every source (always; the sources at once, each in its own hipcc process) and link libmlp.so."""
    os.makedirs(BUILD, exist_ok=True)
    t0 = time.time()
    jobs = []
    for src, is_hip in SOURCES:
        if not os.path.exists(src):
            continue
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        cmd = [HIPCC] + COMMON + (DEVICE + ["-x", "hip"] if is_hip else ["-x", "c++"]) + ["-c", src, "-o", obj]
        jobs.append((obj, cmd))
    with ThreadPoolExecutor(max_workers=len(jobs)) as pool:
        outs = list(pool.map(lambda j: _run(j[1]), jobs))     # a failed step raises here
    for out in outs:
        if verbose and out.strip():
            print(out)
    objs = [obj for obj, _ in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # the C++ drop-in driver, linked against libmlp.so
    demo = os.path.join(CPP, "dropin_demo.cc")
    if os.path.exists(demo):
        _run([HIPCC] + COMMON + ["-x", "c++", demo, "-o", DEMO, "-L" + HERE, "-lmlp",
                                 "-Wl,-rpath,$ORIGIN/.."])
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f}s")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
