"""Build libflock_hip.so in-tree with hipcc for gfx950.

    python -m flock_amd.build          # or flock_amd.build.build()
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libflock_hip.so")
SOURCES = ["api.hip", "field.hip", "ntt.hip", "merkle.hip", "sumcheck.hip", "zerocheck.hip",
           "lincheck.hip", "witness.hip", "prove.hip", "opening.hip", "circuits.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src: str, obj: str, verbose: bool) -> None:
    cmd = [HIPCC, *FLAGS, "-c", "-I", os.path.join(ROOT, "include"), "-o", obj, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed on {os.path.basename(src)}")


def build(verbose: bool = False, incremental: bool = False) -> str:
    """Compile every source to an object in parallel (one hipcc per file), then link the shared
    library. incremental=True skips objects newer than their source and every header."""
    from concurrent.futures import ThreadPoolExecutor

    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    headers.append(os.path.join(ROOT, "include", "flock_hip.h"))
    newest_h = max(os.path.getmtime(h) for h in headers)
    jobs = []
    objs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(objdir, s.replace(".hip", ".o"))
        objs.append(obj)
        fresh = (incremental and os.path.exists(obj)
                 and os.path.getmtime(obj) > max(os.path.getmtime(src), newest_h))
        if not fresh:
            jobs.append((src, obj))
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS") or os.cpu_count() or 1), 16))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for f in [ex.submit(_compile, src, obj, verbose) for src, obj in jobs]:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libflock_hip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
