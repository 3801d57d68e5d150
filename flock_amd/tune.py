"""Autotuner for the library's launch tunables (flock_hip_set_tunable).

Coordinate descent over each tunable's candidate values: every candidate is timed on whole
prove_fast proofs of the bench's workload (2^n_log SHA-256 compressions, inputs resident in HBM),
best of `reps` proofs after a warm-up, and a candidate replaces the current value only when it is
faster by more than `min_gain`. Every proof is checked byte for byte against the proof made with
the defaults (a tunable changes launch shapes and thresholds, never the bytes). The choices go
to flock_amd/tuning.json, which bench.py (and any caller, through lib.apply_tuning) applies.

    python -m flock_amd.tune [--n-log 16] [--reps 5] [--passes 2] [--out PATH]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from flock_amd import lib as F  # noqa: E402

# candidate values per tunable (each within the library's allowed range)
CANDIDATES = {
    "ntt_cols_per_thread": [1, 2, 3, 4],
    "ntt_deep_cols": [64, 128, 256],
    "zc_split_min_vars": [12, 14, 16, 18, 20],
    "zc_la_max_log": [12, 14, 16, 17, 18, 20],
    "sc_la_max_log": [14, 16, 18, 20, 22],
    "open_split_fold_log": [12, 14, 16, 18, 20],
}
DOMAIN = b"flock-readme-bench-v0"
SHA256_IV = [0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A, 0x510E527F, 0x9B05688C, 0x1F83D9AB,
             0x5BE0CD19]


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-log", type=int, default=16)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--min-gain", type=float, default=0.005,
                    help="relative time a candidate must save to replace the current value")
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.json"))
    args = ap.parse_args()

    n = 1 << args.n_log
    rng = np.random.default_rng(0x5A256)
    inputs = np.zeros((n, 24), dtype=np.uint32)
    inputs[:, :8] = SHA256_IV
    inputs[:, 8:] = rng.integers(0, 1 << 32, (n, 16), dtype=np.uint64).astype(np.uint32)
    ctx = F.Context(0)
    d_in = ctx.upload(inputs)
    ctx.sync()
    pr = F.Prover(0)
    req, keep = F.device_request(F.HASH_SHA256, args.n_log, n, d_in, DOMAIN)

    def digest(out) -> str:
        _, proof, cap, ab, c = out
        return hashlib.sha256(proof + cap + ab.to_bytes(16, "little")
                              + c.to_bytes(16, "little")).hexdigest()

    def timed() -> tuple[float, str]:
        d0 = digest(F.prove_fast(pr, req))  # warm-up (buffers grow to the new shapes)
        best = float("inf")
        for _ in range(args.reps):
            t0 = time.perf_counter()
            out = F.prove_fast(pr, req)
            best = min(best, (time.perf_counter() - t0) * 1e3)
            if digest(out) != d0:
                raise RuntimeError("a tunable changed the proof bytes")
        return best, d0

    defaults = {k: F.get_tunable(k) for k in CANDIDATES}
    base_ms, base_digest = timed()
    print(f"defaults {defaults}: {base_ms:.3f} ms", flush=True)
    cur = dict(defaults)
    cur_ms = base_ms
    sweeps: dict[str, dict[str, float]] = {}
    for p in range(args.passes):
        for name, cands in CANDIDATES.items():
            res = {}
            for v in cands:
                F.set_tunable(name, v)
                ms, dg = timed()
                if dg != base_digest:
                    raise RuntimeError(f"{name}={v} changed the proof bytes")
                res[str(v)] = round(ms, 4)
                print(f"pass {p} {name}={v}: {ms:.3f} ms", flush=True)
            best_v = min(cands, key=lambda v: res[str(v)])
            # keep the current value unless the best candidate is clearly faster
            cur_here = res.get(str(cur[name]), cur_ms)
            if res[str(best_v)] < cur_here * (1.0 - args.min_gain):
                cur[name] = best_v
            F.set_tunable(name, cur[name])
            cur_ms = min(cur_here, res[str(cur[name])])
            sweeps[f"pass{p}:{name}"] = res
    tuned_ms, _ = timed()
    out = {
        "workload": f"prove_fast SHA-256 x 2^{args.n_log}, inputs in HBM, best of {args.reps}",
        "values": cur,
        "defaults": defaults,
        "default_ms": round(base_ms, 4),
        "tuned_ms": round(tuned_ms, 4),
        "sweeps": sweeps,
    }
    with open(args.out, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps({"values": cur, "default_ms": out["default_ms"], "tuned_ms": out["tuned_ms"]}),
          flush=True)
    pr.close()
    d_in.free()
    ctx.close()


if __name__ == "__main__":
    main()
