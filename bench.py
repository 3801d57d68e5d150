#!/usr/bin/env python3
"""Benchmark of the prove_fast data-parallel core on one MI355X at the metric's workload:
a batch of 2^16 SHA-256 compressions (IV initial state, seeded uniform 64-byte blocks).

One step = the GPU prove core of that batch, every stage on resident device buffers:
  1. witness   SHA-256 batch-major (z, A_0 z, B_0 z) and the lincheck stripe
  2. compact   the union dense stack q (199 used chunk-columns x 2^16 rows, 2^24 words)
  3. commit    lane-major Ligerito commit of q: m = 31, rate 1/2, log_batch_size 6,
               50 live lanes (LCH additive NTT to 2^19 x 50 words, then the SHA-256 Merkle tree)
  4. zerocheck the boolean zerocheck over (a, b, c = z), m = 31, grinding on
  5. lincheck  the boolean lincheck at x_ab from the zerocheck claim (K = 2^15, nu = 16),
               over a synthetic CSC pair with the SHA-256 circuit's nonzero counts
The PCS opening and the registry digest are not built here, so the step is not a full
proof; `prove_core_khps` is reported as such, never as `value`.

`value` is the commit NTT throughput in compulsory bytes (read the 50 live message lanes +
write the codeword = 16*(50*2^18 + 50*2^19) bytes) over the encode time, measured with HIP
events on the library's stream. `roofline` is the dominant kernel, k_ntt_pass<3>, timed alone.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("prove_fast k hashes/s (SHA-256 2^16); NTT GB/s vs HBM peak; "
          "p50/p99 proof latency")
HBM_PEAK_GBS = 8000.0
SHA256_IV = [0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A, 0x510E527F, 0x9B05688C, 0x1F83D9AB,
             0x5BE0CD19]
SHA256_USEFUL_BITS, SHA256_K_LOG, SHA256_Z_CONST = 25_470, 15, 25_469
SHA256_NNZ = (28_839_569, 18_527_069)  # nonzeros of (A_0, B_0) of the SHA-256 circuit
BLAKE3_K_LOG, BLAKE3_Z_CONST = 14, 11_706
BLAKE3_NNZ = (27_251_212, 17_191_286)  # nonzeros of (A_0, B_0) of the BLAKE3 circuit


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class HipEvents:
    """hipEvent timing on an explicit stream (the library's own)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                               ctypes.c_void_p]

    def new(self):
        e = ctypes.c_void_p()
        assert self.h.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        out = ctypes.c_float()
        assert self.h.hipEventElapsedTime(ctypes.byref(out), a, b) == 0
        return float(out.value)


def synthetic_csc(k: int, nnz: int, seed: int):
    """A CSC pair member with nnz uniformly spread rows (the circuit's own is the caller's)."""
    rng = np.random.default_rng(seed)
    counts = rng.multinomial(nnz, np.full(k, 1.0 / k))
    ptr = np.zeros(k + 1, dtype=np.uint32)
    ptr[1:] = np.cumsum(counts)
    rows = rng.integers(0, k, nnz, dtype=np.uint32)
    return ptr, rows


def request_inputs(F, kind: int, n: int, rng) -> np.ndarray:
    """Seeded inputs of one request: SHA-256 (IV, random block), BLAKE3 (random cv and block,
    block_len 64, flags 11), Keccak-f[1600] (random 1600-bit states)."""
    if kind == F.HASH_SHA256:
        x = np.zeros((n, 24), dtype=np.uint32)
        x[:, :8] = SHA256_IV
        x[:, 8:] = rng.integers(0, 1 << 32, (n, 16), dtype=np.uint64).astype(np.uint32)
    elif kind == F.HASH_BLAKE3:
        x = rng.integers(0, 1 << 32, (n, 28), dtype=np.uint64).astype(np.uint32)
        x[:, 26], x[:, 27] = 64, 11
    else:
        x = rng.integers(0, 1 << 63, (n, 25), dtype=np.uint64) * 2 + \
            rng.integers(0, 2, (n, 25), dtype=np.uint64)
    return x


def mixed_workloads(F, ctx, args) -> dict:
    """The batched form (64 uneven requests at once, kinds uniform over SHA-256 / BLAKE3 /
    Keccak-f[1600], batch log2 uniform in [10, 16]) and the served form (Poisson arrivals,
    batch log2 in [10, 18], a Keccak 2^18 request included), through flock_hip_serve."""
    circ = {}
    keep = []
    for kind, k_log, nnz, pin in ((F.HASH_SHA256, SHA256_K_LOG, SHA256_NNZ, SHA256_Z_CONST),
                                  (F.HASH_BLAKE3, BLAKE3_K_LOG, BLAKE3_NNZ, BLAKE3_Z_CONST)):
        csc = [synthetic_csc(1 << k_log, z, 31 + kind * 2 + i) for i, z in enumerate(nnz)]
        c, kp = ctx.csc(k_log, *csc[0], *csc[1], const_pin=pin)
        circ[kind] = c
        keep.append(kp)
    rng = np.random.default_rng(0xC4C5)

    def specs(n_req, lo, hi, rate, force_big):
        out, t = [], 0.0
        for i in range(n_req):
            kind = int(rng.integers(0, 3))
            nl = int(rng.integers(lo, hi + 1))
            if force_big and i == n_req // 2:
                kind, nl = F.HASH_KECCAK, hi
            if rate:
                t += float(rng.exponential(1.0 / rate))
            out.append((kind, nl, request_inputs(F, kind, 1 << nl, rng),
                        circ.get(kind), t))
        return out

    budget = int(288e9 * 0.9)
    provers = [F.Prover(0) for _ in range(args.workers)]
    for pr in provers:
        pr.set_round_mode({"host": F.ROUNDS_HOST, "device": F.ROUNDS_DEVICE}[args.round_mode])
    res = {}
    # batched: every request at t = 0; one worker (one proof at a time) vs all workers
    b_specs = specs(64, 10, 16, 0.0, False)
    reqs, kb = F.make_requests(b_specs)
    hashes = sum(1 << s[1] for s in b_specs)
    F.serve(provers, reqs, budget)  # warm-up: buffers grow to their largest shapes
    _, wall1 = F.serve(provers[:1], reqs, budget)
    out, walln = F.serve(provers, reqs, budget)
    lat = [r.done_s * 1e3 for r in out]
    res["batched"] = {
        "requests": len(b_specs), "hashes": hashes, "workers": args.workers,
        "one_at_a_time_khps": round(hashes / wall1, 2),
        "batched_khps": round(hashes / walln, 2),
        "speedup": round(wall1 / walln, 3),
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
    }
    # served: Poisson arrivals at args.serve_rate requests/s
    s_specs = specs(32, 10, 18, args.serve_rate, True)
    reqs, ks = F.make_requests(s_specs)
    hashes = sum(1 << s[1] for s in s_specs)
    F.serve(provers, reqs, budget)  # warm-up
    out, wall = F.serve(provers, reqs, budget)
    lat = [r.queued_ms + r.run_ms for r in out]
    res["served"] = {
        "requests": len(s_specs), "rate_per_s": args.serve_rate, "workers": args.workers,
        "hashes": hashes, "khps": round(hashes / wall, 2),
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "largest": "keccak-f[1600] 2^18",
        "hbm_budget_bytes": budget,
    }
    for p in provers:
        p.close()
    return res


def cpu_baseline(t: int, log_dim: int, reps: int) -> dict:
    """The oracle's scalar NTT (one core) on the commit shape of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O  # the checker doubles as the CPU baseline

    k = log_dim + 1
    msg = O.splitmix(0xC0FFEE, t << log_dim)
    cw = np.ascontiguousarray(np.tile(msg, (2, 1)))  # rate 1/2: the replicated message
    best = None
    O.ntt_forward_interleaved(cw[: t << (k - 4)], t, t, k - 4, k - 4, 1)  # warm-up
    for _ in range(reps):
        t0 = time.perf_counter()
        O.ntt_forward_interleaved(cw, t, t, k, k, 1)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    byts = 16 * ((t << log_dim) + (t << k))
    return {"value": round(byts / best / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"oracle scalar interleaved LCH NTT on the full workload transform "
                      f"({t} lanes x 2^{k}, rate 1/2, layers 1..{k}), best of {reps} after "
                      f"1 warm-up, {best * 1e3:.0f} ms per transform"}


def config1_commit(F, ctx, ev, stream, with_cpu: bool) -> dict:
    """The reference's CPU-runnable case: a 2^10 x 2^8 message of SplitMix64 words (seed
    0xC0FFEE, interleaved msg[p * 256 + lane]), rate 1/4, committed with the SHA-256 Merkle
    tree (m = 25, log_batch_size 8). GPU time of the whole commit with the input resident,
    beside the oracle's commit of the same input on one host core."""
    p = F.PcsParams(25, 2, 8, 0, 0)
    s = F.pcs_sizes(p)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O  # the checker doubles as the CPU baseline
    msg = O.splitmix(0xC0FFEE, 1 << 18)
    d_msg, cw = ctx.upload(msg), ctx.alloc(16 * s["cw_len"])
    tree = ctx.alloc(32 * (2 * s["n_leaves"] - 1))
    e = [ev.new(), ev.new()]
    ctx.commit(d_msg, 1 << 18, p, cw, tree)  # warm-up
    ts = []
    for _ in range(10):
        ev.record(e[0], stream)
        root = ctx.commit(d_msg, 1 << 18, p, cw, tree)
        ev.record(e[1], stream)
        ts.append(ev.ms(e[0], e[1]))
    out = {"gpu_ms": round(float(np.median(ts)), 4), "root": root[0].hex()}
    if with_cpu:
        t0 = time.perf_counter()
        _, etree = O.commit(msg, 25, 2, 8, 0)
        out["cpu_oracle_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
        out["cpu_cores"] = 1
        assert etree[-32:] == root[0], "config 1 root differs from the oracle"
    for b in (d_msg, cw, tree):
        b.free()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-log", type=int, default=16, help="log2 of the SHA-256 batch")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sumcheck", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the batched and served mixed-hash workloads")
    ap.add_argument("--workers", type=int, default=8, help="serving workers (HIP streams)")
    ap.add_argument("--serve-rate", type=float, default=60.0,
                    help="Poisson arrival rate of the served workload, requests/s")
    ap.add_argument("--latency-requests", type=int, default=50,
                    help="sequential prove-core requests timed one by one for p50/p99")
    ap.add_argument("--round-mode", choices=["host", "device"], default="host",
                    help="how the sumcheck round loops meet the transcript between rounds")
    ap.add_argument("--ntt-only", action="store_true",
                    help="only the commit (for profiler runs of the NTT)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 1 GiB device-to-device copy (the reference kernel for "
                         "scaling FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is described in DESIGN.md and not launched.", file=sys.stderr)
        sys.exit(2)

    from flock_amd import lib as F

    ctx = F.Context(0)
    round_mode = {"host": F.ROUNDS_HOST, "device": F.ROUNDS_DEVICE}[args.round_mode]
    ctx.set_round_mode(round_mode)
    ev = HipEvents()
    stream = ctx.stream
    nl = args.n_log
    n = 1 << nl
    reg = F.registry_new([(SHA256_K_LOG, SHA256_USEFUL_BITS, False)], nl)
    lb = 6
    sz = F.union_sizes(reg, [n], 0, lb)
    m = sz["dense_m"]
    lanes = sz["commit_lanes"] or 0
    p = F.PcsParams(m, 1, lb, lanes, 0)
    s = F.pcs_sizes(p)
    k = s["k_code"]
    t = lanes or (1 << lb)
    log(f"SHA-256 x 2^{nl}: dense {sz['dense_words']} committed {sz['committed_words']} m={m} "
        f"lanes={t} k_code={k} cw={s['cw_len']} leaves={s['n_leaves']}")

    rng = np.random.default_rng(0x5A256)
    inputs = np.zeros((n, 24), dtype=np.uint32)
    inputs[:, :8] = SHA256_IV
    inputs[:, 8:] = rng.integers(0, 1 << 32, (n, 16), dtype=np.uint64).astype(np.uint32)
    d_in = ctx.upload(inputs)
    words = 256 << nl
    z, a, b = (ctx.alloc(16 * words) for _ in range(3))
    stripe = ctx.alloc(4096 * n)
    q = ctx.alloc(16 * sz["committed_words"])
    cw = ctx.alloc(16 * s["cw_len"])
    tree = ctx.alloc(32 * (2 * s["n_leaves"] - 1))
    csc = [synthetic_csc(1 << SHA256_K_LOG, nnz, 11 + i) for i, nnz in enumerate(SHA256_NNZ)]
    circuit, keep = ctx.csc(SHA256_K_LOG, *csc[0], *csc[1], const_pin=SHA256_Z_CONST)
    m_zc = nl + SHA256_K_LOG
    e = [ev.new() for _ in range(2)]

    def step(times: dict):
        """One prove-core step; per-stage times in ms (events for kernels, wall for drivers)."""
        ev.record(e[0], stream)
        ctx.witness_sha256(d_in, n, nl, z, a, b, stripe)
        ev.record(e[1], stream)
        times.setdefault("witness", []).append(ev.ms(e[0], e[1]))
        ev.record(e[0], stream)
        ctx.union_compact(reg, [n], z, q)
        ev.record(e[1], stream)
        times.setdefault("compact", []).append(ev.ms(e[0], e[1]))
        ev.record(e[0], stream)
        cap = ctx.commit_lane_major(q, sz["committed_words"], p, cw, tree)
        ev.record(e[1], stream)
        times.setdefault("commit", []).append(ev.ms(e[0], e[1]))
        fs = F.FsChallenger(b"flock-readme-bench-v0")
        fs.bind_statement(bytes(32), [n], cap[0])
        ctx.sync()
        t0 = time.perf_counter()
        # the prover captures s_hat_v_c for the PCS opening (prover.rs:1095-1103)
        r1, ml, fin, nonces, claim, _ = ctx.zerocheck_prove(fs, a, b, z, m_zc, True,
                                                            capture=True)
        times.setdefault("zerocheck", []).append((time.perf_counter() - t0) * 1e3)
        mlv = [int(v[0]) | (int(v[1]) << 64) for v in claim[1:1 + (m_zc - 6)]]
        x_rest, x_outer = F.x_ab_from_mlv(reg, mlv)
        z_skip = int(claim[0][0]) | (int(claim[0][1]) << 64)
        t0 = time.perf_counter()
        ctx.lincheck_prove(fs, circuit, stripe, nl, SHA256_USEFUL_BITS, n, z_skip, x_rest,
                           x_outer, True)
        times.setdefault("lincheck", []).append((time.perf_counter() - t0) * 1e3)

    if not args.ntt_only:
        for _ in range(args.warmup):
            step({})
    else:
        ctx.witness_sha256(d_in, n, nl, z, a, b, stripe)
        ctx.union_compact(reg, [n], z, q)
        for _ in range(args.warmup):
            ctx.commit_lane_major(q, sz["committed_words"], p, cw, tree)
    ctx.sync()

    # Timed region: exactly K steps, bracketed by stream syncs.
    times: dict = {}
    ctx.sync()
    wall0 = time.perf_counter()
    for _ in range(args.steps):
        if args.ntt_only:
            ev.record(e[0], stream)
            ctx.commit_lane_major(q, sz["committed_words"], p, cw, tree)
            ev.record(e[1], stream)
            times.setdefault("commit", []).append(ev.ms(e[0], e[1]))
        else:
            step(times)
    ctx.sync()
    wall = time.perf_counter() - wall0
    stage_ms = {kk: float(np.mean(v)) for kk, v in times.items()}
    step_ms = wall * 1e3 / args.steps

    # Stage split of the commit, and the dominant kernel alone.
    enc, mt, mb, pt = [], [], [], []
    for _ in range(args.steps):
        ev.record(e[0], stream)
        ctx.encode_lane_major(q, sz["committed_words"], p, cw)
        ev.record(e[1], stream)
        enc.append(ev.ms(e[0], e[1]))
        ev.record(e[0], stream)
        ctx.merkle_tree(cw, 16 * s["cw_len"], s["n_leaves"], tree)
        ev.record(e[1], stream)
        mt.append(ev.ms(e[0], e[1]))
        ev.record(e[0], stream)  # the same leaves under HashKind::Blake3
        ctx.merkle_tree(cw, 16 * s["cw_len"], s["n_leaves"], tree, kind=F.MERKLE_BLAKE3)
        ev.record(e[1], stream)
        mb.append(ev.ms(e[0], e[1]))
        ev.record(e[0], stream)
        ctx.ntt_pass(cw, t, t, k, 4, 3)  # layers 4..6, an interior 3-layer pass
        ev.record(e[1], stream)
        pt.append(ev.ms(e[0], e[1]))
    ntt_ms, merkle_ms, merkle_b3_ms, pass_ms = (float(np.mean(x)) for x in (enc, mt, mb, pt))
    ntt_bytes = 16 * ((t << (k - 1)) + s["cw_len"])
    ntt_gbs = ntt_bytes / (ntt_ms * 1e-3) / 1e9
    pass_bytes = 32 * s["cw_len"]  # an in-place pass reads and writes the codeword once
    pass_gbs = pass_bytes / (pass_ms * 1e-3) / 1e9
    muls = (k - 1) * s["cw_len"] // 2

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "ntt_pass_traffic.json")
    if os.path.exists(tpath):
        try:
            with open(tpath) as fh:
                tj = json.load(fh)
            if (tj.get("kernel") == "k_ntt_pass<3>" and tj.get("cw_words") == s["cw_len"]
                    and tj.get("lanes") == t):
                traffic = tj.get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

    result = {
        "metric": METRIC,
        "value": round(ntt_gbs, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "gf2_128",
        "data": "synthetic (seeded uniform SHA-256 message blocks, IV state; synthetic CSC "
                "with the SHA-256 circuit's nonzero counts)",
        "config": {
            "workload": (f"SHA-256 x 2^{nl} prove core (witness, dense stack, lane-major commit "
                         f"m={m} rate 1/2 {t} lanes k_code={k}, zerocheck m={m_zc}, lincheck "
                         f"K=2^15 nu={nl}); value = commit NTT GB/s"),
            "batch": n, "m_commit": m, "log_inv_rate": 1, "log_batch_size": lb, "lanes": t,
            "msg_words": t << (k - 1), "codeword_words": s["cw_len"], "leaves": s["n_leaves"],
        },
        "stages_ms": {kk: round(v, 4) for kk, v in stage_ms.items()},
        "ntt_ms": round(ntt_ms, 4),
        "merkle_ms": round(merkle_ms, 4),
        "ntt_compulsory_frac_of_hbm": round(ntt_gbs / HBM_PEAK_GBS, 5),
        "ntt_f128_muls_per_s": round(muls / (ntt_ms * 1e-3), 1),
        "merkle_leaf_gbs": round(16 * s["cw_len"] / (merkle_ms * 1e-3) / 1e9, 3),
        "merkle_blake3_ms": round(merkle_b3_ms, 4),
        "merkle_blake3_leaf_gbs": round(16 * s["cw_len"] / (merkle_b3_ms * 1e-3) / 1e9, 3),
        "roofline": {
            "bound": "hbm", "achieved": round(pass_gbs, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(pass_gbs / HBM_PEAK_GBS, 5), "traffic": traffic,
            "kernel": "k_ntt_pass<3>", "launch_ms": round(pass_ms, 4),
            "bytes_per_launch": pass_bytes,
        },
        "timed_wall_s": round(wall, 3),
    }
    if not args.ntt_only:
        core = sum(stage_ms.values())
        result["prove_core_ms"] = round(core, 3)
        result["prove_core_khps"] = round(n / core, 2)  # k hashes/s = hashes per ms
        result["prove_core_note"] = ("witness+compact+commit+zerocheck+lincheck; excludes the "
                                     "PCS opening and the registry digest (not a full proof)")

    if not args.ntt_only and args.latency_requests > 0:
        # one request = one prove-core step on inputs already resident in HBM, timed from
        # submission to the last host-visible result (the lincheck's final message)
        lat = []
        for _ in range(args.latency_requests):
            ctx.sync()
            t1 = time.perf_counter()
            step({})
            ctx.sync()
            lat.append((time.perf_counter() - t1) * 1e3)
        result["prove_core_latency_ms"] = {
            "p50": round(float(np.percentile(lat, 50)), 4),
            "p99": round(float(np.percentile(lat, 99)), 4),
            "requests": len(lat),
        }

    if args.copy_calibration:
        nbytes = 1 << 30
        src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
        ev.h.hipMemcpyDtoD.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        ctx.sync()
        assert ev.h.hipMemcpyDtoD(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), nbytes) == 0
        ctx.sync()
        result["copy_calibration_bytes"] = nbytes
        src.free()
        dst.free()

    if not args.no_sumcheck and not args.ntt_only:
        for buf in (z, a, b, stripe, q):
            buf.free()
        log_n = 24
        nn = 1 << log_n
        ta = rng.integers(0, 2**63, size=(nn, 2), dtype=np.uint64)
        tb = rng.integers(0, 2**63, size=(nn, 2), dtype=np.uint64)
        da, db = ctx.upload(ta), ctx.upload(tb)
        del ta, tb
        ctx.prove_sumcheck(F.FsChallenger(b"flock-bench-sumcheck"), da, db, log_n)  # warm-up
        st = []
        for _ in range(3):
            ctx.sync()
            t1 = time.perf_counter()
            ctx.prove_sumcheck(F.FsChallenger(b"flock-bench-sumcheck"), da, db, log_n)
            st.append(time.perf_counter() - t1)
        result["sumcheck_2^24_ms"] = round(min(st) * 1e3, 3)
        result["sumcheck_2^24_algorithmic_gbs"] = round((48 + 56) * 2 * nn / min(st) / 1e9, 3)

    if not args.ntt_only:
        result["config1_commit"] = config1_commit(F, ctx, ev, stream, not args.no_cpu_baseline)

    if not args.no_mixed and not args.ntt_only:
        result.update(mixed_workloads(F, ctx, args))

    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(t, k - 1, reps=3)

    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
