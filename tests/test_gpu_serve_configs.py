"""The batched and served workloads as whole proofs on the built-in circuits.

- Batched (64 uneven requests at once): hash kind uniform over SHA-256 / BLAKE3 /
  Keccak-f[1600], batch log2 uniform in [10, 16], seeded inputs, eight workers through
  flock_hip_serve. Every proof equals, byte for byte, the same request proved alone; the
  smallest SHA-256 and BLAKE3 requests also equal the oracle's prove_fast, and the library's
  verifier accepts every proof.
- Served (Poisson arrivals): 24 requests, batch log2 in [10, 18] with a Keccak-f[1600] 2^18
  request in the middle, admitted by HBM footprint against 288 GB; every proof equals its
  single-request run.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_prove as OP
from flock_amd import lib as F

pytestmark = pytest.mark.gpu
DOMAIN = b"flock-readme-bench-v0"
IV = [0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A, 0x510E527F, 0x9B05688C, 0x1F83D9AB,
      0x5BE0CD19]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def inputs(kind, n, rng):
    if kind == F.HASH_SHA256:
        x = np.zeros((n, 24), dtype=np.uint32)
        x[:, :8] = IV
        x[:, 8:] = rng.integers(0, 1 << 32, (n, 16), dtype=np.uint64).astype(np.uint32)
    elif kind == F.HASH_BLAKE3:
        x = rng.integers(0, 1 << 32, (n, 28), dtype=np.uint64).astype(np.uint32)
        x[:, 26], x[:, 27] = 64, 11
    else:
        x = rng.integers(0, 1 << 63, (n, 25), dtype=np.uint64) * 2 + \
            rng.integers(0, 2, (n, 25), dtype=np.uint64)
    return x


def specs(seed, n_req, lo, hi, rate=0.0, big=None):
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    for i in range(n_req):
        kind, nl = int(rng.integers(0, 3)), int(rng.integers(lo, hi + 1))
        if big is not None and i == n_req // 2:
            kind, nl = big
        if rate:
            t += float(rng.exponential(1.0 / rate))
        out.append((kind, nl, inputs(kind, 1 << nl, rng), None, t))
    return out


def alone(reqs):
    pr = F.Prover(0)
    try:
        return [F.prove_fast(pr, r)[1:] for r in reqs]
    finally:
        pr.close()


def test_batched_64_uneven_requests_equal_their_single_runs_and_the_oracle():
    sp = specs(0xC4C5, 64, 10, 16)
    say(f"case: batched 64 requests, {sum(1 << s[1] for s in sp)} hashes, 8 workers")
    reqs, keep = F.make_requests(sp, domain=DOMAIN)
    provers = [F.Prover(0) for _ in range(8)]
    try:
        res, wall, proofs = F.serve(provers, reqs, int(288e9 * 0.9), proofs=True)
    finally:
        for p in provers:
            p.close()
    assert all(r.status == 0 for r in res)
    say(f"batched wall {wall:.1f} ms; proving each alone")
    single = alone(reqs)
    for i, (a, b) in enumerate(zip(proofs, single)):
        assert a == b, f"request {i} (kind {sp[i][0]}, 2^{sp[i][1]}) differs from its single run"
    for kind in (F.HASH_SHA256, F.HASH_BLAKE3):
        i = min((j for j in range(len(sp)) if sp[j][0] == kind), key=lambda j: sp[j][1])
        say(f"case: oracle prove_fast of request {i} (kind {kind}, 2^{sp[i][1]})")
        o = OP.prove_fast(kind, sp[i][2], DOMAIN, n_log=sp[i][1])
        assert proofs[i] == o, f"request {i} differs from the oracle"
    say("verifying every batched proof")
    for i, (proof, cap, ab, c) in enumerate(proofs):
        kind, nl = sp[i][0], sp[i][1]
        assert F.verify_fast(kind, nl, 1 << nl, cap, proof, domain=DOMAIN) == (ab, c), i


def test_served_poisson_with_keccak_2e18_equals_single_runs():
    sp = specs(0x5E7F, 24, 10, 18, rate=60.0, big=(F.HASH_KECCAK, 18))
    say(f"case: served 24 requests with keccak 2^18, {sum(1 << s[1] for s in sp)} hashes")
    reqs, keep = F.make_requests(sp, domain=DOMAIN)
    provers = [F.Prover(0) for _ in range(8)]
    try:
        res, wall, proofs = F.serve(provers, reqs, int(288e9 * 0.9), proofs=True)
    finally:
        for p in provers:
            p.close()
    assert all(r.status == 0 for r in res)
    lat = sorted(r.queued_ms + r.run_ms for r in res)
    say(f"served wall {wall:.1f} ms, p50 {lat[len(lat) // 2]:.1f} ms, max {lat[-1]:.1f} ms")
    single = alone(reqs)
    for i, (a, b) in enumerate(zip(proofs, single)):
        assert a == b, f"request {i} (kind {sp[i][0]}, 2^{sp[i][1]}) differs from its single run"
