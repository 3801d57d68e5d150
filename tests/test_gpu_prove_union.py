"""The whole prove_fast over a registry of several boolean hash types on the GPU
(flock_hip_prove_fast_union) against the reference's mixed-union anchors and the oracle:

- the reference's four mixed anchors (crates/flock-prover/tests/union_m6_fixtures.rs:196-284):
  BLAKE3 + SHA-256 at nu = 10 (the Blake3Sha2Nu10 registry, code 2) over the fixtures'
  SplitMix64 inputs; the digest is SHA-256 over MixedProofBundleLigerito::to_bytes() and the
  two claim values, here with the library's built-in circuits and registry digest. These pin
  the GPU union lincheck (per-slot combs, prefix weights, pins, one matrix-eval pair per type),
  the witness placement of several slots, the jagged heights and the counts in slot order;
  the library's whole-proof verifier accepts each;
- a union of one part equals flock_hip_prove_fast (the single-slot SHA-256 x 128 anchor);
- Keccak-f[1600] + SHA-256 + BLAKE3 at nu = 7 (three slots, the Keccak walker among them, a
  gap-free and a padded region) equals the oracle's prove_fast_union byte for byte.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_prove as OP
from flock_amd import lib as F

pytestmark = pytest.mark.gpu
FIXTURE_DOMAIN = b"flock-m6-fixture-v0"
MIXED = {
    (8, 8): "a5a4c478c8e33fc05ceef8be011075f6618f73cc1e02cedeb18cc938b19acff8",
    (0, 64): "b34ac2ba4425e805cdc248c17a11ff7437d338493ad4b709933e84c55bcbc04b",
    (50, 37): "ed988c189e253a03568b69a208d0527034a75754fa6fbed87b8c17b05483391f",
    (1024, 1024): "44e50df0ac2f4282e8a94e9ce1eb4ef094fda6d7556795c9ba174514b113d3cb",
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def prover():
    pr = F.Prover(0)
    yield pr
    pr.close()


def prove_union(prover, parts, nu, domain):
    req, keep = F.union_request(parts, nu, domain=domain)
    info = F.union_info(req)
    res, proof, cap, ab, c = F.prove_fast_union(prover, req)
    assert res.status == 0
    return proof, cap, ab, c, info


@pytest.mark.parametrize("counts", list(MIXED))
def test_gpu_reproduces_the_mixed_union_anchor(prover, counts):
    n_sha2, n_blake3 = counts
    say(f"case: mixed anchor sha2={n_sha2} blake3={n_blake3} nu=10")
    rng = OP.SplitMix32(0x4D36_0000 ^ (n_sha2 << 16) ^ n_blake3)
    sha = OP.fixture_sha2_inputs(rng, n_sha2)
    b3 = OP.fixture_blake3_inputs(rng, n_blake3)
    # the registry lists BLAKE3 then SHA-256; Registry::new puts SHA-256 (k_log 15) first
    proof, cap, ab, c, info = prove_union(
        prover, [(F.HASH_BLAKE3, b3, n_blake3), (F.HASH_SHA256, sha, n_sha2)], 10, FIXTURE_DOMAIN)
    assert info["slot_part"] == [1, 0] and info["counts"] == [n_sha2, n_blake3]
    bundle = OP.mixed_bundle_bytes(2, info["counts"], cap, info["dense_m"],
                                   info["log_batch_size"], info["lanes"], proof)
    assert OP.mixed_anchor_digest(bundle, ab, c) == MIXED[counts]
    # and the library's verifier accepts it (verify_ligerito_union)
    assert F.verify_fast_union([(F.HASH_BLAKE3, n_blake3), (F.HASH_SHA256, n_sha2)], 10, cap,
                               proof, domain=FIXTURE_DOMAIN) == (ab, c)


def test_gpu_union_of_one_part_is_prove_fast(prover):
    say("case: union of one SHA-256 x 128 part, nu=7")
    rng = OP.SplitMix32(0x4D36_5252)
    inputs = OP.fixture_sha2_inputs(rng, 128)
    proof, cap, ab, c, _ = prove_union(prover, [(F.HASH_SHA256, inputs)], 7, FIXTURE_DOMAIN)
    assert OP.anchor_digest(proof, cap, ab, c) == \
        "d4686311e9887d778c00b7e1b93ddddabb3bf10db74ca9bfdac75f8db55e0b80"


def test_gpu_three_type_union_equals_the_oracle(prover):
    say("case: keccak x 40 + sha2 x 100 + blake3 x 50, nu=7")
    rng = OP.SplitMix32(0x3333)
    sha = OP.fixture_sha2_inputs(rng, 100)
    b3 = OP.fixture_blake3_inputs(rng, 50)
    kc = np.array([[rng.next_u32() | (rng.next_u32() << 32) for _ in range(25)]
                   for _ in range(40)], dtype=np.uint64)
    parts = [(F.HASH_KECCAK, kc), (F.HASH_SHA256, sha), (F.HASH_BLAKE3, b3)]
    domain = b"flock-union-test"
    proof, cap, ab, c, info = prove_union(prover, parts, 7, domain)
    assert info["counts"] == [128, 100, 50]
    say("oracle prove_fast_union")
    o_proof, o_cap, o_ab, o_c, o_info = OP.prove_fast_union(parts, 7, domain)
    assert (info["dense_m"], info["log_batch_size"], info["lanes"]) == \
        (o_info["dense_m"], o_info["log_batch_size"], o_info["lanes"])
    assert cap == o_cap
    assert (ab, c) == (o_ab, o_c)
    assert proof == o_proof
    assert F.verify_fast_union([(F.HASH_KECCAK, 40), (F.HASH_SHA256, 100), (F.HASH_BLAKE3, 50)],
                               7, cap, proof, domain=domain) == (ab, c)
