"""The oracle's `prove_fast` for a registry of two boolean hash types against the reference's
mixed-union anchors (crates/flock-prover/tests/union_m6_fixtures.rs:196-284).

The registry is BLAKE3 + SHA-256 at nu = 10 (the `Blake3Sha2Nu10` tier); slots are sorted by
k_log (SHA-256 first); the inputs are the fixture's SplitMix64 draws (SHA-256 rows, then
BLAKE3 rows, seed 0x4D36_0000 ^ (n_sha2 << 16) ^ n_blake3); the digest is SHA-256 over
`MixedProofBundleLigerito::to_bytes()` and the two claim values. These pin the multi-slot
union lincheck (per-slot combs, prefix weights, pins, one matrix-eval pair per type), the
union witness placement, the jagged heights of several slots and the mixed wire format.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as O
import oracle_prove as OP

DOMAIN = b"flock-m6-fixture-v0"
ANCHORS = {
    (8, 8): "a5a4c478c8e33fc05ceef8be011075f6618f73cc1e02cedeb18cc938b19acff8",
    (0, 64): "b34ac2ba4425e805cdc248c17a11ff7437d338493ad4b709933e84c55bcbc04b",
    (50, 37): "ed988c189e253a03568b69a208d0527034a75754fa6fbed87b8c17b05483391f",
    (1024, 1024): "44e50df0ac2f4282e8a94e9ce1eb4ef094fda6d7556795c9ba174514b113d3cb",
}


def fixture_slots(n_sha2, n_blake3):
    rng = OP.SplitMix32(0x4D36_0000 ^ (n_sha2 << 16) ^ n_blake3)
    sha = OP.fixture_sha2_inputs(rng, n_sha2)
    b3 = OP.fixture_blake3_inputs(rng, n_blake3)
    return [(OP.HASH_BLAKE3, b3), (OP.HASH_SHA256, sha)]


def mixed_digest(slots):
    proof, cap, ab, c, info = OP.prove_fast_union(slots, 10, DOMAIN)
    bundle = OP.mixed_bundle_bytes(2, info["counts"], cap, info["dense_m"], info["log_batch_size"],
                                   info["lanes"], proof)
    return OP.mixed_anchor_digest(bundle, ab, c)


@pytest.mark.parametrize("counts", [(8, 8), (50, 37)])
def test_oracle_reproduces_the_mixed_union_anchor(counts):
    assert mixed_digest(fixture_slots(*counts)) == ANCHORS[counts]


def test_union_prover_of_one_slot_is_the_single_slot_prover():
    # the single-slot anchor through the union path (one slot filling the boolean region)
    rng = OP.SplitMix32(0x4D36_5252)
    inputs = OP.fixture_sha2_inputs(rng, 128)
    proof, cap, ab, c, _ = OP.prove_fast_union([(OP.HASH_SHA256, inputs)], 7, DOMAIN)
    assert OP.anchor_digest(proof, cap, ab, c) == \
        "d4686311e9887d778c00b7e1b93ddddabb3bf10db74ca9bfdac75f8db55e0b80"
