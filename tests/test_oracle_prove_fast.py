"""The oracle's whole prove_fast against the reference's own byte-identity anchors.

crates/flock-prover/tests/union_m6_fixtures.rs:289-323 pins the SHA-256 digest of
`bincode(R1csProofMergedLigerito) || commitment cap || claim values` for two single-slot
`prove_fast` runs over SplitMix64-seeded inputs in the FS domain `flock-m6-fixture-v0`:

- BLAKE3, 256 compressions (m = 22), seed 0x4D36_B3B3 -> 8bd3451e...
- SHA-256, 128 compressions (m = 22), seed 0x4D36_5252 -> d4686311...

Reproducing them fixes every part of the oracle at once: the batch-major witness bit layouts,
the registry digest, the transcript schedule (labels, observe order, PoW bits), the zerocheck,
the union lincheck, the merged opening (ring switch, jagged merged sumcheck, multipoint
twisted assist, F256 Ligerito ladder with stratified queries and capped paths) and the wire
encoding. The jagged evaluator the opening uses is checked separately against its definition.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle_lib as O
import oracle_prove as OP

DOMAIN = b"flock-m6-fixture-v0"


@pytest.mark.parametrize("kind,n,seed,expected", [
    (OP.HASH_BLAKE3, 256, 0x4D36_B3B3,
     "8bd3451eaf94fb8ad433c094bb4a1375f856565546b26fdf303d4a0b48a8325c"),
    (OP.HASH_SHA256, 128, 0x4D36_5252,
     "d4686311e9887d778c00b7e1b93ddddabb3bf10db74ca9bfdac75f8db55e0b80"),
])
def test_single_slot_anchor_digest(kind, n, seed, expected):
    rng = OP.SplitMix32(seed)
    if kind == OP.HASH_BLAKE3:
        inputs = OP.fixture_blake3_inputs(rng, n)
    else:
        inputs = OP.fixture_sha2_inputs(rng, n)
    proof, cap, ab, c = OP.prove_fast(kind, inputs, DOMAIN)
    assert len(cap) == 32 << 8  # the m22 L0 cap: 2^8 nodes (279 queries -> depth 8)
    assert OP.anchor_digest(proof, cap, ab, c) == expected


def test_one_bit_changes_the_digest():
    """A planted mistake in the oracle's inputs moves the digest (the anchor is not vacuous)."""
    rng = OP.SplitMix32(0x4D36_5252)
    inputs = OP.fixture_sha2_inputs(rng, 128)
    inputs[77, 3] ^= 1 << 9
    proof, cap, ab, c = OP.prove_fast(OP.HASH_SHA256, inputs, DOMAIN)
    assert OP.anchor_digest(proof, cap, ab, c) != \
        "d4686311e9887d778c00b7e1b93ddddabb3bf10db74ca9bfdac75f8db55e0b80"


def test_jagged_fhat_is_the_mle_of_the_jagged_weight():
    """or_jag_fhat (branching program, jagged.rs:199-262) equals the brute-force MLE
    sum_{d < area} eq(x, d) eq(z_row, row(d)) eq(z_col, col(d)) (jagged.rs:4782-4800)."""
    rng = np.random.default_rng(5)
    n, k, m = 3, 3, 6
    heights = np.array([8, 8, 3, 0, 8, 5, 0, 0], dtype=np.uint64)
    ps = np.concatenate([[0], np.cumsum(heights)]).astype(np.int64)
    rand = lambda c: [int(x) | (int(y) << 64) for x, y in rng.integers(0, 2**63, (c, 2))]
    z_row, z_col, x = rand(n), rand(k), rand(m)
    er, ec, ex = O.build_eq(z_row), O.build_eq(z_col), O.build_eq(x)
    er_i, ec_i, ex_i = O.arr_to_ints(er), O.arr_to_ints(ec), O.arr_to_ints(ex)
    want = 0
    for col in range(1 << k):
        for d in range(ps[col], ps[col + 1]):
            want ^= O.mul(ex_i[d], O.mul(er_i[d - ps[col]], ec_i[col]))
    out = np.zeros((1, 2), dtype=np.uint64)
    zr, zc, xx = O.ints_to_arr(z_row), O.ints_to_arr(z_col), O.ints_to_arr(x)
    assert O.lib().or_jag_fhat(heights.ctypes.data, len(heights), n, m, O._p(zr), O._p(zc),
                               O._p(xx), O._p(out)) == 0
    assert O.arr_to_ints(out)[0] == want
