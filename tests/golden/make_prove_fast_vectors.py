"""Writes tests/golden/prove_fast_vectors.json: the oracle's whole prove_fast (tests/oracle_prove.py,
pinned by the reference's anchor digests in tests/test_oracle_prove_fast.py) on seeded inputs at
shapes too slow to run inside the GPU tests. Each case stores the seed that regenerates its inputs
(the fixtures' SplitMix64 generator), the SHA-256 of proof || cap || claim values, and the proof
length. Run from the repository root: python tests/golden/make_prove_fast_vectors.py"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [HERE, os.path.dirname(HERE), os.path.dirname(os.path.dirname(HERE))]

import oracle_prove as OP  # noqa: E402

DOMAIN = b"flock-readme-bench-v0"
CASES = [  # (kind, n_decl, seed)
    (OP.HASH_SHA256, 1024, 0xC0DE_0002),   # the 1024-compression SHA-256 batch
    (OP.HASH_BLAKE3, 1000, 0xC0DE_0B3B),   # a partial count: 1000 of 1024 rows
    (OP.HASH_SHA256, 300, 0xC0DE_5A5A),    # dense m 24, partial count
]


def inputs_for(kind, n, seed):
    rng = OP.SplitMix32(seed)
    if kind == OP.HASH_SHA256:
        return OP.fixture_sha2_inputs(rng, n)
    return OP.fixture_blake3_inputs(rng, n)


def main():
    out = {"domain": DOMAIN.decode(), "cases": []}
    for kind, n, seed in CASES:
        proof, cap, ab, c = OP.prove_fast(kind, inputs_for(kind, n, seed), DOMAIN)
        out["cases"].append({"kind": kind, "n_decl": n, "seed": seed,
                             "digest": OP.anchor_digest(proof, cap, ab, c), "proof_len": len(proof)})
        print(out["cases"][-1], flush=True)
    with open(os.path.join(HERE, "prove_fast_vectors.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
