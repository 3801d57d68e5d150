"""The whole prove_fast on the GPU (flock_hip_prove_fast) against the reference's own anchors and
the oracle:

- the reference's byte-identity anchors (crates/flock-prover/tests/union_m6_fixtures.rs:289-323):
  BLAKE3 x 256 and SHA-256 x 128 over the fixtures' SplitMix64 inputs in the FS domain
  `flock-m6-fixture-v0` give the pinned SHA-256 of proof || cap || claim values, here through
  the library's built-in circuits and registry digest;
- the oracle's digests at larger and partial shapes (tests/golden/prove_fast_vectors.json: the
  1024-compression SHA-256 batch, BLAKE3 1000 of 1024 rows, SHA-256 300 of 512 rows);
- byte-for-byte equality with the oracle's proof for Keccak-f[1600] (walker lincheck), whose
  registry carries the empty-stub matrices.
"""
from __future__ import annotations

import json
import os
import sys

import pytest

import oracle_prove as OP
from flock_amd import lib as F
from test_gpu_prove import circuits, ctx, digest_for  # noqa: F401  (fixtures)

pytestmark = pytest.mark.gpu
FIXTURE_DOMAIN = b"flock-m6-fixture-v0"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_prove_fast(kind, inputs, domain, circuits, n_log=None, builtin=False):
    """circuits / digest_for: the oracle's matrices and their registry digest; builtin: the
    library's own circuit and digest (request circuit and digest NULL)."""
    n_log = n_log if n_log is not None else OP.min_n_log(len(inputs))
    circ = circuits[kind][0] if kind != F.HASH_KECCAK and not builtin else None
    dig = None if builtin else digest_for(kind, n_log)
    reqs, keep = F.make_requests([(kind, n_log, inputs, circ, 0.0)], digest=dig, domain=domain)
    pr = F.Prover(0)
    try:
        res, proof, cap, ab, c = F.prove_fast(pr, reqs[0])
    finally:
        pr.close()
    assert res.status == 0
    return proof, cap, ab, c


@pytest.mark.parametrize("kind,n,seed,expected", [
    (F.HASH_BLAKE3, 256, 0x4D36_B3B3,
     "8bd3451eaf94fb8ad433c094bb4a1375f856565546b26fdf303d4a0b48a8325c"),
    (F.HASH_SHA256, 128, 0x4D36_5252,
     "d4686311e9887d778c00b7e1b93ddddabb3bf10db74ca9bfdac75f8db55e0b80"),
])
def test_gpu_reproduces_the_reference_anchor(circuits, kind, n, seed, expected):
    say(f"case: anchor kind={kind} n={n}")
    rng = OP.SplitMix32(seed)
    inputs = OP.fixture_blake3_inputs(rng, n) if kind == F.HASH_BLAKE3 else \
        OP.fixture_sha2_inputs(rng, n)
    # the library's built-in circuit and registry digest: nothing from the oracle
    proof, cap, ab, c = gpu_prove_fast(kind, inputs, FIXTURE_DOMAIN, circuits, builtin=True)
    assert OP.anchor_digest(proof, cap, ab, c) == expected
    n_log = OP.min_n_log(n)
    assert F.verify_fast(kind, n_log, n, cap, proof, domain=FIXTURE_DOMAIN) == (ab, c)


def test_gpu_matches_the_oracle_golden_vectors(circuits):
    path = os.path.join(os.path.dirname(__file__), "golden", "prove_fast_vectors.json")
    with open(path) as fh:
        gold = json.load(fh)
    for case in gold["cases"]:
        say(f"case: golden prove_fast kind={case['kind']} n_decl={case['n_decl']}")
        rng = OP.SplitMix32(case["seed"])
        inputs = OP.fixture_sha2_inputs(rng, case["n_decl"]) if case["kind"] == F.HASH_SHA256 \
            else OP.fixture_blake3_inputs(rng, case["n_decl"])
        proof, cap, ab, c = gpu_prove_fast(case["kind"], inputs, gold["domain"].encode(), circuits)
        assert len(proof) == case["proof_len"]
        assert OP.anchor_digest(proof, cap, ab, c) == case["digest"]


def test_gpu_keccak_proof_bytes_equal_the_oracle(circuits):
    say("case: keccak 64 against the oracle")
    from test_oracle_keccak_r1cs import random_states
    inputs = random_states(64, 0x4B)
    proof, cap, ab, c = gpu_prove_fast(F.HASH_KECCAK, inputs, b"flock-keccak-test", circuits)
    oproof, ocap, oab, oc = OP.prove_fast(OP.HASH_KECCAK, inputs, b"flock-keccak-test")
    assert cap == ocap and (ab, c) == (oab, oc)
    assert len(proof) == len(oproof)
    first = next((i for i in range(len(proof)) if proof[i] != oproof[i]), None)
    assert first is None, f"proof bytes differ from offset {first}"


def test_shapes_below_the_m22_configs_are_rejected(circuits):
    # the reference has no Ligerito config below m = 22 (prover_config_for, ligerito.rs:590-618)
    from test_gpu_prove import inputs_for
    with pytest.raises(F.FlockHipError) as e:
        gpu_prove_fast(F.HASH_SHA256, inputs_for(F.HASH_SHA256, 16, 1), b"x", circuits)
    assert e.value.code == F.E_INVALID


@pytest.mark.parametrize("setting", [
    {"ntt_cols_per_thread": 1, "ntt_deep_cols": 32, "zc_split_min_vars": 10, "zc_la_max_log": 10,
     "sc_la_max_log": 10, "open_split_fold_log": 10},
    {"ntt_cols_per_thread": 4, "ntt_deep_cols": 64, "zc_split_min_vars": 24, "zc_la_max_log": 22,
     "sc_la_max_log": 24, "open_split_fold_log": 22},
])
def test_tunables_keep_the_anchor_bytes(circuits, setting):
    """The launch tunables (flock_hip_set_tunable) at the ends of their ranges change launch
    shapes and thresholds only: the SHA-256 x 128 anchor and the oracle's 1024-compression
    golden proof are unchanged."""
    say(f"case: tunables {setting}")
    defaults = {k: F.get_tunable(k) for k in F.TUNABLES}
    try:
        for k, v in setting.items():
            F.set_tunable(k, v)
        inputs = OP.fixture_sha2_inputs(OP.SplitMix32(0x4D36_5252), 128)
        proof, cap, ab, c = gpu_prove_fast(F.HASH_SHA256, inputs, FIXTURE_DOMAIN, circuits,
                                           builtin=True)
        assert OP.anchor_digest(proof, cap, ab, c) == \
            "d4686311e9887d778c00b7e1b93ddddabb3bf10db74ca9bfdac75f8db55e0b80"
        path = os.path.join(os.path.dirname(__file__), "golden", "prove_fast_vectors.json")
        with open(path) as fh:
            gold = json.load(fh)
        case = gold["cases"][0]  # SHA-256 x 1024 (dense m 25)
        inputs = OP.fixture_sha2_inputs(OP.SplitMix32(case["seed"]), case["n_decl"])
        proof, cap, ab, c = gpu_prove_fast(case["kind"], inputs, gold["domain"].encode(), circuits)
        assert OP.anchor_digest(proof, cap, ab, c) == case["digest"]
    finally:
        for k, v in defaults.items():
            F.set_tunable(k, v)
