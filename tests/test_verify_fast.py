"""The whole-proof verifier (flock_verify_fast / flock_verify_fast_union, host only) against
proofs the oracle writes byte-identical to the reference's prover:

- the reference's own anchors (crates/flock-prover/tests/union_m6_fixtures.rs:289-323; their
  digests are re-checked here, so the bytes verified are the reference's): SHA-256 x 128 and
  BLAKE3 x 256 are accepted and the claims returned equal the prover's;
- a Keccak-f[1600] proof (walker lincheck) and a two-type SHA-256 + BLAKE3 union are
  accepted;
- one flipped bit in each field of the wire format the verifier binds (zerocheck, lincheck,
  ring switch, merged sumcheck, q_eval, the multipoint values and rounds, the Frobenius
  anchor, the Ligerito caps, rows, paths, messages, OOD values and yr) is rejected, a PoW
  nonce is rejected once it fails its leading-zero check, and a proof checked against
  another statement (count, domain, cap, kind) is rejected.

The lincheck's reported matrix evaluations are the one field the full verifier does not read:
MatrixAssertion::check recomputes the combs from the circuit (lincheck/union.rs:579-628), so
those bytes are left out of the flips, as the reference leaves them unchecked.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_prove as OP
from flock_amd import lib as F

DOMAIN = b"flock-m6-fixture-v0"
SHA_ANCHOR = "d4686311e9887d778c00b7e1b93ddddabb3bf10db74ca9bfdac75f8db55e0b80"
B3_ANCHOR = "8bd3451eaf94fb8ad433c094bb4a1375f856565546b26fdf303d4a0b48a8325c"


def layout(proof: bytes) -> dict:
    """Byte ranges of the R1csProofMergedLigerito fields (proof.rs:42-47; bincode 1.3)."""
    off = 0
    out = {}

    def u64():
        nonlocal off
        v = int.from_bytes(proof[off:off + 8], "little")
        off += 8
        return v

    def vec(name, elem):
        nonlocal off
        k = u64()
        out[name] = (off, off + k * elem)
        off += k * elem
        return k

    def fixed(name, n):
        nonlocal off
        out[name] = (off, off + n)
        off += n

    vec("zc_round1_ab", 16)
    vec("zc_round1_c", 16)
    vec("zc_ml", 32)
    fixed("zc_finals", 48)
    vec("zc_nonces", 8)
    vec("lc_rounds", 32)
    vec("lc_z_partial", 16)
    vec("lc_matrix_evals", 32)
    vec("lc_nonces", 8)
    n_rs = u64()
    for i in range(n_rs):
        vec(f"rs{i}_s_hat_v", 16)
        fixed(f"rs{i}_nonce", 8)
    vec("batch_nonce", 8)
    vec("merged_rounds", 32)
    vec("merged_nonces", 8)
    fixed("q_eval", 16)
    for i in range(u64()):
        vec(f"mp_values{i}", 16)
    vec("mp_group_values", 16)
    fixed("mp_gamma_nonce", 8)
    vec("mp_rounds", 32)
    vec("mp_nonces", 8)
    fixed("assist_v", 16)
    vec("assist_rounds", 32)
    vec("assist_nonces", 8)
    u64()  # inner ring switches (none)
    u64()  # inner batching nonces (none)
    vec("lig_initial_cap", 32)

    def opened(name):
        nonlocal off
        start = off
        for _ in range(u64()):
            vec(name + "_row", 16)
        out[name + "_rows"] = (start, off)
        vec(name + "_paths", 32)

    opened("lig_initial")
    start = off
    for _ in range(u64()):
        vec("lig_cap", 32)
    out["lig_caps"] = (start, off)
    start = off
    for _ in range(u64()):
        opened("lig_rec")
    out["lig_recursive"] = (start, off)
    vec("lig_yr", 16)
    opened("lig_final")
    vec("lig_sumcheck_f128", 32)
    vec("lig_msgs", 64)
    vec("lig_grind_nonces", 8)
    vec("lig_ood", 16)
    vec("lig_fold_nonces", 8)
    vec("lig_claim_nonces", 8)
    vec("lig_cons_nonces", 8)
    assert off == len(proof), (off, len(proof))
    return out


@pytest.fixture(scope="module")
def sha_proof():
    rng = OP.SplitMix32(0x4D36_5252)
    x = OP.fixture_sha2_inputs(rng, 128)
    proof, cap, ab, c = OP.prove_fast(OP.HASH_SHA256, x, DOMAIN)
    assert OP.anchor_digest(proof, cap, ab, c) == SHA_ANCHOR
    return proof, cap, ab, c


def rejects(fn, *a, **k):
    with pytest.raises(F.FlockHipError) as e:
        fn(*a, **k)
    assert e.value.code == F.E_REJECT, str(e.value)
    return str(e.value)


def test_the_reference_sha256_anchor_proof_verifies(sha_proof):
    proof, cap, ab, c = sha_proof
    assert F.verify_fast(F.HASH_SHA256, 7, 128, cap, proof, domain=DOMAIN) == (ab, c)


def test_every_bound_field_of_the_wire_format_is_checked(sha_proof):
    proof, cap, _, _ = sha_proof
    lay = layout(proof)
    skip = {"lc_matrix_evals", "mp_group_values", "batch_nonce", "lig_sumcheck_f128",
            "lig_fold_nonces", "lig_caps", "lig_recursive", "lig_initial_rows",
            "lig_rec_rows", "lig_final_rows"}
    seen = 0
    for name, (a, b) in sorted(lay.items(), key=lambda kv: kv[1]):
        if name in skip or b <= a:
            continue
        if name.endswith("nonces") or name.endswith("nonce"):
            # a PoW nonce is checked for its leading zero bits only, so another nonce that
            # passes is another valid proof (the honest rounds hold at any challenge); of
            # eight altered nonces at least one must fail its PoW and be rejected
            out = []
            for t in range(1, 9):
                bad = bytearray(proof)
                bad[a] ^= t
                try:
                    F.verify_fast(F.HASH_SHA256, 7, 128, cap, bytes(bad), domain=DOMAIN)
                    out.append(False)
                except F.FlockHipError as e:
                    assert e.code == F.E_REJECT
                    out.append(True)
            assert any(out), name
            seen += 1
            continue
        for pos in sorted({a, (a + b) // 2, b - 1}):
            bad = bytearray(proof)
            bad[pos] ^= 0x01
            rejects(F.verify_fast, F.HASH_SHA256, 7, 128, cap, bytes(bad), domain=DOMAIN)
            seen += 1
    assert seen > 60
    # the matrix evals are carried but not read by the full check (as in the reference)
    a, b = lay["lc_matrix_evals"]
    bad = bytearray(proof)
    bad[a] ^= 0x01
    assert F.verify_fast(F.HASH_SHA256, 7, 128, cap, bytes(bad), domain=DOMAIN)
    # trailing bytes are rejected
    rejects(F.verify_fast, F.HASH_SHA256, 7, 128, cap, proof + b"\0", domain=DOMAIN)


def test_a_proof_does_not_verify_another_statement(sha_proof):
    proof, cap, _, _ = sha_proof
    rejects(F.verify_fast, F.HASH_SHA256, 7, 127, cap, proof, domain=DOMAIN)
    rejects(F.verify_fast, F.HASH_SHA256, 7, 128, cap, proof, domain=b"another-domain")
    bad = bytearray(cap)
    bad[7] ^= 0x80
    rejects(F.verify_fast, F.HASH_SHA256, 7, 128, bytes(bad), proof, domain=DOMAIN)
    with pytest.raises(F.FlockHipError):
        F.verify_fast(F.HASH_BLAKE3, 7, 128, cap, proof, domain=DOMAIN)


def test_the_reference_blake3_anchor_proof_verifies():
    rng = OP.SplitMix32(0x4D36_B3B3)
    x = OP.fixture_blake3_inputs(rng, 256)
    proof, cap, ab, c = OP.prove_fast(OP.HASH_BLAKE3, x, DOMAIN)
    assert OP.anchor_digest(proof, cap, ab, c) == B3_ANCHOR
    assert F.verify_fast(F.HASH_BLAKE3, 8, 256, cap, proof, domain=DOMAIN) == (ab, c)


def test_a_keccak_proof_verifies():
    rng = np.random.default_rng(77)
    x = rng.integers(0, 1 << 63, (50, 25), dtype=np.uint64)
    proof, cap, ab, c = OP.prove_fast(OP.HASH_KECCAK, x, DOMAIN, n_log=6)
    assert F.verify_fast(F.HASH_KECCAK, 6, 50, cap, proof, domain=DOMAIN) == (ab, c)
    bad = bytearray(proof)
    bad[len(proof) // 3] ^= 0x04
    rejects(F.verify_fast, F.HASH_KECCAK, 6, 50, cap, bytes(bad), domain=DOMAIN)


def test_a_two_type_union_proof_verifies():
    rng = OP.SplitMix32(0x5151)
    sha = OP.fixture_sha2_inputs(rng, 100)
    b3 = OP.fixture_blake3_inputs(rng, 50)
    proof, cap, ab, c, info = OP.prove_fast_union([(OP.HASH_BLAKE3, b3), (OP.HASH_SHA256, sha)], 7,
                                                  DOMAIN)
    parts = [(F.HASH_BLAKE3, 50), (F.HASH_SHA256, 100)]
    assert F.verify_fast_union(parts, 7, cap, proof, domain=DOMAIN) == (ab, c)
    # the counts bind: swapping them is another statement
    rejects(F.verify_fast_union, [(F.HASH_BLAKE3, 100), (F.HASH_SHA256, 50)], 7, cap, proof,
            domain=DOMAIN)
    bad = bytearray(proof)
    bad[len(proof) // 2] ^= 0x02
    rejects(F.verify_fast_union, parts, 7, cap, bytes(bad), domain=DOMAIN)
