"""flock_zerocheck_verify (zerocheck::verify_with_grinding, zerocheck.rs:1032-1225) on the CPU
against proofs the oracle prover writes.

- An honest proof verifies. The verifier's claim equals the prover's, and its transcript ends
  in the prover's state.
- Each single-bit change in any part of the proof is rejected.
- A proof over a witness with a*b != c is rejected (zerocheck.rs:1372-1450 test the same
  properties).
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle_lib as O
from flock_amd import lib as F
from test_oracle_zerocheck import L, _witness

DOM = b"flock-zc-verify-test"


def prove(a, b, c, m, grinding):
    """The oracle's zerocheck proof and its transcript digest after proving."""
    n_mlv = m - 6
    fs = O.Fs(DOM)
    r1 = np.zeros((128, 2), dtype=np.uint64)
    ml = np.zeros((2 * n_mlv, 2), dtype=np.uint64)
    fin = np.zeros((3, 2), dtype=np.uint64)
    nonces = np.zeros(2 + n_mlv, dtype=np.uint64)
    claim = np.zeros((1 + 2 * n_mlv, 2), dtype=np.uint64)
    assert L.or_zerocheck_prove(fs.h, a.ctypes.data, b.ctypes.data, c.ctypes.data, m,
                                int(grinding), O._p(r1), O._p(ml), O._p(fin),
                                nonces.ctypes.data, O._p(claim)) == 0
    dig = ctypes.create_string_buffer(32)
    L.or_fs_digest.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.or_fs_digest(fs.h, dig)
    return (r1, ml, fin, nonces, claim), dig.raw


def verify(proof, m, grinding):
    r1, ml, fin, nonces, _ = proof
    fs = F.FsChallenger(DOM)
    claim = F.zerocheck_verify(fs, m, grinding, r1, ml, fin, nonces if grinding else None)
    return claim, fs.digest()


@pytest.mark.parametrize("m,grinding", [(13, False), (13, True), (16, True), (18, False)])
def test_honest_proof_verifies(m, grinding):
    a, b, c = _witness(m, 700 + m, satisfying=True)
    proof, dig = prove(a, b, c, m, grinding)
    claim, vdig = verify(proof, m, grinding)
    assert np.array_equal(claim, proof[4])
    assert vdig == dig


def _rejected(proof, m, grinding):
    try:
        verify(proof, m, grinding)
    except F.FlockHipError as e:
        assert e.code == F.E_REJECT, e
        return True
    return False


def test_every_part_of_the_proof_is_bound():
    m, grinding = 14, True
    a, b, c = _witness(m, 714, satisfying=True)
    proof, _ = prove(a, b, c, m, grinding)
    # (part index, row) of each tampered word: round1_ab, round1_c, first/last multilinear
    # messages, the three finals, and the initial, skip and a round nonce
    sites = [(0, 3), (0, 100), (1, 0), (1, 2 * (m - 6) - 1), (2, 0), (2, 1), (2, 2)]
    for part, row in sites:
        bad = [x.copy() for x in proof]
        bad[part][row, 0] ^= np.uint64(1 << 17)
        assert _rejected(bad, m, grinding), (part, row)
    for k in (0, 1, 5):
        bad = [x.copy() for x in proof]
        # a nonce that fails its proof of work, or, if it passes, changes every later
        # challenge
        bad[3][k] ^= np.uint64(0x5a5a)
        assert _rejected(bad, m, grinding), ("nonce", k)


def test_dishonest_witness_is_rejected():
    # c != a*b: the prover's round-1 polynomial is not zero on S
    m = 14
    a, b, c = _witness(m, 715, satisfying=False)
    proof, _ = prove(a, b, c, m, False)
    assert _rejected(proof, m, False)


def test_verify_pow_matches_grind():
    fs_p, fs_v = F.FsChallenger(b"pow"), F.FsChallenger(b"pow")
    for bits in (0, 1, 5, 9):
        n = fs_p.grind_pow(bits)
        assert fs_v.verify_pow(n, bits)
    assert fs_p.digest() == fs_v.digest()
    # zero bits accept only the canonical nonce 0; a nonce that fails is still absorbed
    fs = F.FsChallenger(b"pow")
    assert not fs.verify_pow(3, 0)
    fs2 = F.FsChallenger(b"pow")
    fs2.observe_bytes((3).to_bytes(8, "little"))
    assert fs.digest() == fs2.digest()
