"""The library's built-in hash circuits (flock_circuit_get) against the oracle's restated walks.

The product builds (A_0, B_0) with its own word algebra (flock_amd/csrc/circuits.hip); the oracle
restates sha2.rs:548-701 and blake3.rs:659-783 separately (oracle/sha2_r1cs.c,
oracle/blake3_r1cs.c). Both matrices enter the reference-pinned proofs: the lincheck's matrix
evaluations (lincheck.rs:520-549) and the registry digest (schedule.rs:542-580) are inside the
byte-identity anchors of union_m6_fixtures.rs:289-323, which tests/test_gpu_prove_fast.py
reproduces with the built-in circuits. Host only: no GPU call.
"""
from __future__ import annotations

import numpy as np
import pytest

from flock_amd import lib as F


@pytest.fixture(scope="module")
def oracle_matrices():
    import test_oracle_blake3_r1cs as B3
    import test_oracle_sha2_r1cs as S2
    return {F.HASH_SHA256: S2.matrices(), F.HASH_BLAKE3: B3.matrices()}


@pytest.mark.parametrize("kind,k_log,useful,pin,nnz", [
    # sha2.rs:189-194 (USEFUL_BITS 25,470); 47.4 M template nnz (sha2.rs:203-212)
    (F.HASH_SHA256, 15, 25_470, 25_469, (28_839_569, 18_527_069)),
    # blake3.rs:246-247 (USEFUL_BITS 11,707); 48.3 M nnz (blake3.rs:38-44 rounds it)
    (F.HASH_BLAKE3, 14, 11_707, 11_706, (27_251_212, 17_191_286)),
])
def test_builtin_circuit_equals_the_oracle(oracle_matrices, kind, k_log, useful, pin, nnz):
    c = F.Circuit(kind)
    assert (c.k_log, c.useful_bits, c.const_pin, c.nnz) == (k_log, useful, pin, nnz)
    (ap, ac), (bp, bc), csc_a, csc_b = oracle_matrices[kind]
    for which, (ptr, cols) in enumerate(((ap, ac), (bp, bc))):
        gp, gc = c.csr(which)
        assert np.array_equal(gp, ptr) and np.array_equal(gc, cols), which
    for which, (cp, rows) in enumerate((csc_a, csc_b)):
        gp, gr = c.csc(which)
        assert np.array_equal(gp, cp) and np.array_equal(gr, rows), which


def test_builtin_circuit_rows_are_canonical():
    # rows strictly ascending (sorted symmetric differences), padding rows past the pin empty,
    # and the pin row z[ZC] * z[ZC] = z[ZC]
    for kind in (F.HASH_SHA256, F.HASH_BLAKE3):
        c = F.Circuit(kind)
        for which in (0, 1):
            ptr, cols = c.csr(which)
            lens = np.diff(ptr.astype(np.int64))
            assert (lens[c.useful_bits:] == 0).all()
            d = np.diff(cols.astype(np.int64))
            starts = np.zeros(len(cols), dtype=bool)
            starts[ptr[:-1][lens > 0].astype(np.int64)] = True
            assert (d[~starts[1:]] > 0).all()
            pin = c.const_pin
            assert list(cols[ptr[pin]:ptr[pin + 1]]) == [pin]


def test_keccak_registry_stub():
    c = F.Circuit(F.HASH_KECCAK)
    assert (c.k_log, c.useful_bits, c.const_pin, c.nnz) == (16, 42_560, 4096, (0, 0))


def test_builtin_registry_digest_equals_the_oracle_matrices_digest():
    from test_oracle_blake3_hash import lib_type_statement
    keep = []
    for kind, nu in ((F.HASH_SHA256, 16), (F.HASH_SHA256, 7), (F.HASH_BLAKE3, 8),
                     (F.HASH_KECCAK, 6)):
        want = F.registry_digest([lib_type_statement(kind, keep)], nu)
        assert F.circuit_registry_digest(kind, nu) == want
        assert F.registry_digest([F.Circuit(kind).type_statement()], nu) == want
    assert F.circuit_registry_digest(F.HASH_SHA256, 16) != F.circuit_registry_digest(F.HASH_SHA256, 15)
