"""BLAKE3 hashing modes and Registry::digest on the CPU.

Three restatements are held to outside answers:
- the specification-level Python restatement of BLAKE3 (tests/golden/puremath.py);
- the oracle's C hasher (oracle/blake3_hash.c);
- the library's host hasher and registry digest (flock_blake3_hash, flock_registry_digest).

The outside answers are the published BLAKE3 test vectors (tests/golden/blake3_vectors.json:
one- and multi-chunk inputs up to 8 chunks) and the reference's own pinned registry digests
(schedule.rs:1158-1188). The BLAKE3 Merkle leaf and node functions (merkle.rs:195-205) are
the same modes without the ROOT flag; they are checked against the specification restatement.
"""
from __future__ import annotations

import json
import os
import random
import sys

import numpy as np

import oracle_lib as O
from flock_amd import lib as F

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import puremath as M  # noqa: E402

HERE = os.path.dirname(__file__)


def pattern(n: int) -> bytes:
    return bytes(i % 251 for i in range(n))


def test_published_vectors():
    with open(os.path.join(HERE, "golden", "blake3_vectors.json")) as fh:
        cases = json.load(fh)["cases"]
    assert len(cases) >= 9
    for c in cases:
        msg = pattern(c["input_len"])
        assert M.blake3_hash(msg).hex() == c["hash"], c["input_len"]
        assert O.blake3(msg).hex() == c["hash"], c["input_len"]
        assert F.blake3(msg).hex() == c["hash"], c["input_len"]


def test_lengths_and_streaming_agree():
    rng = random.Random(7)
    lengths = [2, 15, 63, 64, 65, 127, 128, 960, 1023, 1026, 1088, 2047, 3071, 3073, 5121,
               6 * 1024, 7 * 1024 + 1, 16 * 1024 + 64]
    for n in lengths:
        msg = rng.randbytes(n)
        want = O.blake3(msg)
        assert F.blake3(msg) == want, n
        if n <= 8 * 1024:
            assert M.blake3_hash(msg) == want, n
        # the same stream in random pieces, including empty ones
        cuts = sorted(rng.randrange(n + 1) for _ in range(5))
        pieces = [b - a for a, b in zip([0] + cuts, cuts + [n])]
        assert O.blake3(msg, pieces) == want, n


def test_merkle_leaf_and_node_modes():
    rng = random.Random(11)
    for n in (16, 64, 800, 1024, 1040, 2048, 3072, 3184, 4096, 5120):
        msg = rng.randbytes(n)
        assert O.blake3_leaf_cv(msg) == M.blake3_leaf_cv(msg), n
        # the root hash differs from the leaf value only by the ROOT flag
        assert O.blake3_leaf_cv(msg) != O.blake3(msg)
    l, r = rng.randbytes(32), rng.randbytes(32)
    assert O.blake3_parent_cv(l, r) == M.blake3_parent_cv(l, r)
    # the flat tree (merkle.rs:15-22) under HashKind::Blake3
    data = rng.randbytes(8 * 48)
    tree = O.merkle_tree(data, 8, kind=1)
    level = [M.blake3_leaf_cv(data[48 * i:48 * i + 48]) for i in range(8)]
    want = list(level)
    while len(level) > 1:
        level = [M.blake3_parent_cv(level[2 * i], level[2 * i + 1]) for i in range(len(level) // 2)]
        want += level
    assert tree == b"".join(want)
    assert O.merkle_tree(data, 8, kind=0) != tree


def _ty_oracle(k_log, useful, a_rows=None, pin=None, keep=None):
    stub = O.sparse([], keep=keep)
    a = O.sparse(a_rows, num_cols=512, keep=keep) if a_rows else stub
    return O.TypeStmt(k_log, int(pin is not None), useful, pin or 0, a, stub, stub)


def _ty_lib(k_log, useful, a_rows=None, pin=None, keep=None):
    stub = F.sparse_matrix()
    if a_rows:
        ptr = np.cumsum([0] + [len(r) for r in a_rows])
        cols = np.array([c for r in a_rows for c in r])
        a = F.sparse_matrix(ptr, cols, num_cols=512, keep=keep)
    else:
        a = stub
    return F.TypeStatement(k_log, int(pin is not None), useful, pin or 0, a, stub, stub)


def test_registry_digest_matches_reference_pins():
    # schedule.rs:1158-1188: Registry::new(vec![ty(10, 700), ty(9, 300)], 3), and the same with
    # the second type carrying A_0 = [[0, 3], [7]] (512 columns) and const_pin Some(2)
    plain = "8010ecf651ca43eadfe415eac6a081c8f9022dd077070da0f22cea19297699ea"
    with_pin = "2662f8c311c8cdf715ed636bc46a5de16daf06e7516a2e299e6f80132ad9e9c2"
    for ty, digest, keep in ((_ty_oracle, O.registry_digest, []), (_ty_lib, F.registry_digest, [])):
        assert digest([ty(10, 700, keep=keep), ty(9, 300, keep=keep)], 3).hex() == plain
        # Registry::new sorts by k_log (descending, stable): the input order does not matter
        assert digest([ty(9, 300, keep=keep), ty(10, 700, keep=keep)], 3).hex() == plain
        assert digest([ty(10, 700, keep=keep),
                       ty(9, 300, [[0, 3], [7]], pin=2, keep=keep)], 3).hex() == with_pin
        # every field binds (schedule.rs:1133-1148)
        assert digest([ty(10, 700, keep=keep), ty(9, 300, keep=keep)], 4).hex() != plain
        assert digest([ty(10, 700, keep=keep), ty(9, 301, keep=keep)], 3).hex() != plain
        assert digest([ty(10, 700, keep=keep), ty(9, 300, pin=0, keep=keep)], 3).hex() != plain


def test_registry_digest_errors():
    keep = []
    try:
        F.registry_digest([_ty_lib(6, 10, keep=keep)], 3)
        raise AssertionError("expected an error")
    except F.FlockHipError as e:
        assert e.code == F.E_INVALID
    bad = F.sparse_matrix(np.array([0, 2, 1]), np.array([1, 2]), num_cols=4, keep=keep)
    try:
        F.registry_digest([F.TypeStatement(8, 0, 10, 0, bad, F.sparse_matrix(),
                                           F.sparse_matrix())], 3)
        raise AssertionError("expected an error")
    except F.FlockHipError as e:
        assert e.code == F.E_INVALID


def _hash_types():
    import test_oracle_blake3_r1cs as B3
    import test_oracle_sha2_r1cs as S2
    out = {}
    for name, mod, k_log, useful, pin in (("sha256", S2, 15, 25_470, 25_469),
                                          ("blake3", B3, 14, 11_707, 11_706)):
        (ap, ac), (bp, bc), _, _ = mod.matrices()
        out[name] = (k_log, useful, pin, (ap, ac), (bp, bc))
    return out


_LIB_TYPES = {}


def lib_type_statement(kind, keep):
    """The library TypeStatement of a hash's registry type: the SHA-256 / BLAKE3 circuits'
    (A_0, B_0, I) with their constant pins (r1cs_hashes/{sha2,blake3}.rs), and Keccak-f[1600]'s
    empty-stub matrices with the walker's pin at 4096 (keccak.rs:880-888, common.rs:222-260)."""
    if kind not in _LIB_TYPES:
        kp = []
        if kind == F.HASH_KECCAK:
            k = 1 << 16
            zeros = np.zeros(k + 1, dtype=np.uint64)
            none = np.zeros(1, dtype=np.uint32)
            ty = F.TypeStatement(16, 1, 42_560, 4096, F.sparse_matrix(zeros, none, num_cols=k, keep=kp),
                                 F.sparse_matrix(zeros, none, num_cols=k, keep=kp),
                                 F.sparse_matrix(num_rows=k, identity=True))
        else:
            name = "sha256" if kind == F.HASH_SHA256 else "blake3"
            k_log, useful, pin, (ap, ac), (bp, bc) = _hash_types()[name]
            k = 1 << k_log
            ty = F.TypeStatement(k_log, 1, useful, pin, F.sparse_matrix(ap, ac, num_cols=k, keep=kp),
                                 F.sparse_matrix(bp, bc, num_cols=k, keep=kp),
                                 F.sparse_matrix(num_rows=k, identity=True))
        _LIB_TYPES[kind] = (ty, kp)
    keep.append(_LIB_TYPES[kind][1])
    return _LIB_TYPES[kind][0]


def test_registry_digest_of_the_hash_circuits():
    types = _hash_types()
    keep = []
    lib_t, or_t = {}, {}
    for name, (k_log, useful, pin, (ap, ac), (bp, bc)) in types.items():
        k = 1 << k_log
        # the reference's rows are strictly ascending: sha2.rs:247-260 (sorted symmetric
        # differences), common.rs:451-467 (xor_dedup sorts)
        for ptr, cols in ((ap, ac), (bp, bc)):
            assert len(ptr) == k + 1
            d = np.diff(cols.astype(np.int64))
            row_start = np.zeros(len(cols), dtype=bool)
            row_start[ptr[:-1][ptr[:-1] < len(cols)].astype(np.int64)] = True
            assert (d[~row_start[1:]] > 0).all()
        lib_t[name] = F.TypeStatement(
            k_log, 1, useful, pin, F.sparse_matrix(ap, ac, num_cols=k, keep=keep),
            F.sparse_matrix(bp, bc, num_cols=k, keep=keep),
            F.sparse_matrix(num_rows=k, identity=True))
        or_t[name] = O.TypeStmt(k_log, 1, useful, pin, O.sparse_csr(ap, ac, k, keep),
                                O.sparse_csr(bp, bc, k, keep),
                                O.sparse(num_rows=k, identity=True))
    for names, nu in ((["sha256"], 16), (["blake3"], 12), (["blake3", "sha256"], 10)):
        got = F.registry_digest([lib_t[n] for n in names], nu)
        assert got == O.registry_digest([or_t[n] for n in names], nu), names
    # an identity C_0 equals the same matrix spelled out row by row
    k = 1 << 14
    ident_rows = np.arange(k + 1, dtype=np.uint64), np.arange(k, dtype=np.uint32)
    spelled = F.TypeStatement(14, 1, 11_707, 11_706, lib_t["blake3"].a0, lib_t["blake3"].b0,
                              F.sparse_matrix(*ident_rows, num_cols=k, keep=keep))
    assert F.registry_digest([spelled], 12) == F.registry_digest([lib_t["blake3"]], 12)
