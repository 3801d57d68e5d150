"""Host logic of a mixed prove_fast request (no GPU call): the slot order and statement counts
the library reports for the bundle (flock_union_request_info), its PCS parameters against the
oracle's union layout, and the built-in registry digest of several types
(flock_circuit_union_digest) against the oracle's Registry::digest."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as O
import oracle_prove as OP
from flock_amd import lib as F
from test_union_glue import Union


def oracle_info(kinds, counts_caller, nu):
    u = Union([(OP.SHAPES[k][0], OP.SHAPES[k][1], False) for k in kinds], nu)
    order = [u.reg.slots[s].type_index for s in range(len(kinds))]
    counts = [(1 << nu) if kinds[ci] == OP.HASH_KECCAK else counts_caller[ci] for ci in order]
    dense_m = u.sizes(counts, 0, 0)["dense_m"]
    lbs = O.lib().or_lig_initial_k(dense_m)
    lanes = u.sizes(counts, 0, lbs)["commit_lanes"] or 0
    return order, counts, dense_m, lbs, lanes


@pytest.mark.parametrize("kinds,counts,nu", [
    ([F.HASH_BLAKE3, F.HASH_SHA256], [37, 50], 10),
    ([F.HASH_BLAKE3, F.HASH_SHA256], [64, 0], 10),
    ([F.HASH_KECCAK, F.HASH_SHA256, F.HASH_BLAKE3], [40, 100, 50], 7),
    ([F.HASH_SHA256, F.HASH_KECCAK], [3, 1], 5),
])
def test_union_info_matches_the_oracle_layout(kinds, counts, nu):
    parts = [(k, np.zeros((max(n, 1), 25), dtype=np.uint64), n) for k, n in zip(kinds, counts)]
    req, keep = F.union_request(parts, nu)
    info = F.union_info(req)
    order, ocounts, dense_m, lbs, lanes = oracle_info(kinds, counts, nu)
    assert info["slot_part"] == order
    assert info["counts"] == ocounts
    assert (info["dense_m"], info["log_batch_size"], info["lanes"]) == (dense_m, lbs, lanes) \
        or not info["has_opening"]
    assert info["device_bytes"] > 0


def test_union_request_rejects_bad_parts():
    x = np.zeros((4, 24), dtype=np.uint32)
    for parts, nu in (([(F.HASH_SHA256, x, 9)], 3),         # n_decl > 2^nu
                      ([(7, x, 1)], 5),                      # unknown kind
                      ([(F.HASH_SHA256, x, 1)], 2)):         # nu < 3
        req, keep = F.union_request(parts, nu)
        with pytest.raises(F.FlockHipError) as e:
            F.union_info(req)
        assert e.value.code == F.E_INVALID


def test_union_digest_equals_the_oracle_registry_digest():
    # BLAKE3 + SHA-256 at nu 10 (the Blake3Sha2Nu10 registry), in either caller order, and a
    # three-type registry with the Keccak-f[1600] stub
    tys = {k: OP.circuit(k)[0] for k in (F.HASH_SHA256, F.HASH_BLAKE3, F.HASH_KECCAK)}
    for kinds, nu in (([F.HASH_BLAKE3, F.HASH_SHA256], 10), ([F.HASH_SHA256, F.HASH_BLAKE3], 10),
                      ([F.HASH_KECCAK, F.HASH_SHA256, F.HASH_BLAKE3], 7)):
        want = O.registry_digest([tys[k] for k in kinds], nu)
        assert F.circuit_union_digest(kinds, nu) == want
    # one type: the single-type digest
    assert F.circuit_union_digest([F.HASH_SHA256], 7) == F.circuit_registry_digest(F.HASH_SHA256, 7)
