"""The prove core's proof verifies: the statement binding, flock_zerocheck_verify and
flock_lincheck_verify chained on one verifier transcript (the order of verify_fast's boolean
part, prover.rs:926-1160 mirrored), over the oracle's prove core. The zerocheck's final
(a, b) evaluations are the lincheck's input claim at x_ab = x_ab_from_mlv(mlv); the verifier's
transcript ends in the prover's digest, which the GPU prove core is tested to reach byte for
byte, so the GPU's proofs are the ones verified here."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as O
from flock_amd import lib as F


def verify_core(kind, n_log, parts, circuits, domain, digest):
    """Verify the parts oracle_prove_core wrote; returns the verifier's final digest."""
    from test_gpu_prove import SHAPES
    k_log, ub, pin = SHAPES[kind]
    fs = F.FsChallenger(domain)
    fs.bind_statement(digest, [parts["count"]], parts["cap"])
    r1, ml, fin, nonces, zclaim = parts["zc"]
    claim = F.zerocheck_verify(fs, parts["m"], True, r1, ml, fin, nonces)
    assert np.array_equal(claim, zclaim)
    reg = F.registry_new([(k_log, ub, False)], n_log)
    n_mlv = parts["m"] - 6
    zc = O.arr_to_ints(claim)
    x_rest, x_outer = F.x_ab_from_mlv(reg, zc[1:1 + n_mlv])
    va, vb = O.arr_to_ints(fin[:2])
    if kind == F.HASH_KECCAK:
        # the Keccak circuit enters through its walker: the host walk, and the deferred form
        fs2 = F.FsChallenger(domain)
        fs2.bind_statement(digest, [parts["count"]], parts["cap"])
        F.zerocheck_verify(fs2, parts["m"], True, r1, ml, fin, nonces)
        dclaim = F.lincheck_verify(fs2, k_log, pin, n_log, parts["count"], zc[0], x_rest,
                                   x_outer, va, vb, True, parts["lc"], None)
        assert np.array_equal(dclaim, parts["lc"]["claim"])
        lclaim = F.lincheck_verify_keccak(fs, n_log, parts["count"], zc[0], x_rest, x_outer, va,
                                          vb, True, parts["lc"])
    else:
        lclaim = F.lincheck_verify(fs, k_log, pin, n_log, parts["count"], zc[0], x_rest,
                                   x_outer, va, vb, True, parts["lc"],
                                   (*circuits[kind][1], *circuits[kind][2]))
    assert np.array_equal(lclaim, parts["lc"]["claim"])
    return fs.digest()


def _host_circuits():
    import test_oracle_blake3_r1cs as B3
    import test_oracle_sha2_r1cs as S2
    out = {}
    for kind, mod in ((F.HASH_SHA256, S2), (F.HASH_BLAKE3, B3)):
        _, _, a, b = mod.matrices()
        out[kind] = (None, a, b)
    return out


@pytest.mark.parametrize("kind,n_log,n_decl", [(F.HASH_SHA256, 3, 8), (F.HASH_SHA256, 4, 11),
                                               (F.HASH_BLAKE3, 4, 16), (F.HASH_KECCAK, 3, 5)])
def test_prove_core_proof_verifies(kind, n_log, n_decl):
    import test_gpu_prove as T
    circuits = _host_circuits()
    parts = {}
    inputs = T.inputs_for(kind, n_decl, 100 + 10 * kind + n_log)
    _, dig = T.oracle_prove_core(kind, n_log, inputs, circuits, parts)
    assert verify_core(kind, n_log, parts, circuits, T.DOMAIN, T.digest_for(kind, n_log)) == dig


def test_keccak_walk_check_rejects_a_wrong_claim():
    # an honest Keccak proof passes the full check only if the host walk reproduces the
    # prover's marginals (test_prove_core_proof_verifies); a proof checked against another
    # input claim is rejected
    import test_gpu_prove as T
    parts = {}
    T.oracle_prove_core(F.HASH_KECCAK, 3, T.inputs_for(F.HASH_KECCAK, 5, 9), {}, parts)
    fs = F.FsChallenger(T.DOMAIN)
    fs.bind_statement(T.digest_for(F.HASH_KECCAK, 3), [parts["count"]], parts["cap"])
    r1, ml, fin, nonces, _ = parts["zc"]
    zc = O.arr_to_ints(F.zerocheck_verify(fs, parts["m"], True, r1, ml, fin, nonces))
    reg = F.registry_new([(16, 42_560, False)], 3)
    x_rest, x_outer = F.x_ab_from_mlv(reg, zc[1:1 + parts["m"] - 6])
    va, vb = O.arr_to_ints(fin[:2])
    with pytest.raises(F.FlockHipError) as e:
        F.lincheck_verify_keccak(fs, 3, parts["count"], zc[0], x_rest, x_outer, va ^ (1 << 77),
                                 vb, True, parts["lc"])
    assert e.value.code == F.E_REJECT


def test_prove_core_with_another_statement_is_rejected():
    # a proof bound to one commitment root does not verify against another
    import test_gpu_prove as T
    circuits = _host_circuits()
    parts = {}
    T.oracle_prove_core(F.HASH_SHA256, 3, T.inputs_for(F.HASH_SHA256, 8, 5), circuits, parts)
    parts["cap"] = bytes(len(parts["cap"]))
    with pytest.raises(F.FlockHipError) as e:
        verify_core(F.HASH_SHA256, 3, parts, circuits, T.DOMAIN, T.digest_for(F.HASH_SHA256, 3))
    assert e.value.code == F.E_REJECT
