"""The prove core (witness -> dense stack -> commit -> bind_statement -> zerocheck -> lincheck)
through flock_hip_prove_core against the same sequence of oracle functions on one oracle
transcript, for SHA-256, BLAKE3 and Keccak-f[1600] batches; and the serving loop: every
result is the same alone and in a concurrent mix, and admission by footprint serialises
requests when the budget holds one at a time."""
from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest

import oracle_lib as O
from flock_amd import lib as F
import test_oracle_lincheck  # noqa: F401  (sets the oracle lincheck signatures)
from test_union_glue import Union

pytestmark = pytest.mark.gpu
L = O.lib()
P = ctypes.c_void_p
L.or_fs_digest.argtypes = [P, P]
L.or_bind_statement.argtypes = [P, P, P, ctypes.c_size_t, P, ctypes.c_size_t]
L.or_x_ab_from_mlv.argtypes = [P, ctypes.c_size_t, ctypes.c_uint, P, P]
L.or_zerocheck_prove.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_int, P, P, P, P, P]

DOMAIN = b"flock-readme-bench-v0"


def digest_for(kind, n_log):
    """Registry::digest (schedule.rs:542-580) of the one-type registry of the hash at nu = n_log,
    through the library (flock_registry_digest) — the value bind_statement absorbs."""
    from test_oracle_blake3_hash import lib_type_statement
    keep = []
    return F.registry_digest([lib_type_statement(kind, keep)], n_log)
# kind: (k_log, useful_bits, const pin)
SHAPES = {F.HASH_SHA256: (15, 25_470, 25_469), F.HASH_BLAKE3: (14, 11_707, 11_706),
          F.HASH_KECCAK: (16, 42_560, 4096)}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def circuits(ctx):
    """Device CSC circuits of SHA-256 and BLAKE3 from the oracle's restated matrices."""
    import test_oracle_blake3_r1cs as B3
    import test_oracle_sha2_r1cs as S2
    out, keep = {}, []
    for kind, mod in ((F.HASH_SHA256, S2), (F.HASH_BLAKE3, B3)):
        _, _, csc_a, csc_b = mod.matrices()
        k_log, _, pin = SHAPES[kind]
        circ, kp = ctx.csc(k_log, *csc_a, *csc_b, const_pin=pin)
        out[kind] = (circ, csc_a, csc_b)
        keep.append(kp)
    yield out


def inputs_for(kind, n, seed):
    if kind == F.HASH_SHA256:
        from test_oracle_sha2_r1cs import random_inputs
    elif kind == F.HASH_BLAKE3:
        from test_oracle_blake3_r1cs import random_inputs
    else:
        from test_oracle_keccak_r1cs import random_states as random_inputs
    return np.ascontiguousarray(random_inputs(n, seed))


def oracle_prove_core(kind, n_log, inputs, circuits, parts=None):
    """The oracle's functions in the order of flock_hip_prove_core, on one transcript. When
    `parts` is a dict it receives the statement and both proofs."""
    k_log, ub, pin = SHAPES[kind]
    n = 1 << n_log
    if kind == F.HASH_SHA256:
        from test_oracle_sha2_r1cs import batch_major
    elif kind == F.HASH_BLAKE3:
        from test_oracle_blake3_r1cs import batch_major
    else:
        from test_oracle_keccak_r1cs import batch_major
    z, a, b, stripe = batch_major(inputs, n_log)
    count = n if kind == F.HASH_KECCAK else len(inputs)
    U = Union([(k_log, ub, False)], n_log)
    dense_m = U.sizes([count], 0, 6)["dense_m"]
    # log_batch_size = the Ligerito config's initial_k and the cap at the L0 cap depth
    # (pcs/commit.rs:266-277); shapes below the m22 configs keep 6 and the root
    lbs = L.or_lig_initial_k(dense_m)
    cd = max(L.or_cap_depth_for(dense_m), 0)
    lbs = lbs if lbs > 0 else 6
    sz = U.sizes([count], 0, lbs)
    q = U.compact([count], z)
    if sz["commit_lanes"]:
        _, tree = O.commit_lane_major(q, sz["dense_m"], 1, lbs, sz["commit_lanes"])
    else:
        _, tree = O.commit(q, sz["dense_m"], 1, lbs, 0)
    root = tree[-32:]
    n_leaves = (len(tree) // 32 + 1) // 2
    cap = tree[32 * (2 * n_leaves - (2 << cd)):][:32 << cd]
    fs = O.Fs(DOMAIN)
    counts = np.array([count], dtype=np.uint64)
    L.or_bind_statement(fs.h, digest_for(kind, n_log), counts.ctypes.data, 1, cap, len(cap))
    m = n_log + k_log
    n_mlv = m - 6
    r1 = np.zeros((128, 2), dtype=np.uint64)
    ml = np.zeros((2 * n_mlv, 2), dtype=np.uint64)
    fin = np.zeros((3, 2), dtype=np.uint64)
    nonces = np.zeros(2 + n_mlv, dtype=np.uint64)
    claim = np.zeros((1 + 2 * n_mlv, 2), dtype=np.uint64)
    zb, ab, bb = (x.view(np.uint8) for x in (z, a, b))
    assert L.or_zerocheck_prove(fs.h, ab.ctypes.data, bb.ctypes.data, zb.ctypes.data, m, 1,
                                O._p(r1), O._p(ml), O._p(fin), nonces.ctypes.data,
                                O._p(claim)) == 0
    Lr = k_log - 6
    xr = np.zeros((max(Lr, 1), 2), dtype=np.uint64)
    xo = np.zeros((n_log, 2), dtype=np.uint64)
    mlv = np.ascontiguousarray(claim[1:1 + n_mlv])
    L.or_x_ab_from_mlv(O._p(mlv), n_mlv, n_log, O._p(xr), O._p(xo))
    rounds = np.zeros((2 * Lr, 2), dtype=np.uint64)
    zp = np.zeros((64, 2), dtype=np.uint64)
    mev = np.zeros((2, 2), dtype=np.uint64)
    lnon = np.zeros(Lr + 3, dtype=np.uint64)
    lcl = np.zeros((2 + Lr, 2), dtype=np.uint64)
    if kind == F.HASH_KECCAK:
        rc = L.or_lincheck_prove_keccak(fs.h, stripe.ctypes.data, n_log, count,
                                        O.f(O.arr_to_ints(claim[:1])[0]), O._p(xr), O._p(xo), 1,
                                        O._p(rounds), O._p(zp), O._p(mev), lnon.ctypes.data,
                                        O._p(lcl), None)
    else:
        _, csc_a, csc_b = circuits[kind]
        rc = L.or_lincheck_prove(fs.h, k_log, csc_a[0].ctypes.data, csc_a[1].ctypes.data,
                                 csc_b[0].ctypes.data, csc_b[1].ctypes.data, pin,
                                 stripe.ctypes.data, n_log, ub, count,
                                 O.f(O.arr_to_ints(claim[:1])[0]), O._p(xr), O._p(xo), 1,
                                 O._p(rounds), O._p(zp), O._p(mev), lnon.ctypes.data,
                                 O._p(lcl), None)
    assert rc == 0
    dig = ctypes.create_string_buffer(32)
    L.or_fs_digest(fs.h, dig)
    if parts is not None:
        parts.update(root=root, cap=cap, count=count, m=m, zc=(r1, ml, fin, nonces, claim),
                     lc=dict(rounds=rounds, z_partial=zp, matrix_evals=mev, nonces=lnon,
                             claim=lcl))
    return root, dig.raw


def spec(kind, n_log, n_decl, seed, circuits, arrival=0.0):
    circ = circuits[kind][0] if kind != F.HASH_KECCAK else None
    return (kind, n_log, inputs_for(kind, n_decl, seed), circ, arrival)


@pytest.mark.parametrize("kind,n_log,n_decl", [(F.HASH_SHA256, 3, 8), (F.HASH_SHA256, 4, 11),
                                               (F.HASH_BLAKE3, 4, 16), (F.HASH_KECCAK, 3, 5)])
def test_prove_core_matches_oracle(circuits, kind, n_log, n_decl):
    say(f"case: prove core kind={kind} n_log={n_log} n_decl={n_decl}")
    sp = spec(kind, n_log, n_decl, 100 + 10 * kind + n_log, circuits)
    reqs, keep = F.make_requests([sp], digest=digest_for(kind, n_log), domain=DOMAIN)
    pr = F.Prover(0)
    res = F.prove_core(pr, reqs[0])
    pr.close()
    root, dig = oracle_prove_core(kind, n_log, sp[2], circuits)
    assert bytes(res.root) == root
    assert bytes(res.transcript) == dig
    assert res.hashes == n_decl


def make_requests(specs):
    """Requests whose statement binds each hash's real registry digest (per kind and nu)."""
    arr, keep = F.make_requests(specs, digest=bytes(32), domain=DOMAIN)
    for i, sp in enumerate(specs):
        d = digest_for(sp[0], sp[1])
        keep.append(d)
        arr[i].digest = d
    return arr, keep


def _mix(circuits):
    specs = []
    kinds = [F.HASH_SHA256, F.HASH_BLAKE3, F.HASH_KECCAK]
    for i in range(12):
        kind = kinds[i % 3]
        n_log = 3 + (i * 5) % 4
        n_decl = (1 << n_log) if kind == F.HASH_KECCAK else (1 << n_log) - (i % 3)
        specs.append(spec(kind, n_log, n_decl, 500 + i, circuits, arrival=0.002 * i))
    return specs


def test_serve_results_are_the_same_alone_and_in_the_mix(circuits):
    say("case: serve mix of 12 requests on 3 workers")
    specs = _mix(circuits)
    reqs, keep = make_requests(specs)
    provers = [F.Prover(0) for _ in range(3)]
    res, wall = F.serve(provers, reqs, 1 << 40)
    alone = F.Prover(0)
    for i in range(len(specs)):
        r = F.prove_core(alone, reqs[i])
        assert res[i].status == 0
        assert bytes(res[i].root) == bytes(r.root), i
        assert bytes(res[i].transcript) == bytes(r.transcript), i
        assert res[i].done_s * 1e3 <= wall + 1e-6
    # one request of each kind against the oracle
    for i in range(3):
        root, dig = oracle_prove_core(specs[i][0], specs[i][1], specs[i][2], circuits)
        assert bytes(res[i].root) == root and bytes(res[i].transcript) == dig
    alone.close()
    for p in provers:
        p.close()


def test_serve_admission_serialises_under_a_small_budget(circuits):
    say("case: serve with a one-request budget")
    specs = _mix(circuits)[:6]
    reqs, keep = make_requests(specs)
    provers = [F.Prover(0) for _ in range(3)]
    res, _ = F.serve(provers, reqs, 1)
    spans = sorted((r.done_s - r.run_ms * 1e-3, r.done_s) for r in res)
    for (s0, e0), (s1, e1) in zip(spans, spans[1:]):
        assert s1 >= e0 - 1e-4  # no two requests ran at once
    ref, _ = F.serve(provers[:1], reqs, 1 << 40)
    for a, b in zip(res, ref):
        assert bytes(a.transcript) == bytes(b.transcript)
    for p in provers:
        p.close()


def test_footprint_grows_with_the_batch(circuits):
    small, k1 = F.make_requests([spec(F.HASH_SHA256, 4, 16, 1, circuits)])
    big, k2 = F.make_requests([spec(F.HASH_SHA256, 8, 256, 2, circuits)])
    assert F.prove_footprint(big[0]) > F.prove_footprint(small[0]) > 0


def test_prove_core_at_config2_size_matches_golden_vectors(circuits):
    # 1024 instances of each hash (config 2's batch); the oracle's roots and transcript
    # digests were written by tests/golden/make_prove_core_vectors.py from the same seeds
    import json
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "prove_core_vectors.json")
    with open(path) as fh:
        gold = json.load(fh)
    assert gold["domain"].encode() == DOMAIN
    pr = F.Prover(0)
    for case in gold["cases"]:
        say(f"case: golden prove core kind={case['kind']} n_log={case['n_log']}")
        sp = spec(case["kind"], case["n_log"], case["n_decl"], case["seed"], circuits)
        assert digest_for(case["kind"], case["n_log"]).hex() == case["digest"]
        reqs, keep = make_requests([sp])
        res = F.prove_core(pr, reqs[0])
        assert bytes(res.root).hex() == case["root"]
        assert bytes(res.transcript).hex() == case["transcript"]
    pr.close()
