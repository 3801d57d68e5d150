"""GPU parity of the lincheck (CSC fold, stripe partial fold, the single-slot union prover)
against the oracle, bit-exact, through the C ABI."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_lib as O
from flock_amd import lib as F
from test_oracle_lincheck import (csc_fold, lincheck_prove, partial_fold, quirky_eq, random_csc,
                                  stripe)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def input_claim(csc_a, csc_b, z_skip, x_rest, z_vec):
    """(v_a, v_b) = (<A_0^T eq, z_vec>, <B_0^T eq, z_vec>): the claim a lincheck proof answers."""
    sa, sb = csc_fold(*csc_a, *csc_b, quirky_eq(z_skip, x_rest))

    def dot(s):
        r = np.bitwise_xor.reduce(O.mul_batch(s, z_vec), axis=0)
        return int(r[0]) | int(r[1]) << 64

    return dot(sa), dot(sb)


def verify_gpu_proof(dom, got, csc_a, csc_b, k_log, pin, nu, nr, z_skip, x_rest, x_outer,
                     grinding, z_vec):
    """The GPU proof passes flock_lincheck_verify in both forms with the prover's claim."""
    va, vb = input_claim(csc_a, csc_b, z_skip, x_rest, z_vec)
    for circuit in ((*csc_a, *csc_b), None):
        fs = F.FsChallenger(dom)
        claim = F.lincheck_verify(fs, k_log, pin, nu, nr, z_skip, x_rest, x_outer, va, vb,
                                  grinding, got, circuit)
        assert np.array_equal(claim, got["claim"])


@pytest.mark.parametrize("k_log,nnz", [(6, 2), (8, 3), (15, 20)])
def test_csc_fold(ctx, k_log, nnz):
    say(f"case: lc csc fold k_log={k_log} nnz/col={nnz}")
    k = 1 << k_log
    pa, ra = random_csc(k, nnz, 100 + k_log)
    pb, rb = random_csc(k, nnz, 200 + k_log)
    eq = O.splitmix(0x700 + k_log, k)
    alpha = O.arr_to_ints(O.splitmix(0x710 + k_log, 1))[0]
    esa, esb = csc_fold(pa, ra, pb, rb, eq)
    circuit, keep = ctx.csc(k_log, pa, ra, pb, rb)
    sa, sb, comb = ctx.alloc(16 * k), ctx.alloc(16 * k), ctx.alloc(16 * k)
    ctx.lc_csc_fold(circuit, ctx.upload(eq), alpha, sa, sb, comb)
    assert np.array_equal(sa.download_f128(k), esa)
    assert np.array_equal(sb.download_f128(k), esb)
    ecomb = O.ints_to_arr([O.mul(alpha, a) ^ b for a, b in zip(O.arr_to_ints(esa),
                                                                O.arr_to_ints(esb))])
    assert np.array_equal(comb.download_f128(k), ecomb)


def test_csc_fold_rejects_bad_indexes(ctx):
    say("case: lc csc fold bad indexes")
    k_log, k = 6, 64
    pa, ra = random_csc(k, 2, 1)
    pb, rb = random_csc(k, 2, 2)
    ra = ra.copy()
    ra[len(ra) // 2] = k + 5  # a row index past the matrix (skipped by the kernel, reported)
    circuit, keep = ctx.csc(k_log, pa, ra, pb, rb)
    eq = ctx.upload(O.splitmix(1, k))
    out = ctx.alloc(16 * k)
    with pytest.raises(F.FlockHipError):
        ctx.lc_csc_fold(circuit, eq, 3, comb=out)


@pytest.mark.parametrize("k_log,nu,ub,nr", [(6, 3, 64, 8), (8, 5, 200, 21), (15, 10, None, None),
                                            (10, 14, 1000, 9000), (7, 3, 128, 0),
                                            (12, 16, None, None)])
def test_partial_fold(ctx, k_log, nu, ub, nr):
    say(f"case: lc partial fold k_log={k_log} nu={nu} ub={ub} nr={nr}")
    k = 1 << k_log
    ub = k if ub is None else ub
    nr = (1 << nu) if nr is None else nr
    z = stripe(k_log, nu, 300 + k_log)
    eq_outer = O.build_eq(O.arr_to_ints(O.splitmix(0x720 + nu, nu)))
    exp = partial_fold(z, k_log, nu, ub, nr, eq_outer)
    out = ctx.alloc(16 * k)
    ctx.lc_partial_fold(ctx.upload(z), k_log, nu, ub, nr, ctx.upload(eq_outer), out)
    assert np.array_equal(out.download_f128(k), exp)


@pytest.mark.parametrize("k_log,nu,pin,grinding,ub,nr",
                         [(6, 3, -1, False, None, None), (8, 5, 3, True, 200, 21),
                          (12, 8, -1, False, None, None), (15, 10, 17, True, None, None),
                          (16, 12, 5, False, 60000, 4000)])
def test_lincheck_prove(ctx, k_log, nu, pin, grinding, ub, nr):
    say(f"case: lincheck prove k_log={k_log} nu={nu} pin={pin} grinding={grinding}")
    k = 1 << k_log
    L = k_log - 6
    ub = k if ub is None else ub
    nr = (1 << nu) if nr is None else nr
    csc_a = random_csc(k, 4, 400 + k_log)
    csc_b = random_csc(k, 3, 500 + k_log)
    z = stripe(k_log, nu, 600 + k_log, ub, nr, pin)
    pts = O.arr_to_ints(O.splitmix(0x800 + k_log, 1 + L + nu))
    z_skip, x_rest, x_outer = pts[0], pts[1:1 + L], pts[1 + L:]
    dom = b"flock-lincheck-gpu-test"
    exp = lincheck_prove(dom, k_log, csc_a, csc_b, pin, z, nu, ub, nr, z_skip, x_rest, x_outer,
                         1 if grinding else 0)
    circuit, keep = ctx.csc(k_log, *csc_a, *csc_b, const_pin=pin)
    zv = ctx.alloc(16 * k)
    got = ctx.lincheck_prove(F.FsChallenger(dom), circuit, ctx.upload(z), nu, ub, nr, z_skip,
                             x_rest, x_outer, grinding, z_vec=zv)
    for name in ("rounds", "z_partial", "matrix_evals", "claim"):
        assert np.array_equal(got[name], exp[name]), name
    if grinding:
        assert np.array_equal(got["nonces"], exp["nonces"])
    assert np.array_equal(zv.download_f128(k), exp["z_vec"])
    verify_gpu_proof(dom, got, csc_a, csc_b, k_log, pin, nu, nr, z_skip, x_rest, x_outer,
                     grinding, exp["z_vec"])


def test_lincheck_sha256_circuit_config2(ctx):
    # the SHA-256 compression circuit (47.4M nonzeros) at 1024 instances (k_log 15, nu 10) with
    # the GPU-built witness stripe, the constant-wire pin at Z_CONST_POS and grinding on
    from test_oracle_sha2_r1cs import matrices, random_inputs
    say("case: lincheck sha256 circuit k_log=15 nu=10")
    _, _, csc_a, csc_b = matrices()
    k_log, nu, n_decl = 15, 10, 1000
    inputs = random_inputs(n_decl, 4242)
    n = 1 << nu
    z, a, b = (ctx.alloc(16 * 256 * n) for _ in range(3))
    stripe = ctx.alloc(4096 * n)
    ctx.witness_sha256(ctx.upload(np.ascontiguousarray(inputs, dtype=np.uint32)), n_decl, nu, z,
                       a, b, stripe)
    zs = np.frombuffer(stripe.download(4096 * n), dtype=np.uint8).copy()
    L = k_log - 6
    pts = O.arr_to_ints(O.splitmix(0x4243, 1 + L + nu))
    z_skip, x_rest, x_outer = pts[0], pts[1:1 + L], pts[1 + L:]
    dom = b"flock-readme-bench-v0"
    pin, ub = 25_469, 25_470
    exp = lincheck_prove(dom, k_log, csc_a, csc_b, pin, zs, nu, ub, n_decl, z_skip, x_rest,
                         x_outer, 1)
    circuit, keep = ctx.csc(k_log, *csc_a, *csc_b, const_pin=pin)
    zv = ctx.alloc(16 << k_log)
    got = ctx.lincheck_prove(F.FsChallenger(dom), circuit, stripe, nu, ub, n_decl, z_skip,
                             x_rest, x_outer, True, z_vec=zv)
    for name in ("rounds", "z_partial", "matrix_evals", "claim", "nonces"):
        assert np.array_equal(got[name], exp[name]), name
    # the proof of the real circuit over the GPU witness verifies (the constant column is 1 on
    # the 1000 declared rows)
    say("case: lincheck sha256 verify")
    verify_gpu_proof(dom, got, csc_a, csc_b, k_log, pin, nu, n_decl, z_skip, x_rest, x_outer,
                     True, zv.download_f128(1 << k_log))
