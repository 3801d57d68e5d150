"""GPU parity of the witness data path: the union dense-stack gather (compact_witness,
crates/flock-core/src/union.rs:418-490) against the oracle, bit-exact."""
from __future__ import annotations

import random
import sys

import numpy as np
import pytest

import oracle_lib as O
from flock_amd import lib as F
from test_union_glue import Union, _random_types

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("seed", range(6))
def test_union_compact(ctx, seed):
    rng = random.Random(100 + seed)
    types = _random_types(rng, rng.randrange(1, 5))
    nu = rng.randrange(0, 8)
    counts = [rng.choice([0, 1, (1 << nu), rng.randrange(0, (1 << nu) + 1)]) for _ in types]
    u = Union(types, nu)
    reg = F.registry_new(types, nu)
    sz = F.union_sizes(reg, counts)
    say(f"case: union compact seed={seed} types={types} nu={nu} counts={counts} "
        f"packed={sz['packed_len']} committed={sz['committed_words']}")
    z = O.splitmix(0x900 + seed, sz["packed_len"])
    exp = u.compact(counts, z)
    q = ctx.alloc(16 * sz["committed_words"])
    ctx.union_compact(reg, counts, ctx.upload(z), q)
    assert np.array_equal(q.download_f128(sz["committed_words"]), exp)


def test_union_compact_sha256_full_size(ctx):
    # the SHA-256 shape at 1024 instances: 199 used columns of 256, committed 2^18
    types = [(15, 25_470, False)]
    nu = 10
    u = Union(types, nu)
    reg = F.registry_new(types, nu)
    counts = [1000]
    sz = F.union_sizes(reg, counts)
    say(f"case: union compact sha256 counts={counts} committed={sz['committed_words']}")
    z = O.splitmix(0x990, sz["packed_len"])
    exp = u.compact(counts, z)
    q = ctx.alloc(16 * sz["committed_words"])
    ctx.union_compact(reg, counts, ctx.upload(z), q)
    assert np.array_equal(q.download_f128(sz["committed_words"]), exp)


@pytest.mark.parametrize("n_log,n_decl", [(3, 8), (3, 5), (6, 64), (8, 200), (10, 1024),
                                          (10, 1000), (12, 3000)])
def test_sha256_witness(ctx, n_log, n_decl):
    from test_oracle_sha2_r1cs import batch_major, random_inputs
    say(f"case: sha256 witness n_log={n_log} n_decl={n_decl}")
    inputs = random_inputs(n_decl, 700 + n_log)
    ez, ea, eb, es = batch_major(inputs, n_log)
    n = 1 << n_log
    z, a, b = (ctx.alloc(16 * 256 * n) for _ in range(3))
    stripe = ctx.alloc(4096 * n)
    dev_in = ctx.upload(np.ascontiguousarray(inputs, dtype=np.uint32))
    ctx.witness_sha256(dev_in, n_decl, n_log, z, a, b, stripe)
    assert np.array_equal(z.download_f128(256 * n), ez)
    assert np.array_equal(a.download_f128(256 * n), ea)
    assert np.array_equal(b.download_f128(256 * n), eb)
    assert np.array_equal(np.frombuffer(stripe.download(4096 * n), dtype=np.uint8), es)


def test_zerocheck_on_sha256_witness(ctx):
    # the zerocheck reads the batch-major buffers as packed bits (m = n_log + 15), with c = z
    from test_oracle_sha2_r1cs import random_inputs
    from test_oracle_zerocheck import zerocheck_prove
    n_log, n_decl = 7, 100
    m = n_log + 15
    say(f"case: zerocheck on sha256 witness m={m}")
    n = 1 << n_log
    z, a, b = (ctx.alloc(16 * 256 * n) for _ in range(3))
    ctx.witness_sha256(ctx.upload(random_inputs(n_decl, 77)), n_decl, n_log, z, a, b)
    hz, ha, hb = (np.frombuffer(x.download(16 * 256 * n), dtype=np.uint8).copy()
                  for x in (z, a, b))
    assert np.array_equal(ha & hb, hz)  # honest witness: a AND b = z = c
    dom = b"flock-zc-sha256-witness"
    exp = zerocheck_prove(dom, ha, hb, hz, m, 1)
    got = ctx.zerocheck_prove(F.FsChallenger(dom), a, b, z, m, True)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


@pytest.mark.parametrize("n_log,n_decl", [(3, 8), (3, 3), (7, 100), (10, 1024), (11, 2000)])
def test_keccak_witness(ctx, n_log, n_decl):
    from test_oracle_keccak_r1cs import batch_major, random_states
    say(f"case: keccak witness n_log={n_log} n_decl={n_decl}")
    inputs = random_states(n_decl, 900 + n_log)
    ez, ea, eb, es = batch_major(inputs, n_log)
    n = 1 << n_log
    z, a, b = (ctx.alloc(16 * 512 * n) for _ in range(3))
    stripe = ctx.alloc(8192 * n)
    ctx.witness_keccak(ctx.upload(np.ascontiguousarray(inputs, dtype=np.uint64)), n_decl, n_log,
                       z, a, b, stripe)
    assert np.array_equal(z.download_f128(512 * n), ez)
    assert np.array_equal(a.download_f128(512 * n), ea)
    assert np.array_equal(b.download_f128(512 * n), eb)
    assert np.array_equal(np.frombuffer(stripe.download(8192 * n), dtype=np.uint8), es)


@pytest.mark.parametrize("n_log,n_decl", [(3, 8), (3, 1), (7, 100), (10, 1024), (12, 4000)])
def test_blake3_witness(ctx, n_log, n_decl):
    from test_oracle_blake3_r1cs import batch_major, random_inputs
    say(f"case: blake3 witness n_log={n_log} n_decl={n_decl}")
    inputs = random_inputs(n_decl, 1100 + n_log)
    ez, ea, eb, es = batch_major(inputs, n_log)
    n = 1 << n_log
    z, a, b = (ctx.alloc(16 * 128 * n) for _ in range(3))
    stripe = ctx.alloc(2048 * n)
    ctx.witness_blake3(ctx.upload(np.ascontiguousarray(inputs, dtype=np.uint32)), n_decl, n_log,
                       z, a, b, stripe)
    assert np.array_equal(z.download_f128(128 * n), ez)
    assert np.array_equal(a.download_f128(128 * n), ea)
    assert np.array_equal(b.download_f128(128 * n), eb)
    assert np.array_equal(np.frombuffer(stripe.download(2048 * n), dtype=np.uint8), es)


def test_lincheck_and_zerocheck_on_blake3_witness(ctx):
    # the BLAKE3 circuit (restated A_0, B_0) with the GPU witness: lincheck at K = 2^14 and the
    # zerocheck over (a, b, c = z), both against the oracle
    from test_oracle_blake3_r1cs import matrices, random_inputs
    from test_oracle_lincheck import lincheck_prove
    from test_oracle_zerocheck import zerocheck_prove
    n_log, n_decl = 7, 120
    say(f"case: lincheck+zerocheck on blake3 witness n_log={n_log}")
    _, _, csc_a, csc_b = matrices()
    n = 1 << n_log
    z, a, b = (ctx.alloc(16 * 128 * n) for _ in range(3))
    stripe = ctx.alloc(2048 * n)
    ctx.witness_blake3(ctx.upload(random_inputs(n_decl, 55)), n_decl, n_log, z, a, b, stripe)
    hz, ha, hb = (np.frombuffer(x.download(16 * 128 * n), dtype=np.uint8).copy()
                  for x in (z, a, b))
    assert np.array_equal(ha & hb, hz)
    m = n_log + 14
    dom = b"flock-blake3-witness"
    exp = zerocheck_prove(dom, ha, hb, hz, m, 1)
    got = ctx.zerocheck_prove(F.FsChallenger(dom), a, b, z, m, True)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)
    k_log = 14
    L = k_log - 6
    pts = O.arr_to_ints(O.splitmix(0xB3, 1 + L + n_log))
    zs = np.frombuffer(stripe.download(2048 * n), dtype=np.uint8).copy()
    pin, ub = 11_706, 11_707
    exp = lincheck_prove(dom, k_log, csc_a, csc_b, pin, zs, n_log, ub, n_decl, pts[0],
                         pts[1:1 + L], pts[1 + L:], 1)
    circuit, keep = ctx.csc(k_log, *csc_a, *csc_b, const_pin=pin)
    got = ctx.lincheck_prove(F.FsChallenger(dom), circuit, stripe, n_log, ub, n_decl, pts[0],
                             pts[1:1 + L], pts[1 + L:], True)
    for name in ("rounds", "z_partial", "matrix_evals", "claim", "nonces"):
        assert np.array_equal(got[name], exp[name]), name


def test_keccak_lincheck_marginals(ctx):
    # the GPU walker against the oracle's restated walker (itself pinned by
    # <A_0^T eq, z> = <eq, a> on FIPS 202 witnesses), on random row weights
    from test_oracle_keccak_r1cs import K, lc_fold_split
    say("case: keccak lincheck marginals")
    rng = np.random.default_rng(0x4B)
    eq = rng.integers(0, 1 << 63, (K, 2), dtype=np.uint64) * 2 + 1
    alpha = int(rng.integers(1, 1 << 62)) << 64 | int(rng.integers(1, 1 << 62))
    esa, esb = lc_fold_split(eq)
    d_eq = ctx.upload(eq)
    sa, sb, comb = (ctx.alloc(16 * K) for _ in range(3))
    ctx.lc_keccak_fold(d_eq, alpha, sa, sb, comb)
    assert np.array_equal(sa.download_f128(K), esa)
    assert np.array_equal(sb.download_f128(K), esb)
    ecomb = O.mul_batch(np.tile(O.ints_to_arr([alpha]), (K, 1)), esa) ^ esb
    assert np.array_equal(comb.download_f128(K), ecomb)


@pytest.mark.parametrize("n_log,n_decl", [(3, 5), (7, 100)])
def test_lincheck_and_zerocheck_on_keccak_witness(ctx, n_log, n_decl):
    # the Keccak-f[1600] witness from the GPU, the lincheck with the walker circuit (k_log 16)
    # and the zerocheck over (a, b, c = z), both against the oracle
    from test_oracle_keccak_r1cs import random_states
    from test_oracle_lincheck import lincheck_prove_keccak
    from test_oracle_zerocheck import zerocheck_prove
    say(f"case: lincheck+zerocheck on keccak witness n_log={n_log} n_decl={n_decl}")
    n = 1 << n_log
    z, a, b = (ctx.alloc(16 * 512 * n) for _ in range(3))
    stripe = ctx.alloc(8192 * n)
    inputs = random_states(n_decl, 77 + n_log)
    ctx.witness_keccak(ctx.upload(np.ascontiguousarray(inputs, dtype=np.uint64)), n_decl, n_log,
                       z, a, b, stripe)
    dom = b"flock-keccak-witness"
    L = 10
    pts = O.arr_to_ints(O.splitmix(0x4C + n_log, 1 + L + n_log))
    zs = np.frombuffer(stripe.download(8192 * n), dtype=np.uint8).copy()
    # padding rows run keccak_f(0), so every row of the batch is a constraint row
    exp = lincheck_prove_keccak(dom, zs, n_log, n, pts[0], pts[1:1 + L], pts[1 + L:], 1)
    zv = ctx.alloc(16 << 16)
    got = ctx.lincheck_prove_keccak(F.FsChallenger(dom), stripe, n_log, n, pts[0],
                                    pts[1:1 + L], pts[1 + L:], True, z_vec=zv)
    for name in ("rounds", "z_partial", "matrix_evals", "claim", "nonces"):
        assert np.array_equal(got[name], exp[name]), name
    assert np.array_equal(zv.download_f128(1 << 16), exp["z_vec"])
    if n_log <= 3:
        hz, ha, hb = (np.frombuffer(x.download(16 * 512 * n), dtype=np.uint8).copy()
                      for x in (z, a, b))
        m = n_log + 16
        exp = zerocheck_prove(dom, ha, hb, hz, m, 1)
        got = ctx.zerocheck_prove(F.FsChallenger(dom), a, b, z, m, True)
        for g, e in zip(got, exp):
            assert np.array_equal(g, e)
