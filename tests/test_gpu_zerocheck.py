"""GPU parity of the zerocheck front (round 1, fold at z, the full prover) against the
oracle, bit-exact, with and without grinding."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_lib as O
from flock_amd import lib as F
from test_oracle_zerocheck import L, _witness, zerocheck_prove

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("m", [6, 7, 13, 15, 18])
def test_round1(ctx, m):
    say(f"case: zc round1 m={m}")
    a, b, c = _witness(m, 100 + m, satisfying=False)
    r = O.arr_to_ints(O.splitmix(0x200 + m, m))
    exp = np.zeros((128, 2), dtype=np.uint64)
    ra = O.ints_to_arr(r)
    L.or_zc_round1(a.ctypes.data, b.ctypes.data, c.ctypes.data, m, O._p(ra), O._p(exp),
                   O._p(exp[64:]))
    eq = ctx.upload(O.build_eq(r[6:]))
    got = ctx.zc_round1(ctx.upload(a), ctx.upload(b), ctx.upload(c), m, eq)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("m", [6, 9, 14, 17])
def test_fold_at_z(ctx, m):
    say(f"case: zc fold at z m={m}")
    a, b, _ = _witness(m, 200 + m, satisfying=False)
    z = O.arr_to_ints(O.splitmix(0x300 + m, 1))[0]
    n = 1 << (m - 6)
    az, bz = ctx.alloc(16 * n), ctx.alloc(16 * n)
    ctx.zc_fold_at_z(ctx.upload(a), ctx.upload(b), m, z, az, bz)
    for bits, dev in ((a, az), (b, bz)):
        exp = np.zeros((n, 2), dtype=np.uint64)
        L.or_fold_at_z(bits.ctypes.data, m, O.f(z), O._p(exp))
        assert np.array_equal(dev.download_f128(n), exp)


@pytest.mark.parametrize("m,grinding,satisfying", [(13, True, True), (14, False, False),
                                                   (17, True, True), (20, True, False),
                                                   (25, True, True)])
def test_zerocheck_prove(ctx, m, grinding, satisfying):
    say(f"case: zerocheck prove m={m} grinding={grinding} satisfying={satisfying}")
    a, b, c = _witness(m, 300 + m, satisfying=satisfying)
    dom = b"flock-zc-gpu-test"
    exp = zerocheck_prove(dom, a, b, c, m, 1 if grinding else 0)
    got = ctx.zerocheck_prove(F.FsChallenger(dom), ctx.upload(a), ctx.upload(b), ctx.upload(c), m,
                              grinding)
    names = ["round1", "multilinear", "finals", "nonces", "claim"]
    for name, g, e in zip(names, got, exp):
        if name == "nonces" and not grinding:
            continue
        assert np.array_equal(g, e), name


@pytest.mark.parametrize("m,grinding", [(13, False), (18, True), (24, False)])
def test_zerocheck_prove_capture_s_hat_v_c(ctx, m, grinding):
    # prove_packed_padded_capture_s_hat_v_c_with_grinding: the same wire output plus
    # s_hat_v_c at r[6..m) (the claim's r_rest)
    say(f"case: zerocheck prove capture m={m} grinding={grinding}")
    a, b, c = _witness(m, 400 + m, satisfying=True)
    dom = b"flock-zc-capture-test"
    exp = zerocheck_prove(dom, a, b, c, m, 1 if grinding else 0)
    got = ctx.zerocheck_prove(F.FsChallenger(dom), ctx.upload(a), ctx.upload(b), ctx.upload(c), m,
                              grinding, capture=True)
    for i in (0, 1, 2, 4):
        assert np.array_equal(got[i], exp[i])
    n_mlv = m - 6
    r_rest = O.arr_to_ints(exp[4])[1 + n_mlv:]
    shv = np.zeros((128, 2), dtype=np.uint64)
    ra = O.ints_to_arr([0] * 6 + r_rest)
    L.or_zc_s_hat_v_c(c.ctypes.data, m, O._p(ra), O._p(shv))
    assert np.array_equal(got[5], shv)


@pytest.mark.parametrize("n", [1, 255, 256, 257, 1000, 131075])
def test_fold_1b_rows(ctx, n):
    say(f"case: fold_1b_rows n={n}")
    rows = O.splitmix(0x500 + n, n)
    t = O.splitmix(0x600 + n, n)
    exp = np.zeros((128, 2), dtype=np.uint64)
    L.or_fold_1b_rows(O._p(rows), O._p(t), n, O._p(exp))
    got = ctx.fold_1b_rows(ctx.upload(rows), n, ctx.upload(t))
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("m,grinding", [(17, True), (24, False)])
def test_gpu_zerocheck_proof_verifies(ctx, m, grinding):
    # the GPU prover's proof through the host verifier (zerocheck.rs:1032-1225): accepted, the
    # verifier's claim is the prover's, both transcripts end in the same state
    say(f"case: zerocheck verify m={m} grinding={grinding}")
    a, b, c = _witness(m, 800 + m, satisfying=True)
    dom = b"flock-zc-gpu-verify"
    fs = F.FsChallenger(dom)
    r1, ml, fin, nonces, claim = ctx.zerocheck_prove(fs, ctx.upload(a), ctx.upload(b),
                                                     ctx.upload(c), m, grinding)
    vfs = F.FsChallenger(dom)
    got = F.zerocheck_verify(vfs, m, grinding, r1, ml, fin, nonces if grinding else None)
    assert np.array_equal(got, claim)
    assert vfs.digest() == fs.digest()
    bad = ml.copy()
    bad[3, 1] ^= np.uint64(1)
    with pytest.raises(F.FlockHipError) as e:
        F.zerocheck_verify(F.FsChallenger(dom), m, grinding, r1, bad, fin,
                           nonces if grinding else None)
    assert e.value.code == F.E_REJECT
