"""Pin the oracle's zerocheck front (GF(2^8), phi_8, the GF(2^8) additive NTT, round 1,
fold-at-z, the full prover) against pure mathematics."""
from __future__ import annotations

import hashlib

import ctypes
import random

import numpy as np
import pytest

import oracle_lib as O
import puremath as M

L = O.lib()
L.or_f8_mul.restype = ctypes.c_uint8
L.or_f8_mul.argtypes = [ctypes.c_uint8, ctypes.c_uint8]
L.or_phi8.restype = O.F128
L.or_phi8.argtypes = [ctypes.c_uint8]
L.or_f8_ntt_forward.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint8]
L.or_f8_ntt_inverse.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint8]
L.or_zc_round1.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                   ctypes.c_void_p]
L.or_lagrange_phi8.argtypes = [ctypes.c_int, O.F128, ctypes.c_void_p]
L.or_fold_at_z.argtypes = [ctypes.c_void_p, ctypes.c_int, O.F128, ctypes.c_void_p]
L.or_zc_inner_constants.argtypes = [ctypes.c_void_p]
L.or_zerocheck_prove.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_int] + \
    [ctypes.c_void_p] * 5


def phi8(v):
    return O.to_int(L.or_phi8(v))


def test_f8_fips197_and_field():
    assert L.or_f8_mul(0x57, 0x13) == 0xFE
    assert L.or_f8_mul(0x57, 0x83) == 0xC1
    rng = random.Random(1)
    for _ in range(500):
        a, b = rng.randrange(256), rng.randrange(256)
        assert L.or_f8_mul(a, b) == M.f8_mul(a, b)
    for a in range(256):  # xtime
        assert L.or_f8_mul(a, 2) == (((a << 1) ^ 0x1B) & 0xFF if a & 0x80 else a << 1)


def test_phi8_is_the_homomorphism():
    alpha = phi8(2)
    assert alpha == 0x0DCB364640A222FE6B8330483C2E9849  # phi8.rs:28-31
    # alpha is a root of the AES polynomial x^8 + x^4 + x^3 + x + 1 in F128
    a2 = M.mul(alpha, alpha)
    a4 = M.mul(a2, a2)
    a8 = M.mul(a4, a4)
    a3 = M.mul(a2, alpha)
    assert a8 ^ a4 ^ a3 ^ alpha ^ 1 == 0
    # table entries of phi8.rs (PHI_8_TABLE[4], [6], [8], [255])
    assert phi8(4) == 0x3D5BD35C94646A247573DA4A5F7710ED
    assert phi8(6) == 0x3090E51AD4C648DA1EF0EA02635988A4
    assert phi8(8) == 0x6D58C4E181F9199F41A12DB1F974F3AC
    assert phi8(255) == 0x71AF641F08DBD1A0990483806BFFBE0D
    rng = random.Random(2)
    for _ in range(300):
        a, b = rng.randrange(256), rng.randrange(256)
        assert phi8(M.f8_mul(a, b)) == M.mul(phi8(a), phi8(b))
        assert phi8(a ^ b) == phi8(a) ^ phi8(b)


def test_f8_ntt_extension_is_lagrange_interpolation():
    # inverse NTT on S = {0..63} then forward on Lambda = 64 + {0..63}: the row polynomial
    # through (s, v_s), s in S, evaluated at 64 + i
    rng = random.Random(3)
    for trial in range(3):
        v = [rng.randrange(2) if trial < 2 else rng.randrange(256) for _ in range(64)]
        buf = (ctypes.c_uint8 * 64)(*v)
        L.or_f8_ntt_inverse(buf, 6, 0)
        L.or_f8_ntt_forward(buf, 6, 64)
        assert list(buf) == M.f8_lagrange_extend(v, list(range(64)), list(range(64, 128)))
    # forward/inverse round trip on several cosets
    for beta in (0, 64, 128):
        v = [rng.randrange(256) for _ in range(64)]
        buf = (ctypes.c_uint8 * 64)(*v)
        L.or_f8_ntt_forward(buf, 6, beta)
        L.or_f8_ntt_inverse(buf, 6, beta)
        assert list(buf) == v


def _witness(m, seed, satisfying=True):
    rng = np.random.default_rng(seed)
    nb = (1 << m) // 8
    a = rng.integers(0, 256, nb, dtype=np.uint8)
    b = rng.integers(0, 256, nb, dtype=np.uint8)
    c = (a & b) if satisfying else rng.integers(0, 256, nb, dtype=np.uint8)
    return a, b, c


def _round1(a, b, c, m, r):
    ra = O.ints_to_arr(r)
    pab = np.zeros((64, 2), dtype=np.uint64)
    pc = np.zeros((64, 2), dtype=np.uint64)
    L.or_zc_round1(a.ctypes.data, b.ctypes.data, c.ctypes.data, m, O._p(ra), O._p(pab), O._p(pc))
    return O.arr_to_ints(pab), O.arr_to_ints(pc)


def test_bucketed_round1_equals_the_naive_one():
    # or_zc_round1 (nibble-table extension, eq sums per value, parallel rows) against
    # round1_naive's per-row F8 NTTs and direct eq products, on satisfying and random witnesses
    L.or_zc_round1_naive.argtypes = L.or_zc_round1.argtypes
    for m, sat, seed in ((6, True, 1), (9, False, 2), (11, True, 3), (12, False, 4)):
        a, b, c = _witness(m, seed, satisfying=sat)
        r = [int.from_bytes(np.random.default_rng(seed).bytes(16), "little") for _ in range(m)]
        ra = O.ints_to_arr(r)
        outs = []
        for fn in (L.or_zc_round1, L.or_zc_round1_naive):
            pab = np.zeros((64, 2), dtype=np.uint64)
            pc = np.zeros((64, 2), dtype=np.uint64)
            fn(a.ctypes.data, b.ctypes.data, c.ctypes.data, m, O._p(ra), O._p(pab), O._p(pc))
            outs.append((pab.tobytes(), pc.tobytes()))
        assert outs[0] == outs[1], (m, sat)


def _fold(bits, m, z):
    out = np.zeros((1 << (m - 6), 2), dtype=np.uint64)
    L.or_fold_at_z(bits.ctypes.data, m, O.f(z), O._p(out))
    return O.arr_to_ints(out)


def test_round1_and_fold_at_z_satisfy_the_vanishing_identity():
    # For a satisfying witness (c = a AND b) the combined round-1 polynomial
    # P = P^AB + P^C vanishes on phi8(S) and has degree < 128, so its values on Lambda
    # determine P(z) = sum_x eq(r[6..], x) (a_z(x) b_z(x) + c_z(x)) with a_z the fold at z.
    m = 9
    a, b, c = _witness(m, 11)
    r = M.splitmix(0x77, m)
    pab, pc = _round1(a, b, c, m, r)
    z = M.splitmix(0x78, 1)[0]
    nodes = [phi8(v) for v in range(128)]
    w = M.lagrange_f128(nodes, z)
    p_z = 0
    for i in range(64):
        p_z ^= M.mul(w[64 + i], pab[i] ^ pc[i])
    az, bz, cz = _fold(a, m, z), _fold(b, m, z), _fold(c, m, z)
    rhs = 0
    for x in range(1 << (m - 6)):
        e = M.eq_point(r[6:], x)
        rhs ^= M.mul(e, M.mul(az[x], bz[x]) ^ cz[x])
    assert p_z == rhs
    # the c-claim alone: P^C has degree < 64, determined by its Lambda values
    wl = M.lagrange_f128(nodes[64:], z)
    c_eval = 0
    for i in range(64):
        c_eval ^= M.mul(wl[i], pc[i])
    c_rhs = 0
    for x in range(1 << (m - 6)):
        c_rhs ^= M.mul(M.eq_point(r[6:], x), cz[x])
    assert c_eval == c_rhs


def test_round1_matches_definition_on_rows():
    # direct formula: p_ab[i] = sum_x eq_x phi8(A_x(64+i) * B_x(64+i)), rows extended by
    # F8 Lagrange interpolation from {0..63}
    m = 7
    a, b, c = _witness(m, 12, satisfying=False)
    r = M.splitmix(0x79, m)
    pab, pc = _round1(a, b, c, m, r)
    bits = lambda p, i: (int(p[i >> 3]) >> (i & 7)) & 1
    exp_ab, exp_c = [0] * 64, [0] * 64
    for x in range(2):
        rows = [[bits(p, x * 64 + s) for s in range(64)] for p in (a, b, c)]
        ext = [M.f8_lagrange_extend(row, list(range(64)), list(range(64, 128))) for row in rows]
        e = M.eq_point(r[6:], x)
        for i in range(64):
            exp_ab[i] ^= M.mul(e, phi8(M.f8_mul(ext[0][i], ext[1][i])))
            exp_c[i] ^= M.mul(e, phi8(ext[2][i]))
    assert pab == exp_ab and pc == exp_c


def test_inner_constants():
    out = np.zeros((7, 2), dtype=np.uint64)
    L.or_zc_inner_constants(O._p(out))
    v = O.arr_to_ints(out)
    assert v[:3] == [phi8(0xF7), phi8(0x53), phi8(0xB5)]
    for i, g in enumerate((2, 4, 16, 256)):  # x^1, x^2, x^4, x^8
        assert M.mul(v[3 + i], 1 ^ g) == g


def zerocheck_prove(dom, a, b, c, m, grinding):
    n_mlv = m - 6
    fs = O.Fs(dom)
    r1 = np.zeros((128, 2), dtype=np.uint64)
    ml = np.zeros((2 * n_mlv, 2), dtype=np.uint64)
    fin = np.zeros((3, 2), dtype=np.uint64)
    nonces = np.zeros(2 + n_mlv, dtype=np.uint64)
    claim = np.zeros((1 + 2 * n_mlv, 2), dtype=np.uint64)
    rc = L.or_zerocheck_prove(fs.h, a.ctypes.data, b.ctypes.data, c.ctypes.data, m, grinding,
                              O._p(r1), O._p(ml), O._p(fin), nonces.ctypes.data, O._p(claim))
    assert rc == 0
    return r1, ml, fin, nonces, claim


def test_zerocheck_prove_final_claims_are_mle_evaluations():
    # final a, b are the multilinear extensions of the fold-at-z tables at the round
    # challenges; final c is the c-claim at (z, r_rest)
    m = 14
    a, b, c = _witness(m, 13)
    r1, ml, fin, nonces, claim = zerocheck_prove(b"zc-test", a, b, c, m, 1)
    cl = O.arr_to_ints(claim)
    n_mlv = m - 6
    z, rhos, r_rest = cl[0], cl[1:1 + n_mlv], cl[1 + n_mlv:]
    az, bz, cz = _fold(a, m, z), _fold(b, m, z), _fold(c, m, z)
    eq_rho = O.arr_to_ints(O.build_eq(rhos))
    fa = fb = 0
    for x in range(1 << n_mlv):
        fa ^= M.mul(eq_rho[x], az[x])
        fb ^= M.mul(eq_rho[x], bz[x])
    f = O.arr_to_ints(fin)
    assert f[0] == fa and f[1] == fb
    eq_r = O.arr_to_ints(O.build_eq(r_rest))
    fc = 0
    for x in range(1 << n_mlv):
        fc ^= M.mul(eq_r[x], cz[x])
    assert f[2] == fc
    # the 7 fixed inner coordinates sit at r_rest[0..7)
    inner = np.zeros((7, 2), dtype=np.uint64)
    L.or_zc_inner_constants(O._p(inner))
    assert r_rest[:7] == O.arr_to_ints(inner)
    # grinding: replay the transcript and check every nonce (below)
    _check_nonce_schedule(b"zc-test", m, r1, ml, nonces)


def _pow_zero_bits(digest: bytes, nonce: int) -> int:
    h = hashlib.sha256(digest + nonce.to_bytes(8, "little")).digest()
    bits = 0
    for byte in h:
        if byte:
            return bits + 8 - byte.bit_length()
        bits += 8
    return bits


def _check_nonce_schedule(dom, m, r1, ml, nonces):
    """Replay the zerocheck transcript on the library's FsChallenger (the order of
    zerocheck.rs:641-728) and check each nonce against the grinding bits of zerocheck.rs:96-139
    — bits_for(m) at the initial point, bits_for(2^7 - 1) = 7 at the skip point, bits_for(2) = 2
    per multilinear round — with hashlib: SHA-256(state digest || nonce LE) has that many
    leading zero bits and no smaller nonce does."""
    from flock_amd import lib as F
    n_mlv = m - 6
    schedule = [m.bit_length(), 7] + [2] * n_mlv
    fs = F.FsChallenger(dom)
    fs.observe_label(b"flock-zerocheck-v0")

    def ground(i):
        bits, nonce = schedule[i], int(nonces[i])
        d = fs.digest()
        assert _pow_zero_bits(d, nonce) >= bits, (i, bits, nonce)
        assert all(_pow_zero_bits(d, k) < bits for k in range(nonce)), (i, "not the smallest")
        assert fs.verify_pow(nonce, bits)

    ground(0)
    fs.sample_f128_vec(6)
    fs.sample_f128_vec(m - 6 - 7)  # the 7 inner coordinates are constants
    fs.observe_f128_slice(O.arr_to_ints(r1[:64]))
    fs.observe_f128_slice(O.arr_to_ints(r1[64:]))
    ground(1)
    fs.sample_f128()
    mlv = O.arr_to_ints(ml)
    for j in range(n_mlv):
        fs.observe_f128(mlv[2 * j])
        fs.observe_f128(mlv[2 * j + 1])
        ground(2 + j)
        fs.sample_f128()
    assert len(nonces) == len(schedule)


def test_nonce_schedule_rejects_a_planted_nonce():
    m = 14
    a, b, c = _witness(m, 13)
    r1, ml, fin, nonces, claim = zerocheck_prove(b"zc-test", a, b, c, m, 1)
    _check_nonce_schedule(b"zc-test", m, r1, ml, nonces)
    bad = nonces.copy()
    bad[3] += 1
    with pytest.raises(AssertionError):
        _check_nonce_schedule(b"zc-test", m, r1, ml, bad)


L.or_zc_s_hat_v_c.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
L.or_fold_1b_rows.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]


def s_hat_v_c(c, m, r):
    out = np.zeros((128, 2), dtype=np.uint64)
    ra = O.ints_to_arr(r)
    L.or_zc_s_hat_v_c(c.ctypes.data, m, O._p(ra), O._p(out))
    return O.arr_to_ints(out)


def test_s_hat_v_c_by_definition():
    # canonical form: weight eq(r[7..m), y) on the 128-bit row y, no r[6] factor
    # (univariate_skip.rs:375-383); the oracle divides banks instead
    m = 9
    _, _, c = _witness(m, 21, satisfying=False)
    r = M.splitmix(0x7A, m)
    got = s_hat_v_c(c, m, r)
    bits = lambda i: (int(c[i >> 3]) >> (i & 7)) & 1
    exp = [0] * 128
    for y in range(1 << (m - 7)):
        e = M.eq_point(r[7:], y)
        for j in range(128):
            if bits(y * 128 + j):
                exp[j] ^= e
    assert got == exp


def test_s_hat_v_c_recombines_to_round1_c():
    # sum_b7 eq(r[6], b7) s_hat_v_c[lane | b7 << 6] is P^C on S; extended to Lambda by
    # Lagrange interpolation on phi8(S) it is round 1's p_c
    m = 8
    a, b, c = _witness(m, 22, satisfying=False)
    r = M.splitmix(0x7B, m)
    shv = s_hat_v_c(c, m, r)
    _, pc = _round1(a, b, c, m, r)
    on_s = [M.mul(1 ^ r[6], shv[s]) ^ M.mul(r[6], shv[64 + s]) for s in range(64)]
    nodes = [phi8(v) for v in range(64)]
    dinv = []
    for i in range(64):
        d = 1
        for j in range(64):
            if j != i:
                d = M.mul(d, nodes[i] ^ nodes[j])
        dinv.append(M.inv(d))
    for i in (0, 37, 63):
        zz = phi8(64 + i)
        acc = 0
        for s in range(64):
            num = 1
            for j in range(64):
                if j != s:
                    num = M.mul(num, zz ^ nodes[j])
            acc ^= M.mul(on_s[s], M.mul(num, dinv[s]))
        assert acc == pc[i]


def test_fold_1b_rows_by_definition():
    rng = random.Random(23)
    n = 37
    rows = [rng.getrandbits(128) for _ in range(n)]
    t = [rng.getrandbits(128) for _ in range(n)]
    out = np.zeros((128, 2), dtype=np.uint64)
    ra, ta = O.ints_to_arr(rows), O.ints_to_arr(t)
    L.or_fold_1b_rows(O._p(ra), O._p(ta), n, O._p(out))
    exp = [0] * 128
    for y in range(n):
        for j in range(128):
            if (rows[y] >> j) & 1:
                exp[j] ^= t[y]
    assert O.arr_to_ints(out) == exp
