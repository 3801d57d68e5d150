"""Pin the oracle's Keccak-f[1600] R1CS witness (crates/flock-prover/src/r1cs_hashes/keccak.rs)
against a FIPS 202 restatement (itself pinned by hashlib's SHA3-256): the permutation, each
round's chi AND term t_r and its operands, R1CS satisfaction a AND b = z (C = I), and the
batch-major / stripe layout with padding rows running keccak_f(0)."""
from __future__ import annotations

import ctypes
import hashlib

import numpy as np

import oracle_lib as O
import puremath as M

L = O.lib()
P, S = ctypes.c_void_p, ctypes.c_size_t
L.or_keccak_f.argtypes = [P]
L.or_keccak_instance.argtypes = [P, P, P, P]
L.or_keccak_witness_batch_major.argtypes = [P, S, ctypes.c_int, P, P, P, P]
K = 1 << 16
USEFUL_BITS = 42_560


def random_states(n, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 1 << 63, (n, 25), dtype=np.uint64) * 2 + \
        rng.integers(0, 2, (n, 25), dtype=np.uint64)


def instance(state):
    z, a, b = (np.zeros(1024, dtype=np.uint64) for _ in range(3))
    s = np.ascontiguousarray(state, dtype=np.uint64)
    L.or_keccak_instance(s.ctypes.data, z.ctypes.data, a.ctypes.data, b.ctypes.data)
    return z, a, b


def batch_major(inputs, n_log):
    n = 1 << n_log
    z, a, b = (np.zeros((n * 512, 2), dtype=np.uint64) for _ in range(3))
    stripe = np.zeros(n * (K // 8), dtype=np.uint8)
    inp = np.ascontiguousarray(inputs, dtype=np.uint64)
    assert L.or_keccak_witness_batch_major(inp.ctypes.data, len(inputs), n_log, O._p(z), O._p(a),
                                           O._p(b), stripe.ctypes.data) == 0
    return z, a, b, stripe


def test_pure_keccak_is_fips202():
    for msg in (b"", b"abc", bytes(range(150))):
        assert M.sha3_256(msg) == hashlib.sha3_256(msg).digest()
    assert M.keccak_f([0] * 25)[0] == 0xF1258F7940E1DDE7


def test_oracle_permutation():
    for st in [np.zeros(25, dtype=np.uint64)] + list(random_states(3, 1)):
        s = np.ascontiguousarray(st, dtype=np.uint64).copy()
        L.or_keccak_f(s.ctypes.data)
        assert [int(v) for v in s] == M.keccak_f([int(v) for v in st])


def test_instance_rows():
    st = random_states(1, 2)[0]
    z, a, b = instance(st)
    A = [int(v) for v in st]
    assert [int(v) for v in z[:25]] == A and [int(v) for v in a[:25]] == A
    assert all(int(v) == M.M64 for v in b[:25])
    for ir in range(24):
        B, t = M.keccak_round_parts(A)
        base = 65 + 25 * ir
        for i in range(25):
            x, y = i % 5, i // 5
            assert int(z[base + i]) == t[i]
            assert int(a[base + i]) == (~B[(x + 1) % 5 + 5 * y]) & M.M64
            assert int(b[base + i]) == B[(x + 2) % 5 + 5 * y]
        A = [B[i] ^ t[i] for i in range(25)]
        A[0] ^= M.keccak_round_constant(ir)
    assert [int(v) for v in z[32:57]] == A == M.keccak_f([int(v) for v in st])
    assert int(z[64]) == int(a[64]) == int(b[64]) == 1
    assert np.array_equal(a & b, z)  # C = I
    for buf in (z, a, b):
        assert not buf[25:32].any() and not buf[57:64].any() and not buf[665:].any()


def test_batch_major_layout_with_padding_rows():
    n_log, n_decl = 3, 5
    inputs = random_states(n_decl, 3)
    z, a, b, stripe = batch_major(inputs, n_log)
    n = 1 << n_log
    for o in range(n):
        iz, ia, ib = instance(inputs[o] if o < n_decl else np.zeros(25, dtype=np.uint64))
        for buf, inst in ((z, iz), (a, ia), (b, ib)):
            assert np.array_equal(buf[np.arange(512) * n + o].reshape(-1), inst)
    zb = np.zeros((n, K), dtype=np.uint8)
    for o in range(n):
        words = z[np.arange(512) * n + o].reshape(-1)
        zb[o] = np.unpackbits(words.view(np.uint8), bitorder="little")
    exp = np.zeros_like(stripe).reshape(-1, K)
    for o in range(n):
        exp[o // 8] |= (zb[o] << (o % 8)).astype(np.uint8)
    assert np.array_equal(stripe.reshape(-1, K), exp)


L.or_keccak_lc_fold_split.argtypes = [P, P, P]


def lc_fold_split(eq):
    sa = np.zeros((K, 2), dtype=np.uint64)
    sb = np.zeros((K, 2), dtype=np.uint64)
    L.or_keccak_lc_fold_split(O._p(eq), O._p(sa), O._p(sb))
    return sa, sb


def _xor_where(vals, bits_u64):
    """XOR of vals[c] over the set bits c of a packed LSB-first bit vector."""
    mask = np.unpackbits(bits_u64.view(np.uint8), bitorder="little").astype(bool)
    sel = vals[:len(mask)][mask]
    return tuple(np.bitwise_xor.reduce(sel, axis=0)) if len(sel) else (0, 0)


def test_lincheck_marginals_are_the_transposed_matrices():
    # A_0 z = a and B_0 z = b for every witness, so for any row weights eq:
    #   <A_0^T eq, z> = <eq, a>  and  <B_0^T eq, z> = <eq, b>.
    # Random states make every t and state_0 column take both values; each column's
    # coefficient is then tested by several independent witnesses.
    rng = np.random.default_rng(0xCC)
    eq = rng.integers(0, 1 << 63, (K, 2), dtype=np.uint64) * 2 + 1
    sa, sb = lc_fold_split(eq)
    for st in [np.zeros(25, dtype=np.uint64)] + list(random_states(7, 4)):
        z, a, b = instance(st)
        assert _xor_where(sa, z) == _xor_where(eq, a)
        assert _xor_where(sb, z) == _xor_where(eq, b)
    # columns that no witness bit occupies carry no coefficient
    for lo, hi in ((1600, 2048), (3648, 4096), (4097, 4160), (USEFUL_BITS, K)):
        assert not sa[lo:hi].any() and not sb[lo:hi].any()


def test_lincheck_marginals_are_linear_in_eq():
    rng = np.random.default_rng(0xCD)
    e1 = rng.integers(0, 1 << 63, (K, 2), dtype=np.uint64)
    e2 = rng.integers(0, 1 << 63, (K, 2), dtype=np.uint64)
    s1, s2, s12 = lc_fold_split(e1), lc_fold_split(e2), lc_fold_split(e1 ^ e2)
    assert np.array_equal(s12[0], s1[0] ^ s2[0]) and np.array_equal(s12[1], s1[1] ^ s2[1])
