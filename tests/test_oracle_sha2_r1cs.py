"""Pin the oracle's SHA-256 compression R1CS (crates/flock-prover/src/r1cs_hashes/sha2.rs):
the layout constants the reference documents, H_out against FIPS 180-4 (hashlib), R1CS
satisfaction a AND b = z (C = I), a = A_0 z and b = B_0 z through the independently restated
symbolic row walk, and the batch-major / stripe layout of the driver (common.rs)."""
from __future__ import annotations

import ctypes
import functools
import hashlib
import struct

import numpy as np
import pytest

import oracle_lib as O

L = O.lib()
P, S = ctypes.c_void_p, ctypes.c_size_t
L.or_sha2_layout.argtypes = [P]
L.or_sha2_instance.argtypes = [P, P, P, P, P]
L.or_sha2_witness_batch_major.argtypes = [P, S, ctypes.c_int, P, P, P, P]
L.or_sha2_matrices.argtypes = [P, P, P, P]
L.or_csr_to_csc.argtypes = [S, P, P, P, P]
L.or_csr_apply_bits.argtypes = [S, P, P, P, P]

IV = [0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A, 0x510E527F, 0x9B05688C, 0x1F83D9AB,
      0x5BE0CD19]
K = 1 << 15
USEFUL_BITS = 25_470


def padded_block(msg: bytes) -> list[int]:
    assert len(msg) <= 55
    blk = msg + b"\x80" + bytes(55 - len(msg)) + struct.pack(">Q", 8 * len(msg))
    return list(struct.unpack(">16I", blk))


def instance(h_in, m):
    z, a, b = (np.zeros(512, dtype=np.uint64) for _ in range(3))
    hi = np.array(h_in, dtype=np.uint32)
    mm = np.array(m, dtype=np.uint32)
    L.or_sha2_instance(hi.ctypes.data, mm.ctypes.data, z.ctypes.data, a.ctypes.data,
                       b.ctypes.data)
    return z, a, b


def bits_word(buf, pos):
    v = 0
    for i in range(32):
        v |= ((int(buf[(pos + i) >> 6]) >> ((pos + i) & 63)) & 1) << i
    return v


@functools.lru_cache(maxsize=1)
def matrices():
    """(A_0, B_0) as CSR and CSC arrays."""
    ap = np.zeros(K + 1, dtype=np.uint64)
    bp = np.zeros(K + 1, dtype=np.uint64)
    L.or_sha2_matrices(ap.ctypes.data, None, bp.ctypes.data, None)
    ac = np.zeros(int(ap[-1]), dtype=np.uint32)
    bc = np.zeros(int(bp[-1]), dtype=np.uint32)
    L.or_sha2_matrices(ap.ctypes.data, ac.ctypes.data, bp.ctypes.data, bc.ctypes.data)
    csc = []
    for ptr, cols in ((ap, ac), (bp, bc)):
        cp = np.zeros(K + 1, dtype=np.uint32)
        rows = np.zeros(len(cols), dtype=np.uint32)
        L.or_csr_to_csc(K, ptr.ctypes.data, cols.ctypes.data, cp.ctypes.data, rows.ctypes.data)
        csc.append((cp, rows))
    return (ap, ac), (bp, bc), csc[0], csc[1]


def random_inputs(n, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 1 << 32, (n, 24), dtype=np.uint64).astype(np.uint32)


def test_layout_matches_reference_constants():
    lay = np.zeros(8, dtype=np.uint32)
    L.or_sha2_layout(lay.ctypes.data)
    # SCHED_CARRY_BASE .. USEFUL_BITS as documented at sha2.rs:189-194
    assert list(lay[:6]) == [18_757, 23_173, 24_197, 25_221, 25_469, 25_470]


def test_h_out_is_the_sha256_compression():
    for msg in (b"", b"abc", b"flock-readme-bench-v0", bytes(range(55))):
        z, _, _ = instance(IV, padded_block(msg))
        h_out = [bits_word(z, 256 + 32 * w) for w in range(8)]
        assert h_out == list(struct.unpack(">8I", hashlib.sha256(msg).digest()))
        h_in = [bits_word(z, 32 * w) for w in range(8)]
        assert h_in == IV


def test_r1cs_is_satisfied_and_padding_is_zero():
    for seed in range(4):
        inp = random_inputs(1, seed)[0]
        z, a, b = instance(list(inp[:8]), list(inp[8:]))
        assert np.array_equal(a & b, z)  # C = I: (A z)(B z) = z row by row
        hi = USEFUL_BITS >> 6
        tail = ~np.uint64((1 << (USEFUL_BITS & 63)) - 1)
        for buf in (z, a, b):
            assert not buf[hi + 1:].any() and not (buf[hi] & tail)
        assert (int(z[25_469 >> 6]) >> (25_469 & 63)) & 1  # Z_CONST


def test_matrices_reproduce_a_and_b():
    (ap, ac), (bp, bc), _, _ = matrices()
    assert int(ap[-1]) + int(bp[-1]) == 47_366_638  # "47.4M template nnz", sha2.rs:178-185
    for seed in range(2):
        inp = random_inputs(1, 10 + seed)[0]
        z, a, b = instance(list(inp[:8]), list(inp[8:]))
        ya = np.zeros(512, dtype=np.uint64)
        yb = np.zeros(512, dtype=np.uint64)
        L.or_csr_apply_bits(K, ap.ctypes.data, ac.ctypes.data, z.ctypes.data, ya.ctypes.data)
        L.or_csr_apply_bits(K, bp.ctypes.data, bc.ctypes.data, z.ctypes.data, yb.ctypes.data)
        assert np.array_equal(ya, a) and np.array_equal(yb, b)


def batch_major(inputs, n_log):
    n = 1 << n_log
    z, a, b = (np.zeros((n * 256, 2), dtype=np.uint64) for _ in range(3))
    stripe = np.zeros(n * (K // 8), dtype=np.uint8)
    inp = np.ascontiguousarray(inputs, dtype=np.uint32)
    assert L.or_sha2_witness_batch_major(inp.ctypes.data, len(inputs), n_log, O._p(z), O._p(a),
                                         O._p(b), stripe.ctypes.data) == 0
    return z, a, b, stripe


def test_batch_major_layout():
    n_log, n_decl = 4, 13
    inputs = random_inputs(n_decl, 20)
    z, a, b, stripe = batch_major(inputs, n_log)
    for o in range(1 << n_log):
        if o < n_decl:
            iz, ia, ib = instance(list(inputs[o][:8]), list(inputs[o][8:]))
        else:
            iz = ia = ib = np.zeros(512, dtype=np.uint64)
        for buf, inst in ((z, iz), (a, ia), (b, ib)):
            rows = buf[np.arange(256) * (1 << n_log) + o]
            assert np.array_equal(rows.reshape(-1), inst)
    # stripe[(o/8) K + i] bit (o % 8) = z bit i of row o
    zb = np.zeros((1 << n_log, K), dtype=np.uint8)
    for o in range(1 << n_log):
        words = z[np.arange(256) * (1 << n_log) + o].reshape(-1)
        zb[o] = np.unpackbits(words.view(np.uint8), bitorder="little")
    exp = np.zeros_like(stripe).reshape(-1, K)
    for o in range(1 << n_log):
        exp[o // 8] |= (zb[o] << (o % 8)).astype(np.uint8)
    assert np.array_equal(stripe.reshape(-1, K), exp)
