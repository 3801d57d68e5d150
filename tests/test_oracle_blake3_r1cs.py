"""Pin the oracle's BLAKE3 compression R1CS (crates/flock-prover/src/r1cs_hashes/blake3.rs): the
compression against the BLAKE3 specification restated in puremath (itself pinned by the
published BLAKE3("") and BLAKE3("abc") vectors), the layout constants, R1CS satisfaction
a AND b = z (C = I), a = A_0 z and b = B_0 z through the restated symbolic walk, and the
batch-major / stripe layout with zero dummy rows."""
from __future__ import annotations

import ctypes
import functools

import numpy as np

import oracle_lib as O
import puremath as M

L = O.lib()
P, S = ctypes.c_void_p, ctypes.c_size_t
L.or_blake3_layout.argtypes = [P]
L.or_blake3_compress.argtypes = [P, P, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, P]
L.or_blake3_instance.argtypes = [P, P, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, P, P,
                                 P]
L.or_blake3_witness_batch_major.argtypes = [P, S, ctypes.c_int, P, P, P, P]
L.or_blake3_matrices.argtypes = [P, P, P, P]
L.or_csr_apply_bits.argtypes = [S, P, P, P, P]
L.or_csr_to_csc.argtypes = [S, P, P, P, P]
K = 1 << 14
USEFUL_BITS = 11_707


def random_inputs(n, seed):
    """(cv[8], m[16], counter_lo, counter_hi, block_len, flags) rows, as the hash_throughput
    inputs: random cv and block, block_len 64, flags 11."""
    rng = np.random.default_rng(seed)
    out = rng.integers(0, 1 << 32, (n, 28), dtype=np.uint64).astype(np.uint32)
    out[:, 26] = 64
    out[:, 27] = 11
    return out


def instance(row):
    z, a, b = (np.zeros(256, dtype=np.uint64) for _ in range(3))
    cv = np.ascontiguousarray(row[:8], dtype=np.uint32)
    m = np.ascontiguousarray(row[8:24], dtype=np.uint32)
    L.or_blake3_instance(cv.ctypes.data, m.ctypes.data, int(row[24]) | (int(row[25]) << 32),
                         int(row[26]), int(row[27]), z.ctypes.data, a.ctypes.data, b.ctypes.data)
    return z, a, b


def word_at(buf, pos):
    return sum(((int(buf[(pos + i) >> 6]) >> ((pos + i) & 63)) & 1) << i for i in range(32))


@functools.lru_cache(maxsize=1)
def matrices():
    ap = np.zeros(K + 1, dtype=np.uint64)
    bp = np.zeros(K + 1, dtype=np.uint64)
    L.or_blake3_matrices(ap.ctypes.data, None, bp.ctypes.data, None)
    ac = np.zeros(int(ap[-1]), dtype=np.uint32)
    bc = np.zeros(int(bp[-1]), dtype=np.uint32)
    L.or_blake3_matrices(ap.ctypes.data, ac.ctypes.data, bp.ctypes.data, bc.ctypes.data)
    csc = []
    for ptr, cols in ((ap, ac), (bp, bc)):
        cp = np.zeros(K + 1, dtype=np.uint32)
        rows = np.zeros(len(cols), dtype=np.uint32)
        L.or_csr_to_csc(K, ptr.ctypes.data, cols.ctypes.data, cp.ctypes.data, rows.ctypes.data)
        csc.append((cp, rows))
    return (ap, ac), (bp, bc), csc[0], csc[1]


def batch_major(inputs, n_log):
    n = 1 << n_log
    z, a, b = (np.zeros((n * 128, 2), dtype=np.uint64) for _ in range(3))
    stripe = np.zeros(n * (K // 8), dtype=np.uint8)
    inp = np.ascontiguousarray(inputs, dtype=np.uint32)
    assert L.or_blake3_witness_batch_major(inp.ctypes.data, len(inputs), n_log, O._p(z), O._p(a),
                                           O._p(b), stripe.ctypes.data) == 0
    return z, a, b, stripe


def test_pure_blake3_known_answers():
    assert M.blake3_hash_short(b"").hex() == \
        "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262"
    assert M.blake3_hash_short(b"abc").hex() == \
        "6437b3ac38465133ffb63b75273a8db548c558465d79db03fd359c6cd5bd9d85"


def test_layout_and_compression():
    lay = np.zeros(4, dtype=np.uint32)
    L.or_blake3_layout(lay.ctypes.data)
    assert int(lay[0]) == 11_706 and int(lay[1]) == USEFUL_BITS  # blake3.rs:246-247
    assert int(lay[3]) - 1408 == 183 and int(lay[2]) - 1408 == 183 + 183 + 182 + 182
    for row in random_inputs(4, 1):
        out = np.zeros(16, dtype=np.uint32)
        cv = np.ascontiguousarray(row[:8])
        m = np.ascontiguousarray(row[8:24])
        ctr = int(row[24]) | (int(row[25]) << 32)
        L.or_blake3_compress(cv.ctypes.data, m.ctypes.data, ctr, int(row[26]), int(row[27]),
                             out.ctypes.data)
        exp = M.blake3_compress([int(x) for x in row[:8]], [int(x) for x in row[8:24]], ctr,
                                int(row[26]), int(row[27]))
        assert [int(x) for x in out] == exp
        z, a, b = instance(row)
        assert [word_at(z, 256 + 32 * w) for w in range(8)] == exp[:8]    # out_lo
        assert [word_at(z, 1152 + 32 * w) for w in range(8)] == exp[8:]   # out_hi
        assert [word_at(z, 32 * w) for w in range(8)] == [int(x) for x in row[:8]]
        assert np.array_equal(a & b, z)  # C = I
        hi = USEFUL_BITS >> 6
        tail = ~np.uint64((1 << (USEFUL_BITS & 63)) - 1)
        for buf in (z, a, b):
            assert not buf[hi + 1:].any() and not (buf[hi] & tail)


def test_matrices_reproduce_a_and_b():
    (ap, ac), (bp, bc), _, _ = matrices()
    # 52*184 + 2*183 + 2*182 = 10,298 AND rows per compression (blake3.rs:25-26): every row
    # of the G region [1408, 11706) has a non-empty A and B side
    rows = np.arange(1408, 11_706)
    assert len(rows) == 10_298
    assert (ap[rows + 1] > ap[rows]).all() and (bp[rows + 1] > bp[rows]).all()
    for row in random_inputs(2, 2):
        z, a, b = instance(row)
        ya = np.zeros(256, dtype=np.uint64)
        yb = np.zeros(256, dtype=np.uint64)
        L.or_csr_apply_bits(K, ap.ctypes.data, ac.ctypes.data, z.ctypes.data, ya.ctypes.data)
        L.or_csr_apply_bits(K, bp.ctypes.data, bc.ctypes.data, z.ctypes.data, yb.ctypes.data)
        assert np.array_equal(ya, a) and np.array_equal(yb, b)


def test_batch_major_layout():
    n_log, n_decl = 4, 11
    inputs = random_inputs(n_decl, 3)
    z, a, b, stripe = batch_major(inputs, n_log)
    n = 1 << n_log
    for o in range(n):
        if o < n_decl:
            iz, ia, ib = instance(inputs[o])
        else:
            iz = ia = ib = np.zeros(256, dtype=np.uint64)
        for buf, inst in ((z, iz), (a, ia), (b, ib)):
            assert np.array_equal(buf[np.arange(128) * n + o].reshape(-1), inst)
    zb = np.zeros((n, K), dtype=np.uint8)
    for o in range(n):
        zb[o] = np.unpackbits(z[np.arange(128) * n + o].reshape(-1).view(np.uint8),
                              bitorder="little")
    exp = np.zeros_like(stripe).reshape(-1, K)
    for o in range(n):
        exp[o // 8] |= (zb[o] << (o % 8)).astype(np.uint8)
    assert np.array_equal(stripe.reshape(-1, K), exp)
