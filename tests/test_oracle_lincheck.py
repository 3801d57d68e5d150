"""Pin the oracle's lincheck (quirky eq, CSC fold, stripe partial fold, top-bit product
sumcheck, the single-slot union prover) against definitions and the sumcheck identities
the verifier checks (crates/flock-core/src/lincheck.rs, lincheck/union.rs)."""
from __future__ import annotations

import ctypes
import random

import numpy as np

import oracle_lib as O
import puremath as M

L = O.lib()
P, S = ctypes.c_void_p, ctypes.c_size_t
L.or_quirky_eq.argtypes = [O.F128, P, S, P]
L.or_csc_fold_split.argtypes = [S, P, P, P, P, P, P, P]
L.or_partial_fold.argtypes = [P, ctypes.c_int, ctypes.c_int, S, S, P, P]
L.or_lc_round_eval.argtypes = [P, P, S, P]
L.or_lc_bind_top.argtypes = [P, S, O.F128]
L.or_lincheck_prove.argtypes = [P, ctypes.c_int, P, P, P, P, ctypes.c_long, P, ctypes.c_int, S,
                                S, O.F128, P, P, ctypes.c_int, P, P, P, P, P, P]
L.or_fs_digest.argtypes = [P, P]
L.or_phi8.restype = O.F128
L.or_phi8.argtypes = [ctypes.c_uint8]

DOMAIN = b"flock-lincheck-test"


def random_csc(k: int, nnz_per_col: int, seed: int):
    """Random boolean k x k matrix in CSC form (col_ptr u32[k+1], rows u32[nnz]), rows
    sorted and distinct within a column (csc_from_rows, lincheck.rs:367-388)."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(0, 2 * nnz_per_col + 1, k)
    counts[rng.integers(0, k, max(1, k // 16))] = 0  # some empty columns
    ptr = np.zeros(k + 1, dtype=np.uint32)
    ptr[1:] = np.cumsum(np.minimum(counts, k))
    rows = np.concatenate([np.sort(rng.choice(k, size=min(int(c), k), replace=False))
                           for c in counts]).astype(np.uint32) if ptr[-1] else \
        np.zeros(0, dtype=np.uint32)
    return ptr, rows


def stripe(k_log: int, nu: int, seed: int, useful_bits: int | None = None,
           n_rows: int | None = None, pin: int = -1, pin_rows: int | None = None) -> np.ndarray:
    """Random lincheck z stripe (z[byte_idx * k + i_inner], bit r = outer row 8*byte_idx + r),
    zero past useful_bits columns and n_rows outer rows when given (honest padding). With
    pin >= 0, column pin is the constant wire: 1 on the first pin_rows (default n_rows) rows
    and 0 below."""
    k = 1 << k_log
    rng = np.random.default_rng(seed)
    z = rng.integers(0, 256, (1 << nu) // 8 * k, dtype=np.uint8).reshape(-1, k)
    if useful_bits is not None:
        z[:, useful_bits:] = 0
    if n_rows is not None:
        for s in range(z.shape[0]):
            for r in range(8):
                if 8 * s + r >= n_rows:
                    z[s] &= np.uint8(~(1 << r) & 0xFF)
    if pin >= 0:
        rows = ((1 << nu) if n_rows is None else n_rows) if pin_rows is None else pin_rows
        z[:, pin] = 0
        for row in range(rows):
            z[row // 8, pin] |= np.uint8(1 << (row % 8))
    return np.ascontiguousarray(z.reshape(-1))


def phi8(v):
    return O.to_int(L.or_phi8(v))


def quirky_eq(z_skip, x_rest):
    out = np.zeros((64 << len(x_rest), 2), dtype=np.uint64)
    xr = O.ints_to_arr(x_rest) if x_rest else np.zeros((1, 2), dtype=np.uint64)
    L.or_quirky_eq(O.f(z_skip), O._p(xr), len(x_rest), O._p(out))
    return out


def csc_fold(ptr_a, rows_a, ptr_b, rows_b, eq):
    k = len(ptr_a) - 1
    sa = np.zeros((k, 2), dtype=np.uint64)
    sb = np.zeros((k, 2), dtype=np.uint64)
    L.or_csc_fold_split(k, ptr_a.ctypes.data, rows_a.ctypes.data, ptr_b.ctypes.data,
                        rows_b.ctypes.data, O._p(eq), O._p(sa), O._p(sb))
    return sa, sb


def partial_fold(z, k_log, nu, useful_bits, n_rows, eq_outer):
    out = np.zeros((1 << k_log, 2), dtype=np.uint64)
    L.or_partial_fold(z.ctypes.data, k_log, nu, useful_bits, n_rows, O._p(eq_outer), O._p(out))
    return out


def lincheck_prove(dom, k_log, csc_a, csc_b, const_pin, z, nu, useful_bits, n_rows, z_skip,
                   x_inner_rest, x_outer, grinding):
    """The oracle prover. Returns dict of numpy arrays."""
    Lr = k_log - 6
    fs = O.Fs(dom)
    out = dict(rounds=np.zeros((max(2 * Lr, 1), 2), dtype=np.uint64),
               z_partial=np.zeros((64, 2), dtype=np.uint64),
               matrix_evals=np.zeros((2, 2), dtype=np.uint64),
               nonces=np.zeros(Lr + 3, dtype=np.uint64),
               claim=np.zeros((2 + Lr, 2), dtype=np.uint64),
               z_vec=np.zeros((1 << k_log, 2), dtype=np.uint64))
    xr = O.ints_to_arr(x_inner_rest) if x_inner_rest else np.zeros((1, 2), dtype=np.uint64)
    xo = O.ints_to_arr(x_outer)
    rc = L.or_lincheck_prove(fs.h, k_log, csc_a[0].ctypes.data, csc_a[1].ctypes.data,
                             csc_b[0].ctypes.data, csc_b[1].ctypes.data, const_pin,
                             z.ctypes.data, nu, useful_bits, n_rows, O.f(z_skip), O._p(xr),
                             O._p(xo), grinding, O._p(out["rounds"]), O._p(out["z_partial"]),
                             O._p(out["matrix_evals"]), out["nonces"].ctypes.data,
                             O._p(out["claim"]), O._p(out["z_vec"]))
    assert rc == 0
    out["rounds"] = out["rounds"][:2 * Lr]
    dig = ctypes.create_string_buffer(32)
    L.or_fs_digest(fs.h, dig)
    out["digest"] = dig.raw
    return out


def test_quirky_eq_by_definition():
    rng = random.Random(31)
    z = rng.getrandbits(128)
    x_rest = [rng.getrandbits(128) for _ in range(2)]
    got = O.arr_to_ints(quirky_eq(z, x_rest))
    w = M.lagrange_f128([phi8(v) for v in range(64)], z)
    for j in range(4):
        e = M.eq_point(x_rest, j)
        for i in range(64):
            assert got[j * 64 + i] == M.mul(w[i], e)
    # at a node phi8(s) the skip weights are the indicator of s
    got = O.arr_to_ints(quirky_eq(phi8(5), []))
    assert got == [1 if i == 5 else 0 for i in range(64)]


def test_csc_fold_split_is_the_transposed_matvec():
    k = 128
    pa, ra = random_csc(k, 3, 1)
    pb, rb = random_csc(k, 2, 2)
    eq = O.splitmix(0x41, k)
    sa, sb = csc_fold(pa, ra, pb, rb, eq)
    e = O.arr_to_ints(eq)
    for (ptr, rows), s in (((pa, ra), sa), ((pb, rb), sb)):
        dense = np.zeros((k, k), dtype=bool)
        for c in range(k):
            dense[rows[ptr[c]:ptr[c + 1]], c] = True
        exp = []
        for c in range(k):
            acc = 0
            for r in np.nonzero(dense[:, c])[0]:
                acc ^= e[r]
            exp.append(acc)
        assert O.arr_to_ints(s) == exp


def test_partial_fold_by_definition():
    k_log, nu = 7, 5
    k = 1 << k_log
    z = stripe(k_log, nu, 3)
    x_outer = M.splitmix(0x42, nu)
    eq_outer = O.ints_to_arr([M.eq_point(x_outer, o) for o in range(1 << nu)])
    bit = lambda i, o: (int(z[(o >> 3) * k + i]) >> (o & 7)) & 1
    # dense: out[i] = z^(i, x_outer) (MLE over the outer rows)
    got = O.arr_to_ints(partial_fold(z, k_log, nu, k, 1 << nu, eq_outer))
    for i in range(k):
        acc = 0
        for o in range(1 << nu):
            if bit(i, o):
                acc ^= M.eq_point(x_outer, o)
        assert got[i] == acc
    # support-proportional: columns < useful_bits, whole stripes below n_rows
    ub, nr = 100, 13
    got = O.arr_to_ints(partial_fold(z, k_log, nu, ub, nr, eq_outer))
    for i in range(k):
        acc = 0
        if i < ub:
            for o in range(8 * ((nr + 7) // 8)):
                if bit(i, o):
                    acc ^= M.eq_point(x_outer, o)
        assert got[i] == acc
    # on an honestly padded witness the restricted fold equals the dense one
    zp = stripe(k_log, nu, 4, useful_bits=ub, n_rows=nr)
    assert np.array_equal(partial_fold(zp, k_log, nu, ub, nr, eq_outer),
                          partial_fold(zp, k_log, nu, k, 1 << nu, eq_outer))


def _mle_low(v, r, lo_bits=6):
    """sum_j eq(r, j) v[i + 2^lo_bits j] for each i < 2^lo_bits."""
    n = 1 << lo_bits
    out = [0] * n
    for j in range(len(v) >> lo_bits):
        e = M.eq_point(r, j)
        for i in range(n):
            out[i] ^= M.mul(e, v[i + n * j])
    return out


def _check_prover(k_log, nu, const_pin, grinding, useful_bits=None, n_rows=None, seed=0):
    k = 1 << k_log
    Lr = k_log - 6
    csc_a = random_csc(k, 3, 10 + seed)
    csc_b = random_csc(k, 2, 20 + seed)
    ub = k if useful_bits is None else useful_bits
    nr = (1 << nu) if n_rows is None else n_rows
    z = stripe(k_log, nu, 30 + seed, useful_bits, n_rows)
    pts = M.splitmix(0x900 + seed, 1 + Lr + nu)
    z_skip, x_rest, x_outer = pts[0], pts[1:1 + Lr], pts[1 + Lr:]
    out = lincheck_prove(DOMAIN, k_log, csc_a, csc_b, const_pin, z, nu, ub, nr, z_skip, x_rest,
                         x_outer, grinding)
    rounds = O.arr_to_ints(out["rounds"])
    zp = O.arr_to_ints(out["z_partial"])
    claim = O.arr_to_ints(out["claim"])
    nonces = [int(v) for v in out["nonces"]]
    # z_vec is the partial fold at x_outer
    eq_outer = O.build_eq(x_outer)
    assert np.array_equal(out["z_vec"], partial_fold(z, k_log, nu, ub, nr, eq_outer))
    zv = O.arr_to_ints(out["z_vec"])
    # replay the transcript (challenger.rs framing, restated in puremath)
    fs = M.FsChallenger(DOMAIN)
    fs.observe_label(b"flock-lincheck-v0")
    ni = 0
    if grinding:
        assert fs.grind_pow(1) == nonces[ni]
        ni += 1
    alpha = fs.sample_f128()
    sa, sb = csc_fold(*csc_a, *csc_b, quirky_eq(z_skip, x_rest))
    sa, sb = O.arr_to_ints(sa), O.arr_to_ints(sb)
    comb = [M.mul(alpha, a) ^ b for a, b in zip(sa, sb)]
    if const_pin >= 0:
        if grinding:
            assert fs.grind_pow(1) == nonces[ni]
            ni += 1
        comb[const_pin] ^= fs.sample_f128()
    # running claim T = sum comb * z_vec; q(0) + q(1) = T each round, T' = q(r)
    T = 0
    for c, v in zip(comb, zv):
        T ^= M.mul(c, v)
    rs = []
    for t in range(Lr):
        e1, ei = rounds[2 * t], rounds[2 * t + 1]
        q0 = T ^ e1
        fs.observe_f128(e1)
        fs.observe_f128(ei)
        if grinding:
            assert fs.grind_pow(2) == nonces[ni]
            ni += 1
        r = fs.sample_f128()
        rs.append(r)
        T = q0 ^ M.mul(q0 ^ e1 ^ ei, r) ^ M.mul(ei, M.mul(r, r))
    r_rest = rs[::-1]
    assert claim[2:] == r_rest
    # z_partial and the bound comb are the low-64 MLEs at r_inner_rest (LSB-first)
    assert zp == _mle_low(zv, r_rest)
    comb_final = _mle_low(comb, r_rest)
    acc = 0
    for c, v in zip(comb_final, zp):
        acc ^= M.mul(c, v)
    assert acc == T
    fs.observe_f128_slice(zp)
    if grinding:
        assert fs.grind_pow(6) == nonces[ni]
        ni += 1
    r_skip = fs.sample_f128()
    assert claim[0] == r_skip
    lam = M.lagrange_f128([phi8(v) for v in range(64)], r_skip)
    w = 0
    for a, b in zip(lam, zp):
        w ^= M.mul(a, b)
    assert claim[1] == w
    # matrix evals: <z_partial (x) eq(r_inner_rest), A-marginal / B-marginal>
    me = O.arr_to_ints(out["matrix_evals"])
    ma = mb = 0
    for j in range(1 << Lr):
        e = M.eq_point(r_rest, j)
        for i in range(64):
            wc = M.mul(e, zp[i])
            ma ^= M.mul(wc, sa[j * 64 + i])
            mb ^= M.mul(wc, sb[j * 64 + i])
    assert me == [ma, mb]
    return out


def test_lincheck_prover_identities_and_transcript():
    _check_prover(8, 4, -1, 0, seed=1)


def test_lincheck_prover_pin_grinding_padded():
    _check_prover(9, 5, 7, 1, useful_bits=300, n_rows=21, seed=2)


def test_lincheck_prover_no_rounds():
    # k_log = k_skip: no multilinear rounds, z_partial = z_vec
    out = _check_prover(6, 3, 0, 1, seed=3)
    assert np.array_equal(out["z_partial"], out["z_vec"])


L.or_lincheck_prove_keccak.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_size_t, O.F128, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int] + [ctypes.c_void_p] * 6


def lincheck_prove_keccak(dom, z, nu, n_rows, z_skip, x_inner_rest, x_outer, grinding):
    """The oracle prover with KeccakLincheckCircuit (k_log 16). Returns dict of numpy arrays."""
    Lr = 10
    fs = O.Fs(dom)
    out = dict(rounds=np.zeros((2 * Lr, 2), dtype=np.uint64),
               z_partial=np.zeros((64, 2), dtype=np.uint64),
               matrix_evals=np.zeros((2, 2), dtype=np.uint64),
               nonces=np.zeros(Lr + 3, dtype=np.uint64),
               claim=np.zeros((2 + Lr, 2), dtype=np.uint64),
               z_vec=np.zeros((1 << 16, 2), dtype=np.uint64))
    xr, xo = O.ints_to_arr(x_inner_rest), O.ints_to_arr(x_outer)
    rc = L.or_lincheck_prove_keccak(fs.h, z.ctypes.data, nu, n_rows, O.f(z_skip), O._p(xr),
                                    O._p(xo), grinding, O._p(out["rounds"]),
                                    O._p(out["z_partial"]), O._p(out["matrix_evals"]),
                                    out["nonces"].ctypes.data, O._p(out["claim"]),
                                    O._p(out["z_vec"]))
    assert rc == 0
    return out
