"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: the parity checker for the HIP path. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
Field elements travel as numpy uint64 arrays of shape (n, 2) = (lo, hi).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


class F128(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_uint64), ("hi", ctypes.c_uint64)]


class PcsParams(ctypes.Structure):
    _fields_ = [
        ("m", ctypes.c_uint32),
        ("log_inv_rate", ctypes.c_uint32),
        ("log_batch_size", ctypes.c_uint32),
        ("num_lanes", ctypes.c_uint32),
        ("merkle_hash", ctypes.c_uint32),
    ]


class Sparse(ctypes.Structure):
    """or_sparse: a SparseBinaryMatrix as CSR, or the identity."""
    _fields_ = [("num_rows", ctypes.c_uint64), ("num_cols", ctypes.c_uint64),
                ("row_ptr", ctypes.c_void_p), ("cols", ctypes.c_void_p),
                ("identity", ctypes.c_uint32)]


class TypeStmt(ctypes.Structure):
    """or_type_stmt: the digested fields of a boolean TableType."""
    _fields_ = [("k_log", ctypes.c_uint32), ("has_const_pin", ctypes.c_uint32),
                ("useful_bits", ctypes.c_uint64), ("const_pin", ctypes.c_uint64),
                ("a", Sparse), ("b", Sparse), ("c", Sparse)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        S = ctypes.c_size_t
        L.or_f128_mul.restype = F128
        L.or_f128_mul.argtypes = [F128, F128]
        L.or_f128_mul_portable.restype = F128
        L.or_f128_mul_portable.argtypes = [F128, F128]
        L.or_f128_inv.restype = F128
        L.or_f128_inv.argtypes = [F128]
        L.or_f128_mul_by_x.restype = F128
        L.or_f128_mul_by_x.argtypes = [F128]
        L.or_f128_mul_batch.argtypes = [P, P, P, S]
        L.or_splitmix_f128.argtypes = [ctypes.c_uint64, P, S]
        L.or_ntt_evals.argtypes = [ctypes.c_int, P]
        L.or_ntt_twiddle.restype = F128
        L.or_ntt_twiddle.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_uint64]
        L.or_ntt_forward_interleaved.argtypes = [P, S, S, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_ntt_inverse_scalar.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.or_sha256.argtypes = [P, S, P]
        L.or_merkle_tree.argtypes = [P, S, S, P]
        L.or_merkle_tree_kind.argtypes = [P, S, S, ctypes.c_uint32, P]
        L.or_blake3_hash.argtypes = [P, S, P]
        L.or_blake3_hash_pieces.argtypes = [P, P, S, P]
        L.or_blake3_leaf_cv.argtypes = [P, S, P]
        L.or_blake3_parent_cv.argtypes = [P, P, P]
        L.or_registry_digest.argtypes = [P, S, ctypes.c_uint32, P]
        L.or_pcs_sizes.argtypes = [ctypes.POINTER(PcsParams), P, P, P, P, P]
        L.or_commit_lane_major.argtypes = [P, S, ctypes.POINTER(PcsParams), P, P]
        L.or_commit.argtypes = [P, S, ctypes.POINTER(PcsParams), P, P]
        L.or_fs_new.restype = P
        L.or_fs_new.argtypes = [ctypes.c_char_p, S]
        L.or_fs_free.argtypes = [P]
        L.or_fs_observe_label.argtypes = [P, ctypes.c_char_p, S]
        L.or_fs_observe_f128.argtypes = [P, F128]
        L.or_fs_observe_f128_slice.argtypes = [P, P, S]
        L.or_fs_observe_bytes.argtypes = [P, ctypes.c_char_p, S]
        L.or_fs_sample_f128.restype = F128
        L.or_fs_sample_f128.argtypes = [P]
        L.or_fs_sample_f128_vec.argtypes = [P, P, S]
        L.or_fs_grind_pow.restype = ctypes.c_uint64
        L.or_fs_grind_pow.argtypes = [P, ctypes.c_uint32]
        L.or_round_msg_lsb.argtypes = [P, P, S, P]
        L.or_fold_and_msg_lsb.argtypes = [P, P, S, F128, P, P, P]
        L.or_build_eq.argtypes = [P, S, P]
        L.or_round_pair.argtypes = [P, P, S, P, P]
        L.or_fold_pair.argtypes = [P, P, S, F128, P, P]
        L.or_sumcheck_transcript.argtypes = [ctypes.c_char_p, S, P, P, ctypes.c_int, P, P]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def f(v: int) -> F128:
    return F128(v & ((1 << 64) - 1), v >> 64)


def to_int(x: F128) -> int:
    return x.lo | (x.hi << 64)


def arr_to_ints(a: np.ndarray) -> list[int]:
    return [int(lo) | (int(hi) << 64) for lo, hi in a.reshape(-1, 2)]


def ints_to_arr(v) -> np.ndarray:
    out = np.zeros((len(v), 2), dtype=np.uint64)
    for i, x in enumerate(v):
        out[i, 0] = x & ((1 << 64) - 1)
        out[i, 1] = x >> 64
    return out


def mul(a: int, b: int) -> int:
    return to_int(lib().or_f128_mul(f(a), f(b)))


def mul_batch(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    out = np.empty_like(a)
    lib().or_f128_mul_batch(_p(a), _p(b), _p(out), a.shape[0])
    return out


def splitmix(seed: int, n: int) -> np.ndarray:
    out = np.empty((n, 2), dtype=np.uint64)
    lib().or_splitmix_f128(seed, _p(out), n)
    return out


def ntt_forward_interleaved(data: np.ndarray, num_ntts: int, live: int, log_d: int, L: int,
                            start_layer: int) -> np.ndarray:
    out = np.ascontiguousarray(data.copy())
    rc = lib().or_ntt_forward_interleaved(_p(out), num_ntts, live, log_d, L, start_layer)
    assert rc == 0
    return out


def ntt_inverse(data: np.ndarray, log_d: int, L: int) -> np.ndarray:
    out = np.ascontiguousarray(data.copy())
    assert lib().or_ntt_inverse_scalar(_p(out), log_d, L) == 0
    return out


def twiddle(L: int, layer: int, block: int) -> int:
    ev = np.zeros((L * L, 2), dtype=np.uint64)
    lib().or_ntt_evals(L, _p(ev))
    return to_int(lib().or_ntt_twiddle(_p(ev), L, layer, block))


def merkle_tree(data: bytes, num_leaves: int, kind: int = 0) -> bytes:
    tree = ctypes.create_string_buffer(32 * (2 * num_leaves - 1))
    buf = ctypes.create_string_buffer(data, len(data))
    assert lib().or_merkle_tree_kind(buf, len(data), num_leaves, kind, tree) == 0
    return tree.raw


def blake3(data: bytes, pieces=None) -> bytes:
    out = ctypes.create_string_buffer(32)
    if pieces is None:
        lib().or_blake3_hash(data, len(data), out)
    else:
        assert sum(pieces) == len(data)
        pl = (ctypes.c_size_t * len(pieces))(*pieces)
        lib().or_blake3_hash_pieces(data, pl, len(pieces), out)
    return out.raw


def blake3_leaf_cv(data: bytes) -> bytes:
    out = ctypes.create_string_buffer(32)
    lib().or_blake3_leaf_cv(data, len(data), out)
    return out.raw


def blake3_parent_cv(left: bytes, right: bytes) -> bytes:
    out = ctypes.create_string_buffer(32)
    lib().or_blake3_parent_cv(left, right, out)
    return out.raw


def sparse(rows=None, num_rows=None, num_cols=0, identity=False, keep=None):
    """An or_sparse from a list of column lists (rows), or the num_rows x num_rows identity;
    the CSR arrays are appended to `keep` so they outlive the struct."""
    if identity:
        return Sparse(num_rows, num_rows, None, None, 1)
    rows = rows or []
    ptr = np.zeros(len(rows) + 1, dtype=np.uint64)
    ptr[1:] = np.cumsum([len(r) for r in rows]) if rows else []
    cols = np.array([c for r in rows for c in r] or [0], dtype=np.uint32)
    keep.extend([ptr, cols])
    return Sparse(len(rows) if num_rows is None else num_rows, num_cols, ptr.ctypes.data,
                  cols.ctypes.data, 0)


def sparse_csr(row_ptr: np.ndarray, cols: np.ndarray, num_cols: int, keep) -> Sparse:
    ptr = np.ascontiguousarray(row_ptr, dtype=np.uint64)
    c = np.ascontiguousarray(cols, dtype=np.uint32)
    keep.extend([ptr, c])
    return Sparse(len(ptr) - 1, num_cols, ptr.ctypes.data, c.ctypes.data, 0)


def registry_digest(types, nu: int) -> bytes:
    """Registry::digest of boolean types given as TypeStmt structs (caller order)."""
    arr = (TypeStmt * len(types))(*types)
    out = ctypes.create_string_buffer(32)
    assert lib().or_registry_digest(arr, len(types), nu, out) == 0
    return out.raw


def pcs_sizes(m, log_inv_rate, log_batch_size, num_lanes=0):
    p = PcsParams(m, log_inv_rate, log_batch_size, num_lanes)
    ml, cl, nl, nn = (ctypes.c_size_t() for _ in range(4))
    k = ctypes.c_int()
    rc = lib().or_pcs_sizes(ctypes.byref(p), ctypes.byref(ml), ctypes.byref(cl), ctypes.byref(nl),
                            ctypes.byref(k), ctypes.byref(nn))
    assert rc == 0
    return dict(msg_len=ml.value, cw_len=cl.value, n_leaves=nl.value, k_code=k.value,
                num_ntts=nn.value)


def commit_lane_major(q: np.ndarray, m, log_inv_rate, log_batch_size, num_lanes=0,
           merkle_hash=0):
    s = pcs_sizes(m, log_inv_rate, log_batch_size, num_lanes)
    p = PcsParams(m, log_inv_rate, log_batch_size, num_lanes, merkle_hash)
    cw = np.empty((s["cw_len"], 2), dtype=np.uint64)
    tree = np.empty((2 * s["n_leaves"] - 1) * 32, dtype=np.uint8)
    rc = lib().or_commit_lane_major(_p(q), q.shape[0], ctypes.byref(p), _p(cw), _p(tree))
    assert rc == 0
    return cw, tree.tobytes()


def commit(msg: np.ndarray, m, log_inv_rate, log_batch_size, num_lanes=0,
           merkle_hash=0):
    s = pcs_sizes(m, log_inv_rate, log_batch_size, num_lanes)
    p = PcsParams(m, log_inv_rate, log_batch_size, num_lanes, merkle_hash)
    cw = np.empty((s["cw_len"], 2), dtype=np.uint64)
    tree = np.empty((2 * s["n_leaves"] - 1) * 32, dtype=np.uint8)
    rc = lib().or_commit(_p(msg), msg.shape[0], ctypes.byref(p), _p(cw), _p(tree))
    assert rc == 0
    return cw, tree.tobytes()


class Fs:
    """The oracle FsChallenger (SHA-256 discipline)."""

    def __init__(self, domain: bytes):
        self.h = lib().or_fs_new(domain, len(domain))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_fs_free(self.h)
            self.h = None

    def observe_label(self, b: bytes):
        lib().or_fs_observe_label(self.h, b, len(b))

    def observe_f128(self, v: int):
        lib().or_fs_observe_f128(self.h, f(v))

    def observe_f128_slice(self, vs):
        a = ints_to_arr(vs)
        lib().or_fs_observe_f128_slice(self.h, _p(a), len(vs))

    def observe_bytes(self, b: bytes):
        lib().or_fs_observe_bytes(self.h, b, len(b))

    def sample_f128(self) -> int:
        return to_int(lib().or_fs_sample_f128(self.h))

    def sample_f128_vec(self, n: int) -> list[int]:
        a = np.zeros((max(n, 1), 2), dtype=np.uint64)
        lib().or_fs_sample_f128_vec(self.h, _p(a), n)
        return arr_to_ints(a[:n])

    def grind_pow(self, bits: int) -> int:
        return int(lib().or_fs_grind_pow(self.h, bits))


def round_msg_lsb(fa: np.ndarray, ba: np.ndarray):
    out = np.zeros((2, 2), dtype=np.uint64)
    lib().or_round_msg_lsb(_p(fa), _p(ba), fa.shape[0], _p(out))
    return arr_to_ints(out)


def fold_and_msg_lsb(fa: np.ndarray, ba: np.ndarray, r: int):
    n = fa.shape[0]
    nf = np.zeros((n // 2, 2), dtype=np.uint64)
    nb = np.zeros((n // 2, 2), dtype=np.uint64)
    out = np.zeros((2, 2), dtype=np.uint64)
    lib().or_fold_and_msg_lsb(_p(fa), _p(ba), n, f(r), _p(nf), _p(nb), _p(out))
    return nf, nb, arr_to_ints(out)


def build_eq(r: list[int]) -> np.ndarray:
    ra = ints_to_arr(r) if r else np.zeros((1, 2), dtype=np.uint64)
    out = np.zeros((1 << len(r), 2), dtype=np.uint64)
    lib().or_build_eq(_p(ra), len(r), _p(out))
    return out


def round_pair(a: np.ndarray, b: np.ndarray, r: list[int]):
    ra = ints_to_arr(r)
    out = np.zeros((2, 2), dtype=np.uint64)
    lib().or_round_pair(_p(a), _p(b), a.shape[0], _p(ra), _p(out))
    return arr_to_ints(out)


def fold_pair(a: np.ndarray, b: np.ndarray, c: int):
    n = a.shape[0]
    ao = np.zeros((n // 2, 2), dtype=np.uint64)
    bo = np.zeros((n // 2, 2), dtype=np.uint64)
    lib().or_fold_pair(_p(a), _p(b), n, f(c), _p(ao), _p(bo))
    return ao, bo


def sumcheck_transcript(domain: bytes, a: np.ndarray, b: np.ndarray, log_n: int):
    msgs = np.zeros((4 * log_n, 2), dtype=np.uint64)
    finals = np.zeros((4, 2), dtype=np.uint64)
    rc = lib().or_sumcheck_transcript(domain, len(domain), _p(a), _p(b), log_n, _p(msgs),
                                      _p(finals))
    assert rc == 0
    return msgs, finals
