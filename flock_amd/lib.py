"""Python binding of libflock_hip.so (the C ABI in include/flock_hip.h).

Mirrors the reference's Rust seams for the `prove_fast` data-parallel core:
`pcs::commit_lane_major` / `pcs::commit` (crates/flock-core/src/pcs/commit.rs:345, 474),
`merkle::merkle_tree` (merkle.rs:543), `round_msg_lsb` / `fold_and_msg_lsb`
(pcs/ligerito.rs:3204, 3372), `round_pair_naive` / `fold_and_compute_round_pair_into` /
`fold_in_place_pair` (zerocheck/multilinear.rs:331, 1197, 1148), `build_eq`
(zerocheck/univariate_skip.rs:31) and `FsChallenger` (challenger.rs:609).

There is no CPU fallback: importing this module without the built library raises.
Field elements cross the boundary as numpy uint64 arrays of shape (n, 2) = (lo, hi),
the same bytes as the Rust F128.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflock_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "flock_hip.h")

OK = 0
E_INVALID = 1
E_NOMEM = 2
E_NODEVICE = 3
E_HIP = 100
E_M_TOO_SMALL = 101
E_INTERNAL = 102
E_REJECT = 103


class FlockHipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"flock_hip error {code}: {msg}")
        self.code = code


class F128(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_uint64), ("hi", ctypes.c_uint64)]


class Csc(ctypes.Structure):
    """flock_csc: the CSC form of (A_0, B_0) on the device (CscCircuit, lincheck.rs:353-431)."""

    _fields_ = [
        ("k_log", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("a_col_ptr", ctypes.c_void_p),
        ("a_rows", ctypes.c_void_p),
        ("b_col_ptr", ctypes.c_void_p),
        ("b_rows", ctypes.c_void_p),
        ("nnz_a", ctypes.c_uint64),
        ("nnz_b", ctypes.c_uint64),
        ("const_pin", ctypes.c_int64),
    ]


class TableType(ctypes.Structure):
    """flock_table_type (TableType, schedule.rs:150-215)."""

    _fields_ = [("k_log", ctypes.c_uint32), ("is_element", ctypes.c_uint32),
                ("useful_bits", ctypes.c_uint64)]


class Slot(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_uint64), ("prefix", ctypes.c_uint64),
                ("m_slot", ctypes.c_uint32), ("prefix_bits", ctypes.c_uint32),
                ("type_index", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class Registry(ctypes.Structure):
    """flock_registry (Registry, schedule.rs:275-495), types in slot order."""

    _fields_ = [("n_types", ctypes.c_uint32), ("num_boolean", ctypes.c_uint32),
                ("nu", ctypes.c_uint32), ("m_total", ctypes.c_uint32),
                ("m_bool", ctypes.c_uint32), ("m_elem", ctypes.c_uint32),
                ("element_base", ctypes.c_uint64), ("slots", Slot * 16),
                ("types", TableType * 16)]


class UnionSizes(ctypes.Structure):
    _fields_ = [("dense_words", ctypes.c_uint64), ("committed_words", ctypes.c_uint64),
                ("packed_len", ctypes.c_uint64), ("commit_lanes", ctypes.c_uint64),
                ("dense_m", ctypes.c_uint32), ("compaction_is_identity", ctypes.c_uint32)]


class PcsParams(ctypes.Structure):
    """PcsParams (pcs/commit.rs:31-65); num_lanes = 0 means None."""

    _fields_ = [
        ("m", ctypes.c_uint32),
        ("log_inv_rate", ctypes.c_uint32),
        ("log_batch_size", ctypes.c_uint32),
        ("num_lanes", ctypes.c_uint32),
        ("merkle_hash", ctypes.c_uint32),
    ]


HASH_SHA256, HASH_BLAKE3, HASH_KECCAK = 0, 1, 2
MERKLE_SHA256, MERKLE_BLAKE3 = 0, 1
ROUNDS_HOST, ROUNDS_DEVICE = 0, 2
MUL_SHOUP, MUL_MFMA_I8 = 0, 1


class SparseMatrix(ctypes.Structure):
    """flock_sparse_matrix: a SparseBinaryMatrix (r1cs.rs:14-18) as CSR, or the identity."""

    _fields_ = [("num_rows", ctypes.c_uint64), ("num_cols", ctypes.c_uint64),
                ("row_ptr", ctypes.c_void_p), ("cols", ctypes.c_void_p),
                ("identity", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class TypeStatement(ctypes.Structure):
    """flock_type_statement: the digested fields of a boolean TableType."""

    _fields_ = [("k_log", ctypes.c_uint32), ("has_const_pin", ctypes.c_uint32),
                ("useful_bits", ctypes.c_uint64), ("const_pin", ctypes.c_uint64),
                ("a0", SparseMatrix), ("b0", SparseMatrix), ("c0", SparseMatrix)]


class ProveRequest(ctypes.Structure):
    """flock_prove_request: one prove-core request (include/flock_hip.h)."""

    _fields_ = [("kind", ctypes.c_uint32), ("n_log", ctypes.c_uint32),
                ("n_decl", ctypes.c_uint64), ("h_inputs", ctypes.c_void_p),
                ("circuit", ctypes.c_void_p), ("digest", ctypes.c_char_p),
                ("domain", ctypes.c_char_p), ("domain_len", ctypes.c_size_t),
                ("arrival_s", ctypes.c_double), ("inputs_on_device", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class ProveResult(ctypes.Structure):
    """flock_prove_result."""

    _fields_ = [("root", ctypes.c_uint8 * 32), ("transcript", ctypes.c_uint8 * 32),
                ("zc_z", F128), ("lc_w", F128), ("hashes", ctypes.c_uint64),
                ("queued_ms", ctypes.c_double), ("run_ms", ctypes.c_double),
                ("done_s", ctypes.c_double), ("stage_ms", ctypes.c_double * 4),
                ("status", ctypes.c_int),
                ("reserved", ctypes.c_int)]


class UnionPart(ctypes.Structure):
    """flock_union_part: one hash type of a mixed request."""

    _fields_ = [("kind", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("n_decl", ctypes.c_uint64), ("inputs", ctypes.c_void_p),
                ("circuit", ctypes.c_void_p)]


class UnionRequest(ctypes.Structure):
    """flock_union_request: prove_fast over a registry of several boolean hash types."""

    _fields_ = [("n_log", ctypes.c_uint32), ("n_parts", ctypes.c_uint32),
                ("parts", ctypes.POINTER(UnionPart)), ("digest", ctypes.c_char_p),
                ("domain", ctypes.c_char_p), ("domain_len", ctypes.c_size_t),
                ("inputs_on_device", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class UnionInfo(ctypes.Structure):
    """flock_union_info: slot order, counts and PCS parameters of a mixed request."""

    _fields_ = [("n_slots", ctypes.c_uint32), ("m_bool", ctypes.c_uint32),
                ("dense_m", ctypes.c_uint32), ("log_batch_size", ctypes.c_uint32),
                ("num_lanes", ctypes.c_uint64), ("has_opening", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("device_bytes", ctypes.c_uint64),
                ("slot_part", ctypes.c_uint32 * 16), ("counts", ctypes.c_uint64 * 16)]


class LcSlot(ctypes.Structure):
    """flock_lc_slot: one boolean slot of the union lincheck."""

    _fields_ = [("circuit", ctypes.c_void_p), ("d_z", ctypes.c_void_p),
                ("prefix", ctypes.c_uint64), ("useful_bits", ctypes.c_uint64),
                ("n_rows", ctypes.c_uint64)]


class ProofBytes(ctypes.Structure):
    """flock_proof_bytes: the wire bytes of one prove_fast."""

    _fields_ = [("proof", ctypes.POINTER(ctypes.c_uint8)), ("proof_len", ctypes.c_size_t),
                ("cap", ctypes.POINTER(ctypes.c_uint8)), ("cap_len", ctypes.c_size_t),
                ("claim_ab", F128), ("claim_c", F128)]


_lib = None


def header_functions() -> list[str]:
    """Every function the public header declares."""
    with open(HEADER) as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    return sorted(set(re.findall(r"\b(flock_[a-z0-9_]+)\s*\(", text)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -m flock_amd.build` "
                          "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    P, S, I = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
    U32, U64 = ctypes.c_uint32, ctypes.c_uint64
    PP = ctypes.POINTER(PcsParams)
    sig = {
        "flock_hip_device_count": (I, []),
        "flock_hip_last_error": (ctypes.c_char_p, []),
        "flock_hip_ctx_create": (I, [I, ctypes.POINTER(P)]),
        "flock_hip_ctx_destroy": (None, [P]),
        "flock_hip_ctx_stream": (P, [P]),
        "flock_hip_sync": (I, [P]),
        "flock_hip_ctx_set_round_mode": (I, [P, I]),
        "flock_hip_malloc": (I, [P, ctypes.POINTER(P), S]),
        "flock_hip_dfree": (I, [P, P]),
        "flock_hip_memcpy_h2d": (I, [P, P, P, S]),
        "flock_hip_memcpy_d2h": (I, [P, P, P, S]),
        "flock_hip_free": (None, [P]),
        "flock_hip_f128_mul_batch": (I, [P, P, P, P, S]),
        "flock_hip_f128_mul_const_batch": (I, [P, P, F128, P, S]),
        "flock_hip_f128_mul_const_pow_batch": (I, [P, P, F128, U32, I, P, S]),
        "flock_hip_set_tunable": (I, [ctypes.c_char_p, ctypes.c_int64]),
        "flock_hip_get_tunable": (I, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]),
        "flock_hip_ntt_forward_interleaved": (I, [P, P, S, S, U32, U32]),
        "flock_hip_ntt_pass": (I, [P, P, S, S, U32, U32, U32]),
        "flock_hip_pcs_sizes": (I, [PP, P, P, P, P, P]),
        "flock_hip_commit_lane_major": (I, [P, P, S, PP, P, P, U32, P]),
        "flock_hip_commit": (I, [P, P, S, PP, P, P, U32, P]),
        "flock_hip_merkle_tree": (I, [P, P, S, S, P]),
        "flock_hip_merkle_tree_kind": (I, [P, P, S, S, U32, P]),
        "flock_registry_digest": (I, [P, S, U32, P]),
        "flock_circuit_get": (I, [U32, ctypes.POINTER(P)]),
        "flock_circuit_info": (I, [P, ctypes.POINTER(U32), ctypes.POINTER(U64),
                                   ctypes.POINTER(U64), ctypes.POINTER(U64), ctypes.POINTER(U64)]),
        "flock_circuit_csr": (I, [P, U32, ctypes.POINTER(P), ctypes.POINTER(P)]),
        "flock_circuit_csc": (I, [P, U32, ctypes.POINTER(P), ctypes.POINTER(P)]),
        "flock_circuit_type_statement": (I, [P, P]),
        "flock_circuit_registry_digest": (I, [U32, U32, P]),
        "flock_circuit_union_digest": (I, [P, S, U32, P]),
        "flock_hip_circuit_upload": (I, [P, U32, P]),
        "flock_hip_circuit_release": (None, [P, P]),
        "flock_blake3_hash": (None, [P, S, P]),
        "flock_hip_encode": (I, [P, P, S, PP, P]),
        "flock_hip_encode_lane_major": (I, [P, P, S, PP, P]),
        "flock_hip_sumcheck_round_msg": (I, [P, P, P, S, P]),
        "flock_hip_sumcheck_fold_msg": (I, [P, P, P, S, F128, P, P, P]),
        "flock_hip_sumcheck_fold_msg_static": (I, [P, P, P, S, F128, P, P, P]),
        "flock_hip_build_eq": (I, [P, P, S, P]),
        "flock_hip_zc_round_pair": (I, [P, P, P, S, P, F128, P, P]),
        "flock_hip_zc_fold_round_pair": (I, [P, P, P, S, F128, P, P, P, F128, P, P]),
        "flock_hip_fold_pair": (I, [P, P, P, S, F128, P, P]),
        "flock_hip_prove_sumcheck": (I, [P, P, P, P, U32, P, P]),
        "flock_hip_zc_round1": (I, [P, P, P, P, U32, P, P]),
        "flock_hip_zc_fold_at_z": (I, [P, P, P, U32, F128, P, P]),
        "flock_hip_zerocheck_prove": (I, [P, P, P, P, P, U32, I, P, P, P, P, P, P]),
        "flock_hip_fold_1b_rows": (I, [P, P, S, P, P]),
        "flock_registry_new": (I, [P, S, U32, ctypes.POINTER(Registry)]),
        "flock_union_sizes_compute": (I, [ctypes.POINTER(Registry), P, U32, U32,
                                          ctypes.POINTER(UnionSizes)]),
        "flock_union_x_ab_from_mlv": (I, [ctypes.POINTER(Registry), P, S, P, P]),
        "flock_union_bind_statement": (I, [P, ctypes.c_char_p, P, S, ctypes.c_char_p, S]),
        "flock_hip_union_compact": (I, [P, ctypes.POINTER(Registry), P, U32, P, P]),
        "flock_hip_witness_sha256": (I, [P, P, S, U32, P, P, P, P]),
        "flock_hip_witness_keccak": (I, [P, P, S, U32, P, P, P, P]),
        "flock_hip_witness_blake3": (I, [P, P, S, U32, P, P, P, P]),
        "flock_hip_lc_csc_fold": (I, [P, ctypes.POINTER(Csc), P, F128, P, P, P]),
        "flock_hip_lc_partial_fold": (I, [P, P, U32, U32, S, S, P, P]),
        "flock_hip_lincheck_prove": (I, [P, P, ctypes.POINTER(Csc), P, U32, S, S, F128, P, P, I,
                                         P, P, P, P, P, P]),
        "flock_hip_lc_keccak_fold": (I, [P, P, F128, P, P, P]),
        "flock_hip_lincheck_prove_keccak": (I, [P, P, P, U32, S, F128, P, P, I, P, P, P, P, P,
                                                P]),
        "flock_hip_lincheck_prove_union": (I, [P, P, P, S, U32, U32, F128, P, P, I, P, P, P, P,
                                               P, P]),
        "flock_fs_new": (P, [ctypes.c_char_p, S]),
        "flock_fs_free": (None, [P]),
        "flock_fs_observe_label": (None, [P, ctypes.c_char_p, S]),
        "flock_fs_observe_f128": (None, [P, F128]),
        "flock_fs_observe_f128_slice": (None, [P, P, S]),
        "flock_fs_observe_bytes": (None, [P, ctypes.c_char_p, S]),
        "flock_fs_sample_f128": (None, [P, ctypes.POINTER(F128)]),
        "flock_fs_sample_f128_vec": (None, [P, P, S]),
        "flock_fs_grind_pow": (U64, [P, U32]),
        "flock_fs_verify_pow": (I, [P, U64, U32]),
        "flock_zerocheck_verify": (I, [P, U32, I, P, P, P, P, P]),
        "flock_lincheck_verify": (I, [P, P, U32, ctypes.c_int64, U32, S, F128, P, P, F128, F128,
                                      I, P, P, P, P, P]),
        "flock_lincheck_verify_keccak": (I, [P, U32, S, F128, P, P, F128, F128, I, P, P, P, P]),
        "flock_fs_digest": (None, [P, ctypes.c_char_p]),
        "flock_hip_prover_create": (I, [I, ctypes.POINTER(P)]),
        "flock_hip_prover_destroy": (None, [P]),
        "flock_hip_prover_ctx": (P, [P]),
        "flock_hip_prove_footprint": (I, [ctypes.POINTER(ProveRequest), ctypes.POINTER(U64)]),
        "flock_hip_prove_core": (I, [P, ctypes.POINTER(ProveRequest), ctypes.POINTER(ProveResult)]),
        "flock_hip_serve": (I, [ctypes.POINTER(P), U32, ctypes.POINTER(ProveRequest), S, U64,
                                ctypes.POINTER(ProveResult), ctypes.POINTER(ProofBytes),
                                ctypes.POINTER(ctypes.c_double)]),
        "flock_open_ws_create": (I, [ctypes.POINTER(P)]),
        "flock_open_ws_destroy": (None, [P]),
        "flock_open_ws_bytes": (U64, [P]),
        "flock_open_ws_stage_ms": (I, [P, ctypes.POINTER(ctypes.c_double), S]),
        "flock_hip_prover_open_ws": (P, [P]),
        "flock_hip_prover_held_bytes": (U64, [P]),
        "flock_hip_prover_release": (None, [P]),
        "flock_lig_initial_k": (I, [U32]),
        "flock_l0_cap_depth": (I, [U32]),
        "flock_hip_s_hat_v_from_z_vec": (I, [P, P, P, S, P, S, P]),
        "flock_hip_open_merged": (I, [P, P, P, P, PP, P, P, S, P, S, P, P, S, U32,
                                      ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8)),
                                      ctypes.POINTER(S), P]),
        "flock_hip_prove_fast": (I, [P, ctypes.POINTER(ProveRequest), ctypes.POINTER(ProveResult),
                                     ctypes.POINTER(ProofBytes)]),
        "flock_union_request_info": (I, [ctypes.POINTER(UnionRequest), ctypes.POINTER(UnionInfo)]),
        "flock_verify_fast": (I, [ctypes.POINTER(ProveRequest), ctypes.c_char_p, S, ctypes.c_char_p, S,
                                  ctypes.POINTER(F128), ctypes.POINTER(F128)]),
        "flock_verify_fast_union": (I, [ctypes.POINTER(UnionRequest), ctypes.c_char_p, S,
                                        ctypes.c_char_p, S, ctypes.POINTER(F128),
                                        ctypes.POINTER(F128)]),
        "flock_hip_prove_fast_union": (I, [P, ctypes.POINTER(UnionRequest),
                                           ctypes.POINTER(ProveResult), ctypes.POINTER(ProofBytes)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc: int):
    if rc != OK:
        raise FlockHipError(rc, load().flock_hip_last_error().decode())


def f128(v: int) -> F128:
    return F128(v & ((1 << 64) - 1), v >> 64)


def to_int(x: F128) -> int:
    return int(x.lo) | (int(x.hi) << 64)


def _np_ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def _ints_arr(v) -> np.ndarray:
    """F128 ints -> (n, 2) uint64 array (one dummy word for an empty list)."""
    out = np.zeros((max(len(v), 1), 2), dtype=np.uint64)
    for i, x in enumerate(v):
        out[i, 0] = x & ((1 << 64) - 1)
        out[i, 1] = x >> 64
    return out


def _ints(a: np.ndarray) -> list[int]:
    return [int(lo) | (int(hi) << 64) for lo, hi in a.reshape(-1, 2)]


def device_count() -> int:
    return int(load().flock_hip_device_count())


# ---------------------------------------------------------------------------
# Context and device buffers
# ---------------------------------------------------------------------------
class DeviceBuffer:
    """A device allocation owned by a Context."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(load().flock_hip_malloc(ctx.h, ctypes.byref(p), self.nbytes))
        self.ptr = p.value

    @property
    def words(self) -> int:
        return self.nbytes // 16

    def upload(self, arr: np.ndarray, offset: int = 0):
        arr = np.ascontiguousarray(arr)
        assert offset + arr.nbytes <= self.nbytes
        _check(load().flock_hip_memcpy_h2d(self.ctx.h, self.ptr + offset, _np_ptr(arr), arr.nbytes))

    def download(self, nbytes: int | None = None, offset: int = 0) -> bytes:
        n = self.nbytes - offset if nbytes is None else nbytes
        out = np.empty(n, dtype=np.uint8)
        _check(load().flock_hip_memcpy_d2h(self.ctx.h, _np_ptr(out), self.ptr + offset, n))
        return out.tobytes()

    def download_f128(self, n: int | None = None, offset_words: int = 0) -> np.ndarray:
        n = self.words - offset_words if n is None else n
        out = np.empty((n, 2), dtype=np.uint64)
        _check(load().flock_hip_memcpy_d2h(self.ctx.h, _np_ptr(out), self.ptr + 16 * offset_words,
                                           16 * n))
        return out

    def free(self):
        if self.ptr:
            _check(load().flock_hip_dfree(self.ctx.h, self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                load().flock_hip_dfree(self.ctx.h, self.ptr)
        except Exception:
            pass


class Context:
    """One HIP stream + scratch on one device (flock_hip_ctx)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(load().flock_hip_ctx_create(device, ctypes.byref(h)))
        self.h = h.value

    def close(self):
        if self.h:
            load().flock_hip_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def upload(self, arr: np.ndarray) -> DeviceBuffer:
        arr = np.ascontiguousarray(arr)
        buf = DeviceBuffer(self, max(arr.nbytes, 16))
        if arr.nbytes:
            buf.upload(arr)
        return buf

    def sync(self):
        _check(load().flock_hip_sync(self.h))

    def set_round_mode(self, mode: int):
        """flock_hip_ctx_set_round_mode: ROUNDS_HOST (default) or ROUNDS_DEVICE."""
        _check(load().flock_hip_ctx_set_round_mode(self.h, mode))

    @property
    def stream(self) -> int:
        return int(load().flock_hip_ctx_stream(self.h) or 0)

    # ---- field ----
    def f128_mul_batch(self, a: DeviceBuffer, b: DeviceBuffer, out: DeviceBuffer, n: int):
        _check(load().flock_hip_f128_mul_batch(self.h, a.ptr, b.ptr, out.ptr, n))

    def f128_mul_const_batch(self, a: DeviceBuffer, t: int, out: DeviceBuffer, n: int):
        _check(load().flock_hip_f128_mul_const_batch(self.h, a.ptr, f128(t), out.ptr, n))

    def f128_mul_const_pow_batch(self, a: DeviceBuffer, t: int, reps: int, method: int,
                                 out: DeviceBuffer, n: int):
        """flock_hip_f128_mul_const_pow_batch: out[i] = a[i] * t^reps (method MUL_SHOUP or
        MUL_MFMA_I8)."""
        _check(load().flock_hip_f128_mul_const_pow_batch(self.h, a.ptr, f128(t), reps, method,
                                                          out.ptr, n))

    # ---- NTT / commit ----
    def ntt_forward_interleaved(self, data: DeviceBuffer, num_ntts: int, live: int, log_d: int,
                                start_layer: int):
        _check(load().flock_hip_ntt_forward_interleaved(self.h, data.ptr, num_ntts, live, log_d,
                                                        start_layer))

    def ntt_pass(self, data: DeviceBuffer, num_ntts: int, live: int, log_d: int,
                 first_layer: int, layers: int):
        _check(load().flock_hip_ntt_pass(self.h, data.ptr, num_ntts, live, log_d, first_layer,
                                         layers))

    def merkle_tree(self, data: DeviceBuffer, data_len: int, n_leaves: int, tree: DeviceBuffer,
                    kind: int = MERKLE_SHA256):
        _check(load().flock_hip_merkle_tree_kind(self.h, data.ptr, data_len, n_leaves, kind,
                                                 tree.ptr))

    def commit_lane_major(self, q: DeviceBuffer, q_len: int, params: PcsParams,
                          codeword: DeviceBuffer, tree: DeviceBuffer, cap_depth: int = 0) -> list[bytes]:
        cap = ctypes.create_string_buffer(32 << cap_depth)
        _check(load().flock_hip_commit_lane_major(self.h, q.ptr, q_len, ctypes.byref(params),
                                                  codeword.ptr, tree.ptr, cap_depth, cap))
        return [cap.raw[32 * i:32 * i + 32] for i in range(1 << cap_depth)]

    def encode(self, msg: DeviceBuffer, msg_len: int, params: PcsParams, codeword: DeviceBuffer):
        _check(load().flock_hip_encode(self.h, msg.ptr, msg_len, ctypes.byref(params), codeword.ptr))

    def encode_lane_major(self, q: DeviceBuffer, q_len: int, params: PcsParams,
                          codeword: DeviceBuffer):
        _check(load().flock_hip_encode_lane_major(self.h, q.ptr, q_len, ctypes.byref(params),
                                                  codeword.ptr))

    def commit(self, msg: DeviceBuffer, msg_len: int, params: PcsParams, codeword: DeviceBuffer,
               tree: DeviceBuffer, cap_depth: int = 0) -> list[bytes]:
        cap = ctypes.create_string_buffer(32 << cap_depth)
        _check(load().flock_hip_commit(self.h, msg.ptr, msg_len, ctypes.byref(params),
                                       codeword.ptr, tree.ptr, cap_depth, cap))
        return [cap.raw[32 * i:32 * i + 32] for i in range(1 << cap_depth)]

    # ---- sumcheck rounds ----
    def round_msg_lsb(self, f: DeviceBuffer, b: DeviceBuffer, n: int) -> tuple[int, int]:
        out = (F128 * 2)()
        _check(load().flock_hip_sumcheck_round_msg(self.h, f.ptr, b.ptr, n, out))
        return to_int(out[0]), to_int(out[1])

    def fold_and_msg_lsb(self, f, b, n, r, nf, nb) -> tuple[int, int]:
        out = (F128 * 2)()
        _check(load().flock_hip_sumcheck_fold_msg(self.h, f.ptr, b.ptr, n, f128(r), nf.ptr, nb.ptr,
                                                  out))
        return to_int(out[0]), to_int(out[1])

    def build_eq(self, r: list[int], out: DeviceBuffer):
        arr = np.zeros((max(len(r), 1), 2), dtype=np.uint64)
        for i, v in enumerate(r):
            arr[i] = (v & ((1 << 64) - 1), v >> 64)
        _check(load().flock_hip_build_eq(self.h, _np_ptr(arr), len(r), out.ptr))

    def round_pair(self, a, b, n, eq, r0, eq_next=None) -> tuple[int, int]:
        out = (F128 * 2)()
        _check(load().flock_hip_zc_round_pair(self.h, a.ptr, b.ptr, n, eq.ptr, f128(r0),
                                              eq_next.ptr if eq_next else None, out))
        return to_int(out[0]), to_int(out[1])

    def fold_and_compute_round_pair(self, a, b, n, rho, ao, bo, eq, r0, eq_next=None):
        out = (F128 * 2)()
        _check(load().flock_hip_zc_fold_round_pair(self.h, a.ptr, b.ptr, n, f128(rho), ao.ptr,
                                                   bo.ptr, eq.ptr, f128(r0),
                                                   eq_next.ptr if eq_next else None, out))
        return to_int(out[0]), to_int(out[1])

    def fold_pair(self, a, b, n, c, ao, bo):
        _check(load().flock_hip_fold_pair(self.h, a.ptr, b.ptr, n, f128(c), ao.ptr, bo.ptr))

    # ---- zerocheck front ----
    def zc_round1(self, a: DeviceBuffer, b: DeviceBuffer, c: DeviceBuffer, m: int,
                  eq: DeviceBuffer) -> np.ndarray:
        out = np.zeros((128, 2), dtype=np.uint64)
        _check(load().flock_hip_zc_round1(self.h, a.ptr, b.ptr, c.ptr, m, eq.ptr, _np_ptr(out)))
        return out

    def zc_fold_at_z(self, a: DeviceBuffer, b: DeviceBuffer, m: int, z: int, az: DeviceBuffer,
                     bz: DeviceBuffer):
        _check(load().flock_hip_zc_fold_at_z(self.h, a.ptr, b.ptr, m, f128(z), az.ptr, bz.ptr))

    def zerocheck_prove(self, fs: "FsChallenger", a: DeviceBuffer, b: DeviceBuffer,
                        c: DeviceBuffer, m: int, grinding: bool, capture: bool = False):
        """prove_packed_padded_inner (zerocheck.rs:618-1011). Returns (round1[128],
        multilinear[2(m-6)], finals[3], nonces, claim[1 + 2(m-6)]), plus s_hat_v_c[128]
        when capture is set (zerocheck.rs:598-615)."""
        n_mlv = m - 6
        r1 = np.zeros((128, 2), dtype=np.uint64)
        ml = np.zeros((2 * n_mlv, 2), dtype=np.uint64)
        fin = np.zeros((3, 2), dtype=np.uint64)
        nonces = np.zeros(2 + n_mlv, dtype=np.uint64)
        claim = np.zeros((1 + 2 * n_mlv, 2), dtype=np.uint64)
        shv = np.zeros((128, 2), dtype=np.uint64)
        _check(load().flock_hip_zerocheck_prove(self.h, fs.h, a.ptr, b.ptr, c.ptr, m,
                                                1 if grinding else 0, _np_ptr(r1), _np_ptr(ml),
                                                _np_ptr(fin), _np_ptr(nonces), _np_ptr(claim),
                                                _np_ptr(shv) if capture else None))
        if capture:
            return r1, ml, fin, nonces, claim, shv
        return r1, ml, fin, nonces, claim

    # ---- witness ----
    def witness_sha256(self, inputs: DeviceBuffer | None, n_decl: int, n_log: int,
                       z: DeviceBuffer, a: DeviceBuffer, b: DeviceBuffer,
                       stripe: DeviceBuffer | None = None):
        """generate_witness_batch_major_partial (r1cs_hashes/sha2.rs:1876-1898)."""
        _check(load().flock_hip_witness_sha256(self.h, inputs.ptr if inputs else None, n_decl,
                                               n_log, z.ptr, a.ptr, b.ptr,
                                               stripe.ptr if stripe else None))

    def witness_blake3(self, inputs: DeviceBuffer | None, n_decl: int, n_log: int,
                       z: DeviceBuffer, a: DeviceBuffer, b: DeviceBuffer,
                       stripe: DeviceBuffer | None = None):
        """generate_witness_batch_major_partial (r1cs_hashes/blake3.rs:2072-2088)."""
        _check(load().flock_hip_witness_blake3(self.h, inputs.ptr if inputs else None, n_decl,
                                               n_log, z.ptr, a.ptr, b.ptr,
                                               stripe.ptr if stripe else None))

    def witness_keccak(self, inputs: DeviceBuffer | None, n_decl: int, n_log: int,
                       z: DeviceBuffer, a: DeviceBuffer, b: DeviceBuffer,
                       stripe: DeviceBuffer | None = None):
        """generate_witness_batch_major (r1cs_hashes/keccak.rs:1228-1285)."""
        _check(load().flock_hip_witness_keccak(self.h, inputs.ptr if inputs else None, n_decl,
                                               n_log, z.ptr, a.ptr, b.ptr,
                                               stripe.ptr if stripe else None))

    # ---- union ----
    def union_compact(self, reg: Registry, counts, z_padded: DeviceBuffer, q: DeviceBuffer,
                      dense_floor_m: int = 0):
        """compact_witness_unchecked (union.rs:449-490)."""
        c = np.ascontiguousarray(counts, dtype=np.uint64)
        _check(load().flock_hip_union_compact(self.h, ctypes.byref(reg), c.ctypes.data,
                                              dense_floor_m, z_padded.ptr, q.ptr))

    # ---- lincheck ----
    def csc(self, k_log: int, a_ptr: np.ndarray, a_rows: np.ndarray, b_ptr: np.ndarray,
            b_rows: np.ndarray, const_pin: int = -1):
        """Upload a CSC pair; returns (Csc, buffers) -- keep the buffers alive while used."""
        bufs = [self.upload(np.ascontiguousarray(x, dtype=np.uint32))
                if len(x) else self.alloc(4) for x in (a_ptr, a_rows, b_ptr, b_rows)]
        c = Csc(k_log, 0, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, len(a_rows),
                len(b_rows), const_pin)
        return c, bufs

    def circuit_csc(self, kind: int) -> "Csc":
        """Upload a built-in circuit's CSC (flock_hip_circuit_upload); release with
        release_circuit."""
        c = Csc()
        _check(load().flock_hip_circuit_upload(self.h, kind, ctypes.byref(c)))
        return c

    def release_circuit(self, c: "Csc"):
        load().flock_hip_circuit_release(self.h, ctypes.byref(c))

    def lc_csc_fold(self, circuit: "Csc", eq: DeviceBuffer, alpha: int, sa=None, sb=None,
                    comb=None):
        _check(load().flock_hip_lc_csc_fold(self.h, ctypes.byref(circuit), eq.ptr, f128(alpha),
                                            sa.ptr if sa else None, sb.ptr if sb else None,
                                            comb.ptr if comb else None))

    def lc_partial_fold(self, z: DeviceBuffer, k_log: int, nu: int, useful_bits: int,
                        n_rows: int, eq_outer: DeviceBuffer, out: DeviceBuffer):
        _check(load().flock_hip_lc_partial_fold(self.h, z.ptr, k_log, nu, useful_bits, n_rows,
                                                eq_outer.ptr, out.ptr))

    @staticmethod
    def _lincheck_out(k_log: int) -> dict:
        L = k_log - 6
        return dict(rounds=np.zeros((max(2 * L, 1), 2), dtype=np.uint64),
                    z_partial=np.zeros((64, 2), dtype=np.uint64),
                    matrix_evals=np.zeros((2, 2), dtype=np.uint64),
                    nonces=np.zeros(L + 3, dtype=np.uint64),
                    claim=np.zeros((2 + L, 2), dtype=np.uint64))

    def lincheck_prove(self, fs: "FsChallenger", circuit: "Csc", z: DeviceBuffer, nu: int,
                       useful_bits: int, n_rows: int, z_skip: int, x_inner_rest, x_outer,
                       grinding: bool, z_vec: DeviceBuffer | None = None) -> dict:
        """prove_union_capture_z_vec_with_grinding (lincheck/union.rs:203-430), one slot."""
        L = circuit.k_log - 6
        out = self._lincheck_out(circuit.k_log)
        xr = _ints_arr(x_inner_rest)
        xo = _ints_arr(x_outer)
        _check(load().flock_hip_lincheck_prove(
            self.h, fs.h, ctypes.byref(circuit), z.ptr, nu, useful_bits, n_rows, f128(z_skip),
            _np_ptr(xr), _np_ptr(xo), 1 if grinding else 0, _np_ptr(out["rounds"]),
            _np_ptr(out["z_partial"]), _np_ptr(out["matrix_evals"]), _np_ptr(out["nonces"]),
            _np_ptr(out["claim"]), z_vec.ptr if z_vec else None))
        out["rounds"] = out["rounds"][:2 * L]
        return out

    def lc_keccak_fold(self, eq: DeviceBuffer, alpha: int, sa: DeviceBuffer, sb: DeviceBuffer,
                       comb: DeviceBuffer | None = None):
        """KeccakLincheckCircuit marginals (r1cs_hashes/keccak.rs:692-835) over 2^16 columns."""
        _check(load().flock_hip_lc_keccak_fold(self.h, eq.ptr, f128(alpha), sa.ptr, sb.ptr,
                                               comb.ptr if comb else None))

    def lincheck_prove_keccak(self, fs: "FsChallenger", z: DeviceBuffer, nu: int, n_rows: int,
                              z_skip: int, x_inner_rest, x_outer, grinding: bool,
                              z_vec: DeviceBuffer | None = None) -> dict:
        """lincheck_prove with KeccakLincheckCircuit (k_log 16, keccak.rs:678-835)."""
        out = self._lincheck_out(16)
        xr = _ints_arr(x_inner_rest)
        xo = _ints_arr(x_outer)
        _check(load().flock_hip_lincheck_prove_keccak(
            self.h, fs.h, z.ptr, nu, n_rows, f128(z_skip), _np_ptr(xr), _np_ptr(xo),
            1 if grinding else 0, _np_ptr(out["rounds"]), _np_ptr(out["z_partial"]),
            _np_ptr(out["matrix_evals"]), _np_ptr(out["nonces"]), _np_ptr(out["claim"]),
            z_vec.ptr if z_vec else None))
        out["rounds"] = out["rounds"][:2 * 10]
        return out

    def lincheck_prove_union(self, fs: "FsChallenger", slots, col_vars: int, nu: int,
                             z_skip: int, x_inner_rest, x_outer, grinding: bool,
                             z_vec: DeviceBuffer | None = None) -> dict:
        """The union lincheck over several boolean slots (lincheck/union.rs:203-430). slots:
        [(circuit Csc or None for the Keccak walker, stripe DeviceBuffer, prefix, useful_bits,
        n_rows)] in slot order; matrix_evals holds one pair per slot."""
        arr = (LcSlot * len(slots))()
        pins = 0
        for i, (circ, z, prefix, ub, n_rows) in enumerate(slots):
            arr[i].circuit = ctypes.addressof(circ) if circ is not None else None
            arr[i].d_z, arr[i].prefix, arr[i].useful_bits, arr[i].n_rows = z.ptr, prefix, ub, n_rows
            pins += 1 if circ is None or circ.const_pin >= 0 else 0
        L = col_vars - 6
        out = dict(rounds=np.zeros((max(2 * L, 1), 2), dtype=np.uint64),
                   z_partial=np.zeros((64, 2), dtype=np.uint64),
                   matrix_evals=np.zeros((2 * len(slots), 2), dtype=np.uint64),
                   nonces=np.zeros(L + 2 + pins, dtype=np.uint64),
                   claim=np.zeros((2 + L, 2), dtype=np.uint64))
        xr = _ints_arr(x_inner_rest)
        xo = _ints_arr(x_outer)
        _check(load().flock_hip_lincheck_prove_union(
            self.h, fs.h, arr, len(slots), col_vars, nu, f128(z_skip), _np_ptr(xr), _np_ptr(xo),
            1 if grinding else 0, _np_ptr(out["rounds"]), _np_ptr(out["z_partial"]),
            _np_ptr(out["matrix_evals"]), _np_ptr(out["nonces"]), _np_ptr(out["claim"]),
            z_vec.ptr if z_vec else None))
        out["rounds"] = out["rounds"][:2 * L]
        return out

    def fold_1b_rows(self, rows: DeviceBuffer, n_rows: int, t: DeviceBuffer) -> np.ndarray:
        """fold_1b_rows (pcs/ring_switch.rs:182-230): out[j] = sum_y t[y] bit_j(rows[y])."""
        out = np.zeros((128, 2), dtype=np.uint64)
        _check(load().flock_hip_fold_1b_rows(self.h, rows.ptr, n_rows, t.ptr, _np_ptr(out)))
        return out

    def prove_sumcheck(self, fs: "FsChallenger", a: DeviceBuffer, b: DeviceBuffer, log_n: int):
        msgs = np.zeros((4 * log_n, 2), dtype=np.uint64)
        finals = np.zeros((4, 2), dtype=np.uint64)
        _check(load().flock_hip_prove_sumcheck(self.h, fs.h, a.ptr, b.ptr, log_n, _np_ptr(msgs),
                                               _np_ptr(finals)))
        return msgs, finals


def pcs_sizes(params: PcsParams) -> dict:
    ml, cl, nl, nn = (ctypes.c_size_t() for _ in range(4))
    k = ctypes.c_uint32()
    _check(load().flock_hip_pcs_sizes(ctypes.byref(params), ctypes.byref(ml), ctypes.byref(cl),
                                      ctypes.byref(nl), ctypes.byref(k), ctypes.byref(nn)))
    return dict(msg_len=ml.value, cw_len=cl.value, n_leaves=nl.value, k_code=k.value,
                num_ntts=nn.value)


def blake3(data: bytes) -> bytes:
    """blake3::hash of `data` (host)."""
    out = ctypes.create_string_buffer(32)
    load().flock_blake3_hash(data, len(data), out)
    return out.raw


def sparse_matrix(row_ptr=None, cols=None, num_cols: int = 0, num_rows=None,
                  identity: bool = False, keep=None) -> SparseMatrix:
    """A flock_sparse_matrix over numpy CSR arrays (kept alive in `keep`), or the identity."""
    if identity:
        return SparseMatrix(num_rows, num_rows, None, None, 1, 0)
    if row_ptr is None:
        return SparseMatrix(0, num_cols, None, None, 0, 0)
    ptr = np.ascontiguousarray(row_ptr, dtype=np.uint64)
    c = np.ascontiguousarray(cols if len(cols) else np.zeros(1), dtype=np.uint32)
    keep.extend([ptr, c])
    return SparseMatrix(len(ptr) - 1 if num_rows is None else num_rows, num_cols,
                        ptr.ctypes.data, c.ctypes.data, 0, 0)


class Circuit:
    """A built-in hash circuit (flock_circuit_get): (A_0, B_0) as CSR and CSC numpy copies."""

    def __init__(self, kind: int):
        h = ctypes.c_void_p()
        _check(load().flock_circuit_get(kind, ctypes.byref(h)))
        self.h, self.kind = h, kind
        U64 = ctypes.c_uint64
        k_log, useful, pin, na, nb = ctypes.c_uint32(), U64(), U64(), U64(), U64()
        _check(load().flock_circuit_info(h, ctypes.byref(k_log), ctypes.byref(useful),
                                         ctypes.byref(pin), ctypes.byref(na), ctypes.byref(nb)))
        self.k_log, self.useful_bits, self.const_pin = k_log.value, useful.value, pin.value
        self.nnz = (na.value, nb.value)

    def _view(self, fn, which, ptr_ty, n_ptr):
        p, c = ctypes.c_void_p(), ctypes.c_void_p()
        _check(fn(self.h, which, ctypes.byref(p), ctypes.byref(c)))
        ptr = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ptr_ty)), (n_ptr,)).copy()
        n = int(ptr[-1])
        cols = (np.ctypeslib.as_array(ctypes.cast(c, ctypes.POINTER(ctypes.c_uint32)), (n,)).copy()
                if n else np.zeros(0, dtype=np.uint32))
        return ptr, cols

    def csr(self, which: int):
        """(row_ptr uint64, cols uint32) of A_0 (which=0) or B_0 (which=1)."""
        return self._view(load().flock_circuit_csr, which, ctypes.c_uint64, (1 << self.k_log) + 1)

    def csc(self, which: int):
        """(col_ptr uint32, rows uint32) of A_0 (which=0) or B_0 (which=1)."""
        return self._view(load().flock_circuit_csc, which, ctypes.c_uint32, (1 << self.k_log) + 1)

    def type_statement(self) -> "TypeStatement":
        t = TypeStatement()
        _check(load().flock_circuit_type_statement(self.h, ctypes.byref(t)))
        return t


def circuit_registry_digest(kind: int, nu: int) -> bytes:
    """Registry::digest of the one-type registry of a built-in circuit at nu."""
    out = ctypes.create_string_buffer(32)
    _check(load().flock_circuit_registry_digest(kind, nu, out))
    return out.raw


def circuit_union_digest(kinds, nu: int) -> bytes:
    """Registry::digest of the registry of several built-in types at nu."""
    arr = (ctypes.c_uint32 * len(kinds))(*kinds)
    out = ctypes.create_string_buffer(32)
    _check(load().flock_circuit_union_digest(arr, len(kinds), nu, out))
    return out.raw


def registry_digest(types, nu: int) -> bytes:
    """Registry::digest (schedule.rs:542-580) of TypeStatement structs in any order."""
    arr = (TypeStatement * len(types))(*types)
    out = ctypes.create_string_buffer(32)
    _check(load().flock_registry_digest(arr, len(types), nu, out))
    return out.raw


def zerocheck_verify(fs: "FsChallenger", m: int, grinding: bool, round1, multilinear, finals,
                     nonces=None) -> np.ndarray:
    """zerocheck::verify_with_grinding (zerocheck.rs:1032-1225) on the host transcript `fs`;
    the arrays are the prover's outputs ((n, 2) uint64). Returns the claim, or raises
    FlockHipError with E_REJECT."""
    n_mlv = m - 6
    claim = np.zeros((1 + 2 * n_mlv, 2), dtype=np.uint64)
    r1 = np.ascontiguousarray(round1, dtype=np.uint64)
    ml = np.ascontiguousarray(multilinear, dtype=np.uint64)
    fin = np.ascontiguousarray(finals, dtype=np.uint64)
    nc = np.ascontiguousarray(nonces if nonces is not None else np.zeros(1), dtype=np.uint64)
    _check(load().flock_zerocheck_verify(fs.h, m, int(bool(grinding)), _np_ptr(r1), _np_ptr(ml),
                                         _np_ptr(fin), nc.ctypes.data, _np_ptr(claim)))
    return claim


def lincheck_verify(fs: "FsChallenger", k_log: int, const_pin: int, nu: int, n_rows: int,
                    z_skip: int, x_inner_rest, x_outer, v_a: int, v_b: int, grinding: bool,
                    proof: dict, circuit=None) -> np.ndarray:
    """verify_union_with_grinding (lincheck/union.rs:662-960) for one boolean type, on the host
    transcript `fs`. proof = the prover's output dict (rounds, z_partial, matrix_evals, nonces).
    circuit = (a_ptr, a_rows, b_ptr, b_rows) host CSC arrays for MatrixAssertion::check, or
    None for the deferred form checked against the reported matrix evals. Returns the claim,
    or raises FlockHipError with E_REJECT."""
    L = k_log - 6
    claim = np.zeros((2 + L, 2), dtype=np.uint64)
    keep = []
    cp = None
    if circuit is not None:
        keep = [np.ascontiguousarray(x, dtype=np.uint32) if len(x) else np.zeros(1, np.uint32)
                for x in circuit]
        c = Csc(k_log, 0, *[x.ctypes.data for x in keep], len(circuit[1]), len(circuit[3]),
                const_pin)
        cp = ctypes.byref(c)
    arr = {k: np.ascontiguousarray(proof[k], dtype=np.uint64) if len(proof[k]) else
           np.zeros((1, 2), dtype=np.uint64) for k in ("rounds", "z_partial", "matrix_evals")}
    nc = np.ascontiguousarray(proof.get("nonces", np.zeros(1)), dtype=np.uint64)
    xr, xo = _ints_arr(x_inner_rest), _ints_arr(x_outer)
    _check(load().flock_lincheck_verify(
        fs.h, cp, k_log, const_pin, nu, n_rows, f128(z_skip), _np_ptr(xr), _np_ptr(xo),
        f128(v_a), f128(v_b), int(bool(grinding)), _np_ptr(arr["rounds"]),
        _np_ptr(arr["z_partial"]), _np_ptr(arr["matrix_evals"]), nc.ctypes.data, _np_ptr(claim)))
    del keep
    return claim


def lincheck_verify_keccak(fs: "FsChallenger", nu: int, n_rows: int, z_skip: int, x_inner_rest,
                           x_outer, v_a: int, v_b: int, grinding: bool, proof: dict) -> np.ndarray:
    """lincheck_verify for a Keccak-f[1600] slot with the full matrix check (host walk of
    KeccakLincheckCircuit, keccak.rs:678-835). Returns the claim, or raises E_REJECT."""
    claim = np.zeros((2 + 10, 2), dtype=np.uint64)
    arr = {k: np.ascontiguousarray(proof[k], dtype=np.uint64) for k in ("rounds", "z_partial")}
    nc = np.ascontiguousarray(proof.get("nonces", np.zeros(1)), dtype=np.uint64)
    xr, xo = _ints_arr(x_inner_rest), _ints_arr(x_outer)
    _check(load().flock_lincheck_verify_keccak(
        fs.h, nu, n_rows, f128(z_skip), _np_ptr(xr), _np_ptr(xo), f128(v_a), f128(v_b),
        int(bool(grinding)), _np_ptr(arr["rounds"]), _np_ptr(arr["z_partial"]), nc.ctypes.data,
        _np_ptr(claim)))
    return claim


def registry_new(types, nu: int) -> Registry:
    """Registry::new (schedule.rs:337-495); types = [(k_log, useful_bits, is_element)]."""
    arr = (TableType * max(len(types), 1))(*[TableType(k, int(e), u) for k, u, e in types])
    reg = Registry()
    _check(load().flock_registry_new(arr, len(types), nu, ctypes.byref(reg)))
    return reg


def union_sizes(reg: Registry, counts, dense_floor_m: int = 0,
                log_batch_size: int = 0) -> dict:
    """UnionInstance sizes (union.rs:325-416); counts in slot order."""
    c = np.ascontiguousarray(counts, dtype=np.uint64)
    out = UnionSizes()
    _check(load().flock_union_sizes_compute(ctypes.byref(reg), c.ctypes.data, dense_floor_m,
                                            log_batch_size, ctypes.byref(out)))
    return dict(dense_words=out.dense_words, committed_words=out.committed_words,
                packed_len=out.packed_len, dense_m=out.dense_m,
                commit_lanes=out.commit_lanes or None,
                compaction_is_identity=bool(out.compaction_is_identity))


def x_ab_from_mlv(reg: Registry, mlv):
    """x_ab_from_mlv (union.rs:541-552) -> (x_inner_rest, x_outer)."""
    m = _ints_arr(mlv)
    xr = np.zeros((max(len(mlv) - reg.nu, 1), 2), dtype=np.uint64)
    xo = np.zeros((max(reg.nu, 1), 2), dtype=np.uint64)
    _check(load().flock_union_x_ab_from_mlv(ctypes.byref(reg), _np_ptr(m), len(mlv), _np_ptr(xr),
                                            _np_ptr(xo)))
    return _ints(xr[:len(mlv) - reg.nu]), _ints(xo[:reg.nu])


class FsChallenger:
    """FsChallenger with the SHA-256 discipline (challenger.rs:609-1044)."""

    def __init__(self, domain: bytes):
        self.h = load().flock_fs_new(domain, len(domain))
        if not self.h:
            raise MemoryError("flock_fs_new")

    def __del__(self):
        if getattr(self, "h", None):
            load().flock_fs_free(self.h)
            self.h = None

    def observe_label(self, b: bytes):
        load().flock_fs_observe_label(self.h, b, len(b))

    def observe_f128(self, v: int):
        load().flock_fs_observe_f128(self.h, f128(v))

    def observe_f128_slice(self, vs):
        arr = np.zeros((max(len(vs), 1), 2), dtype=np.uint64)
        for i, v in enumerate(vs):
            arr[i] = (v & ((1 << 64) - 1), v >> 64)
        load().flock_fs_observe_f128_slice(self.h, _np_ptr(arr), len(vs))

    def observe_bytes(self, b: bytes):
        load().flock_fs_observe_bytes(self.h, b, len(b))

    def sample_f128(self) -> int:
        out = F128()
        load().flock_fs_sample_f128(self.h, ctypes.byref(out))
        return to_int(out)

    def sample_f128_vec(self, n: int) -> list[int]:
        arr = np.zeros((max(n, 1), 2), dtype=np.uint64)
        load().flock_fs_sample_f128_vec(self.h, _np_ptr(arr), n)
        return _ints(arr[:n])

    def grind_pow(self, bits: int) -> int:
        return int(load().flock_fs_grind_pow(self.h, bits))

    def verify_pow(self, nonce: int, bits: int) -> bool:
        return bool(load().flock_fs_verify_pow(self.h, nonce, bits))

    def digest(self) -> bytes:
        """SHA-256 of everything the transcript has absorbed."""
        out = ctypes.create_string_buffer(32)
        load().flock_fs_digest(self.h, out)
        return out.raw

    def bind_statement(self, registry_digest: bytes, counts, cap: bytes):
        """UnionInstance::bind_statement (union.rs:619-628)."""
        c = np.ascontiguousarray(counts, dtype=np.uint64)
        _check(load().flock_union_bind_statement(self.h, registry_digest, c.ctypes.data, len(c),
                                                 cap, len(cap)))


class Prover:
    """flock_hip_prover: a context plus grow-only device buffers for prove-core requests."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(load().flock_hip_prover_create(device, ctypes.byref(h)))
        self.h = h

    def set_round_mode(self, mode: int):
        """flock_hip_ctx_set_round_mode on the prover's context."""
        ctx = load().flock_hip_prover_ctx(self.h)
        _check(load().flock_hip_ctx_set_round_mode(ctx, mode))

    def held_bytes(self) -> int:
        return int(load().flock_hip_prover_held_bytes(self.h))

    def release(self):
        load().flock_hip_prover_release(self.h)

    def open_stage_ms(self) -> list[float]:
        """Host wall times of the last prove_fast's opening stages (flock_open_ws_stage_ms)."""
        ws = load().flock_hip_prover_open_ws(self.h)
        out = (ctypes.c_double * 6)()
        if ws:
            _check(load().flock_open_ws_stage_ms(ws, out, 6))
        return list(out)

    def close(self):
        if self.h:
            load().flock_hip_prover_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


TUNABLES = ("ntt_cols_per_thread", "ntt_deep_cols", "zc_split_min_vars", "zc_la_max_log",
            "sc_la_max_log", "open_split_fold_log")


def set_tunable(name: str, value: int):
    """flock_hip_set_tunable (process-wide launch shape / threshold)."""
    _check(load().flock_hip_set_tunable(name.encode(), int(value)))


def get_tunable(name: str) -> int:
    v = ctypes.c_int64()
    _check(load().flock_hip_get_tunable(name.encode(), ctypes.byref(v)))
    return v.value


def apply_tuning(path: str | None = None) -> dict:
    """Apply the autotuner's choices (flock_amd/tuning.json by default) if the file exists;
    returns the values applied ({} when there is no file)."""
    import json
    path = path or os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        vals = json.load(fh).get("values", {})
    for k, v in vals.items():
        set_tunable(k, v)
    return vals


def make_requests(specs, digest: bytes | None = None, domain: bytes = b"flock-readme-bench-v0"):
    """specs: (kind, n_log, inputs ndarray, circuit Csc or None, arrival_s). circuit None: the
    built-in circuit; digest None: the built-in registry's digest. Returns the ctypes request
    array and the objects it points into (keep them alive while it is used)."""
    arr = (ProveRequest * len(specs))()
    keep = [digest, domain]
    for i, (kind, n_log, inputs, circuit, arrival) in enumerate(specs):
        inp = np.ascontiguousarray(inputs)
        keep.append(inp)
        arr[i].kind, arr[i].n_log, arr[i].n_decl = kind, n_log, inp.shape[0]
        arr[i].h_inputs = inp.ctypes.data
        arr[i].circuit = ctypes.cast(ctypes.pointer(circuit), ctypes.c_void_p) if circuit else None
        if circuit is not None:
            keep.append(circuit)
        arr[i].digest, arr[i].domain, arr[i].domain_len = digest, domain, len(domain)
        arr[i].arrival_s = arrival
    return arr, keep


def device_request(kind: int, n_log: int, n_decl: int, d_inputs: "DeviceBuffer",
                   domain: bytes = b"flock-readme-bench-v0", digest: bytes | None = None):
    """A request whose inputs are already in HBM (inputs_on_device), with the built-in circuit.
    Returns (request, keep)."""
    req = ProveRequest()
    req.kind, req.n_log, req.n_decl = kind, n_log, n_decl
    req.h_inputs = d_inputs.ptr
    req.inputs_on_device = 1
    req.digest, req.domain, req.domain_len = digest, domain, len(domain)
    return req, [domain, digest, d_inputs]


def prove_core(prover: Prover, req: ProveRequest) -> ProveResult:
    res = ProveResult()
    _check(load().flock_hip_prove_core(prover.h, ctypes.byref(req), ctypes.byref(res)))
    return res


def prove_fast(prover: Prover, req: ProveRequest):
    """flock_hip_prove_fast: returns (result, proof bytes, cap bytes, claim_ab, claim_c)."""
    res = ProveResult()
    out = ProofBytes()
    _check(load().flock_hip_prove_fast(prover.h, ctypes.byref(req), ctypes.byref(res),
                                       ctypes.byref(out)))
    proof = ctypes.string_at(out.proof, out.proof_len)
    cap = ctypes.string_at(out.cap, out.cap_len)
    load().flock_hip_free(ctypes.cast(out.proof, ctypes.c_void_p))
    load().flock_hip_free(ctypes.cast(out.cap, ctypes.c_void_p))
    return res, proof, cap, to_int(out.claim_ab), to_int(out.claim_c)


def union_request(parts, n_log: int, domain: bytes = b"flock-readme-bench-v0",
                  digest: bytes | None = None):
    """A mixed request from [(kind, inputs ndarray, n_decl or None)] on the host. Returns
    (request, keep)."""
    arr = (UnionPart * len(parts))()
    keep = [domain, digest, arr]
    for i, part in enumerate(parts):
        kind, x = part[0], part[1]
        n = part[2] if len(part) > 2 and part[2] is not None else len(x)
        x = np.ascontiguousarray(x)
        keep.append(x)
        arr[i].kind, arr[i].n_decl = kind, n
        arr[i].inputs = x.ctypes.data if n else None
    req = UnionRequest()
    req.n_log, req.n_parts, req.parts = n_log, len(parts), arr
    req.digest, req.domain, req.domain_len = digest, domain, len(domain)
    return req, keep


def union_info(req: UnionRequest) -> dict:
    """flock_union_request_info: slot order, counts (slot order), dense_m, log_batch_size,
    num_lanes, m_bool, has_opening, device_bytes."""
    info = UnionInfo()
    _check(load().flock_union_request_info(ctypes.byref(req), ctypes.byref(info)))
    n = info.n_slots
    return dict(slot_part=list(info.slot_part[:n]), counts=list(info.counts[:n]),
                dense_m=info.dense_m, log_batch_size=info.log_batch_size,
                lanes=int(info.num_lanes), m_bool=info.m_bool,
                has_opening=bool(info.has_opening), device_bytes=int(info.device_bytes))


def prove_fast_union(prover: Prover, req: UnionRequest):
    """flock_hip_prove_fast_union: returns (result, proof bytes, cap bytes, claim_ab,
    claim_c)."""
    res = ProveResult()
    out = ProofBytes()
    _check(load().flock_hip_prove_fast_union(prover.h, ctypes.byref(req), ctypes.byref(res),
                                             ctypes.byref(out)))
    proof = ctypes.string_at(out.proof, out.proof_len)
    cap = ctypes.string_at(out.cap, out.cap_len)
    load().flock_hip_free(ctypes.cast(out.proof, ctypes.c_void_p))
    load().flock_hip_free(ctypes.cast(out.cap, ctypes.c_void_p))
    return res, proof, cap, to_int(out.claim_ab), to_int(out.claim_c)


def verify_fast(kind: int, n_log: int, n_decl: int, cap: bytes, proof: bytes,
                domain: bytes = b"flock-readme-bench-v0", digest: bytes | None = None):
    """flock_verify_fast (host only) for a built-in-circuit request: returns (claim_ab,
    claim_c); a rejected proof raises FlockHipError with code E_REJECT."""
    req = ProveRequest()
    req.kind, req.n_log, req.n_decl = kind, n_log, n_decl
    req.digest, req.domain, req.domain_len = digest, domain, len(domain)
    ab, c = F128(), F128()
    _check(load().flock_verify_fast(ctypes.byref(req), cap, len(cap), proof, len(proof),
                                    ctypes.byref(ab), ctypes.byref(c)))
    return to_int(ab), to_int(c)


def verify_fast_union(parts, n_log: int, cap: bytes, proof: bytes,
                      domain: bytes = b"flock-readme-bench-v0", digest: bytes | None = None):
    """flock_verify_fast_union (host only): parts = [(kind, n_decl)] in any order. Returns
    (claim_ab, claim_c); a rejected proof raises FlockHipError with code E_REJECT."""
    arr = (UnionPart * len(parts))()
    for i, (kind, n) in enumerate(parts):
        arr[i].kind, arr[i].n_decl = kind, n
    req = UnionRequest()
    req.n_log, req.n_parts, req.parts = n_log, len(parts), arr
    req.digest, req.domain, req.domain_len = digest, domain, len(domain)
    ab, c = F128(), F128()
    _check(load().flock_verify_fast_union(ctypes.byref(req), cap, len(cap), proof, len(proof),
                                          ctypes.byref(ab), ctypes.byref(c)))
    return to_int(ab), to_int(c)


def lig_initial_k(m: int) -> int:
    return int(load().flock_lig_initial_k(m))


def l0_cap_depth(m: int) -> int:
    return int(load().flock_l0_cap_depth(m))


def prove_footprint(req: ProveRequest) -> int:
    b = ctypes.c_uint64()
    _check(load().flock_hip_prove_footprint(ctypes.byref(req), ctypes.byref(b)))
    return int(b.value)


def serve(provers, reqs, hbm_budget: int, proofs: bool = False):
    """flock_hip_serve over the given provers (one worker each). Returns (results, wall_ms), or
    with proofs=True (whole prove_fast per request) (results, wall_ms, [(proof, cap, claim_ab,
    claim_c)])."""
    hs = (ctypes.c_void_p * len(provers))(*[p.h.value for p in provers])
    res = (ProveResult * len(reqs))()
    wall = ctypes.c_double()
    pb = (ProofBytes * len(reqs))() if proofs else None
    try:
        _check(load().flock_hip_serve(hs, len(provers), reqs, len(reqs), hbm_budget, res, pb,
                                      ctypes.byref(wall)))
        if not proofs:
            return res, wall.value
        out = []
        for b in pb:
            out.append((ctypes.string_at(b.proof, b.proof_len), ctypes.string_at(b.cap, b.cap_len),
                        to_int(b.claim_ab), to_int(b.claim_c)))
        return res, wall.value, out
    finally:
        if pb is not None:
            for b in pb:
                load().flock_hip_free(ctypes.cast(b.proof, ctypes.c_void_p))
                load().flock_hip_free(ctypes.cast(b.cap, ctypes.c_void_p))
