"""The oracle's whole `prove_fast` for one boolean hash slot, and its wire bytes.

Test infrastructure: the CPU checker for the GPU prover. The order follows
`prove_union_with_binding_zc` (crates/flock-prover/src/prover.rs:831-1468) for a single-slot
union (the `*Setup::prove_fast` entries, r1cs_hashes/sha2.rs:1289-1306, blake3.rs:1669-1685):

  witness (batch-major, partial driver) -> dense stack (compact_witness) -> PCS commit at the
  L0 cap depth (pcs/commit.rs:474-510 / 345-397, cap_layer merkle.rs:649-656) ->
  bind_statement (union.rs:619-628) -> zerocheck with s_hat_v_c (zerocheck.rs:598-1011) ->
  x_ab (union.rs:541-556) -> union lincheck capturing z_vec (lincheck/union.rs:203-430) ->
  s_hat_v_ab = s_hat_v_from_z_vec (pcs/ring_switch.rs:1278-1330) -> open_batch_merged
  (pcs.rs:1433-1822, oracle/opening.c).

The proof is `R1csProofMergedLigerito {zerocheck, lincheck, pcs_open}` (proof.rs:42-47) in
bincode 1.3 (fixint, little-endian, u64 lengths); `anchor_digest` is the digest
`merged_bundle_digest` of crates/flock-prover/tests/union_m6_fixtures.rs:156-172.
"""
from __future__ import annotations

import ctypes
import hashlib
import struct

import numpy as np

import oracle_lib as O
from test_union_glue import Union

L = O.lib()
P = ctypes.c_void_p
S = ctypes.c_size_t
L.or_fs_digest.argtypes = [P, P]
L.or_bind_statement.argtypes = [P, P, P, S, P, S]
L.or_x_ab_from_mlv.argtypes = [P, S, ctypes.c_uint, P, P]
L.or_zerocheck_prove.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_int, P, P, P, P, P]
L.or_zc_s_hat_v_c.argtypes = [P, ctypes.c_int, P, P]
L.or_lincheck_prove.argtypes = [P, ctypes.c_int, P, P, P, P, ctypes.c_long, P, ctypes.c_int, S, S,
                                O.F128, P, P, ctypes.c_int, P, P, P, P, P, P]
L.or_lincheck_prove_keccak.argtypes = [P, P, ctypes.c_int, S, O.F128, P, P, ctypes.c_int, P, P, P,
                                       P, P, P]
L.or_lig_initial_k.argtypes = [ctypes.c_uint]
L.or_cap_depth_for.argtypes = [ctypes.c_uint]
L.or_open_merged.argtypes = [P, P, ctypes.POINTER(O.PcsParams), P, P, S, P, S, P, P, S,
                             ctypes.c_uint, ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8)),
                             ctypes.POINTER(S), P]
L.or_free_bytes.argtypes = [P]
L.or_jag_fhat.argtypes = [P, S, ctypes.c_int, ctypes.c_int, P, P, P, P]

HASH_SHA256, HASH_BLAKE3, HASH_KECCAK = 0, 1, 2
# kind: (k_log, useful_bits, const pin): sha2.rs:189-194, blake3.rs:233-247, keccak.rs:678-700
SHAPES = {HASH_SHA256: (15, 25_470, 25_469), HASH_BLAKE3: (14, 11_707, 11_706),
          HASH_KECCAK: (16, 42_560, 4096)}


class _Bin:
    """bincode 1.3 default options: little-endian fixed-width ints, u64 lengths."""

    def __init__(self):
        self.parts = []

    def u64(self, v: int):
        self.parts.append(struct.pack("<Q", v))

    def f128(self, v: int):
        self.parts.append(struct.pack("<QQ", v & (2**64 - 1), v >> 64))

    def vec_f128(self, vs):
        self.u64(len(vs))
        for v in vs:
            self.f128(v)

    def vec_pairs(self, vs):
        self.u64(len(vs) // 2)
        for v in vs:
            self.f128(v)

    def vec_u64(self, vs):
        self.u64(len(vs))
        for v in vs:
            self.u64(int(v))

    def raw(self, b: bytes):
        self.parts.append(b)

    def bytes(self) -> bytes:
        return b"".join(self.parts)


def min_n_log(n: int) -> int:
    """min_n_blocks_log (blake3.rs:1225-1229, keccak3.rs:144-149): the smallest nu >= 3 with
    2^nu >= n."""
    nu = 3
    while (1 << nu) < n:
        nu += 1
    return nu


def hash_types(kind):
    """The registry's TableType for the hash (schedule.rs:150-186) as an oracle TypeStmt,
    plus the CSC lincheck circuit."""
    k_log, ub, pin = SHAPES[kind]
    k = 1 << k_log
    keep = []
    if kind == HASH_KECCAK:
        # KeccakSetup's registry type (keccak.rs:880-888): the empty-stub R1CS
        # (common.rs:222-260: k x k matrices without entries, C_0 = I) with the walker's
        # constant pin copied onto the type; the lincheck takes the walker circuit.
        zeros = np.zeros(k + 1, dtype=np.uint64)
        none = np.zeros(1, dtype=np.uint32)
        ty = O.TypeStmt(k_log, 1, ub, pin, O.sparse_csr(zeros, none, k, keep),
                        O.sparse_csr(zeros, none, k, keep), O.sparse(num_rows=k, identity=True))
        return ty, None, keep
    if kind == HASH_SHA256:
        import test_oracle_sha2_r1cs as M
    else:
        import test_oracle_blake3_r1cs as M
    (ap, ac), (bp, bc), csc_a, csc_b = M.matrices()
    ty = O.TypeStmt(k_log, 1, ub, pin, O.sparse_csr(ap, ac, k, keep), O.sparse_csr(bp, bc, k, keep),
                    O.sparse(num_rows=k, identity=True))
    return ty, (csc_a, csc_b), keep


_CIRCUITS = {}


def circuit(kind):
    if kind not in _CIRCUITS:
        _CIRCUITS[kind] = hash_types(kind)
    return _CIRCUITS[kind]


_DIGESTS = {}


def registry_digest(kind, nu):
    """Registry::digest of the one-type registry (setup work, cached per (kind, nu))."""
    if (kind, nu) not in _DIGESTS:
        ty, _, _ = circuit(kind)
        _DIGESTS[(kind, nu)] = O.registry_digest([ty], nu)
    return _DIGESTS[(kind, nu)]


def witness(kind, inputs, n_log):
    if kind == HASH_SHA256:
        from test_oracle_sha2_r1cs import batch_major
    elif kind == HASH_BLAKE3:
        from test_oracle_blake3_r1cs import batch_major
    else:
        from test_oracle_keccak_r1cs import batch_major
    return batch_major(inputs, n_log)


def s_hat_v_from_z_vec(z_vec: np.ndarray, tail: list[int]) -> list[int]:
    """s_hat_v[b] = sum_k eq(tail, k) z_vec[128 k + b] (pcs/ring_switch.rs:1278-1330)."""
    eq = O.build_eq(tail)
    n_tail = 1 << len(tail)
    assert z_vec.shape[0] == 128 * n_tail
    w = np.repeat(eq, 128, axis=0)
    prod = O.mul_batch(z_vec, w).reshape(n_tail, 128, 2)
    acc = np.bitwise_xor.reduce(prod, axis=0)
    return O.arr_to_ints(acc)


def prove_fast(kind, inputs, domain: bytes, n_log: int | None = None, parts: dict | None = None):
    """Returns (proof_bytes, cap_bytes, claim_ab_value, claim_c_value)."""
    inputs = np.ascontiguousarray(inputs)
    n_decl = len(inputs)
    if n_log is None:
        n_log = min_n_log(n_decl)
    k_log, ub, pin = SHAPES[kind]
    z, a, b, stripe = witness(kind, inputs, n_log)
    # Keccak batches count the full capacity: padding rows run keccak_f(0) (keccak.rs:922-939)
    count = (1 << n_log) if kind == HASH_KECCAK else n_decl
    U = Union([(k_log, ub, False)], n_log)
    sz0 = U.sizes([count], 0, 0)
    dense_m = sz0["dense_m"]
    lbs = L.or_lig_initial_k(dense_m)
    assert lbs > 0, f"no Ligerito config for m={dense_m}"
    sz = U.sizes([count], 0, lbs)
    q = U.compact([count], z)
    t = sz["commit_lanes"] or 0
    if t:
        cw, tree = O.commit_lane_major(q, dense_m, 1, lbs, t)
    else:
        cw, tree = O.commit(q, dense_m, 1, lbs, 0)
    n_leaves = (len(tree) // 32 + 1) // 2
    c0 = L.or_cap_depth_for(dense_m)
    cap = tree[32 * (2 * n_leaves - 2 * (1 << c0)):][:32 << c0]
    fs = O.Fs(domain)
    digest = registry_digest(kind, n_log)
    counts = np.array([count], dtype=np.uint64)
    L.or_bind_statement(fs.h, digest, counts.ctypes.data, 1, cap, len(cap))
    m = n_log + k_log
    n_mlv = m - 6
    r1 = np.zeros((128, 2), dtype=np.uint64)
    ml = np.zeros((2 * n_mlv, 2), dtype=np.uint64)
    fin = np.zeros((3, 2), dtype=np.uint64)
    zc_non = np.zeros(2 + n_mlv, dtype=np.uint64)
    claim = np.zeros((1 + 2 * n_mlv, 2), dtype=np.uint64)
    zb, ab, bb = (x.view(np.uint8) for x in (z, a, b))
    assert L.or_zerocheck_prove(fs.h, ab.ctypes.data, bb.ctypes.data, zb.ctypes.data, m, 1,
                                O._p(r1), O._p(ml), O._p(fin), zc_non.ctypes.data,
                                O._p(claim)) == 0
    r_rest = np.ascontiguousarray(claim[1 + n_mlv:1 + 2 * n_mlv])
    # s_hat_v_c (univariate_skip.rs:384-489) over eq(r_rest[1..]); r[0..6] unused
    rr = np.zeros((m, 2), dtype=np.uint64)
    rr[6:] = r_rest
    shv_c = np.zeros((128, 2), dtype=np.uint64)
    L.or_zc_s_hat_v_c(zb.ctypes.data, m, O._p(rr), O._p(shv_c))
    Lr = k_log - 6
    xr = np.zeros((max(Lr, 1), 2), dtype=np.uint64)
    xo = np.zeros((n_log, 2), dtype=np.uint64)
    mlv = np.ascontiguousarray(claim[1:1 + n_mlv])
    L.or_x_ab_from_mlv(O._p(mlv), n_mlv, n_log, O._p(xr), O._p(xo))
    rounds = np.zeros((2 * Lr, 2), dtype=np.uint64)
    zp = np.zeros((64, 2), dtype=np.uint64)
    mev = np.zeros((2, 2), dtype=np.uint64)
    lnon = np.zeros(Lr + 3, dtype=np.uint64)
    lcl = np.zeros((2 + Lr, 2), dtype=np.uint64)
    z_vec = np.zeros((1 << k_log, 2), dtype=np.uint64)
    z_skip = O.f(O.arr_to_ints(claim[:1])[0])
    if kind == HASH_KECCAK:
        rc = L.or_lincheck_prove_keccak(fs.h, stripe.ctypes.data, n_log, count, z_skip, O._p(xr),
                                        O._p(xo), 1, O._p(rounds), O._p(zp), O._p(mev),
                                        lnon.ctypes.data, O._p(lcl), O._p(z_vec))
    else:
        _, (csc_a, csc_b), _ = circuit(kind)
        rc = L.or_lincheck_prove(fs.h, k_log, csc_a[0].ctypes.data, csc_a[1].ctypes.data,
                                 csc_b[0].ctypes.data, csc_b[1].ctypes.data, pin,
                                 stripe.ctypes.data, n_log, ub, count, z_skip, O._p(xr), O._p(xo),
                                 1, O._p(rounds), O._p(zp), O._p(mev), lnon.ctypes.data,
                                 O._p(lcl), O._p(z_vec))
    assert rc == 0
    r_inner_rest = O.arr_to_ints(lcl[2:2 + Lr])
    shv_ab = s_hat_v_from_z_vec(z_vec, r_inner_rest[1:])
    # claim points as quirky x_outer_full (prover.rs:43-48; union.rs:564-600)
    x_ab = [r_inner_rest[0]] + O.arr_to_ints(xo) + r_inner_rest[1:]
    x_c = O.arr_to_ints(r_rest)
    assert len(x_ab) == len(x_c) == m - 6
    # jagged heights (union.rs:293-310): used columns at the declared count
    col_log = k_log - 7
    used = min((ub + 127) // 128, 1 << col_log)
    heights = np.zeros(1 << col_log, dtype=np.uint64)
    heights[:used] = count
    xs = O.ints_to_arr(x_ab + x_c)
    sv = np.ascontiguousarray(np.concatenate([O.ints_to_arr(shv_ab), shv_c]))
    pp = O.PcsParams(dense_m, 1, lbs, t, 0)
    out_p = ctypes.POINTER(ctypes.c_uint8)()
    out_n = ctypes.c_size_t()
    qev = np.zeros((1, 2), dtype=np.uint64)
    rc = L.or_open_merged(fs.h, O._p(q), ctypes.byref(pp), O._p(cw), tree, 2, O._p(xs), m - 6,
                          O._p(sv), heights.ctypes.data, len(heights), n_log, ctypes.byref(out_p),
                          ctypes.byref(out_n), O._p(qev))
    assert rc == 0, rc
    open_bytes = ctypes.string_at(out_p, out_n.value)
    L.or_free_bytes(out_p)
    w = _Bin()
    # ZerocheckProof (zerocheck.rs:427-447)
    r1i = O.arr_to_ints(r1)
    w.vec_f128(r1i[:64])
    w.vec_f128(r1i[64:])
    w.vec_pairs(O.arr_to_ints(ml))
    for v in O.arr_to_ints(fin):
        w.f128(v)
    w.vec_u64(zc_non)
    # LincheckProof (lincheck.rs:520-549): one matrix-eval pair for the one boolean type
    w.vec_pairs(O.arr_to_ints(rounds))
    w.vec_f128(O.arr_to_ints(zp))
    w.vec_pairs(O.arr_to_ints(mev))
    w.vec_u64(lnon)
    w.raw(open_bytes)
    proof = w.bytes()
    claim_ab = O.arr_to_ints(lcl[1:2])[0]
    claim_c = O.arr_to_ints(fin[2:3])[0]
    if parts is not None:
        dig = ctypes.create_string_buffer(32)
        L.or_fs_digest(fs.h, dig)
        parts.update(cap=cap, dense_m=dense_m, lanes=t, log_batch_size=lbs, q_eval=O.arr_to_ints(qev)[0],
                     transcript=dig.raw, open_len=len(open_bytes))
    return proof, cap, claim_ab, claim_c


class OrLcSlot(ctypes.Structure):
    _fields_ = [("k_log", ctypes.c_int), ("prefix", ctypes.c_uint64),
                ("useful_bits", ctypes.c_size_t), ("n_rows", ctypes.c_size_t),
                ("const_pin", ctypes.c_long), ("z", P), ("a_ptr", P), ("a_rows", P),
                ("b_ptr", P), ("b_rows", P)]


L.or_lincheck_prove_union.argtypes = [P, ctypes.POINTER(OrLcSlot), S, ctypes.c_int, ctypes.c_int,
                                      O.F128, P, P, ctypes.c_int, P, P, P, P, P, P]


def prove_fast_union(slots, nu: int, domain: bytes, parts: dict | None = None):
    """prove_union_with_binding_zc (prover.rs:831-1468) for a registry of several boolean hash
    types. slots: [(kind, inputs)] in any order; the registry sorts them (Registry::new, k_log
    descending, stable). Returns (proof_bytes, cap_bytes, claim_ab, claim_c, info) with info =
    dict(counts in slot order, dense_m, log_batch_size, lanes)."""
    kinds = [k for k, _ in slots]
    U = Union([(SHAPES[k][0], SHAPES[k][1], False) for k in kinds], nu)
    n_slots = len(slots)
    order = [U.reg.slots[s].type_index for s in range(n_slots)]  # caller index of slot s
    m_bool = U.reg.m_bool
    packed_words = 1 << (m_bool - 7)
    Z, A, B = (np.zeros((packed_words, 2), dtype=np.uint64) for _ in range(3))
    counts, stripes = [], []
    for s, ci in enumerate(order):
        kind, inputs = slots[ci]
        inputs = np.ascontiguousarray(inputs)
        z, a, b, stripe = witness(kind, inputs, nu)
        off = U.reg.slots[s].offset >> 7
        Z[off:off + len(z)] = z
        A[off:off + len(a)] = a
        B[off:off + len(b)] = b
        stripes.append(stripe)
        counts.append((1 << nu) if kind == HASH_KECCAK else len(inputs))
    sz0 = U.sizes(counts, 0, 0)
    dense_m = sz0["dense_m"]
    lbs = L.or_lig_initial_k(dense_m)
    assert lbs > 0, f"no Ligerito config for m={dense_m}"
    sz = U.sizes(counts, 0, lbs)
    q = U.compact(counts, Z)
    t = sz["commit_lanes"] or 0
    if t:
        cw, tree = O.commit_lane_major(q, dense_m, 1, lbs, t)
    else:
        cw, tree = O.commit(q, dense_m, 1, lbs, 0)
    n_leaves = (len(tree) // 32 + 1) // 2
    c0 = L.or_cap_depth_for(dense_m)
    cap = tree[32 * (2 * n_leaves - 2 * (1 << c0)):][:32 << c0]
    fs = O.Fs(domain)
    digest = O.registry_digest([circuit(k)[0] for k in kinds], nu)
    cnt = np.array(counts, dtype=np.uint64)
    L.or_bind_statement(fs.h, digest, cnt.ctypes.data, n_slots, cap, len(cap))
    m = m_bool
    n_mlv = m - 6
    r1 = np.zeros((128, 2), dtype=np.uint64)
    ml = np.zeros((2 * n_mlv, 2), dtype=np.uint64)
    fin = np.zeros((3, 2), dtype=np.uint64)
    zc_non = np.zeros(2 + n_mlv, dtype=np.uint64)
    claim = np.zeros((1 + 2 * n_mlv, 2), dtype=np.uint64)
    zb, ab, bb = (x.view(np.uint8) for x in (Z, A, B))
    assert L.or_zerocheck_prove(fs.h, ab.ctypes.data, bb.ctypes.data, zb.ctypes.data, m, 1,
                                O._p(r1), O._p(ml), O._p(fin), zc_non.ctypes.data,
                                O._p(claim)) == 0
    r_rest = np.ascontiguousarray(claim[1 + n_mlv:1 + 2 * n_mlv])
    rr = np.zeros((m, 2), dtype=np.uint64)
    rr[6:] = r_rest
    shv_c = np.zeros((128, 2), dtype=np.uint64)
    L.or_zc_s_hat_v_c(zb.ctypes.data, m, O._p(rr), O._p(shv_c))
    col_vars = m - nu
    Lr = col_vars - 6
    xr = np.zeros((max(Lr, 1), 2), dtype=np.uint64)
    xo = np.zeros((nu, 2), dtype=np.uint64)
    mlv = np.ascontiguousarray(claim[1:1 + n_mlv])
    L.or_x_ab_from_mlv(O._p(mlv), n_mlv, nu, O._p(xr), O._p(xo))
    lc_slots = (OrLcSlot * n_slots)()
    keep = []
    for s, ci in enumerate(order):
        kind = kinds[ci]
        k_log, ub, pin = SHAPES[kind]
        sl = lc_slots[s]
        sl.k_log, sl.prefix, sl.useful_bits, sl.n_rows, sl.const_pin = \
            k_log, U.reg.slots[s].prefix, ub, counts[s], pin
        sl.z = stripes[s].ctypes.data
        if kind != HASH_KECCAK:
            _, (csc_a, csc_b), _ = circuit(kind)
            sl.a_ptr, sl.a_rows = csc_a[0].ctypes.data, csc_a[1].ctypes.data
            sl.b_ptr, sl.b_rows = csc_b[0].ctypes.data, csc_b[1].ctypes.data
            keep.append((csc_a, csc_b))
    n_pins = n_slots
    rounds = np.zeros((max(2 * Lr, 1), 2), dtype=np.uint64)
    zp = np.zeros((64, 2), dtype=np.uint64)
    mev = np.zeros((2 * n_slots, 2), dtype=np.uint64)
    lnon = np.zeros(Lr + 2 + n_pins, dtype=np.uint64)
    lcl = np.zeros((2 + Lr, 2), dtype=np.uint64)
    z_vec = np.zeros((1 << col_vars, 2), dtype=np.uint64)
    z_skip = O.f(O.arr_to_ints(claim[:1])[0])
    assert L.or_lincheck_prove_union(fs.h, lc_slots, n_slots, col_vars, nu, z_skip, O._p(xr),
                                     O._p(xo), 1, O._p(rounds), O._p(zp), O._p(mev),
                                     lnon.ctypes.data, O._p(lcl), O._p(z_vec)) == 0
    r_inner_rest = O.arr_to_ints(lcl[2:2 + Lr])
    shv_ab = s_hat_v_from_z_vec(z_vec, r_inner_rest[1:])
    x_ab = [r_inner_rest[0]] + O.arr_to_ints(xo) + r_inner_rest[1:]
    x_c = O.arr_to_ints(r_rest)
    assert len(x_ab) == len(x_c) == m - 6
    # jagged heights (union.rs:293-310): each slot's used columns at its count
    col_log = m - nu - 7
    heights = np.zeros(1 << col_log, dtype=np.uint64)
    for s, ci in enumerate(order):
        k_log, ub, _ = SHAPES[kinds[ci]]
        used = min((ub + 127) // 128, 1 << (k_log - 7))
        c_off = U.reg.slots[s].offset >> (7 + nu)
        heights[c_off:c_off + used] = counts[s]
    xs = O.ints_to_arr(x_ab + x_c)
    sv = np.ascontiguousarray(np.concatenate([O.ints_to_arr(shv_ab), shv_c]))
    pp = O.PcsParams(dense_m, 1, lbs, t, 0)
    out_p = ctypes.POINTER(ctypes.c_uint8)()
    out_n = ctypes.c_size_t()
    qev = np.zeros((1, 2), dtype=np.uint64)
    assert L.or_open_merged(fs.h, O._p(q), ctypes.byref(pp), O._p(cw), tree, 2, O._p(xs), m - 6,
                            O._p(sv), heights.ctypes.data, len(heights), nu, ctypes.byref(out_p),
                            ctypes.byref(out_n), O._p(qev)) == 0
    open_bytes = ctypes.string_at(out_p, out_n.value)
    L.or_free_bytes(out_p)
    w = _Bin()
    r1i = O.arr_to_ints(r1)
    w.vec_f128(r1i[:64])
    w.vec_f128(r1i[64:])
    w.vec_pairs(O.arr_to_ints(ml))
    for v in O.arr_to_ints(fin):
        w.f128(v)
    w.vec_u64(zc_non)
    # LincheckProof (lincheck.rs:520-549): one matrix-eval pair per boolean type
    w.vec_pairs(O.arr_to_ints(rounds[:2 * Lr]))
    w.vec_f128(O.arr_to_ints(zp))
    w.vec_pairs(O.arr_to_ints(mev))
    w.vec_u64(lnon)
    w.raw(open_bytes)
    proof = w.bytes()
    claim_ab = O.arr_to_ints(lcl[1:2])[0]
    claim_c = O.arr_to_ints(fin[2:3])[0]
    info = dict(counts=counts, dense_m=dense_m, log_batch_size=lbs, lanes=t)
    if parts is not None:
        parts.update(info)
    return proof, cap, claim_ab, claim_c, info


def mixed_bundle_bytes(registry_code: int, counts, cap: bytes, dense_m: int, lbs: int,
                       lanes: int, proof: bytes) -> bytes:
    """MixedProofBundleLigerito::to_bytes (proof_io.rs:267-282): "FLOCK", version 21, flavor 4,
    then bincode of {registry_id (u8 code, mixed.rs:91-98), counts, Commitment {cap, PcsParams
    {m, log_inv_rate, log_batch_size, profile Fast (variant 0), num_lanes, merkle_hash SHA-256
    (variant 0)}}, proof}."""
    w = _Bin()
    w.raw(b"FLOCK" + bytes([21, 4, registry_code]))
    w.vec_u64(counts)
    w.u64(len(cap) // 32)
    w.raw(cap)
    w.u64(dense_m)
    w.u64(1)
    w.u64(lbs)
    w.raw(struct.pack("<I", 0))
    w.raw(b"\x01" + struct.pack("<Q", lanes) if lanes else b"\x00")
    w.raw(struct.pack("<I", 0))
    w.raw(proof)
    return w.bytes()


def mixed_anchor_digest(bundle: bytes, claim_ab: int, claim_c: int) -> str:
    """m6_merged_union_proof_bytes_pinned's digest (union_m6_fixtures.rs:265-272): SHA-256 of
    the bundle bytes and the two claim values."""
    h = hashlib.sha256()
    h.update(bundle)
    for v in (claim_ab, claim_c):
        h.update(struct.pack("<QQ", v & (2**64 - 1), v >> 64))
    return h.hexdigest()


def anchor_digest(proof: bytes, cap: bytes, claim_ab: int, claim_c: int) -> str:
    """merged_bundle_digest (union_m6_fixtures.rs:156-172): SHA-256 over bincode(proof),
    the commitment cap, and the two claim values (lo, hi little-endian)."""
    h = hashlib.sha256()
    h.update(proof)
    h.update(cap)
    for v in (claim_ab, claim_c):
        h.update(struct.pack("<QQ", v & (2**64 - 1), v >> 64))
    return h.hexdigest()


class SplitMix32:
    """The fixtures' SplitMix64 generator, truncated to u32 (union_m6_fixtures.rs:116-128)."""

    def __init__(self, seed: int):
        self.s = seed & (2**64 - 1)

    def next_u32(self) -> int:
        M = 2**64 - 1
        self.s = (self.s + 0x9E3779B97F4A7C15) & M
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return (z ^ (z >> 31)) & 0xFFFFFFFF


def fixture_sha2_inputs(rng: SplitMix32, n: int) -> np.ndarray:
    """random_sha2_inputs (union_m6_fixtures.rs:142-152): (H_in[8], M[16]) per row."""
    out = np.zeros((n, 24), dtype=np.uint32)
    for i in range(n):
        for j in range(24):
            out[i, j] = rng.next_u32()
    return out


def fixture_blake3_inputs(rng: SplitMix32, n: int) -> np.ndarray:
    """random_blake3_inputs (union_m6_fixtures.rs:130-140) as oracle rows
    (cv[8], m[16], counter_lo, counter_hi, block_len 64, flags 11); the counter's high word
    is drawn first."""
    out = np.zeros((n, 28), dtype=np.uint32)
    for i in range(n):
        for j in range(24):
            out[i, j] = rng.next_u32()
        hi = rng.next_u32()
        lo = rng.next_u32()
        out[i, 24] = lo
        out[i, 25] = hi
        out[i, 26] = 64
        out[i, 27] = 11
    return out
