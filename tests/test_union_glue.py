"""Union instance host glue (crates/flock-core/src/{schedule,union}.rs): the oracle pinned by
the reference's own size tests (union.rs:1330-1480), then the C ABI checked against it.
Host-only entry points; no GPU is needed."""
from __future__ import annotations

import ctypes
import random

import numpy as np
import pytest

import oracle_lib as O
import puremath as M
from flock_amd import lib as F

L = O.lib()


class OrTableType(ctypes.Structure):
    _fields_ = [("k_log", ctypes.c_uint32), ("is_element", ctypes.c_uint32),
                ("useful_bits", ctypes.c_uint64)]


class OrSlot(ctypes.Structure):
    _fields_ = [("type_index", ctypes.c_size_t), ("m_slot", ctypes.c_uint32),
                ("prefix_bits", ctypes.c_uint32), ("offset", ctypes.c_uint64),
                ("prefix", ctypes.c_uint64)]


class OrRegistry(ctypes.Structure):
    _fields_ = [("n", ctypes.c_size_t), ("num_boolean", ctypes.c_size_t),
                ("nu", ctypes.c_uint32), ("m_total", ctypes.c_uint32),
                ("m_bool", ctypes.c_uint32), ("m_elem", ctypes.c_uint32),
                ("element_base", ctypes.c_uint64), ("slots", OrSlot * 16)]


class OrSizes(ctypes.Structure):
    _fields_ = [("dense_words", ctypes.c_uint64), ("committed_words", ctypes.c_uint64),
                ("packed_len", ctypes.c_uint64), ("commit_lanes", ctypes.c_uint64),
                ("dense_m", ctypes.c_uint32), ("compaction_is_identity", ctypes.c_uint32)]


P = ctypes.c_void_p
L.or_registry_layout.argtypes = [P, ctypes.c_size_t, ctypes.c_uint, ctypes.POINTER(OrRegistry)]
L.or_union_sizes.argtypes = [P, ctypes.POINTER(OrRegistry), P, ctypes.c_uint, ctypes.c_uint,
                             ctypes.POINTER(OrSizes)]
L.or_compact_witness.argtypes = [P, ctypes.POINTER(OrRegistry), P, P, ctypes.c_size_t, P]
L.or_x_ab_from_mlv.argtypes = [P, ctypes.c_size_t, ctypes.c_uint, P, P]
L.or_bind_statement.argtypes = [P, ctypes.c_char_p, P, ctypes.c_size_t, ctypes.c_char_p,
                                ctypes.c_size_t]


class Union:
    """The oracle's registry + sizes for types [(k_log, useful_bits, is_element)]."""

    def __init__(self, types, nu):
        self.types = (OrTableType * len(types))(*[OrTableType(k, int(e), u) for k, u, e in types])
        self.reg = OrRegistry()
        assert L.or_registry_layout(self.types, len(types), nu, ctypes.byref(self.reg)) == 0

    def sizes(self, counts, floor=0, lbs=0):
        c = np.ascontiguousarray(counts, dtype=np.uint64)
        out = OrSizes()
        assert L.or_union_sizes(self.types, ctypes.byref(self.reg), c.ctypes.data, floor, lbs,
                                ctypes.byref(out)) == 0
        return dict(dense_words=out.dense_words, committed_words=out.committed_words,
                    packed_len=out.packed_len, dense_m=out.dense_m,
                    commit_lanes=out.commit_lanes or None,
                    compaction_is_identity=bool(out.compaction_is_identity))

    def compact(self, counts, z):
        c = np.ascontiguousarray(counts, dtype=np.uint64)
        committed = self.sizes(counts)["committed_words"]
        q = np.zeros((committed, 2), dtype=np.uint64)
        L.or_compact_witness(self.types, ctypes.byref(self.reg), c.ctypes.data, O._p(z),
                             committed, O._p(q))
        return q


def bty(k_log, useful_bits):
    return (k_log, useful_bits, False)


def test_oracle_sizes_match_reference_tests():
    # committed_words_scale_with_counts_and_floor (union.rs:1362-1408)
    u = Union([bty(14, 15_409), bty(15, 31_401)], 7)
    assert u.reg.m_total == 23
    s = u.sizes([32, 32])
    assert s["packed_len"] == 1 << 16
    assert s["dense_words"] == 11_744
    assert s["committed_words"] == 1 << 15 and s["dense_m"] == 22
    assert s["committed_words"] * 2 == s["packed_len"]
    prev = (0, 0)
    for counts in ([8, 8], [32, 32], [50, 37], [128, 128]):
        cur = u.sizes(counts)
        assert cur["dense_words"] >= prev[0] and cur["committed_words"] >= prev[1]
        prev = (cur["dense_words"], cur["committed_words"])
    assert u.sizes([8, 8])["committed_words"] == 1 << 15
    assert u.sizes([50, 37])["dense_words"] == 50 * 246 + 37 * 121
    assert u.sizes([128, 128])["committed_words"] == 1 << 16
    # compaction identity (union.rs:1330-1360)
    for k_log, ub, nu in ((14, 15_409, 3), (15, 31_401, 2)):
        u = Union([bty(k_log, ub)], nu)
        s = u.sizes([1 << nu])
        assert s["committed_words"] == s["packed_len"] and s["dense_m"] == u.reg.m_total
        assert s["compaction_is_identity"]
        assert not u.sizes([(1 << nu) - 1])["compaction_is_identity"]
    u = Union([bty(14, 15_409), bty(15, 31_401)], 3)
    s = u.sizes([8, 8])
    assert s["dense_words"] == (246 + 121) << 3
    assert s["committed_words"] == s["packed_len"] and not s["compaction_is_identity"]


def test_oracle_compact_matches_reference_map():
    # compact_witness_matches_map (union.rs:1410-1470)
    u = Union([bty(10, 512), bty(9, 300)], 3)
    z = np.zeros((128, 2), dtype=np.uint64)
    for cols, n_t in ((range(0, 4), 5), (range(8, 11), 3)):
        for c in cols:
            for i in range(n_t):
                z[(c << 3) + i] = (i, c)
    q = u.compact([5, 3], z)
    assert len(q) == 128 and u.sizes([5, 3])["dense_words"] == 4 * 5 + 3 * 3
    cur = 0
    for col, n_t in [(c, 5) for c in range(4)] + [(c, 3) for c in range(8, 11)]:
        for i in range(n_t):
            assert tuple(q[cur]) == (i, col)
            cur += 1
    assert not q[cur:].any()
    # identity registry: q is the padded buffer
    u1 = Union([bty(10, 700)], 3)
    assert u1.sizes([8])["compaction_is_identity"]
    z1 = O.splitmix(0xDE45E, u1.sizes([8])["packed_len"])
    z1[6 << 3:] = 0
    assert np.array_equal(u1.compact([8], z1), z1)


def _random_types(rng, n):
    types = []
    for _ in range(n):
        k = rng.randrange(7, 13)
        if rng.random() < 0.3:
            types.append((k, 128 * rng.randrange(1, (1 << k) // 128 + 1), True))
        else:
            types.append((k, rng.randrange(1, (1 << k) + 1), False))
    return types


def test_abi_registry_and_sizes_match_oracle():
    rng = random.Random(5)
    for trial in range(60):
        types = _random_types(rng, rng.randrange(1, 6))
        nu = rng.randrange(0, 6)
        u = Union(types, nu)
        reg = F.registry_new(types, nu)
        for f in ("m_total", "m_bool", "m_elem", "element_base", "nu"):
            assert getattr(reg, f) == getattr(u.reg, f), f
        assert reg.num_boolean == u.reg.num_boolean and reg.n_types == len(types)
        for s in range(len(types)):
            a, b = reg.slots[s], u.reg.slots[s]
            assert (a.offset, a.prefix, a.m_slot, a.prefix_bits, a.type_index) == \
                (b.offset, b.prefix, b.m_slot, b.prefix_bits, b.type_index)
        for _ in range(3):
            counts = [rng.randrange(0, (1 << nu) + 1) for _ in types]
            lbs = rng.randrange(0, 3)
            if u.reg.m_total < 7:
                continue
            exp = u.sizes(counts, 0, lbs) if u.sizes(counts)["dense_m"] >= 7 + lbs else None
            if exp is None:
                continue
            assert F.union_sizes(reg, counts, 0, lbs) == exp


def test_abi_sizes_reject_bad_counts():
    reg = F.registry_new([(10, 700, False)], 3)
    with pytest.raises(F.FlockHipError):
        F.union_sizes(reg, [9])
    with pytest.raises(F.FlockHipError):
        F.registry_new([(6, 10, False)], 3)


def test_x_ab_from_mlv_and_bind_statement_match_oracle():
    reg = F.registry_new([(15, 25_470, False)], 10)
    n_mlv = reg.m_bool - 6
    mlv = M.splitmix(0x55, n_mlv)
    xr, xo = F.x_ab_from_mlv(reg, mlv)
    assert xr == [mlv[0]] + mlv[1 + 10:] and xo == mlv[1:11]
    exr = np.zeros((n_mlv - 10, 2), dtype=np.uint64)
    exo = np.zeros((10, 2), dtype=np.uint64)
    ma = O.ints_to_arr(mlv)
    L.or_x_ab_from_mlv(O._p(ma), n_mlv, 10, O._p(exr), O._p(exo))
    assert O.arr_to_ints(exr) == xr and O.arr_to_ints(exo) == xo
    # bind_statement: same transcript as the oracle and as the framing restated in puremath
    digest = bytes(range(32))
    counts = [1024, 7]
    cap = bytes(rng_b for rng_b in range(96))
    fs = F.FsChallenger(b"bind-test")
    fs.bind_statement(digest, counts, cap)
    ofs = O.Fs(b"bind-test")
    c = np.array(counts, dtype=np.uint64)
    L.or_bind_statement(ofs.h, digest, c.ctypes.data, 2, cap, len(cap))
    pm = M.FsChallenger(b"bind-test")
    pm.observe_label(b"flock-mixed-v1")
    pm.observe_bytes(digest)
    pm.observe_bytes(b"".join(n.to_bytes(8, "little") for n in counts))
    pm.observe_bytes(cap)
    a, b, e = fs.sample_f128(), O.to_int(L.or_fs_sample_f128(ofs.h)), pm.sample_f128()
    assert a == b == e
