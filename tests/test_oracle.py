"""Pin the CPU oracle: golden fixtures from pure mathematics, known answers,
field axioms, sumcheck identities, and the reference's own host checkers
(oracle/_ref, built from the reference sources in this container only)."""
from __future__ import annotations

import json
import os
import random
import struct
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import puremath as M

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as fh:
        return json.load(fh)


def H(s):
    return int(s, 16)


# ---------------------------------------------------------------- field ----

def test_f128_golden_vectors():
    g = load("f128_mul.json")
    for a, b, c in g["cases"]:
        assert O.mul(H(a), H(b)) == H(c)
        assert O.to_int(O.lib().or_f128_mul_portable(O.f(H(a)), O.f(H(b)))) == H(c)
    for a, ai in g["inverses"]:
        assert O.to_int(O.lib().or_f128_inv(O.f(H(a)))) == H(ai)


def test_f128_ghash_known_answers():
    # gf2_128.rs:503-531: x * x^127 = x^128 = x^7+x^2+x+1 = 0x87; x^64 * x^64 = 0x87; x * x^63 = x^64
    assert O.mul(2, 1 << 127) == 0x87
    assert O.mul(1 << 64, 1 << 64) == 0x87
    assert O.mul(2, 1 << 63) == 1 << 64
    assert O.to_int(O.lib().or_f128_mul_by_x(O.f(1 << 127))) == 0x87


def test_f128_axioms():
    rng = random.Random(7)
    for _ in range(50):
        a, b, c = (rng.getrandbits(128) for _ in range(3))
        assert O.mul(a, b) == O.mul(b, a)
        assert O.mul(O.mul(a, b), c) == O.mul(a, O.mul(b, c))
        assert O.mul(a, b ^ c) == O.mul(a, b) ^ O.mul(a, c)
        if a:
            assert O.mul(a, O.to_int(O.lib().or_f128_inv(O.f(a)))) == 1


def test_splitmix_matches_pack_witness_words():
    # dump_commit_vectors.rs: z = rng.bits(2^m); pack_witness packs z[128i..128i+128]
    # LSB-first, so packed[i] = {lo: word 2i, hi: word 2i+1}.
    a = O.arr_to_ints(O.splitmix(0xC0FFEE, 8))
    assert a == M.splitmix(0xC0FFEE, 8)


# ------------------------------------------------------------------ NTT ----

def test_ntt_matches_definition():
    for case in load("ntt_definition.json")["cases"]:
        L = case["L"]
        c = O.ints_to_arr([H(v) for v in case["coeffs"]])
        got = O.arr_to_ints(O.ntt_forward_interleaved(c, 1, 1, L, L, 0))
        assert got == [H(v) for v in case["evals"]], L


def test_ntt_roundtrip_linearity_zero():
    # additive_ntt_f128.rs:1344-1416
    for L in (1, 2, 3, 5, 8):
        a = O.splitmix(100 + L, 1 << L)
        b = O.splitmix(200 + L, 1 << L)
        fa = O.ntt_forward_interleaved(a, 1, 1, L, L, 0)
        assert np.array_equal(O.ntt_inverse(fa, L, L), a)
        fb = O.ntt_forward_interleaved(b, 1, 1, L, L, 0)
        fab = O.ntt_forward_interleaved(a ^ b, 1, 1, L, L, 0)
        assert np.array_equal(fab, fa ^ fb)
        z = np.zeros_like(a)
        assert not O.ntt_forward_interleaved(z, 1, 1, L, L, 0).any()
    assert O.twiddle(4, 0, 0) == 0


def test_ntt_interleaved_equals_per_lane():
    # additive_ntt_f128.rs:1438-1508: lane l of the SoA transform == the lane alone
    L, t = 6, 5
    data = O.splitmix(0x1A, (1 << L) * t)
    inter = O.ntt_forward_interleaved(data, t, t, L, L, 0)
    for lane in range(t):
        alone = O.ntt_forward_interleaved(np.ascontiguousarray(data[lane::t]), 1, 1, L, L, 0)
        assert np.array_equal(inter[lane::t], alone)


def test_ntt_start_layer_equals_replicated_zero_padded():
    # commit.rs:377-395: replicating the message 2^r times == r layers on [msg, 0..]
    L, r, t = 7, 2, 3
    msg = O.splitmix(0x1B, (1 << (L - r)) * t)
    padded = np.zeros(((1 << L) * t, 2), dtype=np.uint64)
    padded[: msg.shape[0]] = msg
    full = O.ntt_forward_interleaved(padded, t, t, L, L, 0)
    rep = np.ascontiguousarray(np.tile(msg, (1 << r, 1)))
    assert np.array_equal(O.ntt_forward_interleaved(rep, t, t, L, L, r), full)


# ------------------------------------------------------- Merkle / SHA ----

def test_sha256_matches_hashlib():
    import ctypes
    rng = random.Random(3)
    for n in (0, 1, 55, 56, 63, 64, 65, 800, 4096):
        data = bytes(rng.getrandbits(8) for _ in range(n))
        out = ctypes.create_string_buffer(32)
        O.lib().or_sha256(data, n, out)
        assert out.raw == M.sha256(data)


def test_merkle_tree_layout():
    rng = random.Random(4)
    for n_leaves, leaf in ((1, 16), (2, 16), (8, 800), (16, 64)):
        data = bytes(rng.getrandbits(8) for _ in range(n_leaves * leaf))
        tree = O.merkle_tree(data, n_leaves)
        exp = b"".join(M.merkle_tree(data, n_leaves))
        assert tree == exp
    # merkle.rs:763-779: hand-computed two-leaf tree
    l0, l1 = b"\x00" * 16, b"\x01" * 16
    tree = O.merkle_tree(l0 + l1, 2)
    assert tree[64:] == M.sha256(M.sha256(l0) + M.sha256(l1))


# --------------------------------------------------------------- commit ----

def test_commit_lane_major_matches_definition():
    for c in load("commit_definition.json")["cases"]:
        q = O.ints_to_arr([H(v) for v in c["q"]])
        cw, tree = O.commit_lane_major(q, c["m"], c["log_inv_rate"], c["log_batch_size"],
                                       c["num_lanes"])
        assert O.arr_to_ints(cw) == [H(v) for v in c["codeword"]]
        assert tree == b"".join(bytes.fromhex(h) for h in c["tree"])
        assert tree[-32:].hex() == c["root"]


def test_commit_interleaved_equals_lane_major_transposed():
    # commit.rs:462-473: commit_lane_major(q) == commit(extract) with extract[p*t+l] = q[l*D+p]
    m, rate, lb, t = 7 + 3 + 4, 2, 3, 6
    s = O.pcs_sizes(m, rate, lb, t)
    D = 1 << (m - 7 - lb)
    q = O.splitmix(0x77, 1 << (m - 7))
    q[t * D:] = 0
    ext = np.zeros((t * D, 2), dtype=np.uint64)
    for l in range(t):
        for p in range(D):
            ext[p * t + l] = q[l * D + p]
    a = O.commit_lane_major(q, m, rate, lb, t)
    b = O.commit(ext, m, rate, lb, t)
    assert np.array_equal(a[0], b[0]) and a[1] == b[1]
    assert s["cw_len"] == a[0].shape[0]


# ------------------------------------------------------- FsChallenger ----

def _replay_fs(fs, ops):
    for op in ops:
        kind = op[0]
        if kind == "label":
            fs.observe_label(bytes.fromhex(op[1]))
        elif kind == "f128":
            fs.observe_f128(H(op[1]))
        elif kind == "slice":
            fs.observe_f128_slice([H(v) for v in op[1]])
        elif kind == "bytes":
            fs.observe_bytes(bytes.fromhex(op[1]))
        elif kind == "sample":
            assert fs.sample_f128() == H(op[1])
        elif kind == "sample_vec":
            assert fs.sample_f128_vec(len(op[1])) == [H(v) for v in op[1]]
        elif kind == "grind":
            assert fs.grind_pow(op[1]) == op[2]


def test_fs_transcript_golden():
    g = load("fs_transcript.json")
    _replay_fs(O.Fs(bytes.fromhex(g["domain"])), g["ops"])


# ----------------------------------------------------------- sumcheck ----

def test_sumcheck_round_functions_match_definition():
    g = load("sumcheck_definition.json")
    a = O.ints_to_arr([H(v) for v in g["a"]])
    b = O.ints_to_arr([H(v) for v in g["b"]])
    r = H(g["r"])
    assert O.round_msg_lsb(a, b) == [H(v) for v in g["round_msg"]]
    nf, nb, msg = O.fold_and_msg_lsb(a, b, r)
    assert O.arr_to_ints(nf) == [H(v) for v in g["fold_a"]]
    assert O.arr_to_ints(nb) == [H(v) for v in g["fold_b"]]
    assert msg == [H(v) for v in g["fold_msg"]]
    rv = [H(v) for v in g["eq_r"]]
    assert O.round_pair(a, b, rv) == [H(v) for v in g["round_pair"]]
    assert O.arr_to_ints(O.build_eq(rv)) == [H(v) for v in g["eq_table"]]
    fa, fb = O.fold_pair(a, b, r)
    assert O.arr_to_ints(fa) == [H(v) for v in g["fold_a"]]
    assert O.arr_to_ints(fb) == [H(v) for v in g["fold_b"]]


def _quad_eval(u0, u2, claim, r):
    # RoundQuad::from_msg / eval, ligerito.rs:3175-3190: c + b r + a r^2, b = claim + u2
    return u0 ^ M.mul(r, claim ^ u2) ^ M.mul(M.mul(r, r), u2)


def test_product_sumcheck_identities():
    n = 64
    f = O.splitmix(0x91, n)
    b = O.splitmix(0x92, n)
    fi, bi = O.arr_to_ints(f), O.arr_to_ints(b)
    claim = 0
    for x, y in zip(fi, bi):
        claim ^= M.mul(x, y)
    u0, u2 = O.round_msg_lsb(f, b)
    rng = random.Random(5)
    while f.shape[0] > 1:
        r = rng.getrandbits(128)
        nf, nb, nxt = O.fold_and_msg_lsb(f, b, r)
        new_claim = 0
        for x, y in zip(O.arr_to_ints(nf), O.arr_to_ints(nb)):
            new_claim ^= M.mul(x, y)
        assert new_claim == _quad_eval(u0, u2, claim, r)
        if nf.shape[0] > 1:
            assert nxt == O.round_msg_lsb(nf, nb)
        claim, (u0, u2), f, b = new_claim, nxt, nf, nb


def test_zerocheck_round_identities():
    # G(X) = sum_x eq(r', x) a(X,x) b(X,x) is quadratic: folding at rho gives G(rho),
    # with G(0) by brute force and (G(1), G(inf)) from round_pair (r[0] = ONE).
    n = 32
    a = O.splitmix(0x93, n)
    b = O.splitmix(0x94, n)
    rv = [1] + M.splitmix(0x95, 4)
    g1, ginf = O.round_pair(a, b, rv)
    ai, bi = O.arr_to_ints(a), O.arr_to_ints(b)
    g0 = 0
    for x in range(n // 2):
        g0 ^= M.mul(M.eq_point(rv[1:], x), M.mul(ai[2 * x], bi[2 * x]))
    rho = 0x1234567890ABCDEF1122334455667788
    fa, fb = O.fold_pair(a, b, rho)
    lhs = 0
    for x, (p, q) in enumerate(zip(O.arr_to_ints(fa), O.arr_to_ints(fb))):
        lhs ^= M.mul(M.eq_point(rv[1:], x), M.mul(p, q))
    g1c = g0 ^ g1 ^ ginf
    assert lhs == g0 ^ M.mul(rho, g1c) ^ M.mul(M.mul(rho, rho), ginf)


def test_sumcheck_transcript_is_deterministic_and_sized():
    a = O.splitmix(0x96, 1 << 6)
    b = O.splitmix(0x97, 1 << 6)
    m1, f1 = O.sumcheck_transcript(b"flock-test", a, b, 6)
    m2, f2 = O.sumcheck_transcript(b"flock-test", a, b, 6)
    assert np.array_equal(m1, m2) and np.array_equal(f1, f2)
    m3, _ = O.sumcheck_transcript(b"flock-test2", a, b, 6)
    assert not np.array_equal(m1, m3)
    assert m1.shape == (24, 2)


# ------------------------------------------ reference host checkers ----

REF_NTT = os.path.join(O.ORACLE_DIR, "_ref", "host_check_ntt")
REF_FS = os.path.join(O.ORACLE_DIR, "_ref", "test_challenger")
need_ref = pytest.mark.skipif(not (os.path.exists("/root/reference") and os.path.exists(REF_NTT)),
                              reason="reference checkout / oracle/_ref not present")


def write_cmt1(path, m, rate, lb, msg, cw, root):
    s = O.pcs_sizes(m, rate, lb)
    with open(path, "wb") as fh:
        fh.write(struct.pack("<I", 0x434D5431))
        for v in (m, rate, lb, s["k_code"], s["num_ntts"], s["n_leaves"], s["n_leaves"],
                  s["num_ntts"] * 16, msg.shape[0]):
            fh.write(struct.pack("<I", v))
        fh.write(msg.tobytes())
        fh.write(struct.pack("<I", cw.shape[0]))
        fh.write(cw.tobytes())
        fh.write(root)


@need_ref
def test_oracle_commit_config1_against_reference_host_check_ntt(tmp_path):
    # Config 1: m=25, rate 1/4, 256 lanes, SplitMix64 seed 0xC0FFEE message.
    m, rate, lb = 25, 2, 8
    msg = O.splitmix(0xC0FFEE, 1 << (m - 7))
    cw, tree = O.commit(msg, m, rate, lb)
    p = tmp_path / "c1.bin"
    write_cmt1(p, m, rate, lb, msg, cw, tree[-32:])
    out = subprocess.run([REF_NTT, str(p)], capture_output=True, text=True)
    assert out.returncode == 0 and "NTT OK" in out.stdout, out.stdout
    # and the planted-mistake control: a flipped codeword bit must be caught
    cw2 = cw.copy()
    cw2[12345, 0] ^= 1
    write_cmt1(p, m, rate, lb, msg, cw2, tree[-32:])
    out = subprocess.run([REF_NTT, str(p)], capture_output=True, text=True)
    assert out.returncode != 0


@need_ref
def test_oracle_fs_against_reference_challenger(tmp_path):
    rng = random.Random(11)
    fs = O.Fs(b"flock-readme-bench-v0")
    body = b""
    nops = 0
    for i in range(40):
        kind = rng.choice(["f128", "bytes", "sample", "label", "slice", "sample_vec", "grind"])
        if kind == "f128":
            v = rng.getrandbits(128)
            fs.observe_f128(v)
            body += b"\x01" + v.to_bytes(16, "little")
        elif kind == "bytes":
            b = bytes(rng.getrandbits(8) for _ in range(rng.randrange(0, 40)))
            fs.observe_bytes(b)
            body += b"\x02" + struct.pack("<I", len(b)) + b
        elif kind == "sample":
            body += b"\x03" + fs.sample_f128().to_bytes(16, "little")
        elif kind == "label":
            b = b"flock-zerocheck-v0"[: rng.randrange(1, 19)]
            fs.observe_label(b)
            body += b"\x04" + struct.pack("<I", len(b)) + b
        elif kind == "slice":
            vs = [rng.getrandbits(128) for _ in range(rng.randrange(1, 5))]
            fs.observe_f128_slice(vs)
            body += b"\x06" + struct.pack("<I", len(vs)) + b"".join(v.to_bytes(16, "little") for v in vs)
        elif kind == "sample_vec":
            k = rng.randrange(1, 6)
            vs = fs.sample_f128_vec(k)
            body += b"\x07" + struct.pack("<I", k) + b"".join(v.to_bytes(16, "little") for v in vs)
        else:
            bits = rng.randrange(0, 9)
            nonce = fs.grind_pow(bits)
            body += b"\x05" + struct.pack("<IQ", bits, nonce)
        nops += 1
    dom = b"flock-readme-bench-v0"
    p = tmp_path / "chlg.bin"
    p.write_bytes(struct.pack("<I", 0x43484C47) + struct.pack("<I", len(dom)) + dom
                  + struct.pack("<I", nops) + body)
    out = subprocess.run([REF_FS, str(p)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout


REF_B3W = os.path.join(O.ORACLE_DIR, "_ref", "test_blake3_witness_host")
need_ref_b3w = pytest.mark.skipif(
    not (os.path.exists("/root/reference") and os.path.exists(REF_B3W)),
    reason="reference checkout / oracle/_ref/test_blake3_witness_host not present")


def write_b3wt(path, n_blocks_log, inputs, z, a, b, stripe):
    """The B3WT vector file cuda-ghash/test_blake3_witness_host.cpp reads: magic, n_blocks_log,
    n_blocks, k_log; per block cv[8], m[16], counter u64, block_len, flags; then the per-block
    packed z, a, b (256 u64 each) and the lincheck stripe."""
    with open(path, "wb") as fh:
        fh.write(struct.pack("<IIII", 0x42335754, n_blocks_log, len(inputs), 14))
        for row in inputs:
            fh.write(struct.pack("<8I", *[int(x) for x in row[:8]]))
            fh.write(struct.pack("<16I", *[int(x) for x in row[8:24]]))
            fh.write(struct.pack("<QII", int(row[24]) | (int(row[25]) << 32), int(row[26]),
                                 int(row[27])))
        for t in (z, a, b):
            fh.write(np.ascontiguousarray(t, dtype=np.uint64).tobytes())
        fh.write(np.ascontiguousarray(stripe, dtype=np.uint8).tobytes())


@need_ref_b3w
def test_oracle_blake3_witness_against_reference_host_check(tmp_path):
    # the oracle's per-block (z, a, b) traces and its batch-major lincheck stripe, checked by
    # the reference's own host replica of the BLAKE3 witness builder (blake3_witness.cuh)
    import test_oracle_blake3_r1cs as B3
    n_log = 4
    inputs = B3.random_inputs(1 << n_log, 0xB3B3)
    inputs[3, 24:26] = (0xFFFFFFFF, 7)  # a counter with both halves set
    traces = [B3.instance(row) for row in inputs]
    z = np.stack([t[0] for t in traces])
    a = np.stack([t[1] for t in traces])
    b = np.stack([t[2] for t in traces])
    _, _, _, stripe = B3.batch_major(inputs, n_log)
    p = tmp_path / "b3.bin"
    write_b3wt(p, n_log, inputs, z, a, b, stripe)
    out = subprocess.run([REF_B3W, str(p)], capture_output=True, text=True)
    assert out.returncode == 0 and "BLAKE3 WITNESS HOST CHECK OK" in out.stdout, out.stdout
    # planted mistakes: one bit of one trace, then one stripe bit, must each be caught
    a2 = a.copy()
    a2[5, 77] ^= 1 << 13
    write_b3wt(p, n_log, inputs, z, a2, b, stripe)
    out = subprocess.run([REF_B3W, str(p)], capture_output=True, text=True)
    assert out.returncode != 0 and "a FAIL" in out.stdout
    s2 = stripe.copy()
    s2[9000] ^= 4
    write_b3wt(p, n_log, inputs, z, a, b, s2)
    out = subprocess.run([REF_B3W, str(p)], capture_output=True, text=True)
    assert out.returncode != 0 and "z_lincheck FAIL" in out.stdout
