"""GPU parity: every kernel through the C ABI against the CPU oracle, bit-exact."""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from flock_amd import lib as F

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


def H(s):
    return int(s, 16)


# ------------------------------------------------------------------ field ----

def test_f128_mul_batch(ctx):
    say("case: f128_mul_batch")
    g = json.load(open(os.path.join(GOLD, "f128_mul.json")))
    a = O.ints_to_arr([H(x[0]) for x in g["cases"]])
    b = O.ints_to_arr([H(x[1]) for x in g["cases"]])
    exp = O.ints_to_arr([H(x[2]) for x in g["cases"]])
    n = a.shape[0]
    da, db, do = ctx.upload(a), ctx.upload(b), ctx.alloc(16 * n)
    ctx.f128_mul_batch(da, db, do, n)
    assert np.array_equal(do.download_f128(n), exp)
    # larger random batch against the oracle
    n = 1 << 16
    a, b = O.splitmix(1, n), O.splitmix(2, n)
    da, db, do = ctx.upload(a), ctx.upload(b), ctx.alloc(16 * n)
    ctx.f128_mul_batch(da, db, do, n)
    assert np.array_equal(do.download_f128(n), O.mul_batch(a, b))


@pytest.mark.parametrize("method", [F.MUL_SHOUP, F.MUL_MFMA_I8])
def test_f128_mul_const_pow_batch(ctx, method):
    """out = a * t^reps on the Shoup table and on the int8 matrix cores, against the oracle's
    multiply applied reps times (ragged n: the MFMA form's last 32-element tile is partial)."""
    say(f"case: f128_mul_const_pow_batch method={method}")
    n = (1 << 12) + 37
    a = O.splitmix(5, n)
    da, do = ctx.upload(a), ctx.alloc(16 * n)
    for t in (0, 1, 2, 1 << 127, (1 << 128) - 1, 0x0DCB364640A222FE6B8330483C2E9849):
        tt = O.ints_to_arr([t] * n)
        for reps in (0, 1, 3):
            exp = a.copy()
            for _ in range(reps):
                exp = O.mul_batch(exp, tt)
            ctx.f128_mul_const_pow_batch(da, t, reps, method, do, n)
            assert np.array_equal(do.download_f128(n), exp), (hex(t), reps)


def test_f128_mul_const_batch(ctx):
    say("case: f128_mul_const_batch")
    n = 1 << 14
    a = O.splitmix(3, n)
    da, do = ctx.upload(a), ctx.alloc(16 * n)
    for t in (0, 1, 2, 1 << 127, (1 << 128) - 1, 0x0DCB364640A222FE6B8330483C2E9849):
        ctx.f128_mul_const_batch(da, t, do, n)
        tb = np.tile(O.ints_to_arr([t]), (n, 1))
        assert np.array_equal(do.download_f128(n), O.mul_batch(a, tb)), hex(t)


# -------------------------------------------------------------------- NTT ----

@pytest.mark.parametrize("log_d,t,live,start", [
    (1, 1, 1, 0), (2, 3, 3, 0), (5, 1, 1, 0), (6, 4, 4, 1), (7, 50, 50, 1), (9, 7, 5, 2),
    (10, 64, 64, 2), (12, 256, 256, 2), (13, 50, 50, 1), (11, 33, 20, 3), (16, 8, 8, 1),
])
def test_ntt_forward_interleaved(ctx, log_d, t, live, start):
    say(f"case: ntt log_d={log_d} t={t} live={live} start={start}")
    n = (1 << log_d) * t
    data = O.splitmix(100 + log_d * 7 + t, n)
    if live < t:
        data.reshape(-1, t, 2)[:, live:, :] = 0
    d = ctx.upload(data)
    ctx.ntt_forward_interleaved(d, t, live, log_d, start)
    exp = O.ntt_forward_interleaved(data, t, live, log_d, log_d, start)
    assert np.array_equal(d.download_f128(n), exp)


def test_ntt_single_passes_compose(ctx):
    say("case: ntt single passes 3+3+1+2")
    log_d, t, start = 11, 24, 2
    n = (1 << log_d) * t
    data = O.splitmix(0x5A55, n)
    d = ctx.upload(data)
    l = start
    for k in (3, 3, 1, 2):
        ctx.ntt_pass(d, t, t, log_d, l, k)
        l += k
    assert np.array_equal(d.download_f128(n), O.ntt_forward_interleaved(data, t, t, log_d, log_d, start))
    with pytest.raises(F.FlockHipError):
        ctx.ntt_pass(d, t, t, log_d, 9, 3)  # runs past the last layer


# ----------------------------------------------------------------- Merkle ----

@pytest.mark.parametrize("n_leaves,leaf", [(1, 16), (2, 32), (4, 48), (256, 64), (512, 800),
                                           (1024, 4096), (4096, 16), (8192, 800)])
def test_merkle_tree(ctx, n_leaves, leaf):
    say(f"case: merkle n_leaves={n_leaves} leaf={leaf}")
    rng = np.random.default_rng(n_leaves * 7 + leaf)
    data = rng.integers(0, 256, size=n_leaves * leaf, dtype=np.uint8)
    d = ctx.upload(data)
    tree = ctx.alloc(32 * (2 * n_leaves - 1))
    ctx.merkle_tree(d, n_leaves * leaf, n_leaves, tree)
    assert tree.download() == O.merkle_tree(data.tobytes(), n_leaves)


# BLAKE3 nodes (HashKind::Blake3, merkle.rs:195-232): one-chunk leaves, and leaves of 2, 3, 4
# and 5 chunks of 1024 bytes (the left-balanced chunk tree; 3184 B ends in a partial chunk)
@pytest.mark.parametrize("n_leaves,leaf", [(1, 16), (2, 32), (4, 48), (256, 64), (512, 800),
                                           (256, 1024), (128, 1040), (64, 3072), (32, 3184),
                                           (1024, 4096), (16, 5120), (4096, 16), (8192, 800)])
def test_merkle_tree_blake3(ctx, n_leaves, leaf):
    say(f"case: merkle blake3 n_leaves={n_leaves} leaf={leaf}")
    rng = np.random.default_rng(n_leaves * 11 + leaf)
    data = rng.integers(0, 256, size=n_leaves * leaf, dtype=np.uint8)
    d = ctx.upload(data)
    tree = ctx.alloc(32 * (2 * n_leaves - 1))
    ctx.merkle_tree(d, n_leaves * leaf, n_leaves, tree, kind=F.MERKLE_BLAKE3)
    assert tree.download() == O.merkle_tree(data.tobytes(), n_leaves, kind=1)


def test_merkle_tree_rejects_unknown_hash(ctx):
    d, tree = ctx.alloc(64), ctx.alloc(32)
    with pytest.raises(F.FlockHipError) as e:
        ctx.merkle_tree(d, 64, 1, tree, kind=2)
    assert e.value.code == F.E_INVALID


# ----------------------------------------------------------------- commit ----

def test_commit_lane_major_golden(ctx):
    say("case: commit golden fixtures")
    for c in json.load(open(os.path.join(GOLD, "commit_definition.json")))["cases"]:
        q = O.ints_to_arr([H(v) for v in c["q"]])
        p = F.PcsParams(c["m"], c["log_inv_rate"], c["log_batch_size"], c["num_lanes"], 0)
        s = F.pcs_sizes(p)
        dq, cw, tree = ctx.upload(q), ctx.alloc(16 * s["cw_len"]), ctx.alloc(32 * (2 * s["n_leaves"] - 1))
        cap = ctx.commit_lane_major(dq, q.shape[0], p, cw, tree, 0)
        assert O.arr_to_ints(cw.download_f128(s["cw_len"])) == [H(v) for v in c["codeword"]]
        assert tree.download() == b"".join(bytes.fromhex(h) for h in c["tree"])
        assert cap[0].hex() == c["root"]


def _commit_vs_oracle(ctx, m, rate, lb, nl, seed, lane_major=True, cap_depth=0, kind=0):
    p = F.PcsParams(m, rate, lb, nl, kind)
    s = F.pcs_sizes(p)
    if lane_major:
        q = O.splitmix(seed, 1 << (m - 7))
        d = (1 << (m - 7)) >> lb
        q[s["num_ntts"] * d:] = 0
        exp_cw, exp_tree = O.commit_lane_major(q, m, rate, lb, nl, merkle_hash=kind)
    else:
        q = O.splitmix(seed, s["msg_len"])
        exp_cw, exp_tree = O.commit(q, m, rate, lb, nl, merkle_hash=kind)
    dq, cw, tree = ctx.upload(q), ctx.alloc(16 * s["cw_len"]), ctx.alloc(32 * (2 * s["n_leaves"] - 1))
    fn = ctx.commit_lane_major if lane_major else ctx.commit
    cap = fn(dq, q.shape[0], p, cw, tree, cap_depth)
    assert np.array_equal(cw.download_f128(s["cw_len"]), exp_cw)
    got_tree = tree.download()
    assert got_tree == exp_tree
    n = s["n_leaves"]
    lo = 2 * n - 2 * (1 << cap_depth)
    assert b"".join(cap) == exp_tree[32 * lo: 32 * (lo + (1 << cap_depth))]


def test_commit_config1(ctx):
    # message 2^10 rows x 2^8 columns (m = 25), rate 1/4, SplitMix64 seed 0xC0FFEE
    say("case: commit config1 m=25 rate=1/4 lanes=256")
    _commit_vs_oracle(ctx, 25, 2, 8, 0, 0xC0FFEE, lane_major=False, cap_depth=0)
    _commit_vs_oracle(ctx, 25, 2, 8, 0, 0xC0FFEE, lane_major=True, cap_depth=5)


def test_commit_config2_shape(ctx):
    # SHA-256 x1024 union stack: m=25, log_batch_size=6, t=50 integer lanes, rate 1/2
    say("case: commit config2 shape m=25 lb=6 t=50")
    _commit_vs_oracle(ctx, 25, 1, 6, 50, 0x5A256, lane_major=True, cap_depth=8)


def test_commit_edge_shapes(ctx):
    for m, rate, lb, nl in [(7, 1, 0, 0), (9, 3, 2, 1), (14, 2, 5, 17), (16, 1, 3, 0)]:
        say(f"case: commit edge m={m} rate={rate} lb={lb} nl={nl}")
        _commit_vs_oracle(ctx, m, rate, lb, nl, m * 31 + nl, lane_major=True, cap_depth=1)


def test_commit_blake3_merkle(ctx):
    # PcsParams { merkle_hash: Blake3 }: config 1's 4096-byte leaves (4 chunks), the SHA-256
    # union shape's 800-byte leaves (t = 50), and a small odd shape
    say("case: commit blake3 config1 m=25 rate=1/4 lanes=256")
    _commit_vs_oracle(ctx, 25, 2, 8, 0, 0xC0FFEE, lane_major=False, cap_depth=3, kind=1)
    say("case: commit blake3 m=25 lb=6 t=50")
    _commit_vs_oracle(ctx, 25, 1, 6, 50, 0x5A256, lane_major=True, cap_depth=8, kind=1)
    say("case: commit blake3 m=14 lb=5 t=17")
    _commit_vs_oracle(ctx, 14, 2, 5, 17, 0xB3, lane_major=True, cap_depth=1, kind=1)


def test_encode_matches_commit_codeword(ctx):
    say("case: encode / encode_lane_major")
    m, rate, lb, nl = 18, 2, 4, 11
    p = F.PcsParams(m, rate, lb, nl, 0)
    s = F.pcs_sizes(p)
    q = O.splitmix(0xE1, 1 << (m - 7))
    q[nl * ((1 << (m - 7)) >> lb):] = 0
    cw = ctx.alloc(16 * s["cw_len"])
    ctx.encode_lane_major(ctx.upload(q), q.shape[0], p, cw)
    assert np.array_equal(cw.download_f128(s["cw_len"]), O.commit_lane_major(q, m, rate, lb, nl)[0])
    msg = O.splitmix(0xE2, s["msg_len"])
    ctx.encode(ctx.upload(msg), s["msg_len"], p, cw)
    assert np.array_equal(cw.download_f128(s["cw_len"]), O.commit(msg, m, rate, lb, nl)[0])


def test_commit_rejects_bad_input(ctx):
    p = F.PcsParams(20, 1, 5, 0, 0)
    buf = ctx.alloc(16)
    with pytest.raises(F.FlockHipError) as e:
        ctx.commit_lane_major(buf, 7, p, buf, buf)  # wrong q length
    assert e.value.code == F.E_INVALID


# --------------------------------------------------------------- sumcheck ----

@pytest.mark.parametrize("log_n", [1, 2, 3, 4, 9, 14])
def test_product_sumcheck_rounds(ctx, log_n):
    say(f"case: product sumcheck rounds log_n={log_n}")
    n = 1 << log_n
    f, b = O.splitmix(7 + log_n, n), O.splitmix(8 + log_n, n)
    df, db = ctx.upload(f), ctx.upload(b)
    assert list(ctx.round_msg_lsb(df, db, n)) == O.round_msg_lsb(f, b)
    r = 0x1234_5678_9ABC_DEF0_0FED_CBA9_8765_4321 ^ log_n
    nf, nb = ctx.alloc(8 * n), ctx.alloc(8 * n)
    msg = ctx.fold_and_msg_lsb(df, db, n, r, nf, nb)
    ef, eb, em = O.fold_and_msg_lsb(f, b, r)
    assert list(msg) == em
    assert np.array_equal(nf.download_f128(n // 2), ef)
    assert np.array_equal(nb.download_f128(n // 2), eb)


@pytest.mark.parametrize("log_n", [1, 2, 3, 5, 10, 15])
def test_zerocheck_rounds(ctx, log_n):
    say(f"case: zerocheck rounds log_n={log_n}")
    n = 1 << log_n
    a, b = O.splitmix(30 + log_n, n), O.splitmix(40 + log_n, n)
    r = O.arr_to_ints(O.splitmix(50 + log_n, log_n))
    da, db = ctx.upload(a), ctx.upload(b)
    eq = ctx.alloc(16 * max(n // 2, 1))
    ctx.build_eq(r[1:], eq)
    assert np.array_equal(eq.download_f128(n // 2), O.build_eq(r[1:]))
    eq_next = ctx.alloc(16 * max(n // 4, 1))
    msg = ctx.round_pair(da, db, n, eq, r[0], eq_next)
    assert list(msg) == O.round_pair(a, b, r)
    if n >= 4:
        assert np.array_equal(eq_next.download_f128(n // 4), O.build_eq(r[2:]))
    rho = O.arr_to_ints(O.splitmix(60 + log_n, 1))[0]
    ao, bo = ctx.alloc(8 * n), ctx.alloc(8 * n)
    ea, eb = O.fold_pair(a, b, rho)
    ctx.fold_pair(da, db, n, rho, ao, bo)
    assert np.array_equal(ao.download_f128(n // 2), ea)
    assert np.array_equal(bo.download_f128(n // 2), eb)
    if n >= 4:
        # fused fold + next message with eq over r[2..], r0 = r[1]
        ao2, bo2 = ctx.alloc(8 * n), ctx.alloc(8 * n)
        eq1 = ctx.alloc(16 * (n // 4))
        ctx.build_eq(r[2:], eq1)
        nxt = ctx.alloc(16 * max(n // 8, 1))
        msg = ctx.fold_and_compute_round_pair(da, db, n, rho, ao2, bo2, eq1, r[1], nxt)
        assert list(msg) == O.round_pair(ea, eb, r[1:])
        assert np.array_equal(ao2.download_f128(n // 2), ea)
        if n >= 8:
            assert np.array_equal(nxt.download_f128(n // 8), O.build_eq(r[3:]))


@pytest.mark.parametrize("log_n", [1, 2, 3, 6, 11, 16, 17, 18])
def test_prove_sumcheck_transcript(ctx, log_n):
    say(f"case: prove_sumcheck log_n={log_n}")
    n = 1 << log_n
    a, b = O.splitmix(0xA0 + log_n, n), O.splitmix(0xB0 + log_n, n)
    dom = b"flock-sumcheck-test"
    exp_m, exp_f = O.sumcheck_transcript(dom, a, b, log_n)
    da, db = ctx.upload(a), ctx.upload(b)
    m, fin = ctx.prove_sumcheck(F.FsChallenger(dom), da, db, log_n)
    assert np.array_equal(m, exp_m)
    assert np.array_equal(fin, exp_f)
    # inputs are left untouched
    assert np.array_equal(da.download_f128(n), a)


def test_prove_sumcheck_config3_full_size(ctx):
    # two tables of 2^24 seeded uniform F128, 24 product rounds + 24 zerocheck rounds
    log_n = 24
    say(f"case: prove_sumcheck config3 log_n={log_n}")
    n = 1 << log_n
    a, b = O.splitmix(0xC3A, n), O.splitmix(0xC3B, n)
    dom = b"flock-mi355x-sumcheck-v0"
    da, db = ctx.upload(a), ctx.upload(b)
    m, fin = ctx.prove_sumcheck(F.FsChallenger(dom), da, db, log_n)
    exp_m, exp_f = O.sumcheck_transcript(dom, a, b, log_n)
    assert np.array_equal(m, exp_m)
    assert np.array_equal(fin, exp_f)
