"""Write the golden fixtures under tests/golden/ from pure mathematics.

Nothing here runs or imports the reference; every expected value comes from
tests/golden/puremath.py (integer carry-less products with long-division
reduction, the novel-basis polynomial evaluated by its definition, hashlib
SHA-256). Run:  python tests/golden/gen_golden.py
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import puremath as M  # noqa: E402


def hx(v: int) -> str:
    return f"{v:032x}"


def write(name, obj):
    with open(os.path.join(HERE, name), "w") as fh:
        json.dump(obj, fh, indent=1)
        fh.write("\n")


def gen_f128():
    vals = M.splitmix(0xF128, 128)
    pairs = [(vals[2 * i], vals[2 * i + 1]) for i in range(64)]
    # edge operands: 0, 1, x, x^63, x^64, x^127, all-ones
    edges = [0, 1, 2, 1 << 63, 1 << 64, 1 << 127, (1 << 128) - 1]
    for a in edges:
        for b in edges:
            pairs.append((a, b))
    write("f128_mul.json", {
        "poly": "x^128+x^7+x^2+x+1",
        "cases": [[hx(a), hx(b), hx(M.mul(a, b))] for a, b in pairs],
        "inverses": [[hx(a), hx(M.inv(a))] for a in vals[:8]],
    })


def gen_ntt():
    cases = []
    seed = 0xA110
    for L in range(1, 7):
        c = M.splitmix(seed + L, 1 << L)
        cases.append({"L": L, "coeffs": [hx(v) for v in c],
                      "evals": [hx(v) for v in M.ntt_by_definition(c, L)]})
    write("ntt_definition.json", {"cases": cases})


def commit_case(log_batch, log_dim, rate, num_lanes, seed, zero_lanes=()):
    """Lane-major commit by definition: lane l's codeword is the novel-basis
    polynomial with coefficients q[l*D .. (l+1)*D) (zero-padded to 2^k)
    evaluated at every subspace point; leaf P = the t lane values at P."""
    m = 7 + log_batch + log_dim
    D = 1 << log_dim
    lanes = 1 << log_batch
    t = num_lanes or lanes
    k = log_dim + rate
    q = M.splitmix(seed, lanes * D)
    for l in range(t, lanes):
        for p in range(D):
            q[l * D + p] = 0
    for l in zero_lanes:
        for p in range(D):
            q[l * D + p] = 0
    cw = [0] * ((1 << k) * t)
    for l in range(t):
        coeffs = q[l * D:(l + 1) * D] + [0] * ((1 << k) - D)
        ev = M.ntt_by_definition(coeffs, k)
        for P in range(1 << k):
            cw[P * t + l] = ev[P]
    data = b"".join(M.f128_bytes(v) for v in cw)
    tree = M.merkle_tree(data, 1 << k)
    return {"m": m, "log_inv_rate": rate, "log_batch_size": log_batch, "num_lanes": num_lanes,
            "q": [hx(v) for v in q], "codeword": [hx(v) for v in cw],
            "tree": [n.hex() for n in tree], "root": tree[-1].hex()}


def gen_commit():
    cases = [
        commit_case(2, 3, 2, 0, 0xC0FFEE),        # 4 lanes, rate 1/4
        commit_case(2, 3, 1, 3, 0xC0FFEF),        # integer lanes t=3 of 4, rate 1/2
        commit_case(1, 2, 3, 0, 0xC0FFF0),        # 2 lanes, rate 1/8
        commit_case(2, 2, 2, 4, 0xC0FFF1, zero_lanes=(2, 3)),  # dead-lane skip
    ]
    write("commit_definition.json", {"cases": cases})


def gen_fs():
    domain = b"flock-golden-fs-v0"
    fs = M.FsChallenger(domain)
    ops = []
    fs.observe_label(b"flock-zerocheck-v0")
    ops.append(["label", b"flock-zerocheck-v0".hex()])
    vals = M.splitmix(0xF5, 6)
    fs.observe_f128(vals[0])
    ops.append(["f128", hx(vals[0])])
    ops.append(["sample", hx(fs.sample_f128())])
    fs.observe_f128_slice(vals[1:4])
    ops.append(["slice", [hx(v) for v in vals[1:4]]])
    ops.append(["sample_vec", [hx(v) for v in fs.sample_f128_vec(3)]])
    fs.observe_bytes(b"\x01\x02\x03")
    ops.append(["bytes", "010203"])
    ops.append(["grind", 6, fs.grind_pow(6)])
    ops.append(["sample", hx(fs.sample_f128())])
    ops.append(["grind", 0, fs.grind_pow(0)])
    ops.append(["sample_vec", [hx(v) for v in fs.sample_f128_vec(5)]])
    write("fs_transcript.json", {"domain": domain.hex(), "ops": ops})


def gen_sumcheck():
    n = 16
    a = M.splitmix(0x5C1, n)
    b = M.splitmix(0x5C2, n)
    r = M.splitmix(0x5C3, 1)[0]
    # round_msg_lsb by definition
    u0 = 0
    u2 = 0
    for j in range(n // 2):
        u0 ^= M.mul(a[2 * j], b[2 * j])
        u2 ^= M.mul(a[2 * j] ^ a[2 * j + 1], b[2 * j] ^ b[2 * j + 1])
    # fold: the multilinear extension at LSB = r: (1+r) v0 + r v1
    nf = [M.mul(a[2 * j], r ^ 1) ^ M.mul(a[2 * j + 1], r) for j in range(n // 2)]
    nb = [M.mul(b[2 * j], r ^ 1) ^ M.mul(b[2 * j + 1], r) for j in range(n // 2)]
    v0 = 0
    v2 = 0
    for j in range(n // 4):
        v0 ^= M.mul(nf[2 * j], nb[2 * j])
        v2 ^= M.mul(nf[2 * j] ^ nf[2 * j + 1], nb[2 * j] ^ nb[2 * j + 1])
    # eq-weighted round pair: G(1) = sum eq(r[1..], x) a1 b1, G(inf) = sum eq (a0+a1)(b0+b1)
    rv = M.splitmix(0x5C4, 4)
    g1 = 0
    gi = 0
    for x in range(n // 2):
        e = M.eq_point(rv[1:], x)
        g1 ^= M.mul(e, M.mul(a[2 * x + 1], b[2 * x + 1]))
        gi ^= M.mul(e, M.mul(a[2 * x] ^ a[2 * x + 1], b[2 * x] ^ b[2 * x + 1]))
    write("sumcheck_definition.json", {
        "a": [hx(v) for v in a], "b": [hx(v) for v in b], "r": hx(r),
        "round_msg": [hx(u0), hx(u2)],
        "fold_a": [hx(v) for v in nf], "fold_b": [hx(v) for v in nb],
        "fold_msg": [hx(v0), hx(v2)],
        "eq_r": [hx(v) for v in rv], "round_pair": [hx(M.mul(rv[0], g1)), hx(gi)],
        "eq_table": [hx(M.eq_point(rv, x)) for x in range(16)],
    })


if __name__ == "__main__":
    gen_f128()
    gen_ntt()
    gen_commit()
    gen_fs()
    gen_sumcheck()
    print("fixtures written to", HERE)
