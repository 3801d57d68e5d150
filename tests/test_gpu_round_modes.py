"""The sumcheck round loops in both round modes (flock_hip_ctx_set_round_mode):
- host (the default): one stream synchronisation, host transcript step and launch per round;
- device: a device-resident transcript inside the round kernels, replayed and checked on the
  host afterwards.
Every message, nonce, challenge and final value agrees between the modes, and so does the
transcript digest after the proof, with grinding on and off. The oracle parity tests in the
other files run in the default mode; these tests tie the device mode to it."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle_lib as O
from flock_amd import lib as F
from test_oracle_lincheck import random_csc, stripe
from test_oracle_zerocheck import _witness

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.set_round_mode(F.ROUNDS_HOST)
    c.close()


# device twice: the second loop starts from the state the first one left in the context
MODES = (F.ROUNDS_HOST, F.ROUNDS_DEVICE, F.ROUNDS_DEVICE)


def all_modes(ctx, run):
    """run() in each round mode, host first."""
    out = []
    for mode in MODES:
        ctx.set_round_mode(mode)
        out.append(run())
    ctx.set_round_mode(F.ROUNDS_HOST)
    return out


@pytest.mark.parametrize("m,grinding", [(13, True), (14, False), (17, True), (20, True),
                                        (22, False)])
def test_zerocheck_modes_agree(ctx, m, grinding):
    say(f"case: zerocheck in every round mode m={m} grinding={grinding}")
    a, b, c = _witness(m, 900 + m, satisfying=True)
    da, db, dc = ctx.upload(a), ctx.upload(b), ctx.upload(c)

    def run():
        fs = F.FsChallenger(b"flock-device-transcript-zc")
        got = ctx.zerocheck_prove(fs, da, db, dc, m, grinding, capture=True)
        return got, fs.digest()

    (host, hd), *others = all_modes(ctx, run)
    for mode, (got, dig) in zip(MODES[1:], others):
        for name, x, y in zip(["round1", "multilinear", "finals", "nonces", "claim", "shat"],
                              host, got):
            assert np.array_equal(x, y), (mode, name)
        assert hd == dig, mode


@pytest.mark.parametrize("k_log,nu,grinding", [(7, 3, True), (10, 4, False), (12, 5, True)])
def test_lincheck_modes_agree(ctx, k_log, nu, grinding):
    say(f"case: lincheck in every round mode k_log={k_log} nu={nu} grinding={grinding}")
    k = 1 << k_log
    L = k_log - 6
    csc_a = random_csc(k, 4, 1400 + k_log)
    csc_b = random_csc(k, 3, 1500 + k_log)
    z = stripe(k_log, nu, 1600 + k_log, k, 1 << nu)
    pts = O.arr_to_ints(O.splitmix(0x1800 + k_log, 1 + L + nu))
    circuit, keep = ctx.csc(k_log, *csc_a, *csc_b, const_pin=3)
    dz = ctx.upload(z)

    def run():
        fs = F.FsChallenger(b"flock-device-transcript-lc")
        got = ctx.lincheck_prove(fs, circuit, dz, nu, k, 1 << nu, pts[0], pts[1:1 + L],
                                 pts[1 + L:], grinding)
        return got, fs.digest()

    (host, hd), *others = all_modes(ctx, run)
    for mode, (got, dig) in zip(MODES[1:], others):
        for name in ("rounds", "z_partial", "matrix_evals", "claim", "nonces"):
            assert np.array_equal(host[name], got[name]), (mode, name)
        assert hd == dig, mode


@pytest.mark.parametrize("log_n", [1, 2, 5, 14])
def test_prove_sumcheck_modes_agree(ctx, log_n):
    say(f"case: product sumcheck + zerocheck in every round mode log_n={log_n}")
    n = 1 << log_n
    a, b = O.splitmix(0x51 + log_n, n), O.splitmix(0x52 + log_n, n)
    da, db = ctx.upload(a), ctx.upload(b)

    def run():
        fs = F.FsChallenger(b"flock-device-transcript-sc")
        msgs, finals = ctx.prove_sumcheck(fs, da, db, log_n)
        return msgs, finals, fs.digest()

    (hm, hf, hd), *others = all_modes(ctx, run)
    for mode, (m, f, d) in zip(MODES[1:], others):
        assert np.array_equal(hm, m) and np.array_equal(hf, f) and hd == d, mode


def test_prove_core_modes_agree():
    say("case: prove core in every round mode (Keccak-f[1600] x 2^4)")
    from test_oracle_keccak_r1cs import random_states
    inputs = np.ascontiguousarray(random_states(16, 77))
    reqs, keep = F.make_requests([(F.HASH_KECCAK, 4, inputs, None, 0.0)],
                                 digest=bytes(range(32)), domain=b"flock-readme-bench-v0")
    pr = F.Prover(0)
    try:
        out = []
        for mode in MODES:
            pr.set_round_mode(mode)
            r = F.prove_core(pr, reqs[0])
            out.append((bytes(r.root), bytes(r.transcript)))
    finally:
        pr.close()
    assert all(o == out[0] for o in out)
