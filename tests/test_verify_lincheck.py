"""flock_lincheck_verify (lincheck::union::verify_union_with_grinding, lincheck/union.rs:662-960,
with MatrixAssertion::check / check_reported, :538-628) on the CPU against proofs the oracle
prover writes.

- An honest proof verifies in both forms (the O(nnz) check against the circuit, and the deferred
  check against the reported matrix evals). The verifier's claim equals the prover's, and its
  transcript ends in the prover's state.
- A change to any part of the proof, to the input claim or to the circuit is rejected. The
  reported matrix evals are read by the deferred form only, as in the reference.
- A witness whose constant column is not 1 on every declared row is rejected (the const-wire
  pin, union.rs:815-846); union.rs:1000-1400 test the same properties.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as O
import puremath as M
from flock_amd import lib as F
from test_oracle_lincheck import csc_fold, lincheck_prove, quirky_eq, random_csc, stripe

DOM = b"flock-lc-verify-test"


def _instance(k_log, nu, pin, ub, nr, seed, pin_rows=None):
    """A circuit, a stripe whose pin column is 1 on the declared rows (or on `pin_rows`), and
    a point x_ab."""
    k = 1 << k_log
    L = k_log - 6
    ub = k if ub is None else ub
    nr = (1 << nu) if nr is None else nr
    csc_a = random_csc(k, 3, 40 + seed)
    csc_b = random_csc(k, 2, 50 + seed)
    z = stripe(k_log, nu, 60 + seed, ub, nr, pin, pin_rows)
    pts = M.splitmix(0xA00 + seed, 1 + L + nu)
    return dict(k_log=k_log, nu=nu, pin=pin, ub=ub, nr=nr, csc_a=csc_a, csc_b=csc_b, z=z,
                z_skip=pts[0], x_rest=pts[1:1 + L], x_outer=pts[1 + L:])


def _prove(inst, grinding):
    """The oracle's proof and the input claim (v_a, v_b) = (<A_0^T eq, z_vec>, <B_0^T eq, z_vec>)
    it answers: the MLE of A_0 z and B_0 z at x_ab."""
    out = lincheck_prove(DOM, inst["k_log"], inst["csc_a"], inst["csc_b"], inst["pin"], inst["z"],
                         inst["nu"], inst["ub"], inst["nr"], inst["z_skip"], inst["x_rest"],
                         inst["x_outer"], grinding)
    sa, sb = csc_fold(*inst["csc_a"], *inst["csc_b"], quirky_eq(inst["z_skip"], inst["x_rest"]))
    zv = O.arr_to_ints(out["z_vec"])
    va = vb = 0
    for a, b, v in zip(O.arr_to_ints(sa), O.arr_to_ints(sb), zv):
        va ^= M.mul(a, v)
        vb ^= M.mul(b, v)
    return out, va, vb


def _verify(inst, proof, va, vb, grinding, full, circuit=None):
    fs = F.FsChallenger(DOM)
    circ = None
    if full:
        circ = circuit if circuit is not None else (*inst["csc_a"], *inst["csc_b"])
    claim = F.lincheck_verify(fs, inst["k_log"], inst["pin"], inst["nu"], inst["nr"],
                              inst["z_skip"], inst["x_rest"], inst["x_outer"], va, vb, grinding,
                              proof, circ)
    return claim, fs.digest()


def _rejected(*a, **kw):
    try:
        _verify(*a, **kw)
    except F.FlockHipError as e:
        assert e.code == F.E_REJECT, e
        return True
    return False


CASES = [(8, 4, -1, False, None, None), (9, 5, 7, True, 300, 21), (10, 6, 100, True, None, None),
         (11, 7, 1500, False, 1600, 100), (6, 3, 0, True, None, None)]


@pytest.mark.parametrize("k_log,nu,pin,grinding,ub,nr", CASES)
@pytest.mark.parametrize("full", [True, False])
def test_honest_proof_verifies(k_log, nu, pin, grinding, ub, nr, full):
    inst = _instance(k_log, nu, pin, ub, nr, k_log)
    proof, va, vb = _prove(inst, grinding)
    claim, dig = _verify(inst, proof, va, vb, grinding, full)
    assert np.array_equal(claim, proof["claim"])
    assert dig == proof["digest"]


def test_eq_prefix_sum_target():
    # the const-pin target uses sum_{row < n} eq(x_outer, row) (eq_prefix_sum, union.rs:107-130):
    # at n_rows = 0 the honest pin column is all zero, at n_rows = 2^nu the sum is one
    for nr in (0, 1, 5, 31, 32):
        inst = _instance(7, 5, 3, None, nr, 7)
        proof, va, vb = _prove(inst, True)
        claim, _ = _verify(inst, proof, va, vb, True, True)
        assert np.array_equal(claim, proof["claim"])


def test_every_part_of_the_proof_is_bound():
    grinding = True
    inst = _instance(10, 6, 17, 900, 50, 3)
    proof, va, vb = _prove(inst, grinding)
    L = inst["k_log"] - 6
    for full in (True, False):
        for part, row in [("rounds", 0), ("rounds", 2 * L - 1), ("z_partial", 0),
                          ("z_partial", 63)]:
            bad = {k: v.copy() for k, v in proof.items() if k != "digest"}
            bad[part][row, 1] ^= np.uint64(1 << 40)
            assert _rejected(inst, bad, va, vb, grinding, full), (part, row, full)
        for k in (0, 1, 3, L + 2):
            bad = {k2: v.copy() for k2, v in proof.items() if k2 != "digest"}
            bad["nonces"][k] ^= np.uint64(0x3c3c)
            assert _rejected(inst, bad, va, vb, grinding, full), ("nonce", k, full)
        assert _rejected(inst, proof, va ^ 1, vb, grinding, full)
        assert _rejected(inst, proof, va, vb ^ (1 << 100), grinding, full)
    # the reported matrix evals: the deferred form checks them; the full check recomputes the
    # combination from the circuit and never reads them (verify_union, union.rs:662-680)
    for i in (0, 1):
        bad = {k: v.copy() for k, v in proof.items() if k != "digest"}
        bad["matrix_evals"][i, 0] ^= np.uint64(4)
        assert _rejected(inst, bad, va, vb, grinding, False)
        _verify(inst, bad, va, vb, grinding, True)


def test_other_circuit_is_rejected():
    inst = _instance(9, 5, -1, None, None, 4)
    proof, va, vb = _prove(inst, False)
    other = random_csc(1 << 9, 3, 999)
    assert _rejected(inst, proof, va, vb, False, True,
                     circuit=(*other, *inst["csc_b"]))
    assert _rejected(inst, proof, va, vb, False, True,
                     circuit=(*inst["csc_a"], *inst["csc_a"]))


def test_dishonest_pin_column_is_rejected():
    # the constant column is 1 on one row fewer than declared: sum_rows eq(x_outer, row) z[row,
    # pin] != eq_prefix_sum, so the rounds' claim starts off the prover's
    inst = _instance(9, 5, 12, None, 20, 5, pin_rows=19)
    proof, va, vb = _prove(inst, True)
    assert _rejected(inst, proof, va, vb, True, True)
    assert _rejected(inst, proof, va, vb, True, False)


def test_bad_circuit_arguments():
    inst = _instance(8, 4, -1, None, None, 6)
    proof, va, vb = _prove(inst, False)
    a_ptr, a_rows, b_ptr, b_rows = (*inst["csc_a"], *inst["csc_b"])
    bad_rows = a_rows.copy()
    bad_rows[0] = 1 << 8
    with pytest.raises(F.FlockHipError) as e:
        _verify(inst, proof, va, vb, False, True, circuit=(a_ptr, bad_rows, b_ptr, b_rows))
    assert e.value.code == F.E_INVALID
    with pytest.raises(F.FlockHipError) as e:
        F.lincheck_verify(F.FsChallenger(DOM), 8, 300, 4, 16, 0, [0] * 2, [0] * 4, 0, 0, False,
                          proof, None)
    assert e.value.code == F.E_INVALID
