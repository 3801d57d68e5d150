"""CPU-side checks of the C ABI: the library loads, exports every declared entry point,
and its host-side pieces (transcript, size arithmetic, argument validation) agree with
the oracle. No kernel is launched here."""
from __future__ import annotations

import ctypes
import json
import os
import random

import numpy as np
import pytest

import oracle_lib as O
from flock_amd import lib as F

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_library_exports_every_header_symbol():
    L = F.load()
    names = F.header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n


def test_fs_matches_golden_and_oracle():
    with open(os.path.join(GOLD, "fs_transcript.json")) as fh:
        g = json.load(fh)
    fs = F.FsChallenger(bytes.fromhex(g["domain"]))
    for op in g["ops"]:
        k = op[0]
        if k == "label":
            fs.observe_label(bytes.fromhex(op[1]))
        elif k == "f128":
            fs.observe_f128(int(op[1], 16))
        elif k == "slice":
            fs.observe_f128_slice([int(v, 16) for v in op[1]])
        elif k == "bytes":
            fs.observe_bytes(bytes.fromhex(op[1]))
        elif k == "sample":
            assert fs.sample_f128() == int(op[1], 16)
        elif k == "sample_vec":
            assert fs.sample_f128_vec(len(op[1])) == [int(v, 16) for v in op[1]]
        elif k == "grind":
            assert fs.grind_pow(op[1]) == op[2]
    # random op streams against the oracle transcript
    rng = random.Random(21)
    a, b = F.FsChallenger(b"dom"), O.Fs(b"dom")
    for _ in range(200):
        k = rng.randrange(6)
        if k == 0:
            v = rng.getrandbits(128)
            a.observe_f128(v)
            b.observe_f128(v)
        elif k == 1:
            bs = bytes(rng.getrandbits(8) for _ in range(rng.randrange(40)))
            a.observe_bytes(bs)
            b.observe_bytes(bs)
        elif k == 2:
            assert a.sample_f128() == b.sample_f128()
        elif k == 3:
            n = rng.randrange(1, 9)
            assert a.sample_f128_vec(n) == b.sample_f128_vec(n)
        elif k == 4:
            lb = b"label-%d" % rng.randrange(100)
            a.observe_label(lb)
            b.observe_label(lb)
        else:
            bits = rng.randrange(0, 7)
            assert a.grind_pow(bits) == b.grind_pow(bits)


def test_pcs_sizes_match_oracle():
    for m, rate, lb, nl in [(25, 2, 8, 0), (25, 1, 6, 50), (20, 1, 5, 0), (34, 1, 6, 42),
                            (12, 3, 2, 3)]:
        p = F.PcsParams(m, rate, lb, nl, 0)
        assert F.pcs_sizes(p) == O.pcs_sizes(m, rate, lb, nl)


def test_pcs_sizes_errors():
    try:
        F.pcs_sizes(F.PcsParams(10, 1, 5, 0, 0))
        raise AssertionError("expected an error")
    except F.FlockHipError as e:
        assert e.code == F.E_M_TOO_SMALL
    for bad in (F.PcsParams(20, 0, 5, 0, 0), F.PcsParams(20, 1, 5, 33, 0),
                F.PcsParams(20, 1, 5, 0, 2)):
        try:
            F.pcs_sizes(bad)
            raise AssertionError("expected an error")
        except F.FlockHipError as e:
            assert e.code == F.E_INVALID


def test_null_context_is_an_error_not_a_crash():
    L = F.load()
    out = (F.F128 * 2)()
    rc = L.flock_hip_sumcheck_round_msg(None, None, None, 4, out)
    assert rc == F.E_INVALID
    assert b"context" in L.flock_hip_last_error()
    assert L.flock_hip_device_count() >= 0


def test_ctx_create_without_gpu_reports_nodevice():
    if F.device_count() > 0:
        return  # a GPU is present; covered by the gpu tests
    h = ctypes.c_void_p()
    assert F.load().flock_hip_ctx_create(0, ctypes.byref(h)) == F.E_NODEVICE
    assert h.value is None


def test_f128_struct_layout():
    assert ctypes.sizeof(F.F128) == 16
    assert ctypes.sizeof(F.PcsParams) == 20
    v = F.f128((1 << 127) | 5)
    assert F.to_int(v) == (1 << 127) | 5
    assert np.frombuffer(bytes(v), dtype=np.uint64).tolist() == [5, 1 << 63]


def test_tunables_set_get_and_reject():
    """flock_hip_set_tunable / _get_tunable: defaults, a value in range, and rejections (no GPU
    needed: the registry is host state)."""
    from flock_amd import lib as F

    defaults = {k: F.get_tunable(k) for k in F.TUNABLES}
    assert defaults == {"ntt_cols_per_thread": 2, "ntt_deep_cols": 256, "zc_split_min_vars": 16,
                        "zc_la_max_log": 17, "sc_la_max_log": 18, "open_split_fold_log": 16}
    try:
        F.set_tunable("sc_la_max_log", 14)
        assert F.get_tunable("sc_la_max_log") == 14
        for name, bad in (("sc_la_max_log", 99), ("ntt_cols_per_thread", 0), ("no_such", 1)):
            with pytest.raises(F.FlockHipError):
                F.set_tunable(name, bad)
    finally:
        for k, v in defaults.items():
            F.set_tunable(k, v)
