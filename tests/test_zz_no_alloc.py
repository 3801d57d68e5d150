"""flock_hip_sumcheck_fold_msg_static (no allocation) against flock_hip_sumcheck_fold_msg.

Two calls in a row through the C ABI, the second on a smaller table, each compared word for
word with the context-based entry point on the same seeded inputs (folded tables and the
next message), plus a check that the second call leaves the device's free memory unchanged.
"""
from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest

from flock_amd import lib as F

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def free_bytes() -> int:
    h = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert h.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_static_fold_msg_equals_the_context_entry_point():
    ctx = F.Context(0)
    L = F.load()
    rng = np.random.default_rng(0x57A7)
    n_max = 1 << 12
    bufs = {k: ctx.alloc(16 * n_max) for k in ("f", "b", "nf", "nb", "nf2", "nb2")}
    msg = ctx.alloc(32)
    try:
        for call, log_n in enumerate((12, 9)):
            n = 1 << log_n
            say(f"case: fold_msg_static call {call} n=2^{log_n}")
            f = rng.integers(0, 2**63, (n, 2), dtype=np.uint64) * 2 + 1
            b = rng.integers(0, 2**63, (n, 2), dtype=np.uint64) * 2
            r = int(rng.integers(1, 2**63)) << 64 | int(rng.integers(1, 2**63))
            bufs["f"].upload(f)
            bufs["b"].upload(b)
            want = ctx.fold_and_msg_lsb(bufs["f"], bufs["b"], n, r, bufs["nf"], bufs["nb"])
            ctx.sync()
            before = free_bytes()
            rc = L.flock_hip_sumcheck_fold_msg_static(
                ctypes.c_void_p(ctx.stream), bufs["f"].ptr, bufs["b"].ptr, n, F.f128(r),
                bufs["nf2"].ptr, bufs["nb2"].ptr, msg.ptr)
            assert rc == 0, L.flock_hip_last_error()
            ctx.sync()
            if call == 1:
                assert free_bytes() == before, "the static entry point changed free device memory"
            got = msg.download_f128(2)
            assert (F.to_int(F.F128(*got[0])), F.to_int(F.F128(*got[1]))) == want
            for a, c in (("nf", "nf2"), ("nb", "nb2")):
                assert np.array_equal(bufs[a].download_f128(n // 2), bufs[c].download_f128(n // 2))
        # out-of-range sizes are refused before any launch
        rc = L.flock_hip_sumcheck_fold_msg_static(
            ctypes.c_void_p(ctx.stream), bufs["f"].ptr, bufs["b"].ptr, 3, F.f128(1),
            bufs["nf2"].ptr, bufs["nb2"].ptr, msg.ptr)
        assert rc == F.E_INVALID
    finally:
        for x in (*bufs.values(), msg):
            x.free()
        ctx.close()
