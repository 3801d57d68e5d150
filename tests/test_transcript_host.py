"""The host FsChallenger's edge behaviour: out-of-range proof-of-work bits are refused, and the
portable SHA-256 block function (FLOCK_HIP_PORTABLE_SHA256=1) gives the same transcript as
the SHA-extension one, checked against the oracle's challenger. Host only."""
from __future__ import annotations

import json
import os
import subprocess
import sys

import oracle_lib as O
from flock_amd import lib as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRIPT = """
import json, sys
sys.path.insert(0, {root!r})
from flock_amd import lib as F
fs = F.FsChallenger(b"flock-transcript-test")
out = []
for i in range(40):
    fs.observe_label(b"step-%d" % i)
    fs.observe_f128((0x9E3779B97F4A7C15 * (i + 1)) % (1 << 128))
    fs.observe_bytes(bytes(range(i % 70)))
    out.append(fs.sample_f128())
    out.append(fs.grind_pow(i % 9))
out.append(fs.digest().hex())
print(json.dumps(out))
"""


def _run(env_extra):
    env = dict(os.environ)
    env.update(env_extra)
    res = subprocess.run([sys.executable, "-c", SCRIPT.format(root=ROOT)], capture_output=True,
                         text=True, env=env, timeout=120)
    assert res.returncode == 0, res.stderr
    return json.loads(res.stdout)


def _oracle():
    fs = O.Fs(b"flock-transcript-test")
    out = []
    for i in range(40):
        fs.observe_label(b"step-%d" % i)
        fs.observe_f128((0x9E3779B97F4A7C15 * (i + 1)) % (1 << 128))
        fs.observe_bytes(bytes(range(i % 70)))
        out.append(fs.sample_f128())
        out.append(fs.grind_pow(i % 9))
    return out


def test_portable_sha256_block_gives_the_same_transcript():
    native = _run({"FLOCK_HIP_PORTABLE_SHA256": "0"})
    portable = _run({"FLOCK_HIP_PORTABLE_SHA256": "1"})
    assert native == portable
    assert native[:-1] == _oracle()


def test_grind_pow_refuses_more_than_64_bits():
    fs = F.FsChallenger(b"x")
    before = fs.digest()
    assert fs.grind_pow(65) == (1 << 64) - 1
    assert b"64 bits" in F.load().flock_hip_last_error()
    assert fs.digest() == before  # nothing absorbed
    assert fs.grind_pow(0) == 0
