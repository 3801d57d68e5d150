"""GPU parity of the witness data path: the union dense-stack gather (compact_witness,
crates/flock-core/src/union.rs:418-490) against the oracle, bit-exact."""
from __future__ import annotations

import random
import sys

import numpy as np
import pytest

import oracle_lib as O
from flock_amd import lib as F
from test_union_glue import Union, _random_types

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("seed", range(6))
def test_union_compact(ctx, seed):
    rng = random.Random(100 + seed)
    types = _random_types(rng, rng.randrange(1, 5))
    nu = rng.randrange(0, 8)
    counts = [rng.choice([0, 1, (1 << nu), rng.randrange(0, (1 << nu) + 1)]) for _ in types]
    u = Union(types, nu)
    reg = F.registry_new(types, nu)
    sz = F.union_sizes(reg, counts)
    say(f"case: union compact seed={seed} types={types} nu={nu} counts={counts} "
        f"packed={sz['packed_len']} committed={sz['committed_words']}")
    z = O.splitmix(0x900 + seed, sz["packed_len"])
    exp = u.compact(counts, z)
    q = ctx.alloc(16 * sz["committed_words"])
    ctx.union_compact(reg, counts, ctx.upload(z), q)
    assert np.array_equal(q.download_f128(sz["committed_words"]), exp)


def test_union_compact_sha256_full_size(ctx):
    # the SHA-256 shape at 1024 instances: 199 used columns of 256, committed 2^18
    types = [(15, 25_470, False)]
    nu = 10
    u = Union(types, nu)
    reg = F.registry_new(types, nu)
    counts = [1000]
    sz = F.union_sizes(reg, counts)
    say(f"case: union compact sha256 counts={counts} committed={sz['committed_words']}")
    z = O.splitmix(0x990, sz["packed_len"])
    exp = u.compact(counts, z)
    q = ctx.alloc(16 * sz["committed_words"])
    ctx.union_compact(reg, counts, ctx.upload(z), q)
    assert np.array_equal(q.download_f128(sz["committed_words"]), exp)
