"""Write prove_core_vectors.json: the oracle's prove core (witness, dense stack, commit,
bind_statement of the registry digest and the commitment cap, zerocheck, lincheck on one
transcript) at the config-2 batch size of 1024
instances, for SHA-256, BLAKE3 and Keccak-f[1600] inputs drawn from a seed. The GPU test
regenerates the same inputs from the seed and compares the commitment root and the transcript
digest. Run from the repository root: python tests/golden/make_prove_core_vectors.py"""
from __future__ import annotations

import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))  # the repository root
sys.path.insert(0, os.path.dirname(HERE))                   # tests/
sys.path.insert(0, HERE)

from flock_amd import lib as F  # noqa: E402  (constants only; no GPU is touched)
import test_gpu_prove as T  # noqa: E402
import test_oracle_blake3_r1cs as B3  # noqa: E402
import test_oracle_sha2_r1cs as S2  # noqa: E402

CASES = [(F.HASH_SHA256, 10, 1024, 0x5EED0), (F.HASH_BLAKE3, 10, 1024, 0x5EED1),
         (F.HASH_KECCAK, 10, 1024, 0x5EED2)]


def main():
    circ = {}
    for kind, mod in ((F.HASH_SHA256, S2), (F.HASH_BLAKE3, B3)):
        _, _, a, b = mod.matrices()
        circ[kind] = (None, a, b)
    out = []
    for kind, n_log, n_decl, seed in CASES:
        t = time.time()
        root, dig = T.oracle_prove_core(kind, n_log, T.inputs_for(kind, n_decl, seed), circ)
        out.append(dict(kind=kind, n_log=n_log, n_decl=n_decl, seed=seed,
                        digest=T.digest_for(kind, n_log).hex(), root=root.hex(),
                        transcript=dig.hex()))
        print(kind, n_log, f"{time.time() - t:.1f} s", flush=True)
    with open(os.path.join(HERE, "prove_core_vectors.json"), "w") as fh:
        json.dump(dict(domain=T.DOMAIN.decode(), cases=out), fh, indent=1)


if __name__ == "__main__":
    main()
