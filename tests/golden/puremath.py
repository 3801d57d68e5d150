"""Pure-math restatements used to write and check the golden fixtures.

Independent of the C oracle: F_{2^128} products come from integer carry-less
multiplication followed by textbook polynomial long division by
x^128 + x^7 + x^2 + x + 1; the additive NTT is checked by evaluating the
novel-basis polynomial X_j(z) = prod_{i in bits(j)} W_i(z)/W_i(beta_i) directly
at the subspace points (definition: crates/flock-core/src/ntt/additive_ntt_f128.rs:20-41);
SHA-256 is hashlib (FIPS 180-4).
"""
from __future__ import annotations

import hashlib

POLY = (1 << 128) | 0x87
MASK64 = (1 << 64) - 1


def clmul(a: int, b: int) -> int:
    r = 0
    while b:
        low = b & -b
        r ^= a * low  # a << bit, no carries: a * 2^k
        b ^= low
    return r


def reduce(x: int) -> int:
    """Remainder of the polynomial x modulo POLY (long division)."""
    while x.bit_length() > 128:
        x ^= POLY << (x.bit_length() - 129)
    return x


def mul(a: int, b: int) -> int:
    return reduce(clmul(a, b))


def power(a: int, e: int) -> int:
    r = 1
    while e:
        if e & 1:
            r = mul(r, a)
        a = mul(a, a)
        e >>= 1
    return r


def inv(a: int) -> int:
    return power(a, (1 << 128) - 2)


def splitmix(seed: int, n: int) -> list[int]:
    s = seed
    out = []
    for _ in range(n):
        w = []
        for _ in range(2):
            s = (s + 0x9E3779B97F4A7C15) & MASK64
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
            w.append(z ^ (z >> 31))
        out.append(w[0] | (w[1] << 64))
    return out


# ---------------------------------------------------------------------------
# Novel-basis polynomial evaluation (LCH basis over beta_i = x^i)
# ---------------------------------------------------------------------------

def subspace_W(i: int, z: int, basis: list[int]) -> int:
    """W_0(z) = z; W_i(z) = W_{i-1}(z) * (W_{i-1}(z) + W_{i-1}(beta_{i-1}))."""
    if i == 0:
        return z
    w = subspace_W(i - 1, z, basis)
    return mul(w, w ^ subspace_W(i - 1, basis[i - 1], basis))


def novel_basis_eval(coeffs: list[int], z: int, L: int) -> int:
    basis = [1 << i for i in range(L)]
    norm = [inv(subspace_W(i, basis[i], basis)) for i in range(L)]
    what = [mul(subspace_W(i, z, basis), norm[i]) for i in range(L)]
    acc = 0
    for j, c in enumerate(coeffs):
        if c == 0:
            continue
        x = 1
        for i in range(L):
            if (j >> i) & 1:
                x = mul(x, what[i])
        acc ^= mul(c, x)
    return acc


def ntt_by_definition(coeffs: list[int], L: int) -> list[int]:
    """Evaluations at omega_k = sum_i bit_i(k) beta_i = the integer k."""
    return [novel_basis_eval(coeffs, k, L) for k in range(1 << L)]


# ---------------------------------------------------------------------------
# Sumcheck brute force
# ---------------------------------------------------------------------------

def eq_point(r: list[int], x: int) -> int:
    acc = 1
    for i, ri in enumerate(r):
        acc = mul(acc, ri if (x >> i) & 1 else ri ^ 1)
    return acc


def sha256(b: bytes) -> bytes:
    return hashlib.sha256(b).digest()


def f128_bytes(v: int) -> bytes:
    return v.to_bytes(16, "little")


def merkle_tree(data: bytes, n_leaves: int) -> list[bytes]:
    leaf = len(data) // n_leaves
    level = [sha256(data[i * leaf:(i + 1) * leaf]) for i in range(n_leaves)]
    tree = list(level)
    while len(level) > 1:
        level = [sha256(level[2 * i] + level[2 * i + 1]) for i in range(len(level) // 2)]
        tree += level
    return tree


class FsChallenger:
    """SHA-256 Fiat-Shamir transcript from the framing rules of
    crates/flock-core/src/challenger.rs:690-882 (16-byte headers, zero-padded
    payloads, SHA256(state || ctr_le) squeezes)."""

    def __init__(self, domain: bytes):
        self.h = hashlib.sha256()
        self._hdr(1, 0, len(domain))
        self._pad(domain)

    def _hdr(self, op, kind, n):
        self.h.update(bytes([op, kind]) + bytes(6) + n.to_bytes(8, "little"))

    def _pad(self, b):
        self.h.update(b)
        if len(b) % 16:
            self.h.update(bytes(16 - len(b) % 16))

    def observe_label(self, b):
        self._hdr(2, 0, len(b))
        self._pad(b)

    def observe_f128(self, v):
        self._hdr(3, 1, 1)
        self.h.update(f128_bytes(v))

    def observe_f128_slice(self, vs):
        self._hdr(3, 2, len(vs))
        for v in vs:
            self.h.update(f128_bytes(v))

    def observe_bytes(self, b):
        self._hdr(5, 0, len(b))
        self._pad(b)

    def _squeeze(self, n):
        out = b""
        ctr = 0
        while len(out) < n:
            c = self.h.copy()
            c.update(ctr.to_bytes(8, "little"))
            out += c.digest()
            ctr += 1
        return out[:n]

    def sample_f128(self):
        self._hdr(4, 1, 1)
        return int.from_bytes(self._squeeze(16), "little")

    def sample_f128_vec(self, n):
        self._hdr(4, 2, n)
        raw = self._squeeze(16 * n)
        return [int.from_bytes(raw[16 * i:16 * i + 16], "little") for i in range(n)]

    def grind_pow(self, bits):
        digest = self.h.copy().digest()
        nonce = 0
        if bits:
            while True:
                h = hashlib.sha256(digest + nonce.to_bytes(8, "little")).digest()
                v = int.from_bytes(h, "big")
                if v >> (256 - bits) == 0:
                    break
                nonce += 1
        self.observe_bytes(nonce.to_bytes(8, "little"))
        return nonce


# ---------------------------------------------------------------------------
# GF(2^8) (AES polynomial x^8 + x^4 + x^3 + x + 1) and Lagrange helpers
# ---------------------------------------------------------------------------

def f8_mul(a: int, b: int) -> int:
    r = 0
    for i in range(8):
        if (a >> i) & 1:
            r ^= b << i
    for bit in range(14, 7, -1):
        if (r >> bit) & 1:
            r ^= 0x11B << (bit - 8)
    return r


def f8_inv(a: int) -> int:
    for x in range(1, 256):
        if f8_mul(a, x) == 1:
            return x
    raise ZeroDivisionError


def f8_lagrange_extend(vals: list[int], src: list[int], dst: list[int]) -> list[int]:
    """Evaluate the degree < len(src) polynomial through (src[i], vals[i]) at dst, over F8."""
    out = []
    for x in dst:
        acc = 0
        for i, xi in enumerate(src):
            if vals[i] == 0:
                continue
            num, den = 1, 1
            for j, xj in enumerate(src):
                if j != i:
                    num = f8_mul(num, x ^ xj)
                    den = f8_mul(den, xi ^ xj)
            acc ^= f8_mul(vals[i], f8_mul(num, f8_inv(den)))
        out.append(acc)
    return out


def lagrange_f128(nodes: list[int], z: int) -> list[int]:
    """Lagrange basis values L_i(z) on F128 nodes (product form)."""
    w = []
    for i, xi in enumerate(nodes):
        num, den = 1, 1
        for j, xj in enumerate(nodes):
            if j != i:
                num = mul(num, z ^ xj)
                den = mul(den, xi ^ xj)
        w.append(mul(num, inv(den)))
    return w


# ---------------------------------------------------------------------------
# Keccak-f[1600] from FIPS 202 (lane A[x + 5y], bit z of the 64-bit lane)
# ---------------------------------------------------------------------------

_KECCAK_RHO = [[0, 36, 3, 41, 18], [1, 44, 10, 45, 2], [62, 6, 43, 15, 61],
               [28, 55, 25, 21, 56], [27, 20, 39, 8, 14]]  # r[x][y], FIPS 202 Table 2
M64 = (1 << 64) - 1


def _rc(t: int) -> int:
    """rc(t) of FIPS 202 Algorithm 5 (LFSR x^8 + x^6 + x^5 + x^4 + 1)."""
    if t % 255 == 0:
        return 1
    r = 1
    for _ in range(t % 255):
        r <<= 1
        if r & 0x100:
            r ^= 0x171
    return r & 1


def keccak_round_constant(ir: int) -> int:
    rc = 0
    for j in range(7):
        if _rc(j + 7 * ir):
            rc |= 1 << ((1 << j) - 1)
    return rc


def _rotl64(v, r):
    r %= 64
    return ((v << r) | (v >> (64 - r))) & M64 if r else v


def keccak_round_parts(A):
    """One round's (B = pi(rho(theta(A))), t = chi AND term, A') on 25 lanes."""
    C = [A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20] for x in range(5)]
    D = [C[(x - 1) % 5] ^ _rotl64(C[(x + 1) % 5], 1) for x in range(5)]
    T = [A[i] ^ D[i % 5] for i in range(25)]
    B = [0] * 25
    for x in range(5):
        for y in range(5):
            # pi: A'[x, y] = A[(x + 3y) mod 5, x]; rho offset of the source lane
            sx, sy = (x + 3 * y) % 5, x
            B[x + 5 * y] = _rotl64(T[sx + 5 * sy], _KECCAK_RHO[sx][sy])
    t = [(~B[(i % 5 + 1) % 5 + 5 * (i // 5)] & M64) & B[(i % 5 + 2) % 5 + 5 * (i // 5)]
         for i in range(25)]
    return B, t


def keccak_f(A):
    A = list(A)
    for ir in range(24):
        B, t = keccak_round_parts(A)
        A = [B[i] ^ t[i] for i in range(25)]
        A[0] ^= keccak_round_constant(ir)
    return A


def sha3_256(msg: bytes) -> bytes:
    rate = 136
    m = bytearray(msg) + b"\x06"
    m += bytes(-len(m) % rate)
    m[-1] |= 0x80
    A = [0] * 25
    for off in range(0, len(m), rate):
        for i in range(rate // 8):
            A[i] ^= int.from_bytes(m[off + 8 * i:off + 8 * i + 8], "little")
        A = keccak_f(A)
    return b"".join(A[i].to_bytes(8, "little") for i in range(4))


# ---------------------------------------------------------------------------
# BLAKE3 compression from the BLAKE3 specification (section 2.2)
# ---------------------------------------------------------------------------

BLAKE3_IV = [0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A, 0x510E527F, 0x9B05688C,
             0x1F83D9AB, 0x5BE0CD19]
_B3_PERM = [2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8]
M32 = (1 << 32) - 1


def _ror32(x, n):
    return ((x >> n) | (x << (32 - n))) & M32


def blake3_compress(cv, m, counter, block_len, flags):
    v = list(cv) + BLAKE3_IV[:4] + [counter & M32, counter >> 32, block_len, flags]
    m = list(m)

    def g(a, b, c, d, x, y):
        v[a] = (v[a] + v[b] + x) & M32
        v[d] = _ror32(v[d] ^ v[a], 16)
        v[c] = (v[c] + v[d]) & M32
        v[b] = _ror32(v[b] ^ v[c], 12)
        v[a] = (v[a] + v[b] + y) & M32
        v[d] = _ror32(v[d] ^ v[a], 8)
        v[c] = (v[c] + v[d]) & M32
        v[b] = _ror32(v[b] ^ v[c], 7)

    for r in range(7):
        g(0, 4, 8, 12, m[0], m[1])
        g(1, 5, 9, 13, m[2], m[3])
        g(2, 6, 10, 14, m[4], m[5])
        g(3, 7, 11, 15, m[6], m[7])
        g(0, 5, 10, 15, m[8], m[9])
        g(1, 6, 11, 12, m[10], m[11])
        g(2, 7, 8, 13, m[12], m[13])
        g(3, 4, 9, 14, m[14], m[15])
        m = [m[_B3_PERM[i]] for i in range(16)]
    return [v[i] ^ v[i + 8] for i in range(8)] + [v[i + 8] ^ cv[i] for i in range(8)]


def blake3_hash_short(msg: bytes) -> bytes:
    """BLAKE3 of a message of at most 64 bytes: one chunk, one block, root output."""
    assert len(msg) <= 64
    blk = msg + bytes(64 - len(msg))
    words = [int.from_bytes(blk[4 * i:4 * i + 4], "little") for i in range(16)]
    out = blake3_compress(BLAKE3_IV, words, 0, len(msg), 1 | 2 | 8)  # CHUNK_START|END|ROOT
    return b"".join(w.to_bytes(4, "little") for w in out[:8])


def _b3_words(blk: bytes) -> list[int]:
    blk = blk + bytes(64 - len(blk))
    return [int.from_bytes(blk[4 * i:4 * i + 4], "little") for i in range(16)]


def _b3_chunk(data: bytes, counter: int, root: bool) -> list[int]:
    cv = BLAKE3_IV
    nb = max(1, (len(data) + 63) // 64)
    for i in range(nb):
        blk = data[64 * i:64 * i + 64]
        flags = (1 if i == 0 else 0) | (2 if i == nb - 1 else 0) | (8 if root and i == nb - 1 else 0)
        cv = blake3_compress(cv, _b3_words(blk), counter, len(blk), flags)[:8]
    return cv


def _b3_subtree(data: bytes, first_chunk: int, root: bool) -> list[int]:
    """Chaining value of the subtree over `data` (specification section 2.1: the left
    subtree holds the largest power of two of chunks that leaves the right one non-empty)."""
    n = max(1, (len(data) + 1023) // 1024)
    if n == 1:
        return _b3_chunk(data, first_chunk, root)
    left = 1 << ((n - 1).bit_length() - 1)
    lcv = _b3_subtree(data[:1024 * left], first_chunk, False)
    rcv = _b3_subtree(data[1024 * left:], first_chunk + left, False)
    return blake3_compress(BLAKE3_IV, lcv + rcv, 0, 64, 4 | (8 if root else 0))[:8]


def _cv_bytes(cv: list[int]) -> bytes:
    return b"".join(w.to_bytes(4, "little") for w in cv)


def blake3_hash(msg: bytes) -> bytes:
    """BLAKE3 root hash (32 bytes) of any message."""
    return _cv_bytes(_b3_subtree(msg, 0, True))


def blake3_leaf_cv(msg: bytes) -> bytes:
    """The non-root chaining value of a message (the BLAKE3 Merkle leaf)."""
    return _cv_bytes(_b3_subtree(msg, 0, False))


def blake3_parent_cv(left: bytes, right: bytes) -> bytes:
    """PARENT compression of two chaining values, no ROOT flag (the BLAKE3 Merkle node)."""
    m = _b3_words(left[:32] + right[:32])
    return _cv_bytes(blake3_compress(BLAKE3_IV, m, 0, 64, 4)[:8])
