/*
 * flock_oracle.c — CPU restatement of the Flock `prove_fast` data-parallel core.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker + timed CPU baseline). The shipped
 * library never links or loads this file. Each function cites the reference
 * file:line it restates; paths are relative to the reference checkout.
 *
 * Pinning (what fixes each function independently of this text):
 *   - F128 mul/inv: tests/golden vectors from Python integer clmul, GHASH
 *     known answers (gf2_128.rs:503-531), field axioms.
 *   - NTT: brute-force evaluation of the novel-basis polynomial at the subspace
 *     points (tests), forward/inverse round trip, and the reference's own
 *     host_check_ntt.cpp built from source (oracle/_ref).
 *   - SHA-256 / Merkle: hashlib (FIPS 180-4) over the flat tree layout.
 *   - FsChallenger: a hashlib restatement of the framing and the reference's
 *     own test_challenger.cpp built from source (oracle/_ref).
 *   - Sumcheck rounds: brute-force sums and the sumcheck identities
 *     U(0)+U(1) = claim, Σ fold·fold = U(r) (tests).
 */
#include "flock_oracle.h"

#include <stdlib.h>
#include <string.h>

#if defined(__x86_64__)
#include <cpuid.h>
#include <immintrin.h>
#include <wmmintrin.h>
#include <emmintrin.h>
#endif

/* ------------------------------------------------------------------------- */
/* F128 arithmetic                                                            */
/* ------------------------------------------------------------------------- */

/* clmul64, crates/flock-core/src/field/gf2_128/portable.rs:4-19 */
static void clmul64(uint64_t a, uint64_t b, uint64_t* lo, uint64_t* hi) {
    uint64_t l = 0, h = 0;
    for (int i = 0; i < 64; i++) {
        if ((a >> i) & 1) {
            l ^= b << i;
            if (i) h ^= b >> (64 - i);
        }
    }
    *lo = l;
    *hi = h;
}

/* ghash_reduce, crates/flock-core/src/field/gf2_128.rs:278-297 */
static or_f128 ghash_reduce(uint64_t r0, uint64_t r1, uint64_t r2, uint64_t r3) {
    uint64_t s1_lo = r2 << 1, s1_hi = (r3 << 1) | (r2 >> 63);
    uint64_t s2_lo = r2 << 2, s2_hi = (r3 << 2) | (r2 >> 62);
    uint64_t s7_lo = r2 << 7, s7_hi = (r3 << 7) | (r2 >> 57);
    uint64_t t_lo = r2 ^ s1_lo ^ s2_lo ^ s7_lo;
    uint64_t t_hi = r3 ^ s1_hi ^ s2_hi ^ s7_hi;
    uint64_t ov = (r3 >> 63) ^ (r3 >> 62) ^ (r3 >> 57);
    uint64_t corr = ov ^ (ov << 1) ^ (ov << 2) ^ (ov << 7);
    or_f128 r = {r0 ^ t_lo ^ corr, r1 ^ t_hi};
    return r;
}

/* ghash_mul_unreduced + ghash_mul, gf2_128/portable.rs:21-43 */
or_f128 or_f128_mul_portable(or_f128 a, or_f128 b) {
    uint64_t ll_lo, ll_hi, lh_lo, lh_hi, hl_lo, hl_hi, hh_lo, hh_hi;
    clmul64(a.lo, b.lo, &ll_lo, &ll_hi);
    clmul64(a.lo, b.hi, &lh_lo, &lh_hi);
    clmul64(a.hi, b.lo, &hl_lo, &hl_hi);
    clmul64(a.hi, b.hi, &hh_lo, &hh_hi);
    return ghash_reduce(ll_lo, ll_hi ^ lh_lo ^ hl_lo, hh_lo ^ lh_hi ^ hl_hi, hh_hi);
}

#if defined(__x86_64__)
/* Same schoolbook product as the portable path, with the 64x64 carry-less
 * products from PCLMULQDQ (the reference's x86 Mul also uses PCLMULQDQ,
 * gf2_128/x86_64.rs:130). Reduction is the shared ghash_reduce. */
__attribute__((target("pclmul,sse2"))) static or_f128 mul_pclmul(or_f128 a, or_f128 b) {
    __m128i A = _mm_set_epi64x((long long)a.hi, (long long)a.lo);
    __m128i B = _mm_set_epi64x((long long)b.hi, (long long)b.lo);
    __m128i ll = _mm_clmulepi64_si128(A, B, 0x00);
    __m128i hh = _mm_clmulepi64_si128(A, B, 0x11);
    __m128i lh = _mm_clmulepi64_si128(A, B, 0x10);
    __m128i hl = _mm_clmulepi64_si128(A, B, 0x01);
    __m128i cr = _mm_xor_si128(lh, hl);
    uint64_t l[2], h[2], c[2];
    _mm_storeu_si128((__m128i*)l, ll);
    _mm_storeu_si128((__m128i*)h, hh);
    _mm_storeu_si128((__m128i*)c, cr);
    return ghash_reduce(l[0], l[1] ^ c[0], h[0] ^ c[1], h[1]);
}
#endif

#if defined(__x86_64__)
/* Four products by one fixed multiplier t in a 512-bit register (the 4-lane VPCLMULQDQ form of
 * the reference's x86 multiply, gf2_128/x86_64.rs:244): per 128-bit lane the four 64x64
 * carry-less products of the schoolbook, then the same GHASH reduction as ghash_reduce done
 * by folding the top word r3 (r3 x^192 = (r3 * 0x87) x^64) and then r2 (r2 x^128 = r2 * 0x87)
 * with carry-less products by 0x87 = x^7 + x^2 + x + 1. */
__attribute__((target("avx512f,avx512bw,vpclmulqdq"))) static __m512i mul4_v(__m512i A, __m512i B) {
    const __m512i K = _mm512_set1_epi64(0x87);
    const __m512i ll = _mm512_clmulepi64_epi128(A, B, 0x00);
    const __m512i hh = _mm512_clmulepi64_epi128(A, B, 0x11);
    const __m512i mid = _mm512_xor_si512(_mm512_clmulepi64_epi128(A, B, 0x01),
                                         _mm512_clmulepi64_epi128(A, B, 0x10));
    __m512i lo = _mm512_xor_si512(ll, _mm512_bslli_epi128(mid, 8)); /* r0, r1 */
    __m512i hi = _mm512_xor_si512(hh, _mm512_bsrli_epi128(mid, 8)); /* r2, r3 */
    const __m512i t = _mm512_clmulepi64_epi128(hi, K, 0x01);        /* r3 * 0x87 */
    lo = _mm512_xor_si512(lo, _mm512_bslli_epi128(t, 8));
    hi = _mm512_xor_si512(hi, _mm512_bsrli_epi128(t, 8));            /* r2 += carry of t */
    return _mm512_xor_si512(lo, _mm512_clmulepi64_epi128(hi, K, 0x00)); /* + r2 * 0x87 */
}

/* one NTT butterfly row over 4 consecutive lanes: u = top + v tw, bot = v + u */
__attribute__((target("avx512f,avx512bw,vpclmulqdq"))) static void butterfly4(or_f128* top, or_f128* bot,
                                                                           or_f128 tw) {
    const __m512i T = _mm512_set_epi64((long long)tw.hi, (long long)tw.lo, (long long)tw.hi,
                                       (long long)tw.lo, (long long)tw.hi, (long long)tw.lo,
                                       (long long)tw.hi, (long long)tw.lo);
    const __m512i v = _mm512_loadu_si512((const void*)bot);
    const __m512i u = _mm512_xor_si512(_mm512_loadu_si512((const void*)top), mul4_v(v, T));
    _mm512_storeu_si512((void*)top, u);
    _mm512_storeu_si512((void*)bot, _mm512_xor_si512(v, u));
}

/* a[i] * b[i] for four i */
__attribute__((target("avx512f,avx512bw,vpclmulqdq"))) static void mul4(const or_f128* a, const or_f128* b,
                                                                     or_f128* out) {
    _mm512_storeu_si512((void*)out, mul4_v(_mm512_loadu_si512((const void*)a),
                                           _mm512_loadu_si512((const void*)b)));
}
#endif

static int g_vpclmul = -1;

/* AVX-512F/BW + VPCLMULQDQ with the ZMM state enabled by the OS */
int or_has_vpclmul(void) {
    if (g_vpclmul < 0) {
        g_vpclmul = 0;
#if defined(__x86_64__)
        unsigned a, b, c, d;
        if (__get_cpuid(1, &a, &b, &c, &d) && (c & bit_OSXSAVE) &&
            __get_cpuid_count(7, 0, &a, &b, &c, &d) && (b & bit_AVX512F) && (b & bit_AVX512BW) &&
            (c & bit_VPCLMULQDQ)) {
            unsigned lo, hi;
            __asm__ volatile("xgetbv" : "=a"(lo), "=d"(hi) : "c"(0));
            g_vpclmul = (lo & 0xE6) == 0xE6;
        }
        const char* off = getenv("FLOCK_ORACLE_NO_VPCLMUL");
        if (off && *off && *off != '0') g_vpclmul = 0;
#endif
    }
    return g_vpclmul;
}

static int g_pclmul = -1;

int or_has_pclmul(void) {
    if (g_pclmul < 0) {
#if defined(__x86_64__)
        unsigned a, b, c, d;
        g_pclmul = (__get_cpuid(1, &a, &b, &c, &d) && (c & bit_PCLMUL)) ? 1 : 0;
#else
        g_pclmul = 0;
#endif
    }
    return g_pclmul;
}

or_f128 or_f128_mul(or_f128 a, or_f128 b) {
#if defined(__x86_64__)
    if (or_has_pclmul()) return mul_pclmul(a, b);
#endif
    return or_f128_mul_portable(a, b);
}

/* OpenMP thread count for the parallel loops (n <= 0: the runtime's default) */
void or_set_threads(int n) {
#if defined(_OPENMP)
    extern void omp_set_num_threads(int);
    extern int omp_get_num_procs(void);
    omp_set_num_threads(n > 0 ? n : omp_get_num_procs());
#else
    (void)n;
#endif
}

void or_f128_mul_batch(const or_f128* a, const or_f128* b, or_f128* out, size_t n) {
    size_t i = 0;
#if defined(__x86_64__)
    if (or_has_vpclmul())
        for (; i + 4 <= n; i += 4) mul4(a + i, b + i, out + i);
#endif
    for (; i < n; i++) out[i] = or_f128_mul(a[i], b[i]);
}

static inline or_f128 add(or_f128 a, or_f128 b) {
    or_f128 r = {a.lo ^ b.lo, a.hi ^ b.hi};
    return r;
}
static const or_f128 ZERO = {0, 0};
static const or_f128 ONE = {1, 0};

/* inv (Fermat x^(2^128-2)), gf2_128.rs:67-80 */
or_f128 or_f128_inv(or_f128 x) {
    or_f128 r = ONE;
    or_f128 cur = or_f128_mul(x, x);
    for (int i = 1; i < 128; i++) {
        r = or_f128_mul(r, cur);
        cur = or_f128_mul(cur, cur);
    }
    return r;
}

/* mul_by_x, gf2_128.rs:213-220 */
or_f128 or_f128_mul_by_x(or_f128 z) {
    uint64_t carry = z.hi >> 63;
    uint64_t mask = 0 - carry;
    or_f128 r = {(z.lo << 1) ^ (0x87 & mask), (z.hi << 1) | (z.lo >> 63)};
    return r;
}

/* ------------------------------------------------------------------------- */
/* SplitMix64 (dump_commit_vectors.rs:42-52); F128 = {lo: next, hi: next}, the */
/* same words pack_witness (pcs/pack.rs:40-80) assembles from rng.bits().      */
/* ------------------------------------------------------------------------- */
void or_splitmix_f128(uint64_t seed, or_f128* out, size_t n) {
    uint64_t s = seed;
    for (size_t i = 0; i < n; i++) {
        uint64_t w[2];
        for (int k = 0; k < 2; k++) {
            s += 0x9E3779B97F4A7C15ull;
            uint64_t z = s;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            w[k] = z ^ (z >> 31);
        }
        out[i].lo = w[0];
        out[i].hi = w[1];
    }
}

/* ------------------------------------------------------------------------- */
/* Additive NTT                                                               */
/* ------------------------------------------------------------------------- */

/* generate_evals_from_subspace with the standard basis {1, x, .., x^(L-1)},
 * additive_ntt_f128.rs:67-95 and :127-133. evals[i][k] = W_i(beta_{i+k}) /
 * W_i(beta_i). */
int or_ntt_evals(int L, or_f128* ev) {
    if (L < 1 || L > 64) return -1;
    memset(ev, 0, sizeof(or_f128) * (size_t)L * (size_t)L);
    for (int k = 0; k < L; k++) {
        ev[k].lo = 1ull << k;
        ev[k].hi = 0;
    }
    for (int i = 1; i < L; i++) {
        const or_f128* prev = ev + (size_t)(i - 1) * L;
        or_f128* row = ev + (size_t)i * L;
        int prev_len = L - (i - 1);
        for (int k = 1; k < prev_len; k++) row[k - 1] = or_f128_mul(prev[k], add(prev[k], prev[0]));
    }
    /* normalization uses the pre-normalization row 0 of each row; rows were
     * computed from unnormalized predecessors above, as in the reference */
    for (int i = 0; i < L; i++) {
        or_f128* row = ev + (size_t)i * L;
        or_f128 inv = or_f128_inv(row[0]);
        for (int k = 0; k < L - i; k++) row[k] = or_f128_mul(row[k], inv);
    }
    return 0;
}

/* twiddle(l, b) = span of evals[L-l-1][1..] at the set bits of b, :141-144 + span_get :99-108 */
or_f128 or_ntt_twiddle(const or_f128* ev, int L, int layer, uint64_t block) {
    const or_f128* v = ev + (size_t)(L - layer - 1) * L + 1;
    int len = L - (L - layer - 1) - 1; /* = layer */
    or_f128 acc = ZERO;
    for (int j = 0; j < len; j++)
        if ((block >> j) & 1) acc = add(acc, v[j]);
    return acc;
}

/* interleaved_scalar_live_from_layer, additive_ntt_f128.rs:287-319 */
int or_ntt_forward_interleaved(or_f128* data, size_t num_ntts, size_t live, int log_d, int L,
                               int start_layer) {
    if (log_d > L || start_layer > log_d || live > num_ntts) return -1;
    or_f128* ev = (or_f128*)malloc(sizeof(or_f128) * (size_t)L * L);
    if (!ev) return -2;
    or_ntt_evals(L, ev);
    for (int layer = start_layer; layer < log_d; layer++) {
        size_t num_blocks = (size_t)1 << layer;
        size_t block_size = (size_t)1 << (log_d - layer);
        size_t half = block_size >> 1;
        const int vec = or_has_vpclmul();
        /* every (block, row) butterfly row is independent: split them over the threads */
#pragma omp parallel for schedule(static) if (num_blocks * half * live >= 4096)
        for (size_t br = 0; br < num_blocks * half; br++) {
            const size_t block = br / half, row = br % half;
            or_f128 tw = or_ntt_twiddle(ev, L, layer, block);
            size_t block_start = block * block_size * num_ntts;
            size_t off_top = block_start + row * num_ntts;
            size_t off_bot = off_top + half * num_ntts;
            size_t lane = 0;
#if defined(__x86_64__)
            if (vec)
                for (; lane + 4 <= live; lane += 4) butterfly4(data + off_top + lane, data + off_bot + lane, tw);
#endif
            for (; lane < live; lane++) {
                or_f128 v = data[off_bot + lane];
                or_f128 u = add(data[off_top + lane], or_f128_mul(v, tw));
                data[off_top + lane] = u;
                data[off_bot + lane] = add(v, u);
            }
        }
        (void)vec;
    }
    free(ev);
    return 0;
}

/* inverse_transform_scalar, additive_ntt_f128.rs:651-673 */
int or_ntt_inverse_scalar(or_f128* data, int log_d, int L) {
    if (log_d > L) return -1;
    or_f128* ev = (or_f128*)malloc(sizeof(or_f128) * (size_t)L * L);
    if (!ev) return -2;
    or_ntt_evals(L, ev);
    for (int layer = log_d - 1; layer >= 0; layer--) {
        size_t num_blocks = (size_t)1 << layer;
        size_t half = (size_t)1 << (log_d - layer - 1);
        for (size_t block = 0; block < num_blocks; block++) {
            or_f128 tw = or_ntt_twiddle(ev, L, layer, block);
            size_t block_start = block << (log_d - layer);
            for (size_t i0 = block_start; i0 < block_start + half; i0++) {
                size_t i1 = i0 | half;
                or_f128 u = data[i0];
                or_f128 v = add(data[i1], u);
                data[i0] = add(u, or_f128_mul(v, tw));
                data[i1] = v;
            }
        }
    }
    free(ev);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* SHA-256 (FIPS 180-4)                                                       */
/* ------------------------------------------------------------------------- */
typedef struct {
    uint32_t h[8];
    uint64_t total;
    uint8_t buf[64];
    size_t blen;
} sha_ctx;

static const uint32_t SHA_K[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

static inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static void sha_block(uint32_t h[8], const uint8_t* p) {
    uint32_t w[64];
    for (int i = 0; i < 16; i++)
        w[i] = ((uint32_t)p[4 * i] << 24) | ((uint32_t)p[4 * i + 1] << 16) |
               ((uint32_t)p[4 * i + 2] << 8) | (uint32_t)p[4 * i + 3];
    for (int i = 16; i < 64; i++) {
        uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
        uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
        w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
    for (int i = 0; i < 64; i++) {
        uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
        uint32_t ch = (e & f) ^ (~e & g);
        uint32_t t1 = hh + S1 + ch + SHA_K[i] + w[i];
        uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
        uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
        uint32_t t2 = S0 + mj;
        hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
}

static void sha_init(sha_ctx* c) {
    static const uint32_t IV[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                                   0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    memcpy(c->h, IV, sizeof IV);
    c->total = 0;
    c->blen = 0;
}

static void sha_update(sha_ctx* c, const uint8_t* d, size_t n) {
    c->total += n;
    while (n) {
        size_t take = 64 - c->blen;
        if (take > n) take = n;
        memcpy(c->buf + c->blen, d, take);
        c->blen += take;
        d += take;
        n -= take;
        if (c->blen == 64) {
            sha_block(c->h, c->buf);
            c->blen = 0;
        }
    }
}

static void sha_final(const sha_ctx* in, uint8_t out[32]) {
    sha_ctx c = *in;
    uint64_t bits = c.total * 8;
    uint8_t pad = 0x80;
    sha_update(&c, &pad, 1);
    uint8_t z = 0;
    while (c.blen != 56) sha_update(&c, &z, 1);
    uint8_t len[8];
    for (int i = 0; i < 8; i++) len[i] = (uint8_t)(bits >> (56 - 8 * i));
    sha_update(&c, len, 8);
    for (int i = 0; i < 8; i++) {
        out[4 * i] = (uint8_t)(c.h[i] >> 24);
        out[4 * i + 1] = (uint8_t)(c.h[i] >> 16);
        out[4 * i + 2] = (uint8_t)(c.h[i] >> 8);
        out[4 * i + 3] = (uint8_t)c.h[i];
    }
}

void or_sha256(const uint8_t* data, size_t len, uint8_t out[32]) {
    sha_ctx c;
    sha_init(&c);
    sha_update(&c, data, len);
    sha_final(&c, out);
}

/* merkle_tree_sequential, merkle.rs:593-617; hash_leaf = SHA-256(leaf) :205-216,
 * hash_pair = SHA-256(L || R) :220-232 */
int or_merkle_tree(const uint8_t* data, size_t data_len, size_t num_leaves, uint8_t* tree) {
    if (num_leaves == 0 || (num_leaves & (num_leaves - 1)) || data_len % num_leaves) return -1;
    size_t leaf = data_len / num_leaves;
#pragma omp parallel for schedule(static) if (num_leaves >= 256)
    for (size_t i = 0; i < num_leaves; i++) or_sha256(data + i * leaf, leaf, tree + 32 * i);
    size_t rs = 0, rl = num_leaves;
    while (rl > 1) {
        size_t nl = rl >> 1;
#pragma omp parallel for schedule(static) if (nl >= 256)
        for (size_t i = 0; i < nl; i++)
            or_sha256(tree + 32 * (rs + 2 * i), 64, tree + 32 * (rs + rl + i));
        rs += rl;
        rl = nl;
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* PCS commit                                                                 */
/* ------------------------------------------------------------------------- */

/* PcsParams size helpers and validate(), pcs/commit.rs:187-298 */
int or_pcs_sizes(const or_pcs_params* p, size_t* msg_len, size_t* cw_len, size_t* n_leaves,
                 int* k_code, size_t* num_ntts) {
    if (p->m < 7 + p->log_batch_size || p->log_inv_rate < 1) return -1;
    int log_msg = (int)p->m - 7;
    int log_dim = log_msg - (int)p->log_batch_size;
    int k = log_dim + (int)p->log_inv_rate;
    if (k > 64) return -1;
    size_t t = p->num_lanes ? p->num_lanes : ((size_t)1 << p->log_batch_size);
    if (t < 1 || t > ((size_t)1 << p->log_batch_size)) return -1;
    if (msg_len) *msg_len = t << log_dim;
    if (cw_len) *cw_len = ((size_t)1 << k) * t;
    if (n_leaves) *n_leaves = (size_t)1 << k;
    if (k_code) *k_code = k;
    if (num_ntts) *num_ntts = t;
    return 0;
}

static int finalize(or_f128* cw, size_t t, size_t live, const or_pcs_params* p, uint8_t* tree) {
    size_t cw_len, n_leaves;
    int k;
    or_pcs_sizes(p, NULL, &cw_len, &n_leaves, &k, NULL);
    int rc = or_ntt_forward_interleaved(cw, t, live, k, k, (int)p->log_inv_rate);
    if (rc) return rc;
    if (tree)
        rc = or_merkle_tree_kind((const uint8_t*)cw, cw_len * 16, n_leaves, p->merkle_hash, tree);
    return rc;
}

int or_commit_lane_major(const or_f128* q, size_t q_len, const or_pcs_params* p, or_f128* cw,
                         uint8_t* tree) {
    size_t cw_len, t;
    if (or_pcs_sizes(p, NULL, &cw_len, NULL, NULL, &t)) return -1;
    if (q_len != ((size_t)1 << (p->m - 7))) return -1;
    size_t d = q_len >> p->log_batch_size;
    /* dead-lane skip, commit.rs:494-505 (value-identical either way) */
    size_t live = t;
    while (live > 0) {
        int zero = 1;
        for (size_t i = (live - 1) * d; i < live * d && zero; i++)
            if (q[i].lo | q[i].hi) zero = 0;
        if (!zero) break;
        live--;
    }
    size_t msg_len = t * d;
    for (size_t rep = 0; rep < cw_len / msg_len; rep++)
        for (size_t pp = 0; pp < d; pp++)
            for (size_t l = 0; l < t; l++) cw[rep * msg_len + pp * t + l] = q[l * d + pp];
    if (live == 0) {
        return tree ? or_merkle_tree_kind((const uint8_t*)cw, cw_len * 16, cw_len / t,
                                          p->merkle_hash, tree)
                    : 0;
    }
    return finalize(cw, t, live, p, tree);
}

int or_commit(const or_f128* msg, size_t msg_len, const or_pcs_params* p, or_f128* cw,
              uint8_t* tree) {
    size_t ml, cw_len, t;
    if (or_pcs_sizes(p, &ml, &cw_len, NULL, NULL, &t)) return -1;
    if (msg_len != ml) return -1;
    for (size_t i = 0; i < cw_len; i++) cw[i] = msg[i % msg_len];
    return finalize(cw, t, t, p, tree);
}

/* ------------------------------------------------------------------------- */
/* FsChallenger (SHA-256 discipline), challenger.rs:609-1044                  */
/* ------------------------------------------------------------------------- */
struct or_fs {
    sha_ctx st;
};

enum { OP_DOMAIN = 1, OP_LABEL = 2, OP_OBSERVE = 3, OP_SQUEEZE = 4, OP_BYTES = 5 };
enum { KIND_NONE = 0, KIND_SCALAR = 1, KIND_SLICE = 2 };

/* absorb_header, challenger.rs:708-714 */
static void fs_header(or_fs* fs, uint8_t op, uint8_t kind, uint64_t len) {
    uint8_t h[16] = {0};
    h[0] = op;
    h[1] = kind;
    for (int i = 0; i < 8; i++) h[8 + i] = (uint8_t)(len >> (8 * i));
    sha_update(&fs->st, h, 16);
}

/* absorb_padded, challenger.rs:720-726 */
static void fs_padded(or_fs* fs, const uint8_t* b, size_t n) {
    static const uint8_t zeros[16] = {0};
    sha_update(&fs->st, b, n);
    if (n % 16) sha_update(&fs->st, zeros, 16 - n % 16);
}

static void fs_f128(or_fs* fs, or_f128 v) {
    uint8_t b[16];
    for (int i = 0; i < 8; i++) {
        b[i] = (uint8_t)(v.lo >> (8 * i));
        b[8 + i] = (uint8_t)(v.hi >> (8 * i));
    }
    sha_update(&fs->st, b, 16);
}

or_fs* or_fs_new(const uint8_t* domain, size_t len) {
    or_fs* fs = (or_fs*)malloc(sizeof(or_fs));
    if (!fs) return NULL;
    sha_init(&fs->st);
    fs_header(fs, OP_DOMAIN, KIND_NONE, len); /* with_hash, :648-659 */
    fs_padded(fs, domain, len);
    return fs;
}

void or_fs_free(or_fs* fs) { free(fs); }

void or_fs_observe_label(or_fs* fs, const uint8_t* l, size_t n) { /* :806-809 */
    fs_header(fs, OP_LABEL, KIND_NONE, n);
    fs_padded(fs, l, n);
}

void or_fs_observe_f128(or_fs* fs, or_f128 v) { /* :811-814 */
    fs_header(fs, OP_OBSERVE, KIND_SCALAR, 1);
    fs_f128(fs, v);
}

void or_fs_observe_f128_slice(or_fs* fs, const or_f128* v, size_t n) { /* :816-821 */
    fs_header(fs, OP_OBSERVE, KIND_SLICE, n);
    for (size_t i = 0; i < n; i++) fs_f128(fs, v[i]);
}

void or_fs_observe_bytes(or_fs* fs, const uint8_t* b, size_t n) { /* :823-826 */
    fs_header(fs, OP_BYTES, KIND_NONE, n);
    fs_padded(fs, b, n);
}

/* squeeze_into: SHA256(state || ctr_le) blocks, state not mutated, :745-763 */
static void fs_squeeze(or_fs* fs, uint8_t* out, size_t n) {
    size_t off = 0;
    uint64_t ctr = 0;
    while (off < n) {
        sha_ctx c = fs->st;
        uint8_t cb[8];
        for (int i = 0; i < 8; i++) cb[i] = (uint8_t)(ctr >> (8 * i));
        sha_update(&c, cb, 8);
        uint8_t blk[32];
        sha_final(&c, blk);
        size_t take = n - off < 32 ? n - off : 32;
        memcpy(out + off, blk, take);
        off += take;
        ctr++;
    }
}

static or_f128 le_f128(const uint8_t* b) {
    or_f128 v = {0, 0};
    for (int i = 0; i < 8; i++) {
        v.lo |= (uint64_t)b[i] << (8 * i);
        v.hi |= (uint64_t)b[8 + i] << (8 * i);
    }
    return v;
}

or_f128 or_fs_sample_f128(or_fs* fs) { /* :828-846 */
    fs_header(fs, OP_SQUEEZE, KIND_SCALAR, 1);
    uint8_t buf[16];
    fs_squeeze(fs, buf, 16);
    return le_f128(buf);
}

void or_fs_sample_f128_vec(or_fs* fs, or_f128* out, size_t n) { /* :862-882 */
    fs_header(fs, OP_SQUEEZE, KIND_SLICE, n);
    uint8_t* buf = (uint8_t*)malloc(n * 16 + 1);
    fs_squeeze(fs, buf, n * 16);
    for (size_t i = 0; i < n; i++) out[i] = le_f128(buf + 16 * i);
    free(buf);
}

/* has_leading_zero_bits, challenger.rs:1096-1109 */
static int leading_zero_bits(const uint8_t* h, uint32_t bits) {
    uint32_t full = bits / 8, extra = bits % 8;
    for (uint32_t i = 0; i < full; i++)
        if (h[i]) return 0;
    if (extra && (h[full] >> (8 - extra))) return 0;
    return 1;
}

/* grind_pow: smallest nonce with SHA256(state_digest || nonce_le) having `bits`
 * leading zero bits, then observe_bytes(nonce_le), challenger.rs:884-958, 1127-1135 */
/* SHA-256 of everything absorbed so far (the digest grind_pow hashes, challenger.rs:884-958) */
void or_fs_digest(or_fs* fs, uint8_t out[32]) { sha_final(&fs->st, out); }

uint64_t or_fs_grind_pow(or_fs* fs, uint32_t bits) {
    uint8_t digest[32];
    sha_final(&fs->st, digest);
    uint64_t nonce = 0;
    if (bits) {
        for (;; nonce++) {
            uint8_t pre[40], h[32];
            memcpy(pre, digest, 32);
            for (int i = 0; i < 8; i++) pre[32 + i] = (uint8_t)(nonce >> (8 * i));
            or_sha256(pre, 40, h);
            if (leading_zero_bits(h, bits)) break;
        }
    }
    uint8_t nb[8];
    for (int i = 0; i < 8; i++) nb[i] = (uint8_t)(nonce >> (8 * i));
    or_fs_observe_bytes(fs, nb, 8);
    return nonce;
}

/* ------------------------------------------------------------------------- */
/* Sumcheck rounds                                                            */
/* ------------------------------------------------------------------------- */

/* round_msg_lsb, ligerito.rs:3204-3226 (sequential branch) */
void or_round_msg_lsb(const or_f128* f, const or_f128* b, size_t n, or_f128 msg[2]) {
    or_f128 u0 = ZERO, u2 = ZERO;
    for (size_t j = 0; j < n / 2; j++) {
        or_f128 f0 = f[2 * j], f1 = f[2 * j + 1], b0 = b[2 * j], b1 = b[2 * j + 1];
        u0 = add(u0, or_f128_mul(f0, b0));
        u2 = add(u2, or_f128_mul(add(f0, f1), add(b0, b1)));
    }
    msg[0] = u0;
    msg[1] = u2;
}

/* fold_and_msg_lsb, ligerito.rs:3372-3414: fold v0 + r(v0+v1), then the next
 * round's message over the folded pairs (none when half == 1). */
void or_fold_and_msg_lsb(const or_f128* f, const or_f128* b, size_t n, or_f128 r, or_f128* nf,
                         or_f128* nb, or_f128 msg[2]) {
    size_t half = n / 2;
    for (size_t j = 0; j < half; j++) {
        or_f128 f0 = f[2 * j], b0 = b[2 * j];
        nf[j] = add(f0, or_f128_mul(add(f0, f[2 * j + 1]), r));
        nb[j] = add(b0, or_f128_mul(add(b0, b[2 * j + 1]), r));
    }
    or_f128 u0 = ZERO, u2 = ZERO;
    for (size_t k = 0; k + 1 < half; k += 2) {
        u0 = add(u0, or_f128_mul(nf[k], nb[k]));
        u2 = add(u2, or_f128_mul(add(nf[k], nf[k + 1]), add(nb[k], nb[k + 1])));
    }
    msg[0] = u0;
    msg[1] = u2;
}

/* build_eq, univariate_skip.rs:31-47 */
void or_build_eq(const or_f128* r, size_t n, or_f128* t) {
    t[0] = ONE;
    for (size_t i = 0; i < n; i++) {
        or_f128 ri = r[i], omr = add(ONE, r[i]);
        const size_t len = (size_t)1 << i;
        /* entry x only feeds x and x | len: the level's entries are independent */
#pragma omp parallel for schedule(static) if (len >= 8192)
        for (size_t x = 0; x < len; x++) {
            t[x | len] = or_f128_mul(t[x], ri);
            t[x] = or_f128_mul(t[x], omr);
        }
    }
}

/* round_pair_naive, multilinear.rs:331-355: (r[0]·G(1), G(inf)) */
void or_round_pair(const or_f128* a, const or_f128* b, size_t n, const or_f128* r, or_f128 out[2]) {
    size_t half = n / 2;
    size_t log_n = 0;
    while (((size_t)1 << log_n) < n) log_n++;
    or_f128* eq = (or_f128*)malloc(sizeof(or_f128) * half);
    or_build_eq(r + 1, log_n - 1, eq);
    or_f128 g1 = ZERO, gi = ZERO;
#pragma omp parallel
    {
        or_f128 l1 = ZERO, li = ZERO;
        long long xx;
#pragma omp for schedule(static)
        for (xx = 0; xx < (long long)half; xx++) {
            const size_t x = (size_t)xx;
            or_f128 a0 = a[2 * x], a1 = a[2 * x + 1], b0 = b[2 * x], b1 = b[2 * x + 1];
            l1 = add(l1, or_f128_mul(or_f128_mul(eq[x], a1), b1));
            li = add(li, or_f128_mul(or_f128_mul(eq[x], add(a0, a1)), add(b0, b1)));
        }
#pragma omp critical
        {
            g1 = add(g1, l1);
            gi = add(gi, li);
        }
    }
    free(eq);
    out[0] = or_f128_mul(r[0], g1);
    out[1] = gi;
}

/* fold_in_place_pair, multilinear.rs:1148-1163 */
void or_fold_pair(const or_f128* a, const or_f128* b, size_t n, or_f128 c, or_f128* ao,
                  or_f128* bo) {
    long long xx;
#pragma omp parallel for schedule(static)
    for (xx = 0; xx < (long long)(n / 2); xx++) {
        const size_t x = (size_t)xx;
        or_f128 a0 = a[2 * x], a1 = a[2 * x + 1], b0 = b[2 * x], b1 = b[2 * x + 1];
        ao[x] = add(a0, or_f128_mul(c, add(a1, a0)));
        bo[x] = add(b0, or_f128_mul(c, add(b1, b0)));
    }
}

static const uint8_t LBL_PRODUCT[] = "flock-product-sumcheck";
static const uint8_t LBL_ZC[] = "flock-zerocheck-v0";

int or_sumcheck_transcript(const uint8_t* domain, size_t dlen, const or_f128* a, const or_f128* b,
                           int log_n, or_f128* msgs, or_f128 finals[4]) {
    if (log_n < 1 || log_n > 40) return -1;
    size_t n = (size_t)1 << log_n;
    or_fs* fs = or_fs_new(domain, dlen);
    or_f128* f = (or_f128*)malloc(sizeof(or_f128) * n);
    or_f128* g = (or_f128*)malloc(sizeof(or_f128) * n);
    or_f128* f2 = (or_f128*)malloc(sizeof(or_f128) * (n / 2 + 1));
    or_f128* g2 = (or_f128*)malloc(sizeof(or_f128) * (n / 2 + 1));
    or_f128* rr = (or_f128*)malloc(sizeof(or_f128) * (size_t)(log_n + 1));
    if (!fs || !f || !g || !f2 || !g2 || !rr) return -2;

    /* product sumcheck: SumcheckProver::new then fold loop, ligerito.rs:4986-5011 */
    or_fs_observe_label(fs, LBL_PRODUCT, sizeof(LBL_PRODUCT) - 1);
    memcpy(f, a, sizeof(or_f128) * n);
    memcpy(g, b, sizeof(or_f128) * n);
    or_f128 m2[2];
    or_round_msg_lsb(f, g, n, m2);
    int mi = 0;
    msgs[mi++] = m2[0];
    msgs[mi++] = m2[1];
    or_fs_observe_f128(fs, m2[0]);
    or_fs_observe_f128(fs, m2[1]);
    size_t len = n;
    for (int i = 0; i + 1 < log_n; i++) {
        or_f128 r = or_fs_sample_f128(fs);
        or_fold_and_msg_lsb(f, g, len, r, f2, g2, m2);
        len /= 2;
        memcpy(f, f2, sizeof(or_f128) * len);
        memcpy(g, g2, sizeof(or_f128) * len);
        msgs[mi++] = m2[0];
        msgs[mi++] = m2[1];
        or_fs_observe_f128(fs, m2[0]);
        or_fs_observe_f128(fs, m2[1]);
    }
    {
        or_f128 r = or_fs_sample_f128(fs);
        or_fold_and_msg_lsb(f, g, len, r, f2, g2, m2); /* half == 1: fold only */
        finals[0] = f2[0];
        finals[1] = g2[0];
        or_fs_observe_f128(fs, f2[0]);
        or_fs_observe_f128(fs, g2[0]);
    }

    /* eq-weighted zerocheck tail, zerocheck.rs:790-975 with Convention A r_next[0] = ONE */
    or_fs_observe_label(fs, LBL_ZC, sizeof(LBL_ZC) - 1);
    or_f128* r = (or_f128*)malloc(sizeof(or_f128) * (size_t)log_n);
    or_fs_sample_f128_vec(fs, r, (size_t)log_n);
    memcpy(f, a, sizeof(or_f128) * n);
    memcpy(g, b, sizeof(or_f128) * n);
    len = n;
    rr[0] = ONE;
    for (int k = 1; k < log_n; k++) rr[k] = r[k];
    or_round_pair(f, g, len, rr, m2);
    msgs[mi++] = m2[0];
    msgs[mi++] = m2[1];
    or_fs_observe_f128(fs, m2[0]);
    or_fs_observe_f128(fs, m2[1]);
    or_f128 rho = or_fs_sample_f128(fs);
    for (int i = 0; i + 1 < log_n; i++) {
        or_fold_pair(f, g, len, rho, f2, g2);
        len /= 2;
        memcpy(f, f2, sizeof(or_f128) * len);
        memcpy(g, g2, sizeof(or_f128) * len);
        rr[0] = ONE;
        for (int k = i + 2; k < log_n; k++) rr[k - i - 1] = r[k];
        or_round_pair(f, g, len, rr, m2);
        msgs[mi++] = m2[0];
        msgs[mi++] = m2[1];
        or_fs_observe_f128(fs, m2[0]);
        or_fs_observe_f128(fs, m2[1]);
        rho = or_fs_sample_f128(fs);
    }
    or_fold_pair(f, g, len, rho, f2, g2);
    finals[2] = f2[0];
    finals[3] = g2[0];
    or_fs_observe_f128(fs, f2[0]);
    or_fs_observe_f128(fs, g2[0]);

    free(r);
    free(rr);
    free(f);
    free(g);
    free(f2);
    free(g2);
    or_fs_free(fs);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* GF(2^8), phi_8                                                             */
/* ------------------------------------------------------------------------- */

/* clmul8_software + gf8_reduce, gf2_8.rs:102-137 */
uint8_t or_f8_mul(uint8_t a, uint8_t b) {
    uint16_t acc = 0;
    for (int i = 0; i < 8; i++)
        if ((a >> i) & 1) acc ^= (uint16_t)b << i;
    uint16_t h = acc >> 8;
    uint16_t t = (acc & 0xff) ^ h ^ (h << 1) ^ (h << 3) ^ (h << 4);
    uint16_t h2 = t >> 8;
    return (uint8_t)((t & 0xff) ^ h2 ^ (h2 << 1) ^ (h2 << 3) ^ (h2 << 4));
}

uint8_t or_f8_inv(uint8_t a) { /* a^254 */
    uint8_t r = 1, sq = a;
    for (int i = 0; i < 8; i++) {
        if ((0xFE >> i) & 1) r = or_f8_mul(r, sq);
        sq = or_f8_mul(sq, sq);
    }
    return r;
}

/* phi_8 is the ring homomorphism F8 -> F128 fixed by alpha = phi_8(0x02)
 * (phi8.rs:28-31); phi_8(v) = sum over set bits b of v of alpha^b. */
or_f128 or_phi8(uint8_t v) {
    const or_f128 alpha = {0x6b8330483c2e9849ull, 0x0dcb364640a222feull};
    or_f128 acc = ZERO, p = ONE;
    for (int b = 0; b < 8; b++) {
        if ((v >> b) & 1) acc = add(acc, p);
        p = or_f128_mul(p, alpha);
    }
    return acc;
}

/* compute_twiddles, ntt.rs:28-63 */
static void f8_twiddles(int k, uint8_t beta, uint8_t* tw) {
    if (k == 0) return;
    int len = 1 << (k - 1);
    uint8_t layer[128];
    for (int i = 0; i < len; i++) layer[i] = beta ^ (uint8_t)(2 * i);
    uint8_t s_root = 1;
    int write_at = len;
    for (int i = 0; i < len; i++) tw[write_at - 1 + i] = layer[i];
    for (int it = 1; it < k; it++) {
        write_at >>= 1;
        uint8_t s = layer[1] ^ layer[0];
        uint8_t next_root = or_f8_mul(s, s) ^ or_f8_mul(s_root, s); /* next_s, ntt.rs:22-25 */
        for (int i = 0; i < write_at; i++)
            layer[i] = or_f8_mul(layer[2 * i], layer[2 * i]) ^ or_f8_mul(s_root, layer[2 * i]);
        len = write_at;
        s_root = next_root;
        uint8_t s_inv = or_f8_inv(s_root);
        for (int j = 0; j < len; j++) tw[write_at - 1 + j] = or_f8_mul(s_inv, layer[j]);
    }
}

static void f8_fft_rec(uint8_t* v, int n, const uint8_t* tw, int idx) {
    if (n == 1) return;
    int half = n >> 1;
    uint8_t lam = tw[idx - 1];
    for (int i = 0; i < half; i++) { /* fft_butterfly, ntt.rs:65-74 */
        uint8_t w = v[half + i];
        v[i] ^= or_f8_mul(lam, w);
        v[half + i] = w ^ v[i];
    }
    f8_fft_rec(v, half, tw, 2 * idx);
    f8_fft_rec(v + half, half, tw, 2 * idx + 1);
}

static void f8_ifft_rec(uint8_t* v, int n, const uint8_t* tw, int idx) {
    if (n == 1) return;
    int half = n >> 1;
    f8_ifft_rec(v, half, tw, 2 * idx);
    f8_ifft_rec(v + half, half, tw, 2 * idx + 1);
    uint8_t lam = tw[idx - 1];
    for (int i = 0; i < half; i++) { /* ifft_butterfly, ntt.rs:88-95 */
        v[half + i] ^= v[i];
        v[i] ^= or_f8_mul(lam, v[half + i]);
    }
}

void or_f8_ntt_forward(uint8_t* v, int k, uint8_t beta) {
    uint8_t tw[256];
    f8_twiddles(k, beta, tw);
    if (k > 0) f8_fft_rec(v, 1 << k, tw, 1);
}

void or_f8_ntt_inverse(uint8_t* v, int k, uint8_t beta) {
    uint8_t tw[256];
    f8_twiddles(k, beta, tw);
    if (k > 0) f8_ifft_rec(v, 1 << k, tw, 1);
}

/* ------------------------------------------------------------------------- */
/* Zerocheck front                                                           */
/* ------------------------------------------------------------------------- */
static int bit_at(const uint8_t* p, size_t i) { return (p[i >> 3] >> (i & 7)) & 1; }

/* round1_naive, univariate_skip.rs:68-124: per row the F8 NTT pair extends the 64 bits on S to
 * Lambda, and every product is weighted by eq directly. */
void or_zc_round1_naive(const uint8_t* a, const uint8_t* b, const uint8_t* c, int m, const or_f128* r,
                  or_f128 p_ab[64], or_f128 p_c[64]) {
    const int ks = 6, ell = 64;
    size_t nx = (size_t)1 << (m - ks);
    or_f128* eq = (or_f128*)malloc(sizeof(or_f128) * nx);
    or_build_eq(r + ks, (size_t)(m - ks), eq);
    or_f128 phi[256];
    for (int v = 0; v < 256; v++) phi[v] = or_phi8((uint8_t)v);
    for (int i = 0; i < ell; i++) p_ab[i] = p_c[i] = ZERO;
    uint8_t ac[64], bc[64], cc[64];
    for (size_t x = 0; x < nx; x++) {
        for (int s = 0; s < ell; s++) {
            ac[s] = (uint8_t)bit_at(a, x * ell + s);
            bc[s] = (uint8_t)bit_at(b, x * ell + s);
            cc[s] = (uint8_t)bit_at(c, x * ell + s);
        }
        or_f8_ntt_inverse(ac, ks, 0);
        or_f8_ntt_forward(ac, ks, (uint8_t)ell);
        or_f8_ntt_inverse(bc, ks, 0);
        or_f8_ntt_forward(bc, ks, (uint8_t)ell);
        or_f8_ntt_inverse(cc, ks, 0);
        or_f8_ntt_forward(cc, ks, (uint8_t)ell);
        for (int i = 0; i < ell; i++) {
            p_ab[i] = add(p_ab[i], or_f128_mul(eq[x], phi[or_f8_mul(ac[i], bc[i])]));
            p_c[i] = add(p_c[i], or_f128_mul(eq[x], phi[cc[i]]));
        }
    }
    free(eq);
}

/* The same round 1 as sums over buckets. The S -> Lambda extension of a row's 64 bits is
 * F2-linear in the bits, so ext(row) is the XOR of ext(e_s) over the set bits s, read through
 * 16 nibble tables of 64-byte vectors; and p_ab[i] = sum_x eq(x) phi(ab_i(x)) =
 * sum_v phi(v) B_i[v] with B_i[v] the XOR of eq(x) over the rows where ab_i(x) = v. Rows are
 * split over OpenMP threads, each with its own buckets (XOR sums: the order does not matter). */
static uint64_t g_ext[16][16][8];
static uint8_t g_mul8[256][256];
static void round1_tables(void) {
    static int ready = 0;
    if (ready) return;
    uint8_t unit[64][64];
    for (int s = 0; s < 64; s++) {
        uint8_t v[64];
        memset(v, 0, sizeof(v));
        v[s] = 1;
        or_f8_ntt_inverse(v, 6, 0);
        or_f8_ntt_forward(v, 6, 64);
        memcpy(unit[s], v, 64);
    }
    for (int k = 0; k < 16; k++)
        for (int nib = 0; nib < 16; nib++) {
            uint8_t acc[64];
            memset(acc, 0, sizeof(acc));
            for (int bt = 0; bt < 4; bt++)
                if ((nib >> bt) & 1)
                    for (int i = 0; i < 64; i++) acc[i] ^= unit[4 * k + bt][i];
            memcpy(g_ext[k][nib], acc, 64);
        }
    for (int x = 0; x < 256; x++)
        for (int y = 0; y < 256; y++) g_mul8[x][y] = or_f8_mul((uint8_t)x, (uint8_t)y);
    ready = 1;
}
static void ext_row(const uint8_t* p, uint8_t out[64]) {
    uint64_t bits, acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    memcpy(&bits, p, 8);
    for (int k = 0; k < 16; k++) {
        const uint64_t* t = g_ext[k][(bits >> (4 * k)) & 15];
        for (int i = 0; i < 8; i++) acc[i] ^= t[i];
    }
    memcpy(out, acc, 64);
}

void or_zc_round1(const uint8_t* a, const uint8_t* b, const uint8_t* c, int m, const or_f128* r,
                  or_f128 p_ab[64], or_f128 p_c[64]) {
    const int ks = 6;
    const size_t nx = (size_t)1 << (m - ks);
    or_f128* eq = (or_f128*)malloc(sizeof(or_f128) * nx);
    or_build_eq(r + ks, (size_t)(m - ks), eq);
    round1_tables();
    or_f128* tot = (or_f128*)calloc(2 * 64 * 256, sizeof(or_f128));
#pragma omp parallel
    {
        or_f128* bk = (or_f128*)calloc(2 * 64 * 256, sizeof(or_f128));
        uint8_t ac[64], bc[64], cc[64];
        long long xx;
#pragma omp for schedule(static)
        for (xx = 0; xx < (long long)nx; xx++) {
            const size_t x = (size_t)xx;
            ext_row(a + 8 * x, ac);
            ext_row(b + 8 * x, bc);
            ext_row(c + 8 * x, cc);
            for (int i = 0; i < 64; i++) {
                or_f128* u = &bk[i * 256 + g_mul8[ac[i]][bc[i]]];
                or_f128* w = &bk[(64 + i) * 256 + cc[i]];
                *u = add(*u, eq[x]);
                *w = add(*w, eq[x]);
            }
        }
#pragma omp critical
        for (size_t j = 0; j < 2 * 64 * 256; j++) tot[j] = add(tot[j], bk[j]);
        free(bk);
    }
    or_f128 phi[256];
    for (int v = 0; v < 256; v++) phi[v] = or_phi8((uint8_t)v);
    for (int i = 0; i < 64; i++) {
        or_f128 sab = ZERO, sc = ZERO;
        for (int v = 1; v < 256; v++) {
            sab = add(sab, or_f128_mul(phi[v], tot[i * 256 + v]));
            sc = add(sc, or_f128_mul(phi[v], tot[(64 + i) * 256 + v]));
        }
        p_ab[i] = sab;
        p_c[i] = sc;
    }
    free(tot);
    free(eq);
}

/* round1_extract_c_packed_with_s_hat_v's c side (univariate_skip.rs:384-489): two banks
 * keyed by bit 0 of the row index x, each weighted by eq(r[6..m), x), then each divided by
 * eq(r[6], bank). out[lane | bank << 6]. */
void or_zc_s_hat_v_c(const uint8_t* c, int m, const or_f128* r, or_f128 out[128]) {
    const int ks = 6, ell = 64;
    size_t nx = (size_t)1 << (m - ks);
    or_f128* eq = (or_f128*)malloc(sizeof(or_f128) * nx);
    or_build_eq(r + ks, (size_t)(m - ks), eq);
    or_f128 bank[2][64];
    for (int i = 0; i < ell; i++) bank[0][i] = bank[1][i] = ZERO;
#pragma omp parallel
    {
        or_f128 lb[2][64];
        for (int i = 0; i < ell; i++) lb[0][i] = lb[1][i] = ZERO;
        long long xx;
#pragma omp for schedule(static)
        for (xx = 0; xx < (long long)nx; xx++) {
            const size_t x = (size_t)xx;
            for (int s = 0; s < ell; s++)
                if (bit_at(c, x * ell + s)) lb[x & 1][s] = add(lb[x & 1][s], eq[x]);
        }
#pragma omp critical
        for (int i = 0; i < ell; i++) {
            bank[0][i] = add(bank[0][i], lb[0][i]);
            bank[1][i] = add(bank[1][i], lb[1][i]);
        }
    }
    const or_f128 inv0 = or_f128_inv(add(ONE, r[ks])), inv1 = or_f128_inv(r[ks]);
    for (int i = 0; i < ell; i++) {
        out[i] = or_f128_mul(bank[0][i], inv0);
        out[ell + i] = or_f128_mul(bank[1][i], inv1);
    }
    free(eq);
}

/* fold_1b_rows scalar form (pcs/ring_switch.rs:182-230, 410-470): out[j] = sum_y t[y] bit j
 * of the 128-bit row y. */
void or_fold_1b_rows(const or_f128* rows, const or_f128* t, size_t n, or_f128 out[128]) {
    for (int j = 0; j < 128; j++) out[j] = ZERO;
    for (size_t y = 0; y < n; y++)
        for (int j = 0; j < 128; j++) {
            const uint64_t w = j < 64 ? rows[y].lo : rows[y].hi;
            if ((w >> (j & 63)) & 1) out[j] = add(out[j], t[y]);
        }
}

/* Lagrange basis on nodes phi_8(offset + i), i < 64, at z (product form). */
void or_lagrange_phi8(int offset, or_f128 z, or_f128 w[64]) {
    or_f128 nodes[64];
    for (int i = 0; i < 64; i++) nodes[i] = or_phi8((uint8_t)(offset + i));
    for (int i = 0; i < 64; i++) {
        or_f128 num = ONE, den = ONE;
        for (int j = 0; j < 64; j++) {
            if (j == i) continue;
            num = or_f128_mul(num, add(z, nodes[j]));
            den = or_f128_mul(den, add(nodes[i], nodes[j]));
        }
        w[i] = or_f128_mul(num, or_f128_inv(den));
    }
}

/* fold_at_z_naive, multilinear.rs:290-310 */
void or_fold_at_z(const uint8_t* bits, int m, or_f128 z, or_f128* out) {
    or_f128 w[64];
    or_lagrange_phi8(0, z, w);
    size_t nx = (size_t)1 << (m - 6);
    long long xx;
#pragma omp parallel for schedule(static)
    for (xx = 0; xx < (long long)nx; xx++) {
        const size_t x = (size_t)xx;
        or_f128 acc = ZERO;
        for (int s = 0; s < 64; s++)
            if (bit_at(bits, x * 64 + s)) acc = add(acc, w[s]);
        out[x] = acc;
    }
}

/* small_challenges_ghash / medium_challenges_ghash, univariate_skip_optimized.rs:91-139 */
void or_zc_inner_constants(or_f128 out[7]) {
    const uint8_t small[3] = {0xF7, 0x53, 0xB5};
    for (int i = 0; i < 3; i++) out[i] = or_phi8(small[i]);
    for (int i = 0; i < 4; i++) {
        or_f128 g = {1ull << (1 << i), 0}; /* gamma^(2^i), i.e. x^1, x^2, x^4, x^8 */
        out[3 + i] = or_f128_mul(g, or_f128_inv(add(ONE, g)));
    }
}

/* bits_for_numerator, zerocheck.rs:91-94 */
static uint32_t bits_for(size_t n) {
    uint32_t b = 0;
    while (n) {
        b++;
        n >>= 1;
    }
    return b;
}

int or_zerocheck_prove(or_fs* fs, const uint8_t* a, const uint8_t* b, const uint8_t* c, int m,
                       int grinding, or_f128* round1, or_f128* ml, or_f128 finals[3],
                       uint64_t* nonces, or_f128* claim) {
    const int ks = 6, n_inner = 7;
    if (m < ks + n_inner || m > 40) return -1;
    const int n_mlv = m - ks;
    int ni = 0;
    static const uint8_t LBL[] = "flock-zerocheck-v0";
    or_fs_observe_label(fs, LBL, sizeof(LBL) - 1);
    or_f128* r = (or_f128*)malloc(sizeof(or_f128) * (size_t)m);
    /* r layout: r_skip (6) | 7 inner constants | r_outer (zerocheck.rs:644-669) */
    if (grinding) nonces[ni++] = or_fs_grind_pow(fs, bits_for((size_t)m));
    or_fs_sample_f128_vec(fs, r, ks);
    or_fs_sample_f128_vec(fs, r + ks + n_inner, (size_t)(m - ks - n_inner));
    or_zc_inner_constants(r + ks);
    /* round 1 (naive convention) */
    or_zc_round1(a, b, c, m, r, round1, round1 + 64);
    or_fs_observe_f128_slice(fs, round1, 64);
    or_fs_observe_f128_slice(fs, round1 + 64, 64);
    if (grinding) nonces[ni++] = or_fs_grind_pow(fs, bits_for((1u << (ks + 1)) - 1));
    or_f128 z = or_fs_sample_f128(fs);
    /* c-claim: interpolate round1_c on Lambda at z (multilinear.rs:217-228) */
    or_f128 wl[64];
    or_lagrange_phi8(64, z, wl);
    or_f128 c_eval = ZERO;
    for (int i = 0; i < 64; i++) c_eval = add(c_eval, or_f128_mul(wl[i], round1[64 + i]));
    /* round 2: fold at z, first multilinear message with mlv_arg = [ONE, r[ks+1..]] */
    size_t n = (size_t)1 << n_mlv;
    or_f128* A = (or_f128*)malloc(sizeof(or_f128) * n);
    or_f128* B = (or_f128*)malloc(sizeof(or_f128) * n);
    or_f128* A2 = (or_f128*)malloc(sizeof(or_f128) * (n / 2 + 1));
    or_f128* B2 = (or_f128*)malloc(sizeof(or_f128) * (n / 2 + 1));
    or_f128* rr = (or_f128*)malloc(sizeof(or_f128) * (size_t)(n_mlv + 1));
    or_fold_at_z(a, m, z, A);
    or_fold_at_z(b, m, z, B);
    rr[0] = ONE;
    for (int k = 1; k < n_mlv; k++) rr[k] = r[ks + k];
    or_f128 m2[2];
    or_round_pair(A, B, n, rr, m2);
    int mi = 0;
    ml[mi++] = m2[0];
    ml[mi++] = m2[1];
    or_fs_observe_f128(fs, m2[0]);
    or_fs_observe_f128(fs, m2[1]);
    if (grinding) nonces[ni++] = or_fs_grind_pow(fs, bits_for(2));
    or_f128 rho = or_fs_sample_f128(fs);
    claim[1] = rho;
    size_t len = n;
    for (int i = 0; i + 1 < n_mlv; i++) {
        or_fold_pair(A, B, len, rho, A2, B2);
        len /= 2;
        memcpy(A, A2, sizeof(or_f128) * len);
        memcpy(B, B2, sizeof(or_f128) * len);
        rr[0] = ONE;
        for (int k = ks + i + 2; k < m; k++) rr[k - (ks + i + 2) + 1] = r[k];
        or_round_pair(A, B, len, rr, m2);
        ml[mi++] = m2[0];
        ml[mi++] = m2[1];
        or_fs_observe_f128(fs, m2[0]);
        or_fs_observe_f128(fs, m2[1]);
        if (grinding) nonces[ni++] = or_fs_grind_pow(fs, bits_for(2));
        rho = or_fs_sample_f128(fs);
        claim[2 + i] = rho;
    }
    or_fold_pair(A, B, len, rho, A2, B2);
    finals[0] = A2[0];
    finals[1] = B2[0];
    finals[2] = c_eval;
    or_fs_observe_f128(fs, A2[0]);
    or_fs_observe_f128(fs, B2[0]);
    claim[0] = z;
    for (int k = 0; k < n_mlv; k++) claim[1 + n_mlv + k] = r[ks + k];
    free(r);
    free(A);
    free(B);
    free(A2);
    free(B2);
    free(rr);
    return 0;
}

/* ======================================================================== */
/* Lincheck (crates/flock-core/src/lincheck.rs, lincheck/union.rs)          */
/* ======================================================================== */

/* build_quirky_eq_table_from_weights (lincheck.rs:1355-1371) with phi_8 skip weights
 * (build_quirky_eq_table, :1376-1378): out[i_skip + 64 i_rest] = L_i_skip(z) eq(x_rest, i_rest). */
void or_quirky_eq(or_f128 z_skip, const or_f128* x_rest, size_t n_rest, or_f128* out) {
    or_f128 w[64];
    or_lagrange_phi8(0, z_skip, w);
    size_t nr = (size_t)1 << n_rest;
    or_f128* er = (or_f128*)malloc(sizeof(or_f128) * nr);
    or_build_eq(x_rest, n_rest, er);
    for (size_t j = 0; j < nr; j++)
        for (int i = 0; i < 64; i++) out[j * 64 + i] = or_f128_mul(w[i], er[j]);
    free(er);
}

/* CscCircuit::fold_split (lincheck.rs:463-497): sa[c] = sum of eq over the rows of column c
 * of A, sb likewise for B. */
void or_csc_fold_split(size_t n_cols, const uint32_t* a_ptr, const uint32_t* a_rows,
                       const uint32_t* b_ptr, const uint32_t* b_rows, const or_f128* eq,
                       or_f128* sa, or_f128* sb) {
#pragma omp parallel for schedule(dynamic, 256) if (n_cols >= 1024)
    for (size_t c = 0; c < n_cols; c++) {
        or_f128 x = ZERO, y = ZERO;
        for (uint32_t k = a_ptr[c]; k < a_ptr[c + 1]; k++) x = add(x, eq[a_rows[k]]);
        for (uint32_t k = b_ptr[c]; k < b_ptr[c + 1]; k++) y = add(y, eq[b_rows[k]]);
        sa[c] = x;
        sb[c] = y;
    }
}

/* partial_fold_packed_z (lincheck.rs:804-835) in the support-proportional form of
 * partial_fold_packed_z_rows_padded / _fast_padded (:865-990): stripes s < ceil(n_rows/8)
 * (whole bytes), columns i < useful_bits; out[i] = sum of eq_outer[8s + r] over set bits r of
 * z[s*k + i]. Columns >= useful_bits are zero. */
void or_partial_fold(const uint8_t* z, int k_log, int nu, size_t useful_bits, size_t n_rows,
                     const or_f128* eq_outer, or_f128* out) {
    size_t k = (size_t)1 << k_log;
    size_t n_stripes = (n_rows + 7) / 8;
    (void)nu;
    for (size_t i = 0; i < k; i++) out[i] = ZERO;
    /* each output column i is its own sum over the stripes */
#pragma omp parallel for schedule(static) if (useful_bits * n_stripes >= 65536)
    for (size_t i = 0; i < useful_bits; i++) {
        or_f128 acc = ZERO;
        for (size_t s = 0; s < n_stripes; s++) {
            uint8_t byte = z[s * k + i];
            for (int r = 0; r < 8; r++)
                if ((byte >> r) & 1) acc = add(acc, eq_outer[8 * s + r]);
        }
        out[i] = acc;
    }
}

/* sumcheck_round_eval_par (lincheck.rs:1479-1502): top-bit split. */
void or_lc_round_eval(const or_f128* c, const or_f128* z, size_t len, or_f128 out[2]) {
    size_t h = len / 2;
    or_f128 e1 = ZERO, ei = ZERO;
    for (size_t i = 0; i < h; i++) {
        e1 = add(e1, or_f128_mul(c[h + i], z[h + i]));
        ei = add(ei, or_f128_mul(add(c[h + i], c[i]), add(z[h + i], z[i])));
    }
    out[0] = e1;
    out[1] = ei;
}

/* sumcheck_bind_top_in_place_par (lincheck.rs:1506-1523): v[i] += r (v[i+h] + v[i]). */
void or_lc_bind_top(or_f128* v, size_t len, or_f128 r) {
    size_t h = len / 2;
    for (size_t i = 0; i < h; i++) v[i] = add(v[i], or_f128_mul(r, add(v[h + i], v[i])));
}

/* Single-slot union lincheck prover: prove_union_capture_z_vec_with_grinding
 * (lincheck/union.rs:203-430) for a registry with one boolean type filling the boolean
 * region (the T = 1 path, union.rs:270-300), followed by column_sumcheck_prove
 * (lincheck.rs:1907-2028; each round evaluated then bound, as in the unfused form).
 * Grinding (LincheckGrinding::per_challenge_128, lincheck.rs:565-603): alpha 1 bit, beta 1
 * bit, rounds 2 bits, skip k_skip bits. const_pin < 0 = none.
 * Outputs: rounds[2L] (L = k_log - 6), z_partial[64], matrix_evals[2], nonces,
 * claim[2 + L] = (r_inner_skip, w, r_inner_rest[0..L)), z_vec[2^k_log] (captured). */
/* The circuit enters the lincheck only through its column marginals (LincheckCircuit,
 * lincheck.rs:238-278): a CSC matrix pair or a matrix-free walker. */
typedef struct {
    const uint32_t *a_ptr, *a_rows, *b_ptr, *b_rows;
} csc_ref;
typedef void (*fold_split_fn)(const void* circuit, size_t n_cols, const or_f128* eq,
                              or_f128* sa, or_f128* sb);

static void csc_fold_cb(const void* circuit, size_t n_cols, const or_f128* eq, or_f128* sa,
                        or_f128* sb) {
    const csc_ref* k = (const csc_ref*)circuit;
    or_csc_fold_split(n_cols, k->a_ptr, k->a_rows, k->b_ptr, k->b_rows, eq, sa, sb);
}

static void keccak_fold_cb(const void* circuit, size_t n_cols, const or_f128* eq, or_f128* sa,
                           or_f128* sb) {
    (void)circuit;
    (void)n_cols;
    or_keccak_lc_fold_split(eq, sa, sb);
}

static int lincheck_prove_with(fold_split_fn fold, const void* circuit, or_fs* fs, int k_log,
                               long const_pin, const uint8_t* z, int nu, size_t useful_bits,
                               size_t n_rows, or_f128 z_skip, const or_f128* x_inner_rest,
                               const or_f128* x_outer, int grinding, or_f128* rounds,
                               or_f128 z_partial[64], or_f128 matrix_evals[2],
                               uint64_t* nonces, or_f128* claim, or_f128* z_vec_out);

int or_lincheck_prove(or_fs* fs, int k_log, const uint32_t* a_ptr, const uint32_t* a_rows,
                      const uint32_t* b_ptr, const uint32_t* b_rows, long const_pin,
                      const uint8_t* z, int nu, size_t useful_bits, size_t n_rows,
                      or_f128 z_skip, const or_f128* x_inner_rest, const or_f128* x_outer,
                      int grinding, or_f128* rounds, or_f128 z_partial[64],
                      or_f128 matrix_evals[2], uint64_t* nonces, or_f128* claim,
                      or_f128* z_vec_out) {
    const csc_ref k = {a_ptr, a_rows, b_ptr, b_rows};
    return lincheck_prove_with(csc_fold_cb, &k, fs, k_log, const_pin, z, nu, useful_bits, n_rows,
                               z_skip, x_inner_rest, x_outer, grinding, rounds, z_partial,
                               matrix_evals, nonces, claim, z_vec_out);
}

int or_lincheck_prove_keccak(or_fs* fs, const uint8_t* z, int nu, size_t n_rows,
                             or_f128 z_skip, const or_f128* x_inner_rest, const or_f128* x_outer,
                             int grinding, or_f128* rounds, or_f128 z_partial[64],
                             or_f128 matrix_evals[2], uint64_t* nonces, or_f128* claim,
                             or_f128* z_vec_out) {
    /* K_LOG 16, USEFUL_BITS 42,560, const pin Z_CONST = 4096 (keccak.rs:436-451, 688-690) */
    return lincheck_prove_with(keccak_fold_cb, NULL, fs, 16, 4096, z, nu, 42560, n_rows, z_skip,
                               x_inner_rest, x_outer, grinding, rounds, z_partial, matrix_evals,
                               nonces, claim, z_vec_out);
}

static int lincheck_prove_with(fold_split_fn fold, const void* circuit, or_fs* fs, int k_log,
                               long const_pin, const uint8_t* z, int nu, size_t useful_bits,
                               size_t n_rows, or_f128 z_skip, const or_f128* x_inner_rest,
                               const or_f128* x_outer, int grinding, or_f128* rounds,
                               or_f128 z_partial[64], or_f128 matrix_evals[2],
                               uint64_t* nonces, or_f128* claim, or_f128* z_vec_out) {
    const int ks = 6;
    if (k_log < ks || nu < 3) return 1;
    const size_t K = (size_t)1 << k_log;
    const int L = k_log - ks;
    int ni = 0;
    static const uint8_t label[] = "flock-lincheck-v0";
    or_fs_observe_label(fs, label, sizeof(label) - 1);
    if (grinding) nonces[ni++] = or_fs_grind_pow(fs, 1);
    const or_f128 alpha = or_fs_sample_f128(fs);
    or_f128* eq_inner = (or_f128*)malloc(sizeof(or_f128) * K);
    or_f128* sa = (or_f128*)malloc(sizeof(or_f128) * K);
    or_f128* sb = (or_f128*)malloc(sizeof(or_f128) * K);
    or_f128* comb = (or_f128*)malloc(sizeof(or_f128) * K);
    or_f128* zv = (or_f128*)malloc(sizeof(or_f128) * K);
    or_quirky_eq(z_skip, x_inner_rest, (size_t)L, eq_inner);
    fold(circuit, K, eq_inner, sa, sb);
    for (size_t c = 0; c < K; c++) comb[c] = add(or_f128_mul(alpha, sa[c]), sb[c]);
    if (const_pin >= 0) {
        if (grinding) nonces[ni++] = or_fs_grind_pow(fs, 1);
        const or_f128 beta = or_fs_sample_f128(fs);
        comb[const_pin] = add(comb[const_pin], beta);
    }
    size_t n_outer = (size_t)1 << nu;
    or_f128* eq_outer = (or_f128*)malloc(sizeof(or_f128) * n_outer);
    or_build_eq(x_outer, (size_t)nu, eq_outer);
    or_partial_fold(z, k_log, nu, useful_bits, n_rows, eq_outer, zv);
    if (z_vec_out) memcpy(z_vec_out, zv, sizeof(or_f128) * K);
    /* column_sumcheck_prove */
    or_f128* rr = (or_f128*)malloc(sizeof(or_f128) * (L > 0 ? L : 1));
    size_t len = K;
    for (int t = 0; t < L; t++) {
        or_f128 msg[2];
        or_lc_round_eval(comb, zv, len, msg);
        or_fs_observe_f128(fs, msg[0]);
        or_fs_observe_f128(fs, msg[1]);
        if (grinding) nonces[ni++] = or_fs_grind_pow(fs, 2);
        const or_f128 r = or_fs_sample_f128(fs);
        rounds[2 * t] = msg[0];
        rounds[2 * t + 1] = msg[1];
        rr[t] = r;
        or_lc_bind_top(comb, len, r);
        or_lc_bind_top(zv, len, r);
        len /= 2;
    }
    for (int i = 0; i < 64; i++) z_partial[i] = zv[i];
    or_fs_observe_f128_slice(fs, z_partial, 64);
    if (grinding) nonces[ni++] = or_fs_grind_pow(fs, (uint32_t)ks);
    const or_f128 r_skip = or_fs_sample_f128(fs);
    or_f128 lam[64];
    or_lagrange_phi8(0, r_skip, lam);
    or_f128 w = ZERO;
    for (int i = 0; i < 64; i++) w = add(w, or_f128_mul(lam[i], z_partial[i]));
    claim[0] = r_skip;
    claim[1] = w;
    for (int j = 0; j < L; j++) claim[2 + j] = rr[L - 1 - j]; /* LSB-first, lincheck.rs:2007-2013 */
    /* matrix_evals (union.rs:404-416): W_col = z_partial (x) eq(r_inner_rest) */
    {
        or_f128* e = (or_f128*)malloc(sizeof(or_f128) * ((size_t)1 << L));
        or_build_eq(claim + 2, (size_t)L, e);
        or_f128 ma = ZERO, mb = ZERO;
        for (size_t j = 0; j < ((size_t)1 << L); j++)
            for (int i = 0; i < 64; i++) {
                const or_f128 wc = or_f128_mul(e[j], z_partial[i]);
                ma = add(ma, or_f128_mul(wc, sa[j * 64 + i]));
                mb = add(mb, or_f128_mul(wc, sb[j * 64 + i]));
            }
        matrix_evals[0] = ma;
        matrix_evals[1] = mb;
        free(e);
    }
    free(rr);
    free(eq_outer);
    free(zv);
    free(comb);
    free(sb);
    free(sa);
    free(eq_inner);
    return 0;
}

/* prove_union_capture_z_vec_with_grinding over several boolean slots (lincheck/union.rs:
 * 203-430): per slot the alpha-batched comb through the slot's own quirky table (inner
 * coordinates x_inner_rest[..k_t - 6]) scaled by eq_prefix_weight(x_inner_rest[k_t - 6..],
 * prefix_t) (:90-101) at column offset prefix_t << k_t; then each slot's constant-wire beta in
 * slot order; the slots' partial folds against one eq(x_outer) table at the same offsets; the
 * column sumcheck over the 2^col_vars union columns; and per slot the unscaled matrix evals
 * with W_col = z_partial (x) eq(r_inner_rest[..k_t - 6]). matrix_evals holds two per slot. */
int or_lincheck_prove_union(or_fs* fs, const or_lc_slot* slots, size_t n_slots, int col_vars,
                            int nu, or_f128 z_skip, const or_f128* x_inner_rest,
                            const or_f128* x_outer, int grinding, or_f128* rounds,
                            or_f128 z_partial[64], or_f128* matrix_evals, uint64_t* nonces,
                            or_f128* claim, or_f128* z_vec_out) {
    const int ks = 6;
    if (col_vars < ks || nu < 3 || n_slots == 0) return 1;
    const size_t K = (size_t)1 << col_vars;
    const int L = col_vars - ks;
    int ni = 0;
    static const uint8_t label[] = "flock-lincheck-v0";
    or_fs_observe_label(fs, label, sizeof(label) - 1);
    if (grinding) nonces[ni++] = or_fs_grind_pow(fs, 1);
    const or_f128 alpha = or_fs_sample_f128(fs);
    or_f128* comb = (or_f128*)calloc(K, sizeof(or_f128));
    or_f128* zv = (or_f128*)calloc(K, sizeof(or_f128));
    or_f128** sa = (or_f128**)calloc(n_slots, sizeof(or_f128*));
    or_f128** sb = (or_f128**)calloc(n_slots, sizeof(or_f128*));
    for (size_t t = 0; t < n_slots; t++) {
        const or_lc_slot* sl = &slots[t];
        const size_t Kt = (size_t)1 << sl->k_log;
        const int inner = sl->k_log - ks;
        if (sl->k_log < ks || sl->k_log > col_vars) return 1;
        or_f128* eq_inner = (or_f128*)malloc(sizeof(or_f128) * Kt);
        sa[t] = (or_f128*)malloc(sizeof(or_f128) * Kt);
        sb[t] = (or_f128*)malloc(sizeof(or_f128) * Kt);
        or_quirky_eq(z_skip, x_inner_rest, (size_t)inner, eq_inner);
        if (sl->a_ptr) {
            const csc_ref k = {sl->a_ptr, sl->a_rows, sl->b_ptr, sl->b_rows};
            csc_fold_cb(&k, Kt, eq_inner, sa[t], sb[t]);
        } else {
            keccak_fold_cb(NULL, Kt, eq_inner, sa[t], sb[t]);
        }
        or_f128 w = {1, 0};
        for (int j = 0; j < col_vars - sl->k_log; j++) {
            const or_f128 x = x_inner_rest[inner + j];
            w = or_f128_mul(w, ((sl->prefix >> j) & 1) ? x : add(x, (or_f128){1, 0}));
        }
        const size_t off = (size_t)sl->prefix << sl->k_log;
        for (size_t c = 0; c < Kt; c++)
            comb[off + c] = or_f128_mul(w, add(or_f128_mul(alpha, sa[t][c]), sb[t][c]));
        free(eq_inner);
    }
    for (size_t t = 0; t < n_slots; t++)
        if (slots[t].const_pin >= 0) {
            if (grinding) nonces[ni++] = or_fs_grind_pow(fs, 1);
            const or_f128 beta = or_fs_sample_f128(fs);
            const size_t off = ((size_t)slots[t].prefix << slots[t].k_log) + (size_t)slots[t].const_pin;
            comb[off] = add(comb[off], beta);
        }
    const size_t n_outer = (size_t)1 << nu;
    or_f128* eq_outer = (or_f128*)malloc(sizeof(or_f128) * n_outer);
    or_build_eq(x_outer, (size_t)nu, eq_outer);
    for (size_t t = 0; t < n_slots; t++) {
        const size_t off = (size_t)slots[t].prefix << slots[t].k_log;
        or_partial_fold(slots[t].z, slots[t].k_log, nu, slots[t].useful_bits, slots[t].n_rows, eq_outer,
                        zv + off);
    }
    if (z_vec_out) memcpy(z_vec_out, zv, sizeof(or_f128) * K);
    or_f128* rr = (or_f128*)malloc(sizeof(or_f128) * (L > 0 ? L : 1));
    size_t len = K;
    for (int t = 0; t < L; t++) {
        or_f128 msg[2];
        or_lc_round_eval(comb, zv, len, msg);
        or_fs_observe_f128(fs, msg[0]);
        or_fs_observe_f128(fs, msg[1]);
        if (grinding) nonces[ni++] = or_fs_grind_pow(fs, 2);
        const or_f128 r = or_fs_sample_f128(fs);
        rounds[2 * t] = msg[0];
        rounds[2 * t + 1] = msg[1];
        rr[t] = r;
        or_lc_bind_top(comb, len, r);
        or_lc_bind_top(zv, len, r);
        len /= 2;
    }
    for (int i = 0; i < 64; i++) z_partial[i] = zv[i];
    or_fs_observe_f128_slice(fs, z_partial, 64);
    if (grinding) nonces[ni++] = or_fs_grind_pow(fs, (uint32_t)ks);
    const or_f128 r_skip = or_fs_sample_f128(fs);
    or_f128 lam[64];
    or_lagrange_phi8(0, r_skip, lam);
    or_f128 wv = ZERO;
    for (int i = 0; i < 64; i++) wv = add(wv, or_f128_mul(lam[i], z_partial[i]));
    claim[0] = r_skip;
    claim[1] = wv;
    for (int j = 0; j < L; j++) claim[2 + j] = rr[L - 1 - j];
    for (size_t t = 0; t < n_slots; t++) {
        const int inner = slots[t].k_log - ks;
        or_f128* e = (or_f128*)malloc(sizeof(or_f128) * ((size_t)1 << inner));
        or_build_eq(claim + 2, (size_t)inner, e);
        or_f128 ma = ZERO, mb = ZERO;
        for (size_t j = 0; j < ((size_t)1 << inner); j++)
            for (int i = 0; i < 64; i++) {
                const or_f128 wc = or_f128_mul(e[j], z_partial[i]);
                ma = add(ma, or_f128_mul(wc, sa[t][j * 64 + i]));
                mb = add(mb, or_f128_mul(wc, sb[t][j * 64 + i]));
            }
        matrix_evals[2 * t] = ma;
        matrix_evals[2 * t + 1] = mb;
        free(e);
        free(sa[t]);
        free(sb[t]);
    }
    free(sa);
    free(sb);
    free(rr);
    free(eq_outer);
    free(zv);
    free(comb);
    return 0;
}

/* ======================================================================== */
/* Union instance host glue (crates/flock-core/src/{schedule,union}.rs)     */
/* ======================================================================== */

static unsigned pow2_log(uint64_t n) { /* log2(next_power_of_two(n)) */
    unsigned l = 0;
    while ((1ull << l) < n) l++;
    return l;
}

/* Registry::new (schedule.rs:337-495): class-major stable sort (boolean first, k_log
 * descending), slots packed from 0 (boolean) and element_base (element). order[i] = caller
 * index of slot i. */
int or_registry_layout(const or_table_type* types, size_t n, unsigned nu, or_registry* reg) {
    if (n == 0 || n > OR_MAX_TYPES) return 1;
    for (size_t i = 0; i < n; i++)
        if (types[i].k_log < 7 || types[i].k_log > 40 ||
            types[i].useful_bits > (1ull << types[i].k_log))
            return 1;
    size_t order[OR_MAX_TYPES];
    size_t m = 0;
    for (int cls = 0; cls < 2; cls++)              /* boolean, then element */
        for (unsigned k = 41; k-- > 7;)            /* k_log descending */
            for (size_t i = 0; i < n; i++)         /* stable */
                if ((types[i].is_element != 0) == (cls == 1) && types[i].k_log == k)
                    order[m++] = i;
    uint64_t bool_extent = 0, elem_extent = 0;
    size_t nb = 0;
    for (size_t s = 0; s < n; s++) {
        const or_table_type* t = &types[order[s]];
        if (t->is_element) elem_extent += 1ull << (nu + t->k_log);
        else {
            bool_extent += 1ull << (nu + t->k_log);
            nb++;
        }
    }
    reg->m_bool = bool_extent ? pow2_log(bool_extent) : 0;
    reg->m_elem = elem_extent ? pow2_log(elem_extent) : 0;
    reg->element_base = (elem_extent == 0 || bool_extent == 0)
                            ? 0
                            : 1ull << (reg->m_bool > reg->m_elem ? reg->m_bool : reg->m_elem);
    uint64_t offset = 0;
    for (size_t s = 0; s < n; s++) {
        if (s == nb) offset = reg->element_base;
        const unsigned ms = nu + types[order[s]].k_log;
        if (offset & ((1ull << ms) - 1)) return 1;
        reg->slots[s].type_index = order[s];
        reg->slots[s].m_slot = ms;
        reg->slots[s].offset = offset;
        offset += 1ull << ms;
    }
    reg->m_total = pow2_log(offset);
    for (size_t s = 0; s < n; s++) {
        reg->slots[s].prefix = reg->slots[s].offset >> reg->slots[s].m_slot;
        reg->slots[s].prefix_bits = reg->m_total - reg->slots[s].m_slot;
    }
    reg->n = n;
    reg->nu = nu;
    reg->num_boolean = nb;
    return 0;
}

/* used_cols (union.rs:316-319) */
static uint64_t used_cols(const or_table_type* t) {
    uint64_t u = (t->useful_bits + 127) / 128, cap = 1ull << (t->k_log - 7);
    return u < cap ? u : cap;
}

/* dense_words / committed_words / dense_m / commit_lanes / compaction_is_identity
 * (union.rs:325-416; MIN_DENSE_M = 22, :45; dense_lanes, pcs/commit.rs:428-432). counts in
 * slot order; dense_floor_m = 0 for none; commit_lanes 0 = None. */
int or_union_sizes(const or_table_type* types, const or_registry* reg, const uint64_t* counts,
                   unsigned dense_floor_m, unsigned log_batch_size, or_union_size_info* out) {
    uint64_t dense = 0;
    for (size_t s = 0; s < reg->n; s++) {
        if (counts[s] > (1ull << reg->nu)) return 1;
        dense += used_cols(&types[reg->slots[s].type_index]) * counts[s];
    }
    const uint64_t packed = 1ull << (reg->m_total - 7);
    uint64_t floor = 1ull << (22 - 7);
    if (dense_floor_m && (1ull << (dense_floor_m - 7)) > floor) floor = 1ull << (dense_floor_m - 7);
    uint64_t committed = 1;
    while (committed < dense) committed <<= 1;
    if (committed < floor) committed = floor;
    if (committed > packed) committed = packed;
    unsigned dm = 0;
    while ((1ull << dm) < committed) dm++;
    dm += 7;
    out->dense_words = dense;
    out->committed_words = committed;
    out->packed_len = packed;
    out->dense_m = dm;
    if (dm < 7 + log_batch_size) return 1;
    const unsigned log_dim = dm - 7 - log_batch_size;
    uint64_t t = (dense + (1ull << log_dim) - 1) >> log_dim;
    if (t < 1) t = 1;
    if (t > (1ull << log_batch_size)) t = 1ull << log_batch_size;
    out->commit_lanes = t < (1ull << log_batch_size) ? t : 0;
    int ident = 1;
    uint64_t cursor = 0;
    for (size_t s = 0; s < reg->n; s++) {
        if (cursor != (reg->slots[s].offset >> 7) || counts[s] != (1ull << reg->nu)) {
            ident = 0;
            break;
        }
        cursor += used_cols(&types[reg->slots[s].type_index]) << reg->nu;
    }
    out->compaction_is_identity = ident && committed == packed;
    return 0;
}

/* compact_witness_unchecked (union.rs:449-490): q has committed_words words. */
void or_compact_witness(const or_table_type* types, const or_registry* reg,
                        const uint64_t* counts, const or_f128* z_padded, size_t committed,
                        or_f128* q) {
    size_t cursor = 0;
    for (size_t s = 0; s < reg->n; s++) {
        const uint64_t uc = used_cols(&types[reg->slots[s].type_index]);
        const uint64_t start = reg->slots[s].offset >> 7;
        for (uint64_t c = 0; c < uc; c++)
            for (uint64_t i = 0; i < counts[s]; i++)
                q[cursor++] = z_padded[start + (c << reg->nu) + i];
    }
    for (; cursor < committed; cursor++) q[cursor] = ZERO;
}

/* x_ab_from_mlv (union.rs:541-552): x_inner_rest = [mlv[0], mlv[1+nu..]], x_outer =
 * mlv[1..1+nu]; mlv has m_bool - 6 entries. */
void or_x_ab_from_mlv(const or_f128* mlv, size_t n_mlv, unsigned nu, or_f128* x_inner_rest,
                      or_f128* x_outer) {
    x_inner_rest[0] = mlv[0];
    for (size_t j = 1 + nu; j < n_mlv; j++) x_inner_rest[j - nu] = mlv[j];
    for (unsigned j = 0; j < nu; j++) x_outer[j] = mlv[1 + j];
}

/* bind_statement (union.rs:619-628) given the registry digest bytes. */
void or_bind_statement(or_fs* fs, const uint8_t digest[32], const uint64_t* counts, size_t n,
                       const uint8_t* cap, size_t cap_len) {
    static const uint8_t label[] = "flock-mixed-v1";
    or_fs_observe_label(fs, label, sizeof(label) - 1);
    or_fs_observe_bytes(fs, digest, 32);
    uint8_t* le = (uint8_t*)malloc(8 * (n ? n : 1));
    for (size_t i = 0; i < n; i++)
        for (int b = 0; b < 8; b++) le[8 * i + b] = (uint8_t)(counts[i] >> (8 * b));
    or_fs_observe_bytes(fs, le, 8 * n);
    free(le);
    or_fs_observe_bytes(fs, cap, cap_len);
}
