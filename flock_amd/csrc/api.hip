// Host runtime and C ABI of libflock_hip.so (declared in include/flock_hip.h).
//
// Holds the context (stream + scratch), the NTT basis tables, the host Fiat-Shamir
// transcript, and the protocol drivers that sequence kernels between challenges.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <array>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <condition_variable>
#include <thread>
#include <new>
#include <string>
#include <vector>

#if defined(__x86_64__)
#include <cpuid.h>
#include <emmintrin.h>
#include <immintrin.h>
#include <wmmintrin.h>
#endif

#include "../../include/flock_hip.h"
#include "blake3_common.h"
#include "flock_internal.h"
#include "transcript_device.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            return fail(FLOCK_HIP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

inline bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
inline int log2u(uint64_t x) {
    int l = 0;
    while ((1ull << l) < x) l++;
    return l;
}

// ---------------------------------------------------------------------------
// Host F128 arithmetic (setup work only: NTT basis tables and the r0 factor).
// ---------------------------------------------------------------------------
struct F {
    uint64_t lo, hi;
};

inline void clmul64(uint64_t a, uint64_t b, uint64_t* lo, uint64_t* hi) {
    uint64_t l = 0, h = 0;
    for (int i = 0; i < 64; i++)
        if ((a >> i) & 1) {
            l ^= b << i;
            if (i) h ^= b >> (64 - i);
        }
    *lo = l;
    *hi = h;
}

// ghash reduction identity: U*x^128 = U + U<<1 + U<<2 + U<<7 (mod x^128+x^7+x^2+x+1)
inline F reduce256(uint64_t r0, uint64_t r1, uint64_t r2, uint64_t r3) {
    uint64_t tl = r2 ^ (r2 << 1) ^ (r2 << 2) ^ (r2 << 7);
    uint64_t th = r3 ^ ((r3 << 1) | (r2 >> 63)) ^ ((r3 << 2) | (r2 >> 62)) ^
                  ((r3 << 7) | (r2 >> 57));
    uint64_t ov = (r3 >> 63) ^ (r3 >> 62) ^ (r3 >> 57);
    uint64_t corr = ov ^ (ov << 1) ^ (ov << 2) ^ (ov << 7);
    return F{r0 ^ tl ^ corr, r1 ^ th};
}

#if defined(__x86_64__)
__attribute__((target("pclmul,sse2"))) F mul_pclmul(F a, F b) {
    __m128i A = _mm_set_epi64x((long long)a.hi, (long long)a.lo);
    __m128i B = _mm_set_epi64x((long long)b.hi, (long long)b.lo);
    __m128i ll = _mm_clmulepi64_si128(A, B, 0x00);
    __m128i hh = _mm_clmulepi64_si128(A, B, 0x11);
    __m128i cr = _mm_xor_si128(_mm_clmulepi64_si128(A, B, 0x10), _mm_clmulepi64_si128(A, B, 0x01));
    uint64_t l[2], h[2], c[2];
    _mm_storeu_si128((__m128i*)l, ll);
    _mm_storeu_si128((__m128i*)h, hh);
    _mm_storeu_si128((__m128i*)c, cr);
    return reduce256(l[0], l[1] ^ c[0], h[0] ^ c[1], h[1]);
}
bool have_pclmul() {
    static const bool v = [] {
        unsigned a, b, c, d;
        return __get_cpuid(1, &a, &b, &c, &d) && (c & bit_PCLMUL);
    }();
    return v;
}
#endif

F fmul(F a, F b) {
#if defined(__x86_64__)
    if (have_pclmul()) return mul_pclmul(a, b);
#endif
    uint64_t ll, lh, hl, hh, t0, t1, t2, t3;
    clmul64(a.lo, b.lo, &ll, &t0);
    clmul64(a.lo, b.hi, &lh, &t1);
    clmul64(a.hi, b.lo, &hl, &t2);
    clmul64(a.hi, b.hi, &hh, &t3);
    return reduce256(ll, t0 ^ lh ^ hl, hh ^ t1 ^ t2, t3);
}
inline F fadd(F a, F b) { return F{a.lo ^ b.lo, a.hi ^ b.hi}; }
F finv(F x) {  // x^(2^128 - 2)
    F r{1, 0}, cur = fmul(x, x);
    for (int i = 1; i < 128; i++) {
        r = fmul(r, cur);
        cur = fmul(cur, cur);
    }
    return r;
}

// Normalized subspace-polynomial table for the basis {1, x, .., x^(L-1)}:
// row i = [W_i(b_i), .., W_i(b_{L-1})] / W_i(b_i), rows padded to L entries.
std::vector<F> ntt_evals(int L) {
    std::vector<F> ev((size_t)L * L, F{0, 0});
    for (int k = 0; k < L; k++) ev[k] = F{1ull << k, 0};
    for (int i = 1; i < L; i++) {
        const F* prev = &ev[(size_t)(i - 1) * L];
        F* row = &ev[(size_t)i * L];
        for (int k = 1; k < L - i + 1; k++) row[k - 1] = fmul(prev[k], fadd(prev[k], prev[0]));
    }
    for (int i = 0; i < L; i++) {
        F* row = &ev[(size_t)i * L];
        F inv = finv(row[0]);
        for (int k = 0; k < L - i; k++) row[k] = fmul(row[k], inv);
    }
    return ev;
}

// ---------------------------------------------------------------------------
// Host GF(2^8) (AES polynomial x^8+x^4+x^3+x+1), phi_8 and Lagrange weights (setup work).
// ---------------------------------------------------------------------------
uint8_t f8_mul(uint8_t a, uint8_t b) {
    uint16_t r = 0;
    for (int i = 0; i < 8; i++)
        if ((a >> i) & 1) r ^= (uint16_t)b << i;
    for (int bit = 14; bit >= 8; bit--)
        if ((r >> bit) & 1) r ^= (uint16_t)(0x11B << (bit - 8));
    return (uint8_t)r;
}
uint8_t f8_inv(uint8_t a) {
    uint8_t r = 1, sq = a;
    for (int i = 0; i < 8; i++) {
        if ((0xFE >> i) & 1) r = f8_mul(r, sq);
        sq = f8_mul(sq, sq);
    }
    return r;
}

// phi_8(v) = sum over set bits b of v of alpha^b, alpha = phi_8(0x02) (the F8 -> F128
// embedding of crates/flock-core/src/field/phi8.rs, fixed by its image of x).
const F* phi8_table() {
    static F table[256];
    static std::once_flag once;
    std::call_once(once, [] {
        const F alpha{0x6b8330483c2e9849ull, 0x0dcb364640a222feull};
        F pw[8];
        pw[0] = F{1, 0};
        for (int b = 1; b < 8; b++) pw[b] = fmul(pw[b - 1], alpha);
        for (int v = 0; v < 256; v++) {
            F acc{0, 0};
            for (int b = 0; b < 8; b++)
                if ((v >> b) & 1) acc = fadd(acc, pw[b]);
            table[v] = acc;
        }
    });
    return table;
}

// Lagrange basis L_i(z) on the 64 nodes phi8(offset + i) (offset 0: S, 64: Lambda).
// Closed form L_i(z) = Z(z) / ((z + node_i) * D), Z(z) = prod (z + node_j), D = prod of
// phi8(v), v = 1..63 (the same for S and its coset Lambda); indicator when z is a node.
void lagrange_phi8(int offset, F z, F w[64]) {
    const F* phi = phi8_table();
    F diff[64];
    F Z{1, 0};
    int hit = -1;
    for (int i = 0; i < 64; i++) {
        diff[i] = fadd(z, phi[offset + i]);
        if (diff[i].lo == 0 && diff[i].hi == 0) hit = i;
        Z = fmul(Z, diff[i]);
    }
    if (hit >= 0) {
        for (int i = 0; i < 64; i++) w[i] = F{(uint64_t)(i == hit), 0};
        return;
    }
    // batch inversion of diff[i] * D
    F D{1, 0};
    for (int v = 1; v < 64; v++) D = fmul(D, phi[v]);
    F pre[65];
    pre[0] = F{1, 0};
    for (int i = 0; i < 64; i++) pre[i + 1] = fmul(pre[i], diff[i]);
    F inv_all = finv(fmul(pre[64], D));  // 1 / (D * prod diff)
    F suf{1, 0};
    for (int i = 63; i >= 0; i--) {
        const F inv_i = fmul(fmul(pre[i], suf), inv_all);  // 1 / (diff_i * D)
        w[i] = fmul(Z, inv_i);
        suf = fmul(suf, diff[i]);
    }
}

// ---------------------------------------------------------------------------
// SHA-256 with a copyable running state (the transcript hash).
// ---------------------------------------------------------------------------
const uint32_t K256[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

#if defined(__x86_64__)
// One block on the x86 SHA extensions. The state lives as (A,B,E,F) and (C,D,G,H) halves;
// sha256rnds2 runs two rounds from them and the next two W+K words, msg1 / msg2 extend the
// message schedule four words at a time (W[t] = W[t-16] + s0(W[t-15]) + W[t-7] + s1(W[t-2])).
__attribute__((target("sha,sse4.1,ssse3"))) void sha256_block_ni(uint32_t h[8],
                                                                 const uint8_t* p) {
    const __m128i bswap = _mm_set_epi64x(0x0c0d0e0f08090a0bLL, 0x0405060700010203LL);
    __m128i lo = _mm_shuffle_epi32(_mm_loadu_si128((const __m128i*)&h[0]), 0xB1);  // b a d c
    __m128i hi = _mm_shuffle_epi32(_mm_loadu_si128((const __m128i*)&h[4]), 0x1B);  // h g f e
    __m128i abef = _mm_alignr_epi8(lo, hi, 8);      // f e b a (dword 0 first)
    __m128i cdgh = _mm_blend_epi16(hi, lo, 0xF0);   // h g d c
    const __m128i abef0 = abef, cdgh0 = cdgh;
    __m128i m[4];
    for (int i = 0; i < 4; i++)
        m[i] = _mm_shuffle_epi8(_mm_loadu_si128((const __m128i*)(p + 16 * i)), bswap);
    for (int g = 0; g < 16; g++) {
        __m128i w;
        if (g < 4) {
            w = m[g];
        } else {
            w = _mm_sha256msg2_epu32(
                _mm_add_epi32(_mm_sha256msg1_epu32(m[0], m[1]), _mm_alignr_epi8(m[3], m[2], 4)),
                m[3]);
            m[0] = m[1];
            m[1] = m[2];
            m[2] = m[3];
            m[3] = w;
        }
        __m128i wk = _mm_add_epi32(w, _mm_loadu_si128((const __m128i*)&K256[4 * g]));
        cdgh = _mm_sha256rnds2_epu32(cdgh, abef, wk);  // (A,B,E,F) two rounds on
        wk = _mm_shuffle_epi32(wk, 0x0E);
        abef = _mm_sha256rnds2_epu32(abef, cdgh, wk);
    }
    abef = _mm_add_epi32(abef, abef0);
    cdgh = _mm_add_epi32(cdgh, cdgh0);
    lo = _mm_shuffle_epi32(abef, 0x1B);    // a b e f
    hi = _mm_shuffle_epi32(cdgh, 0xB1);    // g h c d
    _mm_storeu_si128((__m128i*)&h[0], _mm_blend_epi16(lo, hi, 0xF0));  // a b c d
    _mm_storeu_si128((__m128i*)&h[4], _mm_alignr_epi8(hi, lo, 8));    // e f g h
}
// FLOCK_HIP_PORTABLE_SHA256=1 forces the portable block function (to exercise it on hosts
// with the SHA extensions)
bool have_sha_ni() {
    static const bool v = [] {
        const char* force = std::getenv("FLOCK_HIP_PORTABLE_SHA256");
        if (force && force[0] && force[0] != '0') return false;
        unsigned a, b, c, d;
        return __get_cpuid_count(7, 0, &a, &b, &c, &d) && (b & (1u << 29)) &&
               __get_cpuid(1, &a, &b, &c, &d) && (c & bit_SSE4_1) && (c & bit_SSSE3);
    }();
    return v;
}
#endif

void sha256_block(uint32_t h[8], const uint8_t* p);

struct Sha256 {
    uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                     0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    uint64_t total = 0;
    uint8_t buf[64];
    size_t blen = 0;

    static uint32_t rr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
    void block(const uint8_t* p) { sha256_block(h, p); }
    static void block_portable(uint32_t h[8], const uint8_t* p) {
        uint32_t w[64];
        for (int i = 0; i < 16; i++)
            w[i] = (uint32_t)p[4 * i] << 24 | (uint32_t)p[4 * i + 1] << 16 |
                   (uint32_t)p[4 * i + 2] << 8 | p[4 * i + 3];
        for (int i = 16; i < 64; i++)
            w[i] = w[i - 16] + (rr(w[i - 15], 7) ^ rr(w[i - 15], 18) ^ (w[i - 15] >> 3)) +
                   w[i - 7] + (rr(w[i - 2], 17) ^ rr(w[i - 2], 19) ^ (w[i - 2] >> 10));
        uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], k = h[7];
        for (int i = 0; i < 64; i++) {
            uint32_t t1 = k + (rr(e, 6) ^ rr(e, 11) ^ rr(e, 25)) + ((e & f) ^ (~e & g)) + K256[i] + w[i];
            uint32_t t2 = (rr(a, 2) ^ rr(a, 13) ^ rr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
            k = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
        }
        h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += k;
    }
    void update(const uint8_t* d, size_t n) {
        total += n;
        while (n) {
            size_t t = 64 - blen < n ? 64 - blen : n;
            memcpy(buf + blen, d, t);
            blen += t; d += t; n -= t;
            if (blen == 64) { block(buf); blen = 0; }
        }
    }
    void digest(uint8_t out[32]) const {
        uint32_t st[8];
        memcpy(st, h, sizeof(st));
        uint8_t b[64];
        memcpy(b, buf, blen);
        b[blen] = 0x80;
        if (blen >= 56) {
            memset(b + blen + 1, 0, 63 - blen);
            sha256_block(st, b);
            memset(b, 0, 56);
        } else {
            memset(b + blen + 1, 0, 55 - blen);
        }
        const uint64_t bits = total * 8;
        for (int i = 0; i < 8; i++) b[56 + i] = (uint8_t)(bits >> (56 - 8 * i));
        sha256_block(st, b);
        for (int i = 0; i < 8; i++)
            for (int j = 0; j < 4; j++) out[4 * i + j] = (uint8_t)(st[i] >> (24 - 8 * j));
    }
};

void sha256_block(uint32_t h[8], const uint8_t* p) {
#if defined(__x86_64__)
    if (have_sha_ni()) {
        sha256_block_ni(h, p);
        return;
    }
#endif
    Sha256::block_portable(h, p);
}

// PoW predicate (challenger.rs:884-958): SHA256(digest || nonce_le) starts with `bits` zero
// bits, MSB first. `blk` is the padded one-block message with the digest in bytes 0..31; the
// nonce goes into bytes 32..39.
constexpr uint32_t kMaxPowBits = 64;
bool pow_ok(uint8_t blk[64], uint64_t nonce, uint32_t bits) {
    if (bits > 256) return false;
    for (int i = 0; i < 8; i++) blk[32 + i] = (uint8_t)(nonce >> (8 * i));
    uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                     0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    sha256_block(h, blk);
    for (int i = 0; i < 8 && bits; i++) {
        const uint32_t take = bits < 32 ? bits : 32;
        if (take == 32 ? h[i] != 0 : (h[i] >> (32 - take)) != 0) return false;
        bits -= take;
    }
    return true;
}

void pow_block(const uint8_t digest[32], uint8_t blk[64]) {
    memcpy(blk, digest, 32);
    memset(blk + 32, 0, 32);
    blk[40] = 0x80;
    blk[62] = 0x01;  // 320 bits, big-endian in bytes 56..63
    blk[63] = 0x40;
}

// Parallel search for the smallest passing nonce, for the grinds the first kPowSerialTries
// nonces do not settle (waking the workers costs about as much as 500 tries). The caller and whichever pool workers join claim blocks of kPowBlock
// nonces in increasing order and stop claiming once a block starts past the best nonce found;
// every block below the best is searched up to its first hit, so the result is the smallest
// passing nonce, the sequential search's. A worker joins a search only while it is open; the
// caller closes it when its own share is done and waits for the workers that joined, never for
// sleeping ones. One search uses the pool at a time (a caller that finds it busy searches
// alone). Idle workers spin for a while after a search (grinds come in bursts), then sleep.
constexpr uint64_t kPowBlock = 64;
constexpr uint64_t kPowSerialTries = 512;  // a multiple of kPowBlock
// Proofs in flight in this process (flock::ProofScope); with more than one, their own threads
// keep the host cores busy and every grind searches alone.
std::atomic<int> g_proofs_in_flight{0};
class PowPool {
  public:
    static PowPool& get() {
        static PowPool p;
        return p;
    }
    // false when the pool is busy or has no workers; *out = the smallest passing nonce >= start
    // (a multiple of kPowBlock) otherwise
    bool search(const uint8_t blk[64], uint32_t bits, uint64_t start, uint64_t* out) {
        if (th_.empty() || !busy_.try_lock()) return false;
        memcpy(blk_, blk, 64);
        bits_ = bits;
        next_.store(start / kPowBlock);
        best_.store(UINT64_MAX);
        const uint64_t g = gen_.load() + 1;
        open_.store(g, std::memory_order_release);
        {
            std::lock_guard<std::mutex> lk(m_);
            gen_.store(g, std::memory_order_release);
        }
        cv_.notify_all();
        run();
        open_.store(0, std::memory_order_release);  // no new joiners
        while (inflight_.load(std::memory_order_acquire) != 0) _mm_pause();
        *out = best_.load();
        busy_.unlock();
        return true;
    }
    ~PowPool() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_.store(true);
        }
        cv_.notify_all();
        for (std::thread& t : th_) t.join();
    }

  private:
    PowPool() {
        const unsigned hw = std::thread::hardware_concurrency();
        const unsigned n = hw > 1 ? std::min(hw - 1, 7u) : 0u;
        try {
            for (unsigned i = 0; i < n; i++) th_.emplace_back([this] { worker(); });
        } catch (...) {
            // fewer (or no) workers: searches run with what started, or alone
        }
    }
    void run() {
        uint8_t b[64];
        memcpy(b, blk_, 64);
        for (;;) {
            const uint64_t base = next_.fetch_add(1) * kPowBlock;
            if (base >= best_.load()) break;
            for (uint64_t n = base; n < base + kPowBlock; n++)
                if (pow_ok(b, n, bits_)) {
                    uint64_t cur = best_.load();
                    while (n < cur && !best_.compare_exchange_weak(cur, n)) {
                    }
                    break;
                }
        }
    }
    void worker() {
        uint64_t seen = 0;
        for (;;) {
            // spin about 500 us for the next search, then sleep
            const auto t0 = std::chrono::steady_clock::now();
            while (gen_.load(std::memory_order_acquire) == seen && !stop_.load()) {
                _mm_pause();
                if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(500)) {
                    std::unique_lock<std::mutex> lk(m_);
                    cv_.wait(lk, [&] { return gen_.load() != seen || stop_.load(); });
                }
            }
            if (stop_.load()) return;
            seen = gen_.load(std::memory_order_acquire);
            inflight_.fetch_add(1, std::memory_order_acq_rel);
            if (open_.load(std::memory_order_acquire) == seen) run();
            inflight_.fetch_sub(1, std::memory_order_release);
        }
    }
    std::vector<std::thread> th_;
    std::mutex busy_, m_;
    std::condition_variable cv_;
    std::atomic<uint64_t> gen_{0}, open_{0};
    std::atomic<int> inflight_{0};
    std::atomic<bool> stop_{false};
    uint8_t blk_[64] = {};
    uint32_t bits_ = 0;
    std::atomic<uint64_t> next_{0}, best_{UINT64_MAX};
};

}  // namespace

// ---------------------------------------------------------------------------
// Fiat-Shamir transcript: FsChallenger with the SHA-256 discipline
// (crates/flock-core/src/challenger.rs:609-1044). Every op absorbs a 16-byte header
// [op][kind][0 x6][len u64 LE] (:708-714); byte payloads are zero-padded to 16 (:720-726);
// a squeeze is SHA256(state || ctr_le) blocks without mutating the state (:745-763).
// ---------------------------------------------------------------------------
struct flock_fs {
    Sha256 st;
    void header(uint8_t op, uint8_t kind, uint64_t len) {
        uint8_t h[16] = {0};
        h[0] = op;
        h[1] = kind;
        for (int i = 0; i < 8; i++) h[8 + i] = (uint8_t)(len >> (8 * i));
        st.update(h, 16);
    }
    void padded(const uint8_t* b, size_t n) {
        static const uint8_t z[16] = {0};
        st.update(b, n);
        if (n % 16) st.update(z, 16 - n % 16);
    }
    void f128(flock_f128 v) {
        uint8_t b[16];
        for (int i = 0; i < 8; i++) {
            b[i] = (uint8_t)(v.lo >> (8 * i));
            b[8 + i] = (uint8_t)(v.hi >> (8 * i));
        }
        st.update(b, 16);
    }
    void squeeze(uint8_t* out, size_t n) const {
        size_t off = 0;
        uint64_t ctr = 0;
        while (off < n) {
            Sha256 c = st;
            uint8_t cb[8];
            for (int i = 0; i < 8; i++) cb[i] = (uint8_t)(ctr >> (8 * i));
            c.update(cb, 8);
            uint8_t blk[32];
            c.digest(blk);
            size_t t = n - off < 32 ? n - off : 32;
            memcpy(out + off, blk, t);
            off += t;
            ctr++;
        }
    }
};

namespace {
enum : uint8_t { OP_DOMAIN = 1, OP_LABEL = 2, OP_OBSERVE = 3, OP_SQUEEZE = 4, OP_BYTES = 5 };
enum : uint8_t { KIND_NONE = 0, KIND_SCALAR = 1, KIND_SLICE = 2 };

flock_f128 le128(const uint8_t* b) {
    flock_f128 v{0, 0};
    for (int i = 0; i < 8; i++) {
        v.lo |= (uint64_t)b[i] << (8 * i);
        v.hi |= (uint64_t)b[8 + i] << (8 * i);
    }
    return v;
}
}  // namespace

extern "C" {

flock_fs* flock_fs_new(const uint8_t* domain, size_t len) {
    flock_fs* fs = new (std::nothrow) flock_fs();
    if (!fs) return nullptr;
    fs->header(OP_DOMAIN, KIND_NONE, len);
    fs->padded(domain, len);
    return fs;
}
void flock_fs_free(flock_fs* fs) { delete fs; }
void flock_fs_observe_label(flock_fs* fs, const uint8_t* l, size_t n) {
    fs->header(OP_LABEL, KIND_NONE, n);
    fs->padded(l, n);
}
void flock_fs_observe_f128(flock_fs* fs, flock_f128 v) {
    fs->header(OP_OBSERVE, KIND_SCALAR, 1);
    fs->f128(v);
}
void flock_fs_observe_f128_slice(flock_fs* fs, const flock_f128* v, size_t n) {
    fs->header(OP_OBSERVE, KIND_SLICE, n);
    for (size_t i = 0; i < n; i++) fs->f128(v[i]);
}
void flock_fs_observe_bytes(flock_fs* fs, const uint8_t* b, size_t n) {
    fs->header(OP_BYTES, KIND_NONE, n);
    fs->padded(b, n);
}
void flock_fs_sample_f128(flock_fs* fs, flock_f128* out) {
    fs->header(OP_SQUEEZE, KIND_SCALAR, 1);
    uint8_t b[16];
    fs->squeeze(b, 16);
    *out = le128(b);
}
void flock_fs_sample_f128_vec(flock_fs* fs, flock_f128* out, size_t n) {
    fs->header(OP_SQUEEZE, KIND_SLICE, n);
    std::vector<uint8_t> b(n * 16 + 1);
    fs->squeeze(b.data(), n * 16);
    for (size_t i = 0; i < n; i++) out[i] = le128(b.data() + 16 * i);
}
void flock_fs_digest(flock_fs* fs, uint8_t out[32]) { fs->st.digest(out); }
// PoW (challenger.rs:884-958, 1096-1135): smallest nonce with SHA256(digest || nonce_le)
// having `bits` leading zero bits (MSB-first), then observe_bytes(nonce_le).
// Challenger::verify_pow (challenger.rs:960-981): bits = 0 accepts only nonce 0; otherwise
// SHA256(digest || nonce_le) must start with `bits` zero bits. The nonce bytes are absorbed
// either way.
int flock_fs_verify_pow(flock_fs* fs, uint64_t nonce, uint32_t bits) {
    uint8_t digest[32];
    fs->st.digest(digest);
    bool ok = nonce == 0;
    if (bits) {
        uint8_t blk[64];
        pow_block(digest, blk);
        ok = pow_ok(blk, nonce, bits);
    }
    uint8_t nb[8];
    for (int i = 0; i < 8; i++) nb[i] = (uint8_t)(nonce >> (8 * i));
    flock_fs_observe_bytes(fs, nb, 8);
    return ok ? 1 : 0;
}

uint64_t flock_fs_grind_pow(flock_fs* fs, uint32_t bits) {
    if (bits > kMaxPowBits) {
        fail(FLOCK_HIP_E_INVALID, "grind_pow: more than 64 bits");
        return UINT64_MAX;  // nothing absorbed
    }
    uint8_t digest[32];
    fs->st.digest(digest);
    uint64_t nonce = 0;
    if (bits) {
        uint8_t blk[64];
        pow_block(digest, blk);
        while (nonce < kPowSerialTries && !pow_ok(blk, nonce, bits)) nonce++;
        if (nonce == kPowSerialTries &&
            (g_proofs_in_flight.load(std::memory_order_relaxed) > 1 ||
             !PowPool::get().search(blk, bits, nonce, &nonce)))
            while (!pow_ok(blk, nonce, bits)) nonce++;
    }
    uint8_t nb[8];
    for (int i = 0; i < 8; i++) nb[i] = (uint8_t)(nonce >> (8 * i));
    flock_fs_observe_bytes(fs, nb, 8);
    return nonce;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct flock_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // a second stream for phases independent of the main chain, joined through events
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // aux waits for everything enqueued on stream so far
    int fork() {
        HIP_TRY(hipEventRecord(ev_fork, stream));
        HIP_TRY(hipStreamWaitEvent(aux, ev_fork, 0));
        return 0;
    }
    // stream waits for everything enqueued on aux so far
    int join() {
        HIP_TRY(hipEventRecord(ev_join, aux));
        HIP_TRY(hipStreamWaitEvent(stream, ev_join, 0));
        return 0;
    }
    // two-round (lookahead) reductions: 32 device words and a ticket counter (zero between
    // launches), published to the coherent pinned hla (32 words, then the flag word)
    uint32_t* la_acc = nullptr;
    uint32_t* hla = nullptr;
    uint32_t* dla = nullptr;
    uint32_t la_seq = 0;
    flock::LaOut la_out() {
        if (++la_seq == 0) ++la_seq;
        return flock::LaOut{la_acc, dla, dla + 32, la_seq};
    }
    int await_la(flock_f128 out[8]);
    uint4* partials = nullptr;  // kMaxReduceBlocks * 2 words, then the grid-reduce ticket counter
    uint4* dout = nullptr;      // device view of hout: round kernels write their message there
    flock_f128* hout = nullptr; // pinned, coherent: the pair, then a flag word
    std::map<int, uint4*> evals;  // k -> k x k basis table
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    // zerocheck round-1 constants (device): extension table, log/exp, phi8(2^b)
    uint4* zc_ext = nullptr;
    uint8_t* zc_logexp = nullptr;
    uint4* zc_phipow = nullptr;
    uint4* zc_ext_bs = nullptr;    // bit-sliced extension table (same shape as zc_ext)
    uint4* zc_phim = nullptr;      // phi8(L_s(64 + i)) at s * 64 + i
    uint4* zc_partials = nullptr;  // kZcMaxBlocks groups x 1024 accumulators
    uint4* zc_partials2 = nullptr; // the same for a kernel on the aux stream
    uint4* zc_out = nullptr;       // 3072 words: message [0,128), shat [128,256), red [1024,2048),
                                   // red2 [2048,3072)
    flock_f128* zc_host = nullptr;  // pinned, 256 words: message and shat
    // Keccak-f[1600] lincheck walker tables (device): phi^T gathers, chi row offsets, RC_r
    uint16_t* kc_inv = nullptr;
    uint16_t* kc_trow = nullptr;
    uint32_t* kc_rcm = nullptr;
    int ensure_kc();
    uint4* zc_shat() { return zc_out + 128; }
    uint4* zc_red() { return zc_out + 1024; }
    uint4* zc_red2() { return zc_out + 2048; }

    int ensure_zc() {
        if (zc_ext) return 0;
        // Extension matrix: M[s][i] = L_s(64 + i) over F8, nodes S = {0..63}: the S -> Lambda
        // extension of the unit vector e_s (inverse NTT on S, forward on Lambda).
        uint8_t M[64][64];
        for (int s0 = 0; s0 < 64; s0++) {
            uint8_t den = 1;
            for (int j = 0; j < 64; j++)
                if (j != s0) den = f8_mul(den, (uint8_t)(s0 ^ j));
            const uint8_t dinv = f8_inv(den);
            for (int i = 0; i < 64; i++) {
                uint8_t num = 1;
                const int x = 64 + i;
                for (int j = 0; j < 64; j++)
                    if (j != s0) num = f8_mul(num, (uint8_t)(x ^ j));
                M[s0][i] = f8_mul(num, dinv);
            }
        }
        // nibble tables: entry (g, v) = XOR of M[4g + b] over set bits b of v (64 bytes)
        std::vector<uint8_t> ext(16 * 16 * 64, 0);
        for (int g = 0; g < 16; g++)
            for (int v = 0; v < 16; v++)
                for (int bb = 0; bb < 4; bb++)
                    if ((v >> bb) & 1)
                        for (int i = 0; i < 64; i++) ext[(g * 16 + v) * 64 + i] ^= M[4 * g + bb][i];
        // the same entries bit-sliced: word 2b + h, bit q = bit b of byte 32h + q
        std::vector<uint32_t> ext_bs(16 * 16 * 16, 0);
        for (int e = 0; e < 16 * 16; e++)
            for (int i = 0; i < 64; i++)
                for (int bb = 0; bb < 8; bb++)
                    if ((ext[e * 64 + i] >> bb) & 1)
                        ext_bs[e * 16 + 2 * bb + (i >> 5)] |= 1u << (i & 31);
        const F* phi0 = phi8_table();
        std::vector<F> phim(64 * 64);
        for (int s0 = 0; s0 < 64; s0++)
            for (int i = 0; i < 64; i++) phim[s0 * 64 + i] = phi0[M[s0][i]];
        // log/exp with generator 0x03; exp doubled so log[u] + log[v] needs no reduction
        std::vector<uint8_t> le(256 + 512, 0);
        uint8_t x = 1;
        for (int i = 0; i < 255; i++) {
            le[256 + i] = x;
            le[256 + 255 + i] = x;
            le[x] = (uint8_t)i;
            x = f8_mul(x, 3);
        }
        const F* phi = phi8_table();
        F pp[8];
        for (int bb = 0; bb < 8; bb++) pp[bb] = phi[1 << bb];
        if (hipMalloc(&zc_ext, ext.size()) != hipSuccess ||
            hipMalloc(&zc_ext_bs, ext_bs.size() * 4) != hipSuccess ||
            hipMalloc(&zc_phim, phim.size() * 16) != hipSuccess ||
            hipMalloc(&zc_logexp, le.size()) != hipSuccess ||
            hipMalloc(&zc_phipow, sizeof(pp)) != hipSuccess ||
            hipMalloc(&zc_partials, sizeof(uint4) * flock::kZcMaxBlocks * 1024) != hipSuccess ||
            hipMalloc(&zc_partials2, sizeof(uint4) * flock::kZcMaxBlocks * 1024) != hipSuccess ||
            hipMalloc(&zc_out, sizeof(uint4) * 3072) != hipSuccess ||
            hipHostMalloc(&zc_host, sizeof(flock_f128) * 256) != hipSuccess) {
            (void)hipGetLastError();
            return fail(FLOCK_HIP_E_NOMEM, "zerocheck tables");
        }
        HIP_TRY(hipMemcpy(zc_ext, ext.data(), ext.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(zc_ext_bs, ext_bs.data(), ext_bs.size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(zc_phim, phim.data(), phim.size() * 16, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(zc_logexp, le.data(), le.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(zc_phipow, pp, sizeof(pp), hipMemcpyHostToDevice));
        return 0;
    }

    int ensure_scratch(size_t bytes) {
        if (scratch_bytes >= bytes) return 0;
        if (scratch) {
            hipStreamSynchronize(stream);  // kernels may still read the old scratch
            hipFree(scratch);
        }
        scratch = nullptr;
        scratch_bytes = 0;
        if (hipMalloc(&scratch, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return fail(FLOCK_HIP_E_NOMEM, "scratch allocation of " + std::to_string(bytes) +
                                               " bytes failed");
        }
        scratch_bytes = bytes;
        return 0;
    }
    int get_evals(int k, const uint4** out) {
        auto it = evals.find(k);
        if (it != evals.end()) {
            *out = it->second;
            return 0;
        }
        std::vector<F> ev = ntt_evals(k);
        uint4* d = nullptr;
        HIP_TRY(hipMalloc(&d, ev.size() * sizeof(F)));
        HIP_TRY(hipMemcpy(d, ev.data(), ev.size() * sizeof(F), hipMemcpyHostToDevice));
        evals[k] = d;
        *out = d;
        return 0;
    }
    // Device-resident transcript for the sumcheck rounds: the state, then a log of
    // kFsLogRounds rounds (device), mirrored in pinned host memory for upload and read-back.
    static constexpr uint32_t kFsLogRounds = 96;
    flock::DevTranscript* dfs = nullptr;
    uint8_t* hfs = nullptr;
    int round_mode = FLOCK_ROUNDS_HOST;
    // commits build their trees up to the cap layer only (flock::ctx_set_tree_to_cap)
    bool tree_to_cap = false;
    static constexpr size_t fs_bytes() {
        return sizeof(flock::DevTranscript) + kFsLogRounds * sizeof(flock::DevRound);
    }
    const flock::DevRound* fs_log() const {
        return reinterpret_cast<const flock::DevRound*>(hfs + sizeof(flock::DevTranscript));
    }
    // device address of round k's challenge
    const uint4* fs_rho(uint32_t k) const {
        return &reinterpret_cast<const flock::DevRound*>(dfs + 1)[k].rho;
    }
    // Whether `rounds` rounds with `bits` of PoW each can run on the device transcript: the
    // host state must sit on a 16-byte boundary (every framed op keeps it there).
    bool fs_usable(const flock_fs* fs, uint32_t bits, uint32_t rounds) const;
    int fs_begin(const flock_fs* fs, uint32_t bits);
    int fs_end(uint32_t rounds);

    // launch result pair -> host
    // The message kernel wrote its pair through dout straight into the coherent pinned hout,
    // so a round costs one launch and one wait, with no copy.
    int fetch_pair(flock_f128 out[2]) {
        HIP_TRY(hipStreamSynchronize(stream));
        out[0] = hout[0];
        out[1] = hout[1];
        return 0;
    }
    // A round kernel's pair, announced through the flag word after it (hout[2]): the next
    // sequence number goes into the kernel's RoundOut, and the host polls the word instead of
    // synchronising the stream.
    uint32_t pair_seq = 0;
    flock::RoundOut polled_out() {
        if (++pair_seq == 0) ++pair_seq;
        return flock::RoundOut{dout, reinterpret_cast<uint32_t*>(dout + 2), pair_seq, nullptr};
    }
    int await_pair(flock_f128 out[2]);
};

namespace {
inline uint4 U(flock_f128 v) {
    return make_uint4((uint32_t)v.lo, (uint32_t)(v.lo >> 32), (uint32_t)v.hi,
                      (uint32_t)(v.hi >> 32));
}
inline const uint4* U(const flock_f128* p) { return reinterpret_cast<const uint4*>(p); }
inline uint4* U(flock_f128* p) { return reinterpret_cast<uint4*>(p); }
inline flock_f128 mulh(flock_f128 a, flock_f128 b) {
    F r = fmul(F{a.lo, a.hi}, F{b.lo, b.hi});
    return flock_f128{r.lo, r.hi};
}
#define CTX_CHECK(ctx) \
    if (!(ctx)) return fail(FLOCK_HIP_E_INVALID, "null context")
#define LAUNCH(expr) HIP_TRY(expr)
}  // namespace

bool flock_hip_ctx::fs_usable(const flock_fs* fs, uint32_t bits, uint32_t rounds) const {
    return fs && fs->st.blen % 16 == 0 && bits <= flock::kFsMaxGrindBits &&
           rounds <= kFsLogRounds;
}

int flock_hip_ctx::fs_begin(const flock_fs* fs, uint32_t bits) {
    if (!dfs) {
        if (hipMalloc(&dfs, fs_bytes()) != hipSuccess ||
            hipHostMalloc((void**)&hfs, fs_bytes()) != hipSuccess) {
            (void)hipGetLastError();
            return fail(FLOCK_HIP_E_NOMEM, "device transcript");
        }
    }
    // the previous read-back has completed (fs_end synchronises), so the staging is free
    flock::DevTranscript t{};
    for (int i = 0; i < 8; i++) t.h[i] = fs->st.h[i];
    for (size_t i = 0; i < fs->st.blen / 4; i++) {
        const uint8_t* b = fs->st.buf + 4 * i;
        t.buf[i] = (uint32_t)b[0] << 24 | (uint32_t)b[1] << 16 | (uint32_t)b[2] << 8 | b[3];
    }
    t.total = fs->st.total;
    t.blen = (uint32_t)fs->st.blen;
    t.round = 0;
    t.grind_bits = bits;
    t.max_rounds = kFsLogRounds;
    std::memcpy(hfs, &t, sizeof(t));
    HIP_TRY(hipMemcpyAsync(dfs, hfs, sizeof(t), hipMemcpyHostToDevice, stream));
    return 0;
}

int flock_hip_ctx::fs_end(uint32_t rounds) {
    HIP_TRY(hipMemcpyAsync(hfs + sizeof(flock::DevTranscript), dfs + 1,
                           rounds * sizeof(flock::DevRound), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return 0;
}

// Poll the flag word after hout for the pair announced with pair_seq. Bounded: past 10 s, or
// once the stream reports an error, the wait ends with an error.
int flock_hip_ctx::await_pair(flock_f128 out[2]) {
    const uint32_t* flag = reinterpret_cast<const uint32_t*>(hout + 2);
    const auto t0 = std::chrono::steady_clock::now();
    for (uint64_t it = 1;; it++) {
        if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == pair_seq) break;
        if ((it & 4095) == 0) {
            const hipError_t q = hipStreamQuery(stream);
            if (q != hipSuccess && q != hipErrorNotReady)
                return fail(FLOCK_HIP_E_HIP, std::string("round kernel: ") + hipGetErrorString(q));
            if (q == hipSuccess && __atomic_load_n(flag, __ATOMIC_ACQUIRE) != pair_seq)
                return fail(FLOCK_HIP_E_INTERNAL, "round kernel finished without its message");
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(10))
                return fail(FLOCK_HIP_E_INTERNAL, "round kernel: no message within 10 s");
        }
        _mm_pause();
    }
    out[0] = hout[0];
    out[1] = hout[1];
    return 0;
}

// Poll hla's flag word for the eight sums published with la_seq (bounded like await_pair).
int flock_hip_ctx::await_la(flock_f128 out[8]) {
    const uint32_t* flag = hla + 32;
    const auto t0 = std::chrono::steady_clock::now();
    for (uint64_t it = 1;; it++) {
        if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == la_seq) break;
        if ((it & 4095) == 0) {
            const hipError_t q = hipStreamQuery(stream);
            if (q != hipSuccess && q != hipErrorNotReady)
                return fail(FLOCK_HIP_E_HIP, std::string("round kernel: ") + hipGetErrorString(q));
            if (q == hipSuccess && __atomic_load_n(flag, __ATOMIC_ACQUIRE) != la_seq)
                return fail(FLOCK_HIP_E_INTERNAL, "round kernel finished without its sums");
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(10))
                return fail(FLOCK_HIP_E_INTERNAL, "round kernel: no sums within 10 s");
        }
        _mm_pause();
    }
    for (int k = 0; k < 8; k++)
        out[k] = flock_f128{(uint64_t)hla[4 * k + 1] << 32 | hla[4 * k], (uint64_t)hla[4 * k + 3] << 32 | hla[4 * k + 2]};
    return 0;
}

namespace {
// Advance the host transcript through one round the device transcript ran, checking that
// both drew the same nonce and challenge.
int fs_replay(flock_fs* fs, const flock::DevRound& r, uint32_t bits, flock_f128 msg[2],
              uint64_t* nonce, flock_f128* rho) {
    msg[0] = flock_f128{(uint64_t)r.m0.y << 32 | r.m0.x, (uint64_t)r.m0.w << 32 | r.m0.z};
    msg[1] = flock_f128{(uint64_t)r.m1.y << 32 | r.m1.x, (uint64_t)r.m1.w << 32 | r.m1.z};
    flock_fs_observe_f128(fs, msg[0]);
    flock_fs_observe_f128(fs, msg[1]);
    if (bits) {
        const uint64_t n = flock_fs_grind_pow(fs, bits);
        if (r.err || n != r.nonce)
            return fail(FLOCK_HIP_E_INTERNAL, "device transcript: proof-of-work nonce differs");
        if (nonce) *nonce = n;
    }
    flock_fs_sample_f128(fs, rho);
    const flock_f128 d{(uint64_t)r.rho.y << 32 | r.rho.x, (uint64_t)r.rho.w << 32 | r.rho.z};
    if (rho->lo != d.lo || rho->hi != d.hi)
        return fail(FLOCK_HIP_E_INTERNAL, "device transcript: challenge differs");
    return 0;
}
}  // namespace

namespace {

inline flock_f128 from_u4(uint4 v) {
    return flock_f128{(uint64_t)v.y << 32 | v.x, (uint64_t)v.w << 32 | v.z};
}

// A loop of n sumcheck rounds on the context's stream. launch(j, out, rin, rho) enqueues
// round j's kernel: for j < n it emits message j to `out`; for j >= 1 it consumes challenge
// j-1, from the device pointer rin when that is set, else the value rho; launch(n, ...) is
// the closing fold. Each message is observed on `fs`, ground (bits > 0) and answered with a
// challenge, in the context's round mode (flock_hip_ctx_set_round_mode).

struct RoundLoop {
    using Launch = std::function<int(uint32_t, flock::RoundOut, const uint4*, uint4)>;
    flock_hip_ctx* c;
    flock_fs* fs;
    uint32_t n, bits;
    Launch launch;
    std::vector<flock_f128> msgs;  // 2n
    std::vector<flock_f128> rhos;  // n
    std::vector<uint64_t> nonces;  // n, when bits > 0
    int mode = FLOCK_ROUNDS_HOST;
    // Two rounds per launch from round la_from on, in the host round mode: la(j, nfold, r1, r2,
    // out) folds the tables by the nfold (1 or 2) pending challenges and, for j < n, returns
    // round j's pair and round j+1's as quadratics in its challenge (launch_zc_fold_la's eight
    // sums); at j == n it is the closing fold. Round j+1's pair then needs no launch.
    std::function<int(uint32_t, int, uint4, uint4, flock_f128*)> la;
    uint32_t la_from = UINT32_MAX;

    void host_round(uint32_t k) {
        flock_fs_observe_f128(fs, msgs[2 * k]);
        flock_fs_observe_f128(fs, msgs[2 * k + 1]);
        if (bits) nonces[k] = flock_fs_grind_pow(fs, bits);
        flock_fs_sample_f128(fs, &rhos[k]);
    }

    int run() {
        msgs.assign(2 * n, flock_f128{0, 0});
        rhos.assign(n, flock_f128{0, 0});
        nonces.assign(n, 0);
        const uint4 zero = make_uint4(0, 0, 0, 0);
        mode = c->round_mode;
        if (mode != FLOCK_ROUNDS_HOST && !c->fs_usable(fs, bits, n)) mode = FLOCK_ROUNDS_HOST;
        if (mode == FLOCK_ROUNDS_DEVICE) {
            if (int rc = c->fs_begin(fs, bits)) return rc;
            const flock::RoundOut ro{nullptr, nullptr, 0u, c->dfs};
            for (uint32_t j = 0; j <= n; j++)
                if (int rc = launch(j, ro, j ? c->fs_rho(j - 1) : nullptr, zero)) return rc;
            if (int rc = c->fs_end(n)) return rc;
            for (uint32_t k = 0; k < n; k++)
                if (int rc = fs_replay(fs, c->fs_log()[k], bits, &msgs[2 * k], &nonces[k],
                                       &rhos[k]))
                    return rc;
            return 0;
        }
        if (int rc = launch(0, c->polled_out(), nullptr, zero)) return rc;
        flock_f128 s8[8];
        bool from_la = false, look = false;
        uint4 pend[2] = {zero, zero};
        int np = 0;
        for (uint32_t k = 0; k < n; k++) {
            if (look) {  // a quadratic from its values at 0 and 1 and its leading coefficient
                const F r{rhos[k - 1].lo, rhos[k - 1].hi};
                auto quad = [&](const flock_f128* p) {
                    const F p0{p[0].lo, p[0].hi}, p1{p[1].lo, p[1].hi}, pi{p[2].lo, p[2].hi};
                    const F v = fadd(fadd(p0, fmul(fadd(fadd(p0, p1), pi), r)), fmul(pi, fmul(r, r)));
                    return flock_f128{v.lo, v.hi};
                };
                msgs[2 * k] = quad(s8 + 2);
                msgs[2 * k + 1] = quad(s8 + 5);
                look = false;
            } else if (from_la) {
                msgs[2 * k] = s8[0];
                msgs[2 * k + 1] = s8[1];
                from_la = false;
                look = k + 1 < n;
            } else if (int rc = c->await_pair(&msgs[2 * k])) {
                return rc;
            }
            host_round(k);
            pend[np++] = U(rhos[k]);
            if (look) continue;  // round k+1's pair is already known
            const uint32_t j = k + 1;
            if (la && j >= la_from) {
                if (int rc = la(j, np, pend[0], pend[1], s8)) return rc;
                from_la = j < n;
            } else if (int rc = launch(j, j < n ? c->polled_out() : flock::RoundOut{}, nullptr, pend[0])) {
                return rc;
            }
            np = 0;
        }
        return 0;
    }
};
}  // namespace

// ---------------------------------------------------------------------------
// Launch tunables: name, default, allowed range, current value (process-wide)
// ---------------------------------------------------------------------------
namespace {
struct TuneEntry {
    const char* name;
    int64_t def, lo, hi;
    std::atomic<int64_t> v;
};
TuneEntry g_tune[flock::kTuneCount] = {
    {"ntt_cols_per_thread", 2, 1, 4, {2}},
    {"ntt_deep_cols", 256, 32, 256, {256}},
    {"zc_split_min_vars", 16, 10, 24, {16}},
    {"zc_la_max_log", 17, 10, 22, {17}},
    {"sc_la_max_log", 18, 10, 24, {18}},
    {"open_split_fold_log", 16, 10, 22, {16}},
};
}  // namespace

int64_t flock::tunable(flock::Tunable t) { return g_tune[t].v.load(std::memory_order_relaxed); }

flock::ProofScope::ProofScope() { g_proofs_in_flight.fetch_add(1); }
flock::ProofScope::~ProofScope() { g_proofs_in_flight.fetch_sub(1); }

extern "C" {

int flock_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

const char* flock_hip_last_error(void) { return g_err.c_str(); }

}  // extern "C"

int flock::set_error(int code, const char* msg) { return fail(code, msg ? msg : ""); }

uint64_t flock::ctx_scratch_bytes(const flock_hip_ctx* c) { return c ? c->scratch_bytes : 0; }

void flock::sha256_host(const void* data, size_t n, uint8_t out[32]) {
    Sha256 h;
    h.update(static_cast<const uint8_t*>(data), n);
    h.digest(out);
}

void flock::ctx_release_scratch(flock_hip_ctx* c) {
    if (!c || !c->scratch) return;
    hipStreamSynchronize(c->stream);
    hipFree(c->scratch);
    c->scratch = nullptr;
    c->scratch_bytes = 0;
}

extern "C" {

int flock_hip_set_tunable(const char* name, int64_t value) {
    if (!name) return fail(FLOCK_HIP_E_INVALID, "null tunable name");
    for (TuneEntry& e : g_tune)
        if (std::strcmp(e.name, name) == 0) {
            if (value < e.lo || value > e.hi)
                return fail(FLOCK_HIP_E_INVALID, std::string("tunable ") + name + " out of range");
            e.v.store(value);
            return 0;
        }
    return fail(FLOCK_HIP_E_INVALID, std::string("unknown tunable ") + name);
}

int flock_hip_get_tunable(const char* name, int64_t* value) {
    if (!name || !value) return fail(FLOCK_HIP_E_INVALID, "null tunable argument");
    for (TuneEntry& e : g_tune)
        if (std::strcmp(e.name, name) == 0) {
            *value = e.v.load();
            return 0;
        }
    return fail(FLOCK_HIP_E_INVALID, std::string("unknown tunable ") + name);
}

int flock_hip_ctx_set_round_mode(flock_hip_ctx* c, int mode) {
    CTX_CHECK(c);
    if (mode != FLOCK_ROUNDS_HOST && mode != FLOCK_ROUNDS_DEVICE)
        return fail(FLOCK_HIP_E_INVALID, "unknown round mode");
    c->round_mode = mode;
    return 0;
}

int flock_hip_ctx_create(int device, flock_hip_ctx** out) {
    if (!out) return fail(FLOCK_HIP_E_INVALID, "null out");
    *out = nullptr;
    int n = flock_hip_device_count();
    if (device < 0 || device >= n) return fail(FLOCK_HIP_E_NODEVICE, "no such device");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(FLOCK_HIP_E_NODEVICE, std::string("device is ") + prop.gcnArchName +
                                              ", this library is built for gfx950");
    HIP_TRY(hipSetDevice(device));
    flock_hip_ctx* c = new (std::nothrow) flock_hip_ctx();
    if (!c) return fail(FLOCK_HIP_E_NOMEM, "ctx");
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipMalloc(&c->partials, sizeof(uint4) * (flock::kMaxReduceBlocks * 2 + 1)) != hipSuccess ||
        hipHostMalloc((void**)&c->hout, sizeof(flock_f128) * 3,
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void**)&c->dout, c->hout, 0) != hipSuccess ||
        hipMalloc(&c->la_acc, 4 * 64) != hipSuccess ||
        hipHostMalloc((void**)&c->hla, 4 * 40, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void**)&c->dla, c->hla, 0) != hipSuccess) {
        hipError_t e = hipGetLastError();
        flock_hip_ctx_destroy(c);
        return fail(FLOCK_HIP_E_HIP, std::string("context setup: ") + hipGetErrorString(e));
    }
    std::memset(c->hla, 0, 4 * 40);
    if (hipMemset(c->la_acc, 0, 4 * 64) != hipSuccess) {
        flock_hip_ctx_destroy(c);
        return fail(FLOCK_HIP_E_HIP, "context setup: lookahead accumulators");
    }
    // the polled flag word after the pair starts at zero, below every sequence number
    std::memset(c->hout, 0, sizeof(flock_f128) * 3);
    // the grid-reduce ticket counter after the partials starts at zero (each reduction that
    // uses it leaves it at zero)
    if (hipMemset(c->partials + 2 * flock::kMaxReduceBlocks, 0, sizeof(uint4)) != hipSuccess) {
        flock_hip_ctx_destroy(c);
        return fail(FLOCK_HIP_E_HIP, "context setup: counter");
    }
    *out = c;
    return 0;
}

void flock_hip_ctx_destroy(flock_hip_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto& kv : c->evals) hipFree(kv.second);
    if (c->dfs) hipFree(c->dfs);
    if (c->hfs) hipHostFree(c->hfs);
    if (c->zc_ext) hipFree(c->zc_ext);
    if (c->zc_ext_bs) hipFree(c->zc_ext_bs);
    if (c->zc_phim) hipFree(c->zc_phim);
    if (c->zc_logexp) hipFree(c->zc_logexp);
    if (c->zc_phipow) hipFree(c->zc_phipow);
    if (c->zc_partials) hipFree(c->zc_partials);
    if (c->zc_partials2) hipFree(c->zc_partials2);
    if (c->zc_out) hipFree(c->zc_out);
    if (c->zc_host) hipHostFree(c->zc_host);
    if (c->kc_inv) hipFree(c->kc_inv);
    if (c->kc_trow) hipFree(c->kc_trow);
    if (c->kc_rcm) hipFree(c->kc_rcm);
    if (c->scratch) hipFree(c->scratch);
    if (c->partials) hipFree(c->partials);
    if (c->hout) hipHostFree(c->hout);
    if (c->la_acc) hipFree(c->la_acc);
    if (c->hla) hipHostFree(c->hla);
    if (c->aux) hipStreamSynchronize(c->aux);
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    if (c->aux) hipStreamDestroy(c->aux);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

void* flock_hip_ctx_stream(flock_hip_ctx* c) { return c ? (void*)c->stream : nullptr; }

int flock_hip_sync(flock_hip_ctx* c) {
    CTX_CHECK(c);
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int flock_hip_malloc(flock_hip_ctx* c, void** p, size_t bytes) {
    CTX_CHECK(c);
    if (!p) return fail(FLOCK_HIP_E_INVALID, "null out");
    if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return fail(FLOCK_HIP_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    return 0;
}

int flock_hip_dfree(flock_hip_ctx* c, void* p) {
    CTX_CHECK(c);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipFree(p));
    return 0;
}

int flock_hip_memcpy_h2d(flock_hip_ctx* c, void* d, const void* h, size_t n) {
    CTX_CHECK(c);
    HIP_TRY(hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int flock_hip_memcpy_d2h(flock_hip_ctx* c, void* h, const void* d, size_t n) {
    CTX_CHECK(c);
    HIP_TRY(hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

void flock_hip_free(void* p) { std::free(p); }

// ---- field ----
int flock_hip_f128_mul_batch(flock_hip_ctx* c, const flock_f128* a, const flock_f128* b,
                             flock_f128* out, size_t n) {
    CTX_CHECK(c);
    if (n && (!a || !b || !out)) return fail(FLOCK_HIP_E_INVALID, "null buffer");
    LAUNCH(flock::launch_f128_mul_batch(U(a), U(b), U(out), n, c->stream));
    return 0;
}

int flock_hip_f128_mul_const_batch(flock_hip_ctx* c, const flock_f128* a, flock_f128 t,
                                   flock_f128* out, size_t n) {
    CTX_CHECK(c);
    if (n && (!a || !out)) return fail(FLOCK_HIP_E_INVALID, "null buffer");
    LAUNCH(flock::launch_f128_mul_const_batch(U(a), U(t), U(out), n, c->stream));
    return 0;
}

int flock_hip_f128_mul_const_pow_batch(flock_hip_ctx* c, const flock_f128* a, flock_f128 t,
                                       uint32_t reps, int method, flock_f128* out, size_t n) {
    CTX_CHECK(c);
    if (n && (!a || !out)) return fail(FLOCK_HIP_E_INVALID, "null buffer");
    if (method != FLOCK_MUL_SHOUP && method != FLOCK_MUL_MFMA_I8)
        return fail(FLOCK_HIP_E_INVALID, "unknown multiply method");
    LAUNCH(flock::launch_f128_mul_const_pow(U(a), U(t), reps, method, U(out), n, c->stream));
    return 0;
}

// ---- NTT ----
int flock_hip_ntt_forward_interleaved(flock_hip_ctx* c, flock_f128* data, size_t num_ntts,
                                      size_t live, uint32_t log_d, uint32_t start_layer) {
    CTX_CHECK(c);
    if (!data || num_ntts == 0 || live > num_ntts || log_d > 40 || start_layer > log_d ||
        num_ntts > 0xffffffffull)
        return fail(FLOCK_HIP_E_INVALID, "bad NTT shape");
    if (live == 0 || start_layer == log_d) return 0;
    const uint4* ev = nullptr;
    int rc = c->get_evals((int)log_d, &ev);
    if (rc) return rc;
    LAUNCH(flock::launch_ntt_forward(U(data), (uint32_t)num_ntts, (uint32_t)live, (int)log_d,
                                     (int)start_layer, ev, nullptr, c->stream));
    return 0;
}

int flock_hip_ntt_pass(flock_hip_ctx* c, flock_f128* data, size_t num_ntts, size_t live,
                       uint32_t log_d, uint32_t first_layer, uint32_t layers) {
    CTX_CHECK(c);
    if (!data || num_ntts == 0 || live > num_ntts || log_d > 40 || layers < 1 || layers > 3 ||
        first_layer + layers > log_d || num_ntts > 0xffffffffull)
        return fail(FLOCK_HIP_E_INVALID, "bad NTT pass shape");
    if (live == 0) return 0;
    const uint4* ev = nullptr;
    int rc = c->get_evals((int)log_d, &ev);
    if (rc) return rc;
    LAUNCH(flock::launch_ntt_pass(U(data), (uint32_t)num_ntts, (uint32_t)live, (int)log_d,
                                  (int)first_layer, (int)layers, ev, c->stream));
    return 0;
}

// ---- PCS ----
int flock_hip_pcs_sizes(const flock_pcs_params* p, size_t* msg_len, size_t* cw_len,
                        size_t* n_leaves, uint32_t* k_code, size_t* num_ntts) {
    if (!p) return fail(FLOCK_HIP_E_INVALID, "null params");
    // PcsParams::validate, pcs/commit.rs:270-297
    if (p->m < 7 + p->log_batch_size)
        return fail(FLOCK_HIP_E_M_TOO_SMALL, "m too small (need m >= 7 + log_batch_size)");
    if (p->log_inv_rate < 1) return fail(FLOCK_HIP_E_INVALID, "log_inv_rate must be >= 1");
    if (p->merkle_hash != FLOCK_MERKLE_SHA256 && p->merkle_hash != FLOCK_MERKLE_BLAKE3)
        return fail(FLOCK_HIP_E_INVALID, "merkle_hash must be SHA-256 (0) or BLAKE3 (1)");
    const int log_dim = (int)p->m - 7 - (int)p->log_batch_size;
    const int k = log_dim + (int)p->log_inv_rate;
    if (k > 40 || p->log_batch_size > 20) return fail(FLOCK_HIP_E_INVALID, "size out of range");
    const size_t full = (size_t)1 << p->log_batch_size;
    const size_t t = p->num_lanes ? p->num_lanes : full;
    if (t < 1 || t > full) return fail(FLOCK_HIP_E_INVALID, "num_lanes out of range");
    if (msg_len) *msg_len = t << log_dim;
    if (cw_len) *cw_len = ((size_t)1 << k) * t;
    if (n_leaves) *n_leaves = (size_t)1 << k;
    if (k_code) *k_code = (uint32_t)k;
    if (num_ntts) *num_ntts = t;
    return 0;
}

int flock_hip_merkle_tree_kind(flock_hip_ctx* c, const uint8_t* data, size_t len,
                               size_t n_leaves, uint32_t kind, uint8_t* tree) {
    CTX_CHECK(c);
    if (!data || !tree || !is_pow2(n_leaves) || len % n_leaves || len == 0 ||
        (len / n_leaves) % 16)
        return fail(FLOCK_HIP_E_INVALID, "bad Merkle shape");
    if (kind != FLOCK_MERKLE_SHA256 && kind != FLOCK_MERKLE_BLAKE3)
        return fail(FLOCK_HIP_E_INVALID, "unknown Merkle hash");
    if (kind == FLOCK_MERKLE_BLAKE3 &&
        (len / n_leaves + flock::kB3ChunkLen - 1) / flock::kB3ChunkLen >
            (1ull << flock::kMerkleB3MaxDepth))
        return fail(FLOCK_HIP_E_INVALID, "BLAKE3 leaf too long");
    LAUNCH(flock::launch_merkle_tree(data, n_leaves, len / n_leaves, kind, tree, c->stream));
    return 0;
}

int flock_hip_merkle_tree(flock_hip_ctx* c, const uint8_t* data, size_t len, size_t n_leaves,
                          uint8_t* tree) {
    return flock_hip_merkle_tree_kind(c, data, len, n_leaves, FLOCK_MERKLE_SHA256, tree);
}

static int encode_common(flock_hip_ctx* c, const flock_f128* src, size_t src_len,
                         const flock_pcs_params* p, bool lane_major, flock_f128* cw) {
    CTX_CHECK(c);
    size_t msg_len, t;
    uint32_t k;
    int rc = flock_hip_pcs_sizes(p, &msg_len, nullptr, nullptr, &k, &t);
    if (rc) return rc;
    if (!src || !cw) return fail(FLOCK_HIP_E_INVALID, "null buffer");
    const uint32_t reps = 1u << p->log_inv_rate;
    const uint4* ev = nullptr;
    rc = c->get_evals((int)k, &ev);
    if (rc) return rc;
    if (lane_major) {
        // commit_lane_major takes the full padded stack (pcs/commit.rs:477-482)
        if (src_len != ((size_t)1 << (p->m - 7)))
            return fail(FLOCK_HIP_E_INVALID, "lane-major message must be the full 2^(m-7) stack");
        const uint64_t d = src_len >> p->log_batch_size;
        if (k > p->log_inv_rate) {
            // replicate_lane_major_fill fused into the first NTT pass
            LAUNCH(flock::launch_ntt_forward(U(cw), (uint32_t)t, (uint32_t)t, (int)k,
                                             (int)p->log_inv_rate, ev, U(src), c->stream, d));
        } else {
            LAUNCH(flock::launch_fill_lane_major(U(src), U(cw), d, (uint32_t)t, reps, c->stream));
        }
    } else {
        if (src_len != msg_len) return fail(FLOCK_HIP_E_INVALID, "message length mismatch");
        if (k > p->log_inv_rate) {
            // replicate_message_fill fused into the first NTT pass
            LAUNCH(flock::launch_ntt_forward(U(cw), (uint32_t)t, (uint32_t)t, (int)k,
                                             (int)p->log_inv_rate, ev, U(src), c->stream));
        } else {
            LAUNCH(flock::launch_fill_replicate(U(src), U(cw), msg_len, reps, c->stream));
        }
    }
    return 0;
}

static int commit_common(flock_hip_ctx* c, const flock_f128* src, size_t src_len,
                         const flock_pcs_params* p, bool lane_major, flock_f128* cw,
                         uint8_t* tree, uint32_t cap_depth, uint8_t* h_cap) {
    CTX_CHECK(c);
    size_t n_leaves, t;
    uint32_t k;
    int rc = flock_hip_pcs_sizes(p, nullptr, nullptr, &n_leaves, &k, &t);
    if (rc) return rc;
    if (!tree) return fail(FLOCK_HIP_E_INVALID, "null buffer");
    if (cap_depth > k) return fail(FLOCK_HIP_E_INVALID, "cap depth exceeds tree depth");
    rc = encode_common(c, src, src_len, p, lane_major, cw);
    if (rc) return rc;
    if (p->merkle_hash == FLOCK_MERKLE_BLAKE3 &&
        (t * 16 + flock::kB3ChunkLen - 1) / flock::kB3ChunkLen > (1ull << flock::kMerkleB3MaxDepth))
        return fail(FLOCK_HIP_E_INVALID, "BLAKE3 leaf too long");
    LAUNCH(flock::launch_merkle_tree(reinterpret_cast<const uint8_t*>(cw), n_leaves, t * 16,
                                     p->merkle_hash, tree, c->stream,
                                     c->tree_to_cap ? 1ull << cap_depth : 1ull));
    if (h_cap) {
        const size_t l = (size_t)1 << cap_depth;  // cap_layer, merkle.rs:649-656
        HIP_TRY(hipMemcpyAsync(h_cap, tree + (2 * n_leaves - 2 * l) * 32, l * 32,
                               hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return 0;
}

int flock_hip_encode(flock_hip_ctx* c, const flock_f128* msg, size_t msg_len,
                     const flock_pcs_params* p, flock_f128* cw) {
    return encode_common(c, msg, msg_len, p, false, cw);
}

int flock_hip_encode_lane_major(flock_hip_ctx* c, const flock_f128* q, size_t q_len,
                                const flock_pcs_params* p, flock_f128* cw) {
    return encode_common(c, q, q_len, p, true, cw);
}

int flock_hip_commit_lane_major(flock_hip_ctx* c, const flock_f128* q, size_t q_len,
                                const flock_pcs_params* p, flock_f128* cw, uint8_t* tree,
                                uint32_t cap_depth, uint8_t* h_cap) {
    return commit_common(c, q, q_len, p, true, cw, tree, cap_depth, h_cap);
}

int flock_hip_commit(flock_hip_ctx* c, const flock_f128* msg, size_t msg_len,
                     const flock_pcs_params* p, flock_f128* cw, uint8_t* tree,
                     uint32_t cap_depth, uint8_t* h_cap) {
    return commit_common(c, msg, msg_len, p, false, cw, tree, cap_depth, h_cap);
}

// ---- sumcheck rounds ----
static int check_n(size_t n, size_t min) {
    if (!is_pow2(n) || n < min)
        return fail(FLOCK_HIP_E_INVALID, "table length must be a power of two >= " +
                                             std::to_string(min));
    return 0;
}

int flock_hip_sumcheck_round_msg(flock_hip_ctx* c, const flock_f128* f, const flock_f128* b,
                                 size_t n, flock_f128 msg[2]) {
    CTX_CHECK(c);
    if (int rc = check_n(n, 2)) return rc;
    if (!f || !b || !msg) return fail(FLOCK_HIP_E_INVALID, "null buffer");
    LAUNCH(flock::launch_sc_round_msg(U(f), U(b), n, c->partials, flock::round_out(c->dout),
                                      c->stream));
    return c->fetch_pair(msg);
}

int flock_hip_sumcheck_fold_msg(flock_hip_ctx* c, const flock_f128* f, const flock_f128* b,
                                size_t n, flock_f128 r, flock_f128* nf, flock_f128* nb,
                                flock_f128 msg[2]) {
    CTX_CHECK(c);
    if (int rc = check_n(n, 2)) return rc;
    if (!f || !b || !nf || !nb || !msg) return fail(FLOCK_HIP_E_INVALID, "null buffer");
    LAUNCH(flock::launch_sc_fold_msg(U(f), U(b), n, U(r), U(nf), U(nb), c->partials,
                                     flock::round_out(c->dout), c->stream));
    return c->fetch_pair(msg);
}

extern "C++" {
namespace flock {
int ctx_device(const flock_hip_ctx* c) { return c ? c->device : -1; }

int build_eq_async(flock_hip_ctx* c, const flock_f128* h_r, size_t nv, flock_f128* out) {
    CTX_CHECK(c);
    if (nv > (size_t)kEqArgMax || !out || (nv && !h_r)) return fail(FLOCK_HIP_E_INVALID, "bad eq arguments");
    const size_t scratch_words = nv + ((size_t)1 << (nv / 2)) + ((size_t)1 << (nv - nv / 2));
    if (int rc = c->ensure_scratch(scratch_words * 16)) return rc;
    uint4* dr = reinterpret_cast<uint4*>(c->scratch);
    LAUNCH(flock::launch_eq_coords(U(h_r), (uint32_t)nv, dr, c->stream));
    LAUNCH(flock::launch_build_eq(dr, (int)nv, U(out), dr + nv, c->stream));
    return 0;
}
void ctx_set_tree_to_cap(flock_hip_ctx* c, bool on) {
    if (c) c->tree_to_cap = on;
}

int build_eq_halves_async(flock_hip_ctx* c, const flock_f128* h_r, size_t nv, flock_f128* lo,
                          flock_f128* hi) {
    CTX_CHECK(c);
    if (nv > (size_t)kEqArgMax || !lo || !hi || (nv && !h_r)) return fail(FLOCK_HIP_E_INVALID, "bad eq arguments");
    if (int rc = c->ensure_scratch(nv * 16 + 16)) return rc;
    uint4* dr = reinterpret_cast<uint4*>(c->scratch);
    LAUNCH(flock::launch_eq_coords(U(h_r), (uint32_t)nv, dr, c->stream));
    LAUNCH(flock::launch_eq_halves(dr, (int)nv, U(lo), U(hi), c->stream));
    return 0;
}
}  // namespace flock
}

int flock_hip_build_eq(flock_hip_ctx* c, const flock_f128* h_r, size_t nv, flock_f128* out) {
    CTX_CHECK(c);
    if (nv > 40 || !out || (nv && !h_r)) return fail(FLOCK_HIP_E_INVALID, "bad eq arguments");
    const size_t scratch_words = nv + ((size_t)1 << (nv / 2)) + ((size_t)1 << (nv - nv / 2));
    if (int rc = c->ensure_scratch(scratch_words * 16)) return rc;
    uint4* dr = reinterpret_cast<uint4*>(c->scratch);
    if (nv) HIP_TRY(hipMemcpyAsync(dr, h_r, nv * 16, hipMemcpyHostToDevice, c->stream));
    LAUNCH(flock::launch_build_eq(dr, (int)nv, U(out), dr + nv, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));  // h_r may be pageable and short-lived
    return 0;
}

int flock_hip_zc_round_pair(flock_hip_ctx* c, const flock_f128* a, const flock_f128* b, size_t n,
                            const flock_f128* eq, flock_f128 r0, flock_f128* eq_next,
                            flock_f128 msg[2]) {
    CTX_CHECK(c);
    if (int rc = check_n(n, 2)) return rc;
    if (!a || !b || !eq || !msg) return fail(FLOCK_HIP_E_INVALID, "null buffer");
    LAUNCH(flock::launch_zc_round(U(a), U(b), n, U(eq), n >= 4 ? U(eq_next) : nullptr,
                                  c->partials, flock::round_out(c->dout), c->stream));
    if (int rc = c->fetch_pair(msg)) return rc;
    msg[0] = mulh(r0, msg[0]);
    return 0;
}

int flock_hip_zc_fold_round_pair(flock_hip_ctx* c, const flock_f128* a, const flock_f128* b,
                                 size_t n, flock_f128 rho, flock_f128* ao, flock_f128* bo,
                                 const flock_f128* eq, flock_f128 r0, flock_f128* eq_next,
                                 flock_f128 msg[2]) {
    CTX_CHECK(c);
    if (int rc = check_n(n, 4)) return rc;
    if (!a || !b || !ao || !bo || !eq || !msg) return fail(FLOCK_HIP_E_INVALID, "null buffer");
    LAUNCH(flock::launch_zc_fold_round(U(a), U(b), n, U(rho), U(ao), U(bo), U(eq),
                                       n >= 8 ? U(eq_next) : nullptr, c->partials,
                                       flock::round_out(c->dout), c->stream));
    if (int rc = c->fetch_pair(msg)) return rc;
    msg[0] = mulh(r0, msg[0]);
    return 0;
}

int flock_hip_fold_pair(flock_hip_ctx* c, const flock_f128* a, const flock_f128* b, size_t n,
                        flock_f128 cc, flock_f128* ao, flock_f128* bo) {
    CTX_CHECK(c);
    if (int rc = check_n(n, 2)) return rc;
    if (!a || !b || !ao || !bo) return fail(FLOCK_HIP_E_INVALID, "null buffer");
    LAUNCH(flock::launch_fold_pair(U(a), U(b), n, U(cc), U(ao), U(bo), c->stream));
    return 0;
}

// ---- zerocheck front ----
static int zc_round1_impl(flock_hip_ctx* c, const uint8_t* a, const uint8_t* b, const uint8_t* cc,
                          uint32_t m, const uint4* eq, flock_f128 out[128]) {
    if (int rc = c->ensure_zc()) return rc;
    const uint64_t rows = 1ull << (m - 6);
    LAUNCH(flock::launch_zc_round1(a, b, cc, eq, rows, c->zc_ext, c->zc_logexp, c->zc_phipow,
                                   c->zc_partials, c->zc_red(), c->zc_out, c->stream));
    HIP_TRY(hipMemcpyAsync(out, c->zc_out, 128 * 16, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int flock_hip_zc_round1(flock_hip_ctx* c, const uint8_t* a, const uint8_t* b, const uint8_t* cc,
                        uint32_t m, const flock_f128* eq, flock_f128 out[128]) {
    CTX_CHECK(c);
    if (!a || !b || !cc || !eq || !out || m < 6 || m > 40)
        return fail(FLOCK_HIP_E_INVALID, "bad zerocheck round-1 arguments (need 6 <= m <= 40)");
    return zc_round1_impl(c, a, b, cc, m, U(eq), out);
}

// 8 x 256 fold table at z: T[j][v] = sum of L_{8j+b}(z) over set bits b of v.
static void fold_table_at(F z, std::vector<F>& t) {
    F w[64];
    lagrange_phi8(0, z, w);
    t.assign(8 * 256, F{0, 0});
    for (int j = 0; j < 8; j++)
        for (int v = 1; v < 256; v++) {
            const int low = v & -v;
            const int bit = __builtin_ctz(v);
            t[j * 256 + v] = fadd(t[j * 256 + (v ^ low)], w[8 * j + bit]);
        }
}

int flock_hip_zc_fold_at_z(flock_hip_ctx* c, const uint8_t* a, const uint8_t* b, uint32_t m,
                           flock_f128 z, flock_f128* az, flock_f128* bz) {
    CTX_CHECK(c);
    if (!a || !b || !az || !bz || m < 6 || m > 40)
        return fail(FLOCK_HIP_E_INVALID, "bad fold-at-z arguments (need 6 <= m <= 40)");
    std::vector<F> t;
    fold_table_at(F{z.lo, z.hi}, t);
    if (int rc = c->ensure_scratch(t.size() * 16)) return rc;
    HIP_TRY(hipMemcpyAsync(c->scratch, t.data(), t.size() * 16, hipMemcpyHostToDevice, c->stream));
    LAUNCH(flock::launch_zc_fold_at_z(a, b, 1ull << (m - 6), reinterpret_cast<uint4*>(c->scratch),
                                      U(az), U(bz), c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int flock_hip_fold_1b_rows(flock_hip_ctx* c, const flock_f128* rows, size_t n_rows,
                           const flock_f128* t, flock_f128 out[128]) {
    CTX_CHECK(c);
    if (!rows || !t || !out || n_rows == 0)
        return fail(FLOCK_HIP_E_INVALID, "bad fold_1b_rows arguments");
    if (int rc = c->ensure_zc()) return rc;
    LAUNCH(flock::launch_fold_1b_rows(U(rows), U(t), n_rows, c->zc_partials, c->zc_red(),
                                      c->zc_out, c->stream));
    HIP_TRY(hipMemcpyAsync(out, c->zc_out, 128 * 16, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// Tail rounds whose eq table has at least this many variables use the split-eq kernel.

static uint32_t bits_for(uint64_t n) {  // floor(log2 n) + 1, zerocheck.rs:91-94
    uint32_t b = 0;
    while (n) {
        b++;
        n >>= 1;
    }
    return b;
}

static const uint8_t kLabelZcDriver[] = "flock-zerocheck-v0";

// The 7 fixed inner challenges r[6..13) of the zerocheck (zerocheck.rs:644-669): phi8 of
// F7, 53, B5 (univariate_skip_optimized.rs:81), then gamma^(2^i) / (1 + gamma^(2^i)) for
// gamma = x (:109-139).
static void zc_inner_constants(flock_f128* out) {
    const F* phi = phi8_table();
    const uint8_t small[3] = {0xF7, 0x53, 0xB5};
    for (int i = 0; i < 3; i++) out[i] = flock_f128{phi[small[i]].lo, phi[small[i]].hi};
    for (int i = 0; i < 4; i++) {
        const F g{1ull << (1 << i), 0};
        const F v = fmul(g, finv(fadd(F{1, 0}, g)));
        out[3 + i] = flock_f128{v.lo, v.hi};
    }
}

int flock_hip_zerocheck_prove(flock_hip_ctx* c, flock_fs* fs, const uint8_t* d_a,
                              const uint8_t* d_b, const uint8_t* d_c, uint32_t m, int grinding,
                              flock_f128* h_round1, flock_f128* h_ml, flock_f128 h_finals[3],
                              uint64_t* h_nonces, flock_f128* h_claim,
                              flock_f128* h_s_hat_v_c) {
    CTX_CHECK(c);
    const int ks = 6, n_inner = 7;
    if (!fs || !d_a || !d_b || !d_c || !h_round1 || !h_ml || !h_finals || !h_claim ||
        (grinding && !h_nonces) || m < (uint32_t)(ks + n_inner) || m > 40)
        return fail(FLOCK_HIP_E_INVALID, "bad zerocheck arguments (need 13 <= m <= 40)");
    if (int rc = c->ensure_zc()) return rc;
    const uint32_t n_mlv = m - ks;
    const uint64_t N = 1ull << n_mlv;  // rows = round-2 table length
    // scratch: Az, Bz (N each), ping-pong P0 (2 x N/2), P1 (2 x N/4), eq EB (N/2), EC (N/4),
    // E9 (N/8), fold table (2048), r (64), build_eq scratch
    const uint64_t h = N / 2, q = N / 4 > 0 ? N / 4 : 1, e9n = N / 8;
    const uint64_t eqs = 2 * (1ull << (n_mlv / 2 + 1));
    // Tail rounds j < j0, whose eq table (over r[7+j..m), nv = n_mlv-1-j variables) has
    // nv >= 16, run split: eq = e_lo (8 variables, EL) x e_hi (nv-8 variables, level j of EH0).
    const uint32_t nv0 = n_mlv - 1;
    const uint32_t split_min = (uint32_t)flock::tunable(flock::kTuneZcSplitMinVars);
    const uint32_t j0 = nv0 >= split_min ? nv0 - split_min + 1 : 0;
    const uint64_t eh0 = j0 ? 1ull << (nv0 - 8) : 0;
    const uint64_t words = 2 * N + 2 * h + 2 * q + h + q + e9n + 2048 + 64 + eqs + 2 * eh0 +
                           256ull * j0;
    if (int rc = c->ensure_scratch(words * 16)) return rc;
    uint4* base = reinterpret_cast<uint4*>(c->scratch);
    uint4* Az = base;
    uint4* Bz = Az + N;
    uint4* P0f = Bz + N;
    uint4* P0b = P0f + h;
    uint4* P1f = P0b + h;
    uint4* P1b = P1f + q;
    uint4* EB = P1b + q;
    uint4* EC = EB + h;
    uint4* E9 = EC + q;
    uint4* TBL = E9 + e9n;
    uint4* R = TBL + 2048;
    uint4* EQS = R + 64;
    uint4* EH0 = EQS + eqs;   // e_hi of round 0: eq over r[15..m)
    uint4* EHL = EH0 + eh0;   // its levels 1..j0-1 (e_hi of rounds 1..j0-1)
    uint4* EL = EHL + eh0;    // e_lo of rounds 0..j0-1: eq over r[7+j .. 15+j)
    uint4* bufs[2][2] = {{P0f, P0b}, {P1f, P1b}};
    uint4* eqb[2] = {EB, EC};
    auto e_hi = [&](uint32_t j) -> const uint4* {
        if (j == 0) return EH0;
        uint64_t off = 0;
        for (uint32_t i = 1; i < j; i++) off += eh0 >> i;
        return EHL + off;
    };

    // ---- 1. r: r_skip | 7 fixed inner constants | r_outer (zerocheck.rs:644-669) ----
    int ni = 0;
    flock_fs_observe_label(fs, kLabelZcDriver, sizeof(kLabelZcDriver) - 1);
    std::vector<flock_f128> r(m);
    if (grinding) h_nonces[ni++] = flock_fs_grind_pow(fs, bits_for(m));
    flock_fs_sample_f128_vec(fs, r.data(), ks);
    flock_fs_sample_f128_vec(fs, r.data() + ks + n_inner, m - ks - n_inner);
    zc_inner_constants(r.data() + ks);
    // eq over r[7..m) (2^(m-7) words: round 2 and shat) and over r[9..m) (round 1's groups)
    HIP_TRY(hipMemcpyAsync(R, r.data() + ks, n_mlv * 16, hipMemcpyHostToDevice, c->stream));
    LAUNCH(flock::launch_build_eq(R + 1, (int)n_mlv - 1, EB, EQS, c->stream));
    LAUNCH(flock::launch_build_eq(R + 3, (int)n_mlv - 3, E9, EQS, c->stream));
    if (j0) {
        // the split tables of rounds 0..j0-1 and the full table of round j0
        LAUNCH(flock::launch_build_eq(R + 9, (int)(nv0 - 8), EH0, EQS, c->stream));
        LAUNCH(flock::launch_eq_levels(EH0, nv0 - 8, j0, EHL, c->stream));
        LAUNCH(flock::launch_eq_windows(R + 1, j0, EL, c->stream));
    }

    // ---- 2. round 1 ----
    // p_ab over groups of 8 rows (r[6..9) are phi8 constants); shat[lane | b7 << 6] =
    // sum_y eq(r[7..m), y) c_bit(y*128 + b7*64 + lane) (univariate_skip.rs:354-383), from which
    // p_c follows.
    // shat and p_c on the aux stream, beside p_ab on the main stream
    if (int rc = c->fork()) return rc;
    LAUNCH(flock::launch_fold_1b_rows(reinterpret_cast<const uint4*>(d_c), EB, h,
                                      c->zc_partials2, c->zc_red2(), c->zc_shat(), c->aux));
    LAUNCH(flock::launch_zc_round1_c(c->zc_shat(), c->zc_phim, U(r[ks]), c->zc_out + 64, c->aux));
    LAUNCH(flock::launch_zc_round1_groups(d_a, d_b, E9, N / 8, c->zc_ext_bs, c->zc_phipow,
                                          c->zc_partials, c->zc_red(), c->zc_out, c->stream));
    if (int rc = c->join()) return rc;
    HIP_TRY(hipMemcpyAsync(c->zc_host, c->zc_out, 256 * 16, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::memcpy(h_round1, c->zc_host, 128 * 16);
    if (h_s_hat_v_c) std::memcpy(h_s_hat_v_c, c->zc_host + 128, 128 * 16);
    flock_fs_observe_f128_slice(fs, h_round1, 64);
    flock_fs_observe_f128_slice(fs, h_round1 + 64, 64);
    if (grinding) h_nonces[ni++] = flock_fs_grind_pow(fs, bits_for((1u << (ks + 1)) - 1));
    flock_f128 z;
    flock_fs_sample_f128(fs, &z);
    // c-claim: round1_c interpolated on Lambda at z (multilinear.rs:217-228)
    {
        F w[64];
        lagrange_phi8(64, F{z.lo, z.hi}, w);
        F acc{0, 0};
        for (int i = 0; i < 64; i++) acc = fadd(acc, fmul(w[i], F{h_round1[64 + i].lo, h_round1[64 + i].hi}));
        h_finals[2] = flock_f128{acc.lo, acc.hi};
    }

    // ---- 3. round 2: fold at z, first multilinear message with eq over r[7..m) ----
    {
        std::vector<F> t;
        fold_table_at(F{z.lo, z.hi}, t);
        HIP_TRY(hipMemcpyAsync(TBL, t.data(), t.size() * 16, hipMemcpyHostToDevice, c->stream));
        LAUNCH(flock::launch_zc_fold_at_z(d_a, d_b, N, TBL, Az, Bz, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));  // t is host memory
    }
    // ---- 3-5. round 2's message, the tail (zerocheck.rs:790-947) and the final binding
    // (:955-975): round j < n_mlv folds the previous tables at challenge j-1 (j > 0) and emits
    // message j; launch n_mlv is the closing fold ----
    // round j0's eq table (the first unsplit round): eq over r[7+j0..m), into eqb[j0 & 1]
    if (j0)
        LAUNCH(flock::launch_build_eq(R + 1 + j0, (int)(nv0 - j0), eqb[j0 & 1], EQS, c->stream));
    // the closing fold writes bufs[(n_mlv - 1) & 1]
    uint4* last_f = bufs[(n_mlv - 1) & 1][0];
    uint4* last_b = bufs[(n_mlv - 1) & 1][1];
    RoundLoop loop{c, fs, n_mlv, grinding ? bits_for(2) : 0u,
                   [&](uint32_t j, flock::RoundOut ro, const uint4* rin, uint4 rho) -> int {
                       if (j == 0) {
                           if (j0)
                               LAUNCH(flock::launch_zc_round_split(Az, Bz, N, EL, e_hi(0),
                                                                   c->partials, ro, c->stream));
                           else
                               LAUNCH(flock::launch_zc_round(Az, Bz, N, EB,
                                                             N >= 4 ? EC : nullptr, c->partials,
                                                             ro, c->stream));
                           return 0;
                       }
                       const uint32_t i = j - 1;  // the fold's index
                       const uint4* cf = i == 0 ? Az : bufs[(i - 1) & 1][0];
                       const uint4* cb = i == 0 ? Bz : bufs[(i - 1) & 1][1];
                       const uint64_t len = N >> i;
                       uint4* nf = bufs[i & 1][0];
                       uint4* nb = bufs[i & 1][1];
                       if (j < j0) {
                           LAUNCH(flock::launch_zc_fold_round_split(
                               cf, cb, len, rho, nf, nb, EL + 256ull * j, e_hi(j), c->partials,
                               ro, c->stream, rin));
                       } else if (j < n_mlv) {
                           const int ep = 1 ^ (int)(i & 1);
                           LAUNCH(flock::launch_zc_fold_round(
                               cf, cb, len, rho, nf, nb, eqb[ep], len >= 8 ? eqb[ep ^ 1] : nullptr,
                               c->partials, ro, c->stream, rin));
                       } else {
                           LAUNCH(flock::launch_fold_pair(cf, cb, len, rho, nf, nb, c->stream, rin));
                       }
                       return 0;
                   }};
    // the unsplit rounds on tables of at most 2^zc_la_max_log words: two rounds per launch
    // (launch_zc_fold_la), tracking the tables and round eq tables the launches hand on
    const uint4* la_a = nullptr;
    const uint4* la_b = nullptr;
    uint4* la_e[2] = {nullptr, nullptr};  // [current round's eq, the free one]
    loop.la_from = std::max<uint32_t>(std::max<uint32_t>(j0, 1u), n_mlv);
    for (uint32_t j = std::max<uint32_t>(j0, 1u); j < n_mlv; j++)
        if ((N >> (j - 1)) <= (1ull << flock::tunable(flock::kTuneZcLaMaxLog))) {
            loop.la_from = j;
            break;
        }
    loop.la = [&](uint32_t j, int nf, uint4 r1, uint4 r2, flock_f128* out) -> int {
        if (!la_a) {  // the first: the tables and eq table the round-j launch would read
            const uint32_t i = j - 1;
            la_a = i == 0 ? Az : bufs[(i - 1) & 1][0];
            la_b = i == 0 ? Bz : bufs[(i - 1) & 1][1];
            la_e[0] = eqb[j & 1];
            la_e[1] = eqb[(j & 1) ^ 1];
        }
        const uint64_t len = N >> (j - (uint32_t)nf);  // current table length
        uint4* oa = la_a == bufs[0][0] ? bufs[1][0] : bufs[0][0];
        uint4* ob = la_a == bufs[0][0] ? bufs[1][1] : bufs[0][1];
        LAUNCH(flock::launch_zc_fold_la(la_a, la_b, len, nf, r1, r2, oa, ob, la_e[0], la_e[1],
                                        c->la_out(), c->stream));
        if (int rc = c->await_la(out)) return rc;
        la_a = oa;
        la_b = ob;
        std::swap(la_e[0], la_e[1]);  // the launch wrote round j+2's eq into the free one
        last_f = oa;
        last_b = ob;
        return 0;
    };
    if (int rc = loop.run()) return rc;
    HIP_TRY(hipMemcpyAsync(c->hout, last_f, 16, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(c->hout + 1, last_b, 16, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (uint32_t k = 0; k < n_mlv; k++) {
        h_ml[2 * k] = loop.msgs[2 * k];  // mlv_arg[0] = ONE
        h_ml[2 * k + 1] = loop.msgs[2 * k + 1];
        if (grinding) h_nonces[ni++] = loop.nonces[k];
        h_claim[1 + k] = loop.rhos[k];
    }
    h_finals[0] = c->hout[0];
    h_finals[1] = c->hout[1];
    flock_fs_observe_f128(fs, h_finals[0]);
    flock_fs_observe_f128(fs, h_finals[1]);
    h_claim[0] = z;
    for (uint32_t k = 0; k < n_mlv; k++) h_claim[1 + n_mlv + k] = r[ks + k];
    return 0;
}

// ---- zerocheck verifier (zerocheck.rs:1032-1225) ----

namespace {
bool feq(F a, F b) { return a.lo == b.lo && a.hi == b.hi; }
F fo(flock_f128 v) { return F{v.lo, v.hi}; }
flock_f128 to(F v) { return flock_f128{v.lo, v.hi}; }

int reject(const std::string& what) {
    return fail(FLOCK_HIP_E_REJECT, "proof rejected: " + what);
}

// interpolate_at_z_combined (zerocheck/multilinear.rs:238-265): the polynomial of degree < 128
// that is zero on S = phi8(0..63) and takes `lambda_vals` on Lambda = phi8(64..127), at z
// (plain Lagrange over the 128 nodes).
F interpolate_combined(const flock_f128* lambda_vals, F z) {
    const F* phi = phi8_table();
    F acc{0, 0};
    for (int i = 64; i < 128; i++) {
        F num{1, 0}, den{1, 0};
        for (int j = 0; j < 128; j++) {
            if (j == i) continue;
            num = fmul(num, fadd(z, phi[j]));
            den = fmul(den, fadd(phi[i], phi[j]));
        }
        acc = fadd(acc, fmul(fmul(num, finv(den)), fo(lambda_vals[i - 64])));
    }
    return acc;
}
}  // namespace

}  // extern "C"

void flock::lagrange_phi8_host(int offset, flock_f128 z, flock_f128 w[64]) {
    F fw[64];
    lagrange_phi8(offset, fo(z), fw);
    for (int i = 0; i < 64; i++) w[i] = to(fw[i]);
}

extern "C" {

int flock_zerocheck_verify(flock_fs* fs, uint32_t m, int grinding, const flock_f128* h_round1,
                           const flock_f128* h_ml, const flock_f128 h_finals[3],
                           const uint64_t* h_nonces, flock_f128* h_claim) {
    const int ks = 6, n_inner = 7;
    if (!fs || !h_round1 || !h_ml || !h_finals || (grinding && !h_nonces) ||
        m < (uint32_t)(ks + n_inner) || m > 40)
        return fail(FLOCK_HIP_E_INVALID, "bad zerocheck verify arguments (need 13 <= m <= 40)");
    const uint32_t n_mlv = m - ks;
    int ni = 0;
    // r: r_skip | inner constants | r_outer, in the prover's order
    flock_fs_observe_label(fs, kLabelZcDriver, sizeof(kLabelZcDriver) - 1);
    if (grinding && !flock_fs_verify_pow(fs, h_nonces[ni++], bits_for(m)))
        return reject("initial proof of work");
    std::vector<flock_f128> r(m);
    flock_fs_sample_f128_vec(fs, r.data(), ks);
    flock_fs_sample_f128_vec(fs, r.data() + ks + n_inner, m - ks - n_inner);
    zc_inner_constants(r.data() + ks);
    // round 1 messages, z
    flock_fs_observe_f128_slice(fs, h_round1, 64);
    flock_fs_observe_f128_slice(fs, h_round1 + 64, 64);
    if (grinding && !flock_fs_verify_pow(fs, h_nonces[ni++], bits_for((1u << (ks + 1)) - 1)))
        return reject("skip proof of work");
    flock_f128 zv;
    flock_fs_sample_f128(fs, &zv);
    const F z = fo(zv);
    // c(z, r_rest) from round1_c on Lambda (interpolate_at_z_on_lambda, multilinear.rs:215)
    F w[64];
    lagrange_phi8(64, z, w);
    F c_eval{0, 0};
    for (int i = 0; i < 64; i++) c_eval = fadd(c_eval, fmul(w[i], fo(h_round1[64 + i])));
    if (!feq(c_eval, fo(h_finals[2]))) return reject("round-1 c evaluation");
    // the AB running claim from the S-zero assumption
    std::vector<flock_f128> comb(64);
    for (int i = 0; i < 64; i++) comb[i] = to(fadd(fo(h_round1[i]), fo(h_round1[64 + i])));
    F run = fadd(interpolate_combined(comb.data(), z), c_eval);
    // multilinear chain: G(0) from (1 + r_eq) G(0) + r_eq G(1) = claim, then claim <- G(rho)
    std::vector<flock_f128> rho(n_mlv);
    for (uint32_t i = 0; i < n_mlv; i++) {
        const F r_eq = fo(r[ks + i]);
        const F g1 = fo(h_ml[2 * i]), gi = fo(h_ml[2 * i + 1]);
        const F g0 = fmul(fadd(run, fmul(r_eq, g1)), finv(fadd(F{1, 0}, r_eq)));
        flock_fs_observe_f128(fs, h_ml[2 * i]);
        flock_fs_observe_f128(fs, h_ml[2 * i + 1]);
        if (grinding && !flock_fs_verify_pow(fs, h_nonces[ni++], bits_for(2)))
            return reject("multilinear round proof of work");
        flock_fs_sample_f128(fs, &rho[i]);
        const F x = fo(rho[i]), x1 = fadd(F{1, 0}, x);
        run = fadd(fadd(fmul(g0, x1), fmul(g1, x)), fmul(gi, fmul(x, x1)));
    }
    if (!feq(run, fmul(fo(h_finals[0]), fo(h_finals[1])))) return reject("final a*b claim");
    flock_fs_observe_f128(fs, h_finals[0]);
    flock_fs_observe_f128(fs, h_finals[1]);
    if (h_claim) {
        h_claim[0] = zv;
        for (uint32_t k = 0; k < n_mlv; k++) h_claim[1 + k] = rho[k];
        for (uint32_t k = 0; k < n_mlv; k++) h_claim[1 + n_mlv + k] = r[ks + k];
    }
    return 0;
}

// ---- protocol driver ----
static const uint8_t kLabelProduct[] = "flock-product-sumcheck";
static const uint8_t kLabelZc[] = "flock-zerocheck-v0";

int flock_hip_prove_sumcheck(flock_hip_ctx* c, flock_fs* fs, const flock_f128* d_a,
                             const flock_f128* d_b, uint32_t log_n, flock_f128* msgs,
                             flock_f128 finals[4]) {
    CTX_CHECK(c);
    if (!fs || !d_a || !d_b || !msgs || !finals || log_n < 1 || log_n > 36)
        return fail(FLOCK_HIP_E_INVALID, "bad sumcheck arguments");
    const size_t n = (size_t)1 << log_n;
    // scratch: two ping-pong table pairs (n/2 and n/4 words each), two eq tables (n/2,
    // n/4), eq challenges and build_eq scratch.
    const size_t h = n / 2 > 0 ? n / 2 : 1, q = n / 4 > 0 ? n / 4 : 1;
    const size_t eq_scratch = 64 + ((size_t)1 << (log_n / 2 + 1)) + ((size_t)1 << (log_n / 2 + 1));
    // zerocheck rounds j < j0 (eq over r[1+j..log_n) with >= 16 variables) run split, as in
    // flock_hip_zerocheck_prove: e_lo windows EL, e_hi = level j of EH0 = eq over r[9..log_n)
    const uint32_t nv0 = log_n - 1;
    const uint32_t split_min = (uint32_t)flock::tunable(flock::kTuneZcSplitMinVars);
    const uint32_t j0 = nv0 >= split_min ? nv0 - split_min + 1 : 0;
    const size_t eh0 = j0 ? (size_t)1 << (nv0 - 8) : 0;
    const size_t words = 2 * h + 2 * q + h + q + eq_scratch + 2 * eh0 + 256 * (size_t)j0;
    if (int rc = c->ensure_scratch(words * 16)) return rc;
    uint4* base = reinterpret_cast<uint4*>(c->scratch);
    uint4* P0f = base;
    uint4* P0b = P0f + h;
    uint4* P1f = P0b + h;
    uint4* P1b = P1f + q;
    uint4* E0 = P1b + q;
    uint4* E1 = E0 + h;
    uint4* R = E1 + q;  // up to 64 challenges
    uint4* EQS = R + 64;
    uint4* EH0 = EQS + (eq_scratch - 64);
    uint4* EHL = EH0 + eh0;
    uint4* EL = EHL + eh0;
    uint4* bufs[2][2] = {{P0f, P0b}, {P1f, P1b}};
    uint4* eqb[2] = {E0, E1};
    auto e_hi = [&](uint32_t j) -> const uint4* {
        if (j == 0) return EH0;
        size_t off = 0;
        for (uint32_t i = 1; i < j; i++) off += eh0 >> i;
        return EHL + off;
    };
    int mi = 0;

    // ---- product sumcheck: round j < log_n folds at challenge j-1 (j > 0) and emits message j;
    // launch log_n is the closing fold ----
    flock_fs_observe_label(fs, kLabelProduct, sizeof(kLabelProduct) - 1);
    // both loops close with a fold into bufs[(log_n - 1) & 1]
    uint4* last_f = bufs[(log_n - 1) & 1][0];
    uint4* last_b = bufs[(log_n - 1) & 1][1];
    RoundLoop prod{c, fs, log_n, 0u,
                   [&](uint32_t j, flock::RoundOut ro, const uint4* rin, uint4 r) -> int {
                       if (j == 0) {
                           LAUNCH(flock::launch_sc_round_msg(U(d_a), U(d_b), n, c->partials, ro,
                                                             c->stream));
                           return 0;
                       }
                       const uint32_t i = j - 1;
                       const uint4* cf = i == 0 ? U(d_a) : bufs[(i - 1) & 1][0];
                       const uint4* cb = i == 0 ? U(d_b) : bufs[(i - 1) & 1][1];
                       const size_t len = n >> i;
                       uint4* nf = bufs[i & 1][0];
                       uint4* nb = bufs[i & 1][1];
                       if (j < log_n) {
                           LAUNCH(flock::launch_sc_fold_msg(cf, cb, len, r, nf, nb, c->partials,
                                                            ro, c->stream, rin));
                       } else {
                           LAUNCH(flock::launch_fold_pair(cf, cb, len, r, nf, nb, c->stream, rin));
                       }
                       return 0;
                   }};
    if (int rc = prod.run()) return rc;
    HIP_TRY(hipMemcpyAsync(c->hout, last_f, 16, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(c->hout + 1, last_b, 16, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (uint32_t k = 0; k < 2 * log_n; k++) msgs[mi++] = prod.msgs[k];
    finals[0] = c->hout[0];
    finals[1] = c->hout[1];
    flock_fs_observe_f128(fs, finals[0]);
    flock_fs_observe_f128(fs, finals[1]);

    // ---- eq-weighted zerocheck tail ----
    flock_fs_observe_label(fs, kLabelZc, sizeof(kLabelZc) - 1);
    std::vector<flock_f128> r(log_n);
    flock_fs_sample_f128_vec(fs, r.data(), log_n);
    if (log_n > 1) HIP_TRY(hipMemcpyAsync(R, r.data() + 1, (log_n - 1) * 16, hipMemcpyHostToDevice,
                                          c->stream));
    if (j0) {
        // split tables of rounds 0..j0-1, and the full table of round j0 (eq over r[1+j0..))
        LAUNCH(flock::launch_build_eq(R + 8, (int)(nv0 - 8), EH0, EQS, c->stream));
        LAUNCH(flock::launch_eq_levels(EH0, nv0 - 8, j0, EHL, c->stream));
        LAUNCH(flock::launch_eq_windows(R, j0, EL, c->stream));
        LAUNCH(flock::launch_build_eq(R + j0, (int)(nv0 - j0), eqb[j0 & 1], EQS, c->stream));
    } else {
        LAUNCH(flock::launch_build_eq(R, (int)log_n - 1, E0, EQS, c->stream));
    }
    RoundLoop zc{c, fs, log_n, 0u,
                 [&](uint32_t j, flock::RoundOut ro, const uint4* rin, uint4 rho) -> int {
                     if (j == 0) {
                         if (j0)
                             LAUNCH(flock::launch_zc_round_split(U(d_a), U(d_b), n, EL, e_hi(0),
                                                                 c->partials, ro, c->stream));
                         else
                             LAUNCH(flock::launch_zc_round(U(d_a), U(d_b), n, E0,
                                                           n >= 4 ? E1 : nullptr, c->partials,
                                                           ro, c->stream));
                         return 0;
                     }
                     const uint32_t i = j - 1;
                     const uint4* cf = i == 0 ? U(d_a) : bufs[(i - 1) & 1][0];
                     const uint4* cb = i == 0 ? U(d_b) : bufs[(i - 1) & 1][1];
                     const size_t len = n >> i;
                     uint4* nf = bufs[i & 1][0];
                     uint4* nb = bufs[i & 1][1];
                     if (j < j0) {
                         LAUNCH(flock::launch_zc_fold_round_split(cf, cb, len, rho, nf, nb,
                                                                  EL + 256ull * j, e_hi(j),
                                                                  c->partials, ro, c->stream,
                                                                  rin));
                     } else if (j < log_n) {
                         const int ep = 1 ^ (int)(i & 1);  // eq table of the next message
                         LAUNCH(flock::launch_zc_fold_round(
                             cf, cb, len, rho, nf, nb, eqb[ep], len >= 8 ? eqb[ep ^ 1] : nullptr,
                             c->partials, ro, c->stream, rin));
                     } else {
                         LAUNCH(flock::launch_fold_pair(cf, cb, len, rho, nf, nb, c->stream, rin));
                     }
                     return 0;
                 }};
    if (int rc = zc.run()) return rc;  // the pageable r copy completed when its call returned
    HIP_TRY(hipMemcpyAsync(c->hout, last_f, 16, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(c->hout + 1, last_b, 16, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (uint32_t k = 0; k < 2 * log_n; k++) msgs[mi++] = zc.msgs[k];  // r_next[0] = ONE
    finals[2] = c->hout[0];
    finals[3] = c->hout[1];
    flock_fs_observe_f128(fs, finals[2]);
    flock_fs_observe_f128(fs, finals[3]);
    return 0;
}

// ---------------------------------------------------------------------------
// Lincheck (crates/flock-core/src/lincheck.rs, lincheck/union.rs)
// ---------------------------------------------------------------------------
static int csc_check(const flock_csc* k) {
    if (!k || !k->a_col_ptr || !k->b_col_ptr || (k->nnz_a && !k->a_rows) ||
        (k->nnz_b && !k->b_rows) || k->k_log < 6 || k->k_log > 24 || k->nnz_a > 0xffffffffull ||
        k->nnz_b > 0xffffffffull || k->const_pin >= ((int64_t)1 << k->k_log))
        return fail(FLOCK_HIP_E_INVALID, "bad CSC circuit (need 6 <= k_log <= 24, nnz < 2^32, "
                                         "const_pin < 2^k_log)");
    return 0;
}

static int csc_fold_impl(flock_hip_ctx* c, const flock_csc* k, const uint4* eq, uint4 alpha,
                         uint4* sa, uint4* sb, uint4* comb, int* d_err) {
    const uint64_t K = 1ull << k->k_log;
    LAUNCH(flock::launch_lc_csc_fold(K, k->a_col_ptr, k->a_rows, k->nnz_a, k->b_col_ptr,
                                     k->b_rows, k->nnz_b, eq, alpha, sa, sb, comb, d_err,
                                     c->stream));
    int herr = 0;
    HIP_TRY(hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (herr) return fail(FLOCK_HIP_E_INVALID, "CSC circuit has a row index >= n_cols or a "
                                               "col_ptr entry > nnz");
    return 0;
}

int flock_hip_lc_csc_fold(flock_hip_ctx* c, const flock_csc* k, const flock_f128* eq,
                          flock_f128 alpha, flock_f128* sa, flock_f128* sb, flock_f128* comb) {
    CTX_CHECK(c);
    if (int rc = csc_check(k)) return rc;
    if (!eq || (!sa && !sb && !comb)) return fail(FLOCK_HIP_E_INVALID, "bad CSC fold arguments");
    if (int rc = c->ensure_scratch(16)) return rc;
    return csc_fold_impl(c, k, U(eq), U(alpha), U(sa), U(sb), U(comb),
                         reinterpret_cast<int*>(c->scratch));
}

static int lc_fold_args(const uint8_t* z, uint32_t k_log, uint32_t nu, size_t useful_bits,
                        size_t n_rows) {
    if (!z || (reinterpret_cast<uintptr_t>(z) & 3) || k_log < 6 || k_log > 24 || nu < 3 ||
        nu > 24 || useful_bits > (1ull << k_log) || n_rows > (1ull << nu))
        return fail(FLOCK_HIP_E_INVALID, "bad partial-fold arguments (need 4-byte aligned z, "
                                         "6 <= k_log <= 24, 3 <= nu <= 24, useful_bits <= 2^k_log, "
                                         "n_rows <= 2^nu)");
    return 0;
}

static uint64_t lc_fold_partial_words(uint64_t K) {
    return (uint64_t)flock::lc_partial_fold_rows(K) * K;
}

int flock_hip_lc_partial_fold(flock_hip_ctx* c, const uint8_t* z, uint32_t k_log, uint32_t nu,
                              size_t useful_bits, size_t n_rows, const flock_f128* eq_outer,
                              flock_f128* out) {
    CTX_CHECK(c);
    if (int rc = lc_fold_args(z, k_log, nu, useful_bits, n_rows)) return rc;
    if (!eq_outer || !out) return fail(FLOCK_HIP_E_INVALID, "bad partial-fold arguments");
    const uint64_t K = 1ull << k_log;
    if (int rc = c->ensure_scratch(16 * lc_fold_partial_words(K))) return rc;
    LAUNCH(flock::launch_lc_partial_fold(z, K, useful_bits, (n_rows + 7) / 8, U(eq_outer),
                                         reinterpret_cast<uint4*>(c->scratch), U(out), c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// ---- Keccak-f[1600] lincheck walker (r1cs_hashes/keccak.rs:136-176, 436-510, 692-835) ----
// Host tables for k_kc_walk. State bit s = x + 5y + 25z; phi = pi o rho o theta and
// B[x, y, z] is the XOR of the 11 state bits of its preimage:
//   pi:    B[x, y, z] = R[(x + 3y) mod 5, x, z]
//   rho:   R[a, b, c] = T[a, b, (c - rho[a][b]) mod 64]
//   theta: T[a, b, c] = S[a, b, c] + sum_y' S[a - 1, y', c] + sum_y' S[a + 1, y', c - 1].
static const unsigned kKcRho[5][5] = {{0, 36, 3, 41, 18}, {1, 44, 10, 45, 2},
                                      {62, 6, 43, 15, 61}, {28, 55, 25, 21, 56},
                                      {27, 20, 39, 8, 14}};  // FIPS 202 Table 2, [x][y]
static const uint64_t kKcRc[24] = {
    0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
    0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
    0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
    0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
    0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
    0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};

static int kc_state(int x, int y, int z) { return x + 5 * y + 25 * z; }
static int kc_lane_pos(int s) { return 64 * (s % 25) + s / 25; }
static void kc_preimage(int s_out, int out[11]) {
    const int x = s_out % 5, y = (s_out % 25) / 5, z = s_out / 25;
    const int a = (x + 3 * y) % 5, b = x;
    const int cz = (z + 64 - (int)(kKcRho[a][b] % 64)) % 64, cp = (cz + 63) % 64;
    out[0] = kc_state(a, b, cz);
    for (int yp = 0; yp < 5; yp++) {
        out[1 + yp] = kc_state((a + 4) % 5, yp, cz);
        out[6 + yp] = kc_state((a + 1) % 5, yp, cp);
    }
}

// The walker's tables: inv[s * 11 + i] = the 11 state bits whose preimage holds s (phi^T),
// trow = the chi A / B row positions each of them feeds, rcm[r] = the round-constant vector
// of state_r as 50 words.
static int kc_build_tables(std::vector<uint16_t>& inv, std::vector<uint16_t>& trow,
                           std::vector<uint32_t>& rcm) {
    constexpr int S = 1600;
    inv.assign(S * 11, 0);
    trow.assign(2 * S * 11, 0);
    std::vector<int> fill(S, 0);
    std::vector<std::array<int, 11>> pre(S);
    for (int so = 0; so < S; so++) {
        kc_preimage(so, pre[so].data());
        for (int i = 0; i < 11; i++) {
            const int si = pre[so][i];
            if (fill[si] >= 11) return fail(FLOCK_HIP_E_INVALID, "Keccak phi table");
            inv[si * 11 + fill[si]++] = (uint16_t)so;
        }
    }
    for (int si = 0; si < S; si++) {
        if (fill[si] != 11) return fail(FLOCK_HIP_E_INVALID, "Keccak phi table");
        for (int i = 0; i < 11; i++) {
            const int u = inv[si * 11 + i];
            const int ux = u % 5, uy = (u % 25) / 5, uz = u / 25;
            // chi A operand of row x reads B[x + 1]; the B operand reads B[x + 2]
            trow[si * 11 + i] = (uint16_t)kc_lane_pos(kc_state((ux + 4) % 5, uy, uz));
            trow[(S + si) * 11 + i] = (uint16_t)kc_lane_pos(kc_state((ux + 3) % 5, uy, uz));
        }
    }
    // RC_0 = 0, RC_{r+1} = phi(RC_r) + iota's constant on lane (0, 0)
    rcm.assign(25 * 50, 0);
    std::vector<uint8_t> rc(S, 0), nrc(S);
    for (int r = 0;; r++) {
        for (int s0 = 0; s0 < S; s0++)
            if (rc[s0]) rcm[r * 50 + (s0 >> 5)] |= 1u << (s0 & 31);
        if (r == 24) break;
        for (int so = 0; so < S; so++) {
            uint8_t bit = 0;
            for (int i = 0; i < 11; i++) bit ^= rc[pre[so][i]];
            nrc[so] = bit;
        }
        rc = nrc;
        for (int z = 0; z < 64; z++)
            if ((kKcRc[r] >> z) & 1) rc[kc_state(0, 0, z)] ^= 1;
    }
    return 0;
}

int flock_hip_ctx::ensure_kc() {
    if (kc_inv) return 0;
    std::vector<uint16_t> inv, trow;
    std::vector<uint32_t> rcm;
    if (int rc = kc_build_tables(inv, trow, rcm)) return rc;
    if (hipMalloc(&kc_inv, inv.size() * 2) != hipSuccess ||
        hipMalloc(&kc_trow, trow.size() * 2) != hipSuccess ||
        hipMalloc(&kc_rcm, rcm.size() * 4) != hipSuccess) {
        (void)hipGetLastError();
        return fail(FLOCK_HIP_E_NOMEM, "Keccak lincheck tables");
    }
    HIP_TRY(hipMemcpy(kc_inv, inv.data(), inv.size() * 2, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(kc_trow, trow.data(), trow.size() * 2, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(kc_rcm, rcm.data(), rcm.size() * 4, hipMemcpyHostToDevice));
    return 0;
}

// The walk of k_kc_walk on the host, for the verifier: sa = A_0^T eq and sb = B_0^T eq over the
// 2^16 columns (KeccakLincheckCircuit::fold_split, keccak.rs:692-835). Per side, the
// coefficient vector on state_r is carried from K_24 (the pin rows) down to K_0 through
// phi^T plus the chi rows of round r - 1, and dropped on the t_(r-2) columns; the constant
// parts collect on the constant column.
static void kc_host_fold(const F* eq, F* sa, F* sb, const std::vector<uint16_t>& inv,
                         const std::vector<uint16_t>& trow, const std::vector<uint32_t>& rcm) {
    constexpr int S = 1600;
    constexpr uint32_t kZc = 4096, kS24 = 2048, kT = 4160;
    std::fill(sa, sa + (1u << 16), F{0, 0});
    std::fill(sb, sb + (1u << 16), F{0, 0});
    auto rc_bit = [&](int r, int st) { return (rcm[r * 50 + (st >> 5)] >> (st & 31)) & 1u; };
    std::vector<F> kv(S), nk(S);
    for (int side = 0; side < 2; side++) {
        F* out = side ? sb : sa;
        const uint16_t* tr = trow.data() + (size_t)side * S * 11;
        F zc = eq[kZc];  // the constant row
        for (int st = 0; st < S; st++) {
            const uint32_t p = (uint32_t)kc_lane_pos(st);
            const F pin = eq[kS24 + p];
            if (side == 0) {
                kv[st] = pin;                       // K_24
                out[kT + 23u * 1600u + p] = pin;    // t_23 takes K_24
                if (rc_bit(24, st)) zc = fadd(zc, pin);
                for (int r = 0; r < 24; r++) zc = fadd(zc, eq[kT + 1600u * r + p]);  // NOT of A
            } else {
                kv[st] = F{0, 0};
                zc = fadd(zc, fadd(pin, eq[p]));  // pin and state_0 rows: B = [1]
            }
        }
        for (int r = 24; r >= 1; r--) {
            const F* erow = eq + kT + 1600u * (r - 1);
            for (int st = 0; st < S; st++) {
                F v{0, 0}, c{0, 0};
                for (int i = 0; i < 11; i++) {
                    v = fadd(v, kv[inv[st * 11 + i]]);
                    c = fadd(c, erow[tr[st * 11 + i]]);
                }
                if (rc_bit(r - 1, st)) zc = fadd(zc, c);
                const F k = fadd(v, c);  // K_(r-1)
                nk[st] = k;
                const uint32_t p = (uint32_t)kc_lane_pos(st);
                if (r >= 2)
                    out[kT + 1600u * (r - 2) + p] = k;
                else
                    out[p] = side == 0 ? fadd(k, eq[p]) : k;  // state_0 self-loop on A
            }
            kv.swap(nk);
        }
        out[kZc] = zc;
    }
}

int flock_hip_lc_keccak_fold(flock_hip_ctx* c, const flock_f128* eq, flock_f128 alpha,
                             flock_f128* sa, flock_f128* sb, flock_f128* comb) {
    CTX_CHECK(c);
    if (!eq || !sa || !sb) return fail(FLOCK_HIP_E_INVALID, "bad Keccak fold arguments");
    if (int rc = c->ensure_kc()) return rc;
    LAUNCH(flock::launch_kc_fold(U(eq), c->kc_inv, c->kc_trow, c->kc_rcm, U(alpha), U(sa), U(sb),
                                 comb ? U(comb) : nullptr, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

static const uint8_t kLabelLincheck[] = "flock-lincheck-v0";

// The circuit enters only through its column marginals (LincheckCircuit, lincheck.rs:238-278):
// fold(eq, alpha) -> sa = A_0^T eq, sb = B_0^T eq, comb = alpha sa + sb.
using LcFold = std::function<int(const uint4* eq, uint4 alpha, uint4* sa, uint4* sb, uint4* comb,
                                 int* d_err)>;

// One boolean slot of the union lincheck: its circuit fold, k_log, column prefix (column
// offset prefix << k_log), constant-wire pin (-1: none) and lincheck stripe.
struct LcSlotRun {
    LcFold fold;
    uint32_t k_log;
    uint64_t prefix;
    int64_t const_pin;
    const uint8_t* d_z;
    size_t useful_bits, n_rows;
};

static int lincheck_prove_slots(flock_hip_ctx* c, flock_fs* fs, const std::vector<LcSlotRun>& slots,
                                uint32_t col_vars, uint32_t nu, flock_f128 z_skip,
                                const flock_f128* h_x_inner_rest, const flock_f128* h_x_outer,
                                int grinding, flock_f128* h_rounds, flock_f128 h_z_partial[64],
                                flock_f128* h_matrix_evals, uint64_t* h_nonces,
                                flock_f128* h_claim, flock_f128* d_z_vec);

int flock_hip_lincheck_prove(flock_hip_ctx* c, flock_fs* fs, const flock_csc* k,
                             const uint8_t* d_z, uint32_t nu, size_t useful_bits, size_t n_rows,
                             flock_f128 z_skip, const flock_f128* h_x_inner_rest,
                             const flock_f128* h_x_outer, int grinding, flock_f128* h_rounds,
                             flock_f128 h_z_partial[64], flock_f128 h_matrix_evals[2],
                             uint64_t* h_nonces, flock_f128* h_claim, flock_f128* d_z_vec) {
    CTX_CHECK(c);
    if (int rc = csc_check(k)) return rc;
    const LcFold fold = [c, k](const uint4* eq, uint4 alpha, uint4* sa, uint4* sb, uint4* comb,
                               int* d_err) {
        return csc_fold_impl(c, k, eq, alpha, sa, sb, comb, d_err);
    };
    const std::vector<LcSlotRun> one{{fold, k->k_log, 0, k->const_pin, d_z, useful_bits, n_rows}};
    return lincheck_prove_slots(c, fs, one, k->k_log, nu, z_skip, h_x_inner_rest, h_x_outer,
                                grinding, h_rounds, h_z_partial, h_matrix_evals, h_nonces, h_claim,
                                d_z_vec);
}

int flock_hip_lincheck_prove_keccak(flock_hip_ctx* c, flock_fs* fs, const uint8_t* d_z,
                                    uint32_t nu, size_t n_rows, flock_f128 z_skip,
                                    const flock_f128* h_x_inner_rest, const flock_f128* h_x_outer,
                                    int grinding, flock_f128* h_rounds,
                                    flock_f128 h_z_partial[64], flock_f128 h_matrix_evals[2],
                                    uint64_t* h_nonces, flock_f128* h_claim, flock_f128* d_z_vec) {
    CTX_CHECK(c);
    if (int rc = c->ensure_kc()) return rc;
    const LcFold fold = [c](const uint4* eq, uint4 alpha, uint4* sa, uint4* sb, uint4* comb,
                            int*) {
        LAUNCH(flock::launch_kc_fold(eq, c->kc_inv, c->kc_trow, c->kc_rcm, alpha, sa, sb, comb,
                                     c->stream));
        return 0;
    };
    // K_LOG 16, USEFUL_BITS 42,560, const pin Z_CONST = 4096 (keccak.rs:436-451, 688-690)
    const std::vector<LcSlotRun> one{{fold, 16, 0, 4096, d_z, 42560, n_rows}};
    return lincheck_prove_slots(c, fs, one, 16, nu, z_skip, h_x_inner_rest, h_x_outer, grinding,
                                h_rounds, h_z_partial, h_matrix_evals, h_nonces, h_claim, d_z_vec);
}

static int lincheck_prove_slots(flock_hip_ctx* c, flock_fs* fs, const std::vector<LcSlotRun>& slots,
                                uint32_t col_vars, uint32_t nu, flock_f128 z_skip,
                                const flock_f128* h_x_inner_rest, const flock_f128* h_x_outer,
                                int grinding, flock_f128* h_rounds, flock_f128 h_z_partial[64],
                                flock_f128* h_matrix_evals, uint64_t* h_nonces,
                                flock_f128* h_claim, flock_f128* d_z_vec) {
    const int ks = 6;
    if (slots.empty() || slots.size() > FLOCK_MAX_TYPES || col_vars < 6 || col_vars > 24)
        return fail(FLOCK_HIP_E_INVALID, "bad lincheck arguments (need 1..16 slots, 6 <= "
                                         "col_vars <= 24)");
    // every slot inside the 2^col_vars columns, the slots disjoint
    std::vector<std::pair<uint64_t, uint64_t>> spans;
    uint64_t pf_words = 1;
    for (const LcSlotRun& sl : slots) {
        if (int rc = lc_fold_args(sl.d_z, sl.k_log, nu, sl.useful_bits, sl.n_rows)) return rc;
        if (sl.k_log > col_vars || sl.prefix >= (1ull << (col_vars - sl.k_log)) ||
            sl.const_pin >= ((int64_t)1 << sl.k_log))
            return fail(FLOCK_HIP_E_INVALID, "lincheck slot outside the union columns (need "
                                             "k_log <= col_vars, prefix < 2^(col_vars - k_log), "
                                             "const_pin < 2^k_log)");
        spans.emplace_back(sl.prefix << sl.k_log, (sl.prefix + 1) << sl.k_log);
        pf_words = std::max<uint64_t>(pf_words, lc_fold_partial_words(1ull << sl.k_log));
    }
    std::sort(spans.begin(), spans.end());
    for (size_t i = 1; i < spans.size(); i++)
        if (spans[i].first < spans[i - 1].second)
            return fail(FLOCK_HIP_E_INVALID, "lincheck slots overlap");
    const uint32_t L = col_vars - ks;
    if (!fs || (L && (!h_x_inner_rest || !h_rounds)) || !h_x_outer || !h_z_partial ||
        !h_matrix_evals || !h_claim || (grinding && !h_nonces))
        return fail(FLOCK_HIP_E_INVALID, "bad lincheck arguments");
    const uint64_t K = 1ull << col_vars;
    const uint64_t n_outer = 1ull << nu;
    // scratch (words): EQI | SA | SB | COMB | ZV (K each; slot t's sa, sb at its column
    // offset) | EO (2^nu) | ER (2^L) | EQS, EQS2 (build_eq) | PF (partial-fold rows of the
    // largest slot) | RV (64 + L + nu) | ERR (1)
    const uint32_t vmax = nu > L ? nu : L;
    const uint64_t eqs_words = 2ull << ((vmax + 1) / 2);
    const uint64_t words =
        5 * K + n_outer + (1ull << L) + 2 * eqs_words + pf_words + 64 + L + nu + 1;
    if (int rc = c->ensure_scratch(words * 16)) return rc;
    uint4* EQI = reinterpret_cast<uint4*>(c->scratch);
    uint4* SA = EQI + K;
    uint4* SB = SA + K;
    uint4* COMB = SB + K;
    uint4* ZV = COMB + K;
    uint4* EO = ZV + K;
    uint4* ER = EO + n_outer;
    uint4* EQS = ER + (1ull << L);
    uint4* EQS2 = EQS + eqs_words;  // build_eq scratch of the aux stream
    uint4* PF = EQS2 + eqs_words;
    uint4* RV = PF + pf_words;
    int* ERR = reinterpret_cast<int*>(RV + 64 + L + nu);
    const bool whole = slots.size() == 1 && slots[0].k_log == col_vars;  // one slot, all columns

    int ni = 0;
    flock_fs_observe_label(fs, kLabelLincheck, sizeof(kLabelLincheck) - 1);
    // 1. alpha (lincheck/union.rs:252-260)
    if (grinding) h_nonces[ni++] = flock_fs_grind_pow(fs, 1);
    flock_f128 alpha;
    flock_fs_sample_f128(fs, &alpha);
    // 2. per slot, the quirky eq over its inner coordinates x_inner_rest[..k_t - 6] and the
    // circuit comb at its column offset, scaled by eq_prefix_weight(x_inner_rest[k_t - 6..],
    // prefix_t) (union.rs:90-101, 270-300)
    std::vector<flock_f128> rv(64 + L + nu);
    {
        F w[64];
        lagrange_phi8(0, F{z_skip.lo, z_skip.hi}, w);
        for (int i = 0; i < 64; i++) rv[i] = flock_f128{w[i].lo, w[i].hi};
        for (uint32_t j = 0; j < L; j++) rv[64 + j] = h_x_inner_rest[j];
        for (uint32_t j = 0; j < nu; j++) rv[64 + L + j] = h_x_outer[j];
    }
    HIP_TRY(hipMemcpyAsync(RV, rv.data(), rv.size() * 16, hipMemcpyHostToDevice, c->stream));
    // 3. (on the aux stream, beside the combs) the slots' partial folds of their stripes at
    // x_outer into z_vec at their offsets (union.rs:354-384); 3b. capture z_vec
    if (int rc = c->fork()) return rc;
    LAUNCH(flock::launch_build_eq(RV + 64 + L, (int)nu, EO, EQS2, c->aux));
    if (!whole) LAUNCH(flock::launch_zero_words(ZV, K, c->aux));
    for (const LcSlotRun& sl : slots)
        LAUNCH(flock::launch_lc_partial_fold(sl.d_z, 1ull << sl.k_log, sl.useful_bits,
                                             (sl.n_rows + 7) / 8, EO, PF,
                                             ZV + (sl.prefix << sl.k_log), c->aux));
    if (d_z_vec) HIP_TRY(hipMemcpyAsync(d_z_vec, ZV, K * 16, hipMemcpyDeviceToDevice, c->aux));
    if (!whole) LAUNCH(flock::launch_zero_words(COMB, K, c->stream));
    for (const LcSlotRun& sl : slots) {
        const uint32_t inner = sl.k_log - ks;
        const uint64_t off = sl.prefix << sl.k_log;
        LAUNCH(flock::launch_build_eq(RV + 64, (int)inner, ER, EQS, c->stream));
        LAUNCH(flock::launch_outer_product(RV, 6, ER, (int)inner, EQI, c->stream));
        if (int rc = sl.fold(EQI, U(alpha), SA + off, SB + off, COMB + off, ERR)) {
            hipStreamSynchronize(c->aux);
            return rc;
        }
        F wt{1, 0};
        for (uint32_t j = 0; j < col_vars - sl.k_log; j++) {
            const F x = fo(h_x_inner_rest[inner + j]);
            wt = fmul(wt, ((sl.prefix >> j) & 1) ? x : fadd(x, F{1, 0}));
        }
        if (!feq(wt, F{1, 0}))
            LAUNCH(flock::launch_lc_scale(COMB + off, 1ull << sl.k_log, U(to(wt)), c->stream));
    }
    // 2b. each slot's constant-wire pin, in slot order (union.rs:318-335)
    for (const LcSlotRun& sl : slots) {
        if (sl.const_pin < 0) continue;
        if (grinding) h_nonces[ni++] = flock_fs_grind_pow(fs, 1);
        flock_f128 beta;
        flock_fs_sample_f128(fs, &beta);
        LAUNCH(flock::launch_lc_add_at(COMB, (sl.prefix << sl.k_log) + (uint64_t)sl.const_pin,
                                       U(beta), c->stream));
    }
    if (int rc = c->join()) return rc;
    // 4-5. product sumcheck over the top L bits (column_sumcheck_prove, lincheck.rs:1907-1980)
    // round t < L binds the top bit at challenge t-1 (t > 0) and emits message t; launch L
    // is the last binding
    std::vector<flock_f128> rr(L ? L : 1);
    if (L) {
        RoundLoop loop{c, fs, L, grinding ? 2u : 0u,
                       [&](uint32_t j, flock::RoundOut ro, const uint4* rin, uint4 r) -> int {
                           if (j == 0) {
                               LAUNCH(flock::launch_lc_eval(COMB, ZV, K, c->partials, ro,
                                                            c->stream));
                           } else if (j < L) {
                               LAUNCH(flock::launch_lc_bind_eval(COMB, ZV, K >> (j - 1), r,
                                                                 c->partials, ro, c->stream, rin));
                           } else {
                               LAUNCH(flock::launch_lc_bind(COMB, ZV, K >> (j - 1), r, c->stream,
                                                            rin));
                           }
                           return 0;
                       }};
        if (int rc = loop.run()) return rc;
        for (uint32_t t = 0; t < L; t++) {
            h_rounds[2 * t] = loop.msgs[2 * t];
            h_rounds[2 * t + 1] = loop.msgs[2 * t + 1];
            rr[t] = loop.rhos[t];
            if (grinding) h_nonces[ni++] = loop.nonces[t];
        }
    }
    // 6-8. z_partial = the first 64 words of the bound z, fresh skip point, w
    // (lincheck.rs:1983-2005)
    HIP_TRY(hipMemcpyAsync(h_z_partial, ZV, 64 * 16, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    flock_fs_observe_f128_slice(fs, h_z_partial, 64);
    if (grinding) h_nonces[ni++] = flock_fs_grind_pow(fs, (uint32_t)ks);
    flock_f128 r_skip;
    flock_fs_sample_f128(fs, &r_skip);
    {
        F lam[64];
        lagrange_phi8(0, F{r_skip.lo, r_skip.hi}, lam);
        F w{0, 0};
        for (int i = 0; i < 64; i++)
            w = fadd(w, fmul(lam[i], F{h_z_partial[i].lo, h_z_partial[i].hi}));
        h_claim[0] = r_skip;
        h_claim[1] = flock_f128{w.lo, w.hi};
    }
    // 9. r_inner_rest = reversed round challenges (LSB-first)
    for (uint32_t j = 0; j < L; j++) h_claim[2 + j] = rr[L - 1 - j];
    // 10. per slot, the unscaled matrix evals with W_col = z_partial (x)
    // eq(r_inner_rest[..k_t - 6]) (union.rs:404-430)
    HIP_TRY(hipMemcpyAsync(RV + 64, h_claim + 2, L * 16, hipMemcpyHostToDevice, c->stream));
    for (size_t t = 0; t < slots.size(); t++) {
        const uint32_t inner = slots[t].k_log - ks;
        const uint64_t off = slots[t].prefix << slots[t].k_log;
        LAUNCH(flock::launch_build_eq(RV + 64, (int)inner, ER, EQS, c->stream));
        LAUNCH(flock::launch_outer_product(ZV, 6, ER, (int)inner, EQI, c->stream));
        LAUNCH(flock::launch_dot2(EQI, SA + off, SB + off, 1ull << slots[t].k_log, c->partials,
                                  c->dout, c->stream));
        if (int rc = c->fetch_pair(h_matrix_evals + 2 * t)) return rc;
    }
    return 0;
}

int flock_hip_lincheck_prove_union(flock_hip_ctx* c, flock_fs* fs, const flock_lc_slot* slots,
                                   size_t n_slots, uint32_t col_vars, uint32_t nu,
                                   flock_f128 z_skip, const flock_f128* h_x_inner_rest,
                                   const flock_f128* h_x_outer, int grinding, flock_f128* h_rounds,
                                   flock_f128 h_z_partial[64], flock_f128* h_matrix_evals,
                                   uint64_t* h_nonces, flock_f128* h_claim, flock_f128* d_z_vec) {
    CTX_CHECK(c);
    if (!slots || n_slots == 0 || n_slots > FLOCK_MAX_TYPES)
        return fail(FLOCK_HIP_E_INVALID, "bad union lincheck slots (need 1..16)");
    std::vector<LcSlotRun> run;
    for (size_t t = 0; t < n_slots; t++) {
        const flock_lc_slot& s = slots[t];
        if (s.circuit) {
            const flock_csc* k = s.circuit;
            if (int rc = csc_check(k)) return rc;
            const LcFold fold = [c, k](const uint4* eq, uint4 alpha, uint4* sa, uint4* sb,
                                       uint4* comb, int* d_err) {
                return csc_fold_impl(c, k, eq, alpha, sa, sb, comb, d_err);
            };
            run.push_back({fold, k->k_log, s.prefix, k->const_pin, s.d_z, (size_t)s.useful_bits,
                           (size_t)s.n_rows});
        } else {
            if (int rc = c->ensure_kc()) return rc;
            const LcFold fold = [c](const uint4* eq, uint4 alpha, uint4* sa, uint4* sb,
                                    uint4* comb, int*) {
                LAUNCH(flock::launch_kc_fold(eq, c->kc_inv, c->kc_trow, c->kc_rcm, alpha, sa, sb,
                                             comb, c->stream));
                return 0;
            };
            run.push_back({fold, 16, s.prefix, 4096, s.d_z, 42560, (size_t)s.n_rows});
        }
    }
    return lincheck_prove_slots(c, fs, run, col_vars, nu, z_skip, h_x_inner_rest, h_x_outer,
                                grinding, h_rounds, h_z_partial, h_matrix_evals, h_nonces, h_claim,
                                d_z_vec);
}

// ---- lincheck verifier (lincheck/union.rs:662-960 with MatrixAssertion, :467-628) ----

namespace {
// eq(x, j) for every j < 2^n, coordinate k at bit k of j (build_eq_table)
std::vector<F> eq_table(const flock_f128* x, uint32_t n) {
    std::vector<F> t((size_t)1 << n, F{0, 0});
    t[0] = F{1, 0};
    for (uint32_t k = 0; k < n; k++) {
        const F xk = fo(x[k]), xk1 = fadd(F{1, 0}, xk);
        for (size_t j = (size_t)1 << k; j-- > 0;) {
            t[j | ((size_t)1 << k)] = fmul(t[j], xk);
            t[j] = fmul(t[j], xk1);
        }
    }
    return t;
}

// eq_prefix_sum (union.rs:107-130): sum of eq(p, i) over i < n, in O(d). The indices below n
// split by the highest bit where i drops below n; each such block contributes the eq factors
// of n's bits above it times eq(p_j, 0), its free low bits summing to one.
F eq_prefix_sum(const flock_f128* p, uint32_t d, uint64_t n) {
    if (n == 1ull << d) return F{1, 0};
    F acc{0, 0}, high{1, 0};
    for (uint32_t j = d; j-- > 0;) {
        const F x = fo(p[j]), x1 = fadd(F{1, 0}, x);
        if ((n >> j) & 1) {
            acc = fadd(acc, fmul(high, x1));
            high = fmul(high, x);
        } else {
            high = fmul(high, x1);
        }
    }
    return acc;
}

// CSC column sums of eq (CscCircuit::fold_split, lincheck.rs:463-497) on host arrays, with the
// structure checked as it is read.
bool host_csc_fold(uint64_t K, const uint32_t* ptr, const uint32_t* rows, uint64_t nnz,
                   const std::vector<F>& eq, std::vector<F>& out) {
    if (ptr[0] != 0 || ptr[K] != nnz) return false;
    for (uint64_t col = 0; col < K; col++) {
        if (ptr[col] > ptr[col + 1]) return false;
        F acc{0, 0};
        for (uint32_t k = ptr[col]; k < ptr[col + 1]; k++) {
            if (rows[k] >= K) return false;
            acc = fadd(acc, eq[rows[k]]);
        }
        out[col] = acc;
    }
    return true;
}
}  // namespace

// The circuit's column marginals on the host: (eq over the rows) -> (A_0^T eq, B_0^T eq); false
// when the circuit is malformed.
using HostFold = std::function<bool(const std::vector<F>& eq, std::vector<F>& sa,
                                    std::vector<F>& sb)>;

// One slot of the union lincheck verifier: its host fold (null: the reported matrix evals are
// checked instead), k_log, constant-wire pin (-1: none), Slot::prefix and statement count.
struct LcVSlot {
    const HostFold* fold;
    uint32_t k_log;
    int64_t const_pin;
    uint64_t prefix, n_rows;
};

// eq_prefix_weight (union.rs:90-101): prod_j eq(x_j, bit j of prefix).
static F eq_prefix_weight(const flock_f128* x, uint32_t n, uint64_t prefix) {
    F w{1, 0};
    for (uint32_t j = 0; j < n; j++)
        w = fmul(w, ((prefix >> j) & 1) ? fo(x[j]) : fadd(F{1, 0}, fo(x[j])));
    return w;
}

static int lincheck_verify_slots(flock_fs* fs, const std::vector<LcVSlot>& slots,
                                 uint32_t col_vars, uint32_t nu, flock_f128 z_skip,
                                 const flock_f128* h_x_inner_rest, const flock_f128* h_x_outer,
                                 flock_f128 v_a, flock_f128 v_b, int grinding,
                                 const flock_f128* h_rounds, const flock_f128 h_z_partial[64],
                                 const flock_f128* h_matrix_evals, const uint64_t* h_nonces,
                                 flock_f128* h_claim);

int flock_lincheck_verify(flock_fs* fs, const flock_csc* h_circuit, uint32_t k_log,
                          int64_t const_pin, uint32_t nu, size_t n_rows, flock_f128 z_skip,
                          const flock_f128* h_x_inner_rest, const flock_f128* h_x_outer,
                          flock_f128 v_a, flock_f128 v_b, int grinding,
                          const flock_f128* h_rounds, const flock_f128 h_z_partial[64],
                          const flock_f128 h_matrix_evals[2], const uint64_t* h_nonces,
                          flock_f128* h_claim) {
    if (!h_circuit) {
        const std::vector<LcVSlot> one{{nullptr, k_log, const_pin, 0, n_rows}};
        return lincheck_verify_slots(fs, one, k_log, nu, z_skip, h_x_inner_rest, h_x_outer, v_a,
                                     v_b, grinding, h_rounds, h_z_partial, h_matrix_evals,
                                     h_nonces, h_claim);
    }
    if (int rc = csc_check(h_circuit)) return rc;
    if (h_circuit->k_log != k_log || h_circuit->const_pin != const_pin)
        return fail(FLOCK_HIP_E_INVALID, "circuit k_log / const_pin differ from the type's");
    const flock_csc* k = h_circuit;
    const HostFold fold = [k](const std::vector<F>& eq, std::vector<F>& sa, std::vector<F>& sb) {
        const uint64_t K = 1ull << k->k_log;
        return host_csc_fold(K, k->a_col_ptr, k->a_rows, k->nnz_a, eq, sa) &&
               host_csc_fold(K, k->b_col_ptr, k->b_rows, k->nnz_b, eq, sb);
    };
    const std::vector<LcVSlot> one{{&fold, k_log, const_pin, 0, n_rows}};
    return lincheck_verify_slots(fs, one, k_log, nu, z_skip, h_x_inner_rest, h_x_outer, v_a, v_b,
                                 grinding, h_rounds, h_z_partial, h_matrix_evals, h_nonces,
                                 h_claim);
}

// The Keccak-f[1600] walk on the host (k_log 16), tables built once per process.
static const HostFold& keccak_host_fold() {
    static std::vector<uint16_t> inv, trow;
    static std::vector<uint32_t> rcm;
    static const bool built = kc_build_tables(inv, trow, rcm) == 0;
    static const HostFold fold = [](const std::vector<F>& eq, std::vector<F>& sa,
                                    std::vector<F>& sb) {
        if (!built) return false;
        kc_host_fold(eq.data(), sa.data(), sb.data(), inv, trow, rcm);
        return true;
    };
    return fold;
}

int flock_lincheck_verify_keccak(flock_fs* fs, uint32_t nu, size_t n_rows, flock_f128 z_skip,
                                 const flock_f128* h_x_inner_rest, const flock_f128* h_x_outer,
                                 flock_f128 v_a, flock_f128 v_b, int grinding,
                                 const flock_f128* h_rounds, const flock_f128 h_z_partial[64],
                                 const uint64_t* h_nonces, flock_f128* h_claim) {
    // K_LOG 16, const pin Z_CONST = 4096 (keccak.rs:436-451, 688-690)
    const std::vector<LcVSlot> one{{&keccak_host_fold(), 16, 4096, 0, n_rows}};
    return lincheck_verify_slots(fs, one, 16, nu, z_skip, h_x_inner_rest, h_x_outer, v_a, v_b,
                                 grinding, h_rounds, h_z_partial, nullptr, h_nonces, h_claim);
}

int flock_lincheck_verify_union(flock_fs* fs, const flock_lc_vslot* slots, size_t n_slots,
                                uint32_t col_vars, uint32_t nu, flock_f128 z_skip,
                                const flock_f128* h_x_inner_rest, const flock_f128* h_x_outer,
                                flock_f128 v_a, flock_f128 v_b, int grinding,
                                const flock_f128* h_rounds, const flock_f128 h_z_partial[64],
                                const flock_f128* h_matrix_evals, const uint64_t* h_nonces,
                                flock_f128* h_claim) {
    if (!slots || n_slots == 0 || n_slots > FLOCK_MAX_TYPES)
        return fail(FLOCK_HIP_E_INVALID, "bad union lincheck verify slots (need 1..16)");
    std::vector<HostFold> folds(n_slots);
    std::vector<LcVSlot> run(n_slots);
    for (size_t t = 0; t < n_slots; t++) {
        const flock_lc_vslot& s = slots[t];
        if (s.h_circuit) {
            const flock_csc* k = s.h_circuit;
            if (int rc = csc_check(k)) return rc;
            folds[t] = [k](const std::vector<F>& eq, std::vector<F>& sa, std::vector<F>& sb) {
                const uint64_t K = 1ull << k->k_log;
                return host_csc_fold(K, k->a_col_ptr, k->a_rows, k->nnz_a, eq, sa) &&
                       host_csc_fold(K, k->b_col_ptr, k->b_rows, k->nnz_b, eq, sb);
            };
            run[t] = {&folds[t], k->k_log, k->const_pin, s.prefix, s.n_rows};
        } else if (s.walker) {
            run[t] = {&keccak_host_fold(), 16, 4096, s.prefix, s.n_rows};
        } else {
            run[t] = {nullptr, s.k_log, s.const_pin, s.prefix, s.n_rows};
        }
    }
    return lincheck_verify_slots(fs, run, col_vars, nu, z_skip, h_x_inner_rest, h_x_outer, v_a,
                                 v_b, grinding, h_rounds, h_z_partial, h_matrix_evals, h_nonces,
                                 h_claim);
}

static int lincheck_verify_slots(flock_fs* fs, const std::vector<LcVSlot>& slots,
                                 uint32_t col_vars, uint32_t nu, flock_f128 z_skip,
                                 const flock_f128* h_x_inner_rest, const flock_f128* h_x_outer,
                                 flock_f128 v_a, flock_f128 v_b, int grinding,
                                 const flock_f128* h_rounds, const flock_f128 h_z_partial[64],
                                 const flock_f128* h_matrix_evals, const uint64_t* h_nonces,
                                 flock_f128* h_claim) {
    const int ks = 6;
    if (!fs || col_vars < (uint32_t)ks || col_vars > 24 || nu > 40 || slots.empty())
        return fail(FLOCK_HIP_E_INVALID, "bad lincheck verify arguments (need 6 <= col_vars <= "
                                         "24, nu <= 40, at least one slot)");
    const bool folded = slots[0].fold != nullptr;
    for (const LcVSlot& s : slots)
        if (s.k_log < (uint32_t)ks || s.k_log > col_vars || s.n_rows > (1ull << nu) ||
            s.const_pin >= ((int64_t)1 << s.k_log) || s.const_pin < -1 ||
            s.prefix >= (1ull << (col_vars - s.k_log)) || (s.fold != nullptr) != folded)
            return fail(FLOCK_HIP_E_INVALID, "bad lincheck verify slot (need 6 <= k_log <= "
                                             "col_vars, n_rows <= 2^nu, const_pin < 2^k_log, "
                                             "prefix < 2^(col_vars - k_log), every slot with "
                                             "a circuit or none)");
    const uint32_t L = col_vars - ks;
    if ((L && (!h_x_inner_rest || !h_rounds)) || (nu && !h_x_outer) || !h_z_partial ||
        (!folded && !h_matrix_evals) || (grinding && !h_nonces))
        return fail(FLOCK_HIP_E_INVALID, "bad lincheck verify arguments (null buffer)");
    int ni = 0;
    flock_fs_observe_label(fs, kLabelLincheck, sizeof(kLabelLincheck) - 1);
    // 1. alpha (union.rs:795-804)
    if (grinding && !flock_fs_verify_pow(fs, h_nonces[ni++], 1))
        return reject("alpha proof of work");
    flock_f128 alpha_v;
    flock_fs_sample_f128(fs, &alpha_v);
    const F alpha = fo(alpha_v);
    // 2b. each slot's const-wire pin in slot order: beta_t after alpha; the target gains
    // beta_t times the count-derived prefix sum of eq over the slot's declared rows
    // (union.rs:815-846)
    F target = fadd(fmul(alpha, fo(v_a)), fo(v_b));
    std::vector<F> beta(slots.size(), F{0, 0});
    for (size_t t = 0; t < slots.size(); t++) {
        if (slots[t].const_pin < 0) continue;
        if (grinding && !flock_fs_verify_pow(fs, h_nonces[ni++], 1))
            return reject("beta proof of work");
        flock_f128 b;
        flock_fs_sample_f128(fs, &b);
        beta[t] = fo(b);
        target = fadd(target, fmul(beta[t], eq_prefix_sum(h_x_outer, nu, slots[t].n_rows)));
    }
    // 3. the product-sumcheck rounds: q(0) = claim + q(1), q(X) = einf X^2 + c1 X + e0
    // (union.rs:851-873)
    F run = target;
    std::vector<flock_f128> rounds_r(L);
    for (uint32_t t = 0; t < L; t++) {
        const F e1 = fo(h_rounds[2 * t]), einf = fo(h_rounds[2 * t + 1]);
        flock_fs_observe_f128(fs, h_rounds[2 * t]);
        flock_fs_observe_f128(fs, h_rounds[2 * t + 1]);
        if (grinding && !flock_fs_verify_pow(fs, h_nonces[ni++], 2))
            return reject("sumcheck round proof of work");
        flock_fs_sample_f128(fs, &rounds_r[t]);
        const F r = fo(rounds_r[t]);
        const F e0 = fadd(run, e1);
        const F c1 = fadd(fadd(e0, e1), einf);
        run = fadd(fadd(fmul(einf, fmul(r, r)), fmul(c1, r)), e0);
    }
    // 4. z_partial after the rounds; rr = the challenges LSB-first (union.rs:876-883)
    flock_fs_observe_f128_slice(fs, h_z_partial, 64);
    std::vector<flock_f128> rr(L);
    for (uint32_t j = 0; j < L; j++) rr[j] = rounds_r[L - 1 - j];
    // 6-7. fresh skip challenge and w = <lambda(r_skip), z_partial> (union.rs:903-946)
    if (grinding && !flock_fs_verify_pow(fs, h_nonces[ni++], (uint32_t)ks))
        return reject("inner-skip proof of work");
    flock_f128 r_skip;
    flock_fs_sample_f128(fs, &r_skip);
    F w{0, 0};
    {
        F lam[64];
        lagrange_phi8(0, fo(r_skip), lam);
        for (int i = 0; i < 64; i++) w = fadd(w, fmul(lam[i], fo(h_z_partial[i])));
    }
    // the matrix assertion (union.rs:467-628), per slot with the row-side prefix weight w_t
    // (at x_inner_rest) and the column-side p_t (at rr)
    F acc{0, 0};
    if (!folded) {
        // check_reported: sum_t p_t w_t (alpha <W_t,A_0> + <W_t,B_0>) + p_t beta_t W_col_t[pin],
        // W_col_t = z_partial (x) eq(rr[..k_t - 6])
        for (size_t t = 0; t < slots.size(); t++) {
            const uint32_t inner = slots[t].k_log - ks, hi = col_vars - slots[t].k_log;
            const F wt = eq_prefix_weight(h_x_inner_rest + inner, hi, slots[t].prefix);
            const F pt = eq_prefix_weight(rr.data() + inner, hi, slots[t].prefix);
            acc = fadd(acc, fmul(fmul(pt, wt), fadd(fmul(alpha, fo(h_matrix_evals[2 * t])),
                                                    fo(h_matrix_evals[2 * t + 1]))));
            if (slots[t].const_pin >= 0) {
                const int64_t pin = slots[t].const_pin;
                F e{1, 0};
                for (uint32_t j = 0; j < inner; j++)
                    e = fmul(e, ((pin >> (6 + j)) & 1) ? fo(rr[j]) : fadd(F{1, 0}, fo(rr[j])));
                acc = fadd(acc, fmul(fmul(pt, beta[t]), fmul(fo(h_z_partial[pin & 63]), e)));
            }
        }
        if (!feq(acc, run)) return reject("matrix evals");
    } else {
        // check: per slot comb_t = w_t (alpha A_0^T eq + B_0^T eq) at the quirky eq of
        // (z_skip, x_inner_rest[..k_t - 6]), + beta_t at the pin; collapsed at rr and p_t
        // against z_partial
        F part[64];
        for (int i = 0; i < 64; i++) part[i] = F{0, 0};
        F lam[64];
        lagrange_phi8(0, fo(z_skip), lam);
        for (size_t t = 0; t < slots.size(); t++) {
            const uint32_t inner = slots[t].k_log - ks, hi = col_vars - slots[t].k_log;
            const uint64_t K = 1ull << slots[t].k_log;
            std::vector<F> eqi(K), sa(K), sb(K);
            const std::vector<F> ex = eq_table(h_x_inner_rest, inner);
            for (uint64_t j = 0; j < ((uint64_t)1 << inner); j++)
                for (int i = 0; i < 64; i++) eqi[j * 64 + i] = fmul(lam[i], ex[j]);
            if (!(*slots[t].fold)(eqi, sa, sb))
                return fail(FLOCK_HIP_E_INVALID, "CSC circuit has a row index >= n_cols or a "
                                                 "col_ptr entry out of order or > nnz");
            const F wt = eq_prefix_weight(h_x_inner_rest + inner, hi, slots[t].prefix);
            const F pt = eq_prefix_weight(rr.data() + inner, hi, slots[t].prefix);
            const std::vector<F> err = eq_table(rr.data(), inner);
            for (uint64_t j = 0; j < ((uint64_t)1 << inner); j++) {
                const F ej = fmul(pt, err[j]);
                for (int i = 0; i < 64; i++) {
                    F cmb = fmul(wt, fadd(fmul(alpha, sa[j * 64 + i]), sb[j * 64 + i]));
                    if ((int64_t)(j * 64 + i) == slots[t].const_pin) cmb = fadd(cmb, beta[t]);
                    part[i] = fadd(part[i], fmul(ej, cmb));
                }
            }
        }
        for (int i = 0; i < 64; i++) acc = fadd(acc, fmul(part[i], fo(h_z_partial[i])));
        if (!feq(acc, run)) return reject("sumcheck final consistency");
    }
    if (h_claim) {
        h_claim[0] = r_skip;
        h_claim[1] = to(w);
        for (uint32_t j = 0; j < L; j++) h_claim[2 + j] = rr[j];
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Union instance host glue (crates/flock-core/src/{schedule,union}.rs)
// ---------------------------------------------------------------------------
static unsigned pow2_log(uint64_t n) {
    unsigned l = 0;
    while ((1ull << l) < n) l++;
    return l;
}

int flock_registry_new(const flock_table_type* types, size_t n, uint32_t nu,
                       flock_registry* out) {
    if (!types || !out || n == 0 || n > FLOCK_MAX_TYPES || nu > 40)
        return fail(FLOCK_HIP_E_INVALID, "registry needs 1..16 types and nu <= 40");
    for (size_t i = 0; i < n; i++) {
        const flock_table_type& t = types[i];
        if (t.k_log < 7 || t.k_log > 40 || t.useful_bits > (1ull << t.k_log) ||
            (t.is_element && t.useful_bits % 128) || nu + t.k_log > 62)
            return fail(FLOCK_HIP_E_INVALID, "bad table type (7 <= k_log <= 40, useful_bits <= "
                                             "2^k_log, element useful_bits = 128 k)");
    }
    std::memset(out, 0, sizeof(*out));
    // class-major (boolean first), then k_log descending, stable
    size_t order[FLOCK_MAX_TYPES];
    size_t m = 0;
    for (int cls = 0; cls < 2; cls++)
        for (int k = 40; k >= 7; k--)
            for (size_t i = 0; i < n; i++)
                if ((types[i].is_element != 0) == (cls == 1) && (int)types[i].k_log == k)
                    order[m++] = i;
    uint64_t bool_extent = 0, elem_extent = 0;
    uint32_t nb = 0;
    for (size_t s = 0; s < n; s++) {
        const flock_table_type& t = types[order[s]];
        if (t.is_element) {
            elem_extent += 1ull << (nu + t.k_log);
        } else {
            bool_extent += 1ull << (nu + t.k_log);
            nb++;
        }
    }
    out->m_bool = bool_extent ? pow2_log(bool_extent) : 0;
    out->m_elem = elem_extent ? pow2_log(elem_extent) : 0;
    out->element_base = (elem_extent == 0 || bool_extent == 0)
                            ? 0
                            : 1ull << (out->m_bool > out->m_elem ? out->m_bool : out->m_elem);
    uint64_t offset = 0;
    for (size_t s = 0; s < n; s++) {
        if (s == nb) offset = out->element_base;
        const uint32_t ms = nu + types[order[s]].k_log;
        if (offset & ((1ull << ms) - 1)) return fail(FLOCK_HIP_E_INVALID, "slot misaligned");
        out->slots[s].type_index = (uint32_t)order[s];
        out->slots[s].m_slot = ms;
        out->slots[s].offset = offset;
        out->types[s] = types[order[s]];
        offset += 1ull << ms;
    }
    out->m_total = pow2_log(offset);
    for (size_t s = 0; s < n; s++) {
        out->slots[s].prefix = out->slots[s].offset >> out->slots[s].m_slot;
        out->slots[s].prefix_bits = out->m_total - out->slots[s].m_slot;
    }
    out->n_types = (uint32_t)n;
    out->num_boolean = nb;
    out->nu = nu;
    return 0;
}

static uint64_t union_used_cols(const flock_table_type& t) {
    const uint64_t u = (t.useful_bits + 127) / 128, cap = 1ull << (t.k_log - 7);
    return u < cap ? u : cap;
}

int flock_union_sizes_compute(const flock_registry* reg, const uint64_t* counts,
                              uint32_t dense_floor_m, uint32_t log_batch_size,
                              flock_union_sizes* out) {
    if (!reg || !counts || !out || reg->n_types == 0 || reg->n_types > FLOCK_MAX_TYPES)
        return fail(FLOCK_HIP_E_INVALID, "bad union size arguments");
    const unsigned kMinDenseM = 22;  // union.rs:45
    uint64_t dense = 0;
    for (uint32_t s = 0; s < reg->n_types; s++) {
        if (counts[s] > (1ull << reg->nu))
            return fail(FLOCK_HIP_E_INVALID, "count exceeds the row capacity 2^nu");
        dense += union_used_cols(reg->types[s]) * counts[s];
    }
    const uint64_t packed = 1ull << (reg->m_total - 7);
    if (dense_floor_m && (dense_floor_m < kMinDenseM || (1ull << (dense_floor_m - 7)) > packed))
        return fail(FLOCK_HIP_E_INVALID, "dense floor below 22 or above the padded domain");
    uint64_t floor = 1ull << (kMinDenseM - 7);
    if (dense_floor_m && (1ull << (dense_floor_m - 7)) > floor) floor = 1ull << (dense_floor_m - 7);
    uint64_t committed = 1;
    while (committed < dense) committed <<= 1;
    if (committed < floor) committed = floor;
    if (committed > packed) committed = packed;
    out->dense_words = dense;
    out->committed_words = committed;
    out->packed_len = packed;
    out->dense_m = pow2_log(committed) + 7;
    if (out->dense_m < 7 + log_batch_size)
        return fail(FLOCK_HIP_E_INVALID, "log_batch_size exceeds dense_m - 7");
    const unsigned log_dim = out->dense_m - 7 - log_batch_size;
    uint64_t t = (dense + (1ull << log_dim) - 1) >> log_dim;  // dense_lanes, pcs/commit.rs:428
    if (t < 1) t = 1;
    if (t > (1ull << log_batch_size)) t = 1ull << log_batch_size;
    out->commit_lanes = t < (1ull << log_batch_size) ? t : 0;
    bool ident = committed == packed;
    uint64_t cursor = 0;
    for (uint32_t s = 0; s < reg->n_types && ident; s++) {
        if (cursor != (reg->slots[s].offset >> 7) || counts[s] != (1ull << reg->nu))
            ident = false;
        cursor += union_used_cols(reg->types[s]) << reg->nu;
    }
    out->compaction_is_identity = ident ? 1u : 0u;
    return 0;
}

int flock_union_x_ab_from_mlv(const flock_registry* reg, const flock_f128* mlv, size_t n_mlv,
                              flock_f128* x_inner_rest, flock_f128* x_outer) {
    if (!reg || !mlv || !x_inner_rest || !x_outer || reg->num_boolean == 0 ||
        n_mlv != (size_t)reg->m_bool - 6 || n_mlv < 1 + (size_t)reg->nu)
        return fail(FLOCK_HIP_E_INVALID, "mlv must have m_bool - 6 entries");
    x_inner_rest[0] = mlv[0];
    for (size_t j = 1 + reg->nu; j < n_mlv; j++) x_inner_rest[j - reg->nu] = mlv[j];
    for (uint32_t j = 0; j < reg->nu; j++) x_outer[j] = mlv[1 + j];
    return 0;
}

void flock_blake3_hash(const void* data, size_t len, uint8_t out[32]) {
    flock::Blake3Host h;
    if (len) h.update(data, len);
    h.finish(out);
}

// absorb_matrix, r1cs.rs:419-428: (num_rows, num_cols, then per row its length and columns),
// every integer u64 little-endian
static int absorb_matrix(flock::Blake3Host& h, const flock_sparse_matrix& m) {
    if (m.identity) {
        if (m.num_rows != m.num_cols) return fail(FLOCK_HIP_E_INVALID, "identity must be square");
        h.u64(m.num_rows);
        h.u64(m.num_cols);
        for (uint64_t r = 0; r < m.num_rows; r++) {
            h.u64(1);
            h.u64(r);
        }
        return 0;
    }
    if (m.num_rows && (!m.row_ptr || (m.row_ptr[m.num_rows] > m.row_ptr[0] && !m.cols)))
        return fail(FLOCK_HIP_E_INVALID, "null CSR arrays");
    for (uint64_t r = 0; r < m.num_rows; r++)
        if (m.row_ptr[r + 1] < m.row_ptr[r]) return fail(FLOCK_HIP_E_INVALID, "row_ptr decreases");
    h.u64(m.num_rows);
    h.u64(m.num_cols);
    std::vector<uint8_t> row;
    for (uint64_t r = 0; r < m.num_rows; r++) {
        const uint64_t a = m.row_ptr[r], b = m.row_ptr[r + 1];
        row.resize(8 * (b - a + 1));
        const uint64_t len = b - a;
        for (int k = 0; k < 8; k++) row[k] = (uint8_t)(len >> (8 * k));
        for (uint64_t i = a; i < b; i++)
            for (int k = 0; k < 8; k++) row[8 * (i - a + 1) + k] = (uint8_t)((uint64_t)m.cols[i] >> (8 * k));
        h.update(row.data(), row.size());
    }
    return 0;
}

int flock_registry_digest(const flock_type_statement* types, size_t n, uint32_t nu,
                          uint8_t out[32]) {
    if (!types || !out || n == 0 || n > FLOCK_MAX_TYPES)
        return fail(FLOCK_HIP_E_INVALID, "bad registry digest arguments");
    for (size_t i = 0; i < n; i++) {
        const flock_type_statement& t = types[i];
        if (t.k_log < 7 || t.k_log > 40 || t.useful_bits > (1ull << t.k_log))
            return fail(FLOCK_HIP_E_INVALID, "table type out of range (schedule.rs:337-355)");
    }
    // Registry::new's order: stable sort by k_log, descending (schedule.rs:386-389)
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return types[a].k_log > types[b].k_log; });
    static const uint8_t kLabel[] = "flock-registry-v1";
    flock::Blake3Host h;
    h.update(kLabel, sizeof(kLabel) - 1);
    const uint8_t version = 1;
    h.update(&version, 1);
    h.u32(nu);
    h.u32((uint32_t)n);
    for (size_t i : order) {
        const flock_type_statement& t = types[i];
        h.u32(t.k_log);
        h.u64(t.useful_bits);
        const uint8_t present = t.has_const_pin ? 1 : 0;
        h.update(&present, 1);
        h.u64(t.has_const_pin ? t.const_pin : 0);
        for (const flock_sparse_matrix* m : {&t.a0, &t.b0, &t.c0})
            if (int rc = absorb_matrix(h, *m)) return rc;
    }
    h.finish(out);
    return 0;
}

static const uint8_t kLabelMixed[] = "flock-mixed-v1";

int flock_union_bind_statement(flock_fs* fs, const uint8_t digest[32], const uint64_t* counts,
                               size_t n, const uint8_t* cap, size_t cap_len) {
    if (!fs || !digest || (n && !counts) || (cap_len && !cap))
        return fail(FLOCK_HIP_E_INVALID, "bad bind_statement arguments");
    flock_fs_observe_label(fs, kLabelMixed, sizeof(kLabelMixed) - 1);
    flock_fs_observe_bytes(fs, digest, 32);
    std::vector<uint8_t> le(8 * n);
    for (size_t i = 0; i < n; i++)
        for (int b = 0; b < 8; b++) le[8 * i + b] = (uint8_t)(counts[i] >> (8 * b));
    flock_fs_observe_bytes(fs, le.data(), le.size());
    flock_fs_observe_bytes(fs, cap, cap_len);
    return 0;
}

int flock_hip_union_compact(flock_hip_ctx* c, const flock_registry* reg, const uint64_t* counts,
                            uint32_t dense_floor_m, const flock_f128* z, flock_f128* q) {
    CTX_CHECK(c);
    if (!z || !q) return fail(FLOCK_HIP_E_INVALID, "bad compact arguments");
    flock_union_sizes sz;
    if (int rc = flock_union_sizes_compute(reg, counts, dense_floor_m, 0, &sz)) return rc;
    uint64_t cursor = 0;
    for (uint32_t s = 0; s < reg->n_types; s++) {
        const uint64_t cols = union_used_cols(reg->types[s]);
        LAUNCH(flock::launch_compact_slot(U(z), reg->slots[s].offset >> 7, reg->nu, cols,
                                          counts[s], U(q) + cursor, c->stream));
        cursor += cols * counts[s];
    }
    LAUNCH(flock::launch_zero_words(U(q) + cursor, sz.committed_words - cursor, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// Witness generation (crates/flock-prover/src/r1cs_hashes)
// ---------------------------------------------------------------------------
int flock_hip_witness_sha256(flock_hip_ctx* c, const uint32_t* d_inputs, size_t n_decl,
                             uint32_t n_log, flock_f128* d_z, flock_f128* d_a, flock_f128* d_b,
                             uint8_t* d_stripe) {
    CTX_CHECK(c);
    if (n_log < 3 || n_log > 26 || n_decl > (1ull << n_log) || (n_decl && !d_inputs) || !d_z ||
        !d_a || !d_b || (reinterpret_cast<uintptr_t>(d_stripe) & 7))
        return fail(FLOCK_HIP_E_INVALID, "bad SHA-256 witness arguments (3 <= n_log <= 26, "
                                         "n_decl <= 2^n_log, 8-byte aligned stripe)");
    LAUNCH(flock::launch_sha2_witness(d_inputs, n_decl, n_log, U(d_z), U(d_a), U(d_b), d_stripe,
                                      c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int flock_hip_witness_keccak(flock_hip_ctx* c, const uint64_t* d_inputs, size_t n_decl,
                             uint32_t n_log, flock_f128* d_z, flock_f128* d_a, flock_f128* d_b,
                             uint8_t* d_stripe) {
    CTX_CHECK(c);
    if (n_log < 3 || n_log > 24 || n_decl > (1ull << n_log) || (n_decl && !d_inputs) || !d_z ||
        !d_a || !d_b || (reinterpret_cast<uintptr_t>(d_stripe) & 7))
        return fail(FLOCK_HIP_E_INVALID, "bad Keccak witness arguments (3 <= n_log <= 24, "
                                         "n_decl <= 2^n_log, 8-byte aligned stripe)");
    LAUNCH(flock::launch_keccak_witness(d_inputs, n_decl, n_log, U(d_z), U(d_a), U(d_b),
                                        d_stripe, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int flock_hip_witness_blake3(flock_hip_ctx* c, const uint32_t* d_inputs, size_t n_decl,
                             uint32_t n_log, flock_f128* d_z, flock_f128* d_a, flock_f128* d_b,
                             uint8_t* d_stripe) {
    CTX_CHECK(c);
    if (n_log < 3 || n_log > 26 || n_decl > (1ull << n_log) || (n_decl && !d_inputs) || !d_z ||
        !d_a || !d_b || (reinterpret_cast<uintptr_t>(d_stripe) & 7))
        return fail(FLOCK_HIP_E_INVALID, "bad BLAKE3 witness arguments (3 <= n_log <= 26, "
                                         "n_decl <= 2^n_log, 8-byte aligned stripe)");
    LAUNCH(flock::launch_blake3_witness(d_inputs, n_decl, n_log, U(d_z), U(d_a), U(d_b),
                                        d_stripe, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

}  // extern "C"
