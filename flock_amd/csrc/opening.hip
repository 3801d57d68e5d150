// The merged PCS opening of prove_fast on the GPU (crates/flock-core/src/pcs.rs:1433-1822).
//
// Host orchestration over device-resident tables, in the reference's transcript order:
//   1. ring switch of the ring-switched claims (pcs/ring_switch.rs:2723-2800) — host, 128 words;
//   2. the jagged merged weight W[d] = sum_c Phi_c(eq_row_c[row(d)] eq_col_c[col(d)]) over the
//      dense cube with its round-0 message fused (pcs/jagged.rs:1521-1650), then the dense merged
//      sumcheck <q, W> (pcs.rs:1640-1700) on the product-sumcheck kernels;
//   3. on a forked transcript, the multipoint twisted assist (pcs/jagged.rs:2884-3290): the
//      128 dual values per claim by the jagged branching program, the two-product sumcheck
//      sum_d abar_d g_d with g = L_gamma(eq(rho, .)), and the Frobenius anchor (host);
//   4. on the parent transcript, the eq-basis Ligerito opening of q^(rho): the F256 ladder with
//      code switches, level commits (NTT + Merkle), OOD samples, stratified queries and capped
//      paths (pcs.rs:214-560, pcs/ligerito/extension.rs:1636-2290);
//   5. the bincode 1.3 encoding of MergedOpenProof (pcs.rs:1362-1377).
//
// Every sum the protocol sends is an XOR reduction, so kernels reduce by block and combine
// block results with atomicXor: the value does not depend on the order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <condition_variable>
#include <mutex>
#include <vector>

#if defined(__x86_64__)
#include <cpuid.h>
#include <emmintrin.h>
#include <wmmintrin.h>
#endif

#include "../../include/flock_hip.h"
#include "f128_device.h"
#include "flock_internal.h"

namespace {

// ---------------------------------------------------------------------------
// host field arithmetic
// ---------------------------------------------------------------------------
using H = flock_f128;
inline H hz() { return H{0, 0}; }
inline H h1() { return H{1, 0}; }
inline H ha(H a, H b) { return H{a.lo ^ b.lo, a.hi ^ b.hi}; }
inline bool hzero(H a) { return !(a.lo | a.hi); }
inline H hbit(int t) {
    return H{t < 64 ? 1ull << t : 0, t >= 64 ? 1ull << (t - 64) : 0};
}
inline int hgetbit(H x, int t) { return (int)(((t < 64 ? x.lo : x.hi) >> (t & 63)) & 1); }

inline H hreduce(uint64_t r0, uint64_t r1, uint64_t r2, uint64_t r3) {
    uint64_t tl = r2 ^ (r2 << 1) ^ (r2 << 2) ^ (r2 << 7);
    uint64_t th = r3 ^ ((r3 << 1) | (r2 >> 63)) ^ ((r3 << 2) | (r2 >> 62)) ^ ((r3 << 7) | (r2 >> 57));
    uint64_t ov = (r3 >> 63) ^ (r3 >> 62) ^ (r3 >> 57);
    uint64_t corr = ov ^ (ov << 1) ^ (ov << 2) ^ (ov << 7);
    return H{r0 ^ tl ^ corr, r1 ^ th};
}
#if defined(__x86_64__)
__attribute__((target("pclmul,sse2"))) H hmul_clmul(H a, H b) {
    __m128i A = _mm_set_epi64x((long long)a.hi, (long long)a.lo);
    __m128i B = _mm_set_epi64x((long long)b.hi, (long long)b.lo);
    __m128i ll = _mm_clmulepi64_si128(A, B, 0x00), hh = _mm_clmulepi64_si128(A, B, 0x11);
    __m128i cr = _mm_xor_si128(_mm_clmulepi64_si128(A, B, 0x10), _mm_clmulepi64_si128(A, B, 0x01));
    uint64_t l[2], h[2], c[2];
    _mm_storeu_si128((__m128i*)l, ll);
    _mm_storeu_si128((__m128i*)h, hh);
    _mm_storeu_si128((__m128i*)c, cr);
    return hreduce(l[0], l[1] ^ c[0], h[0] ^ c[1], h[1]);
}
bool has_clmul() {
    static const bool v = [] {
        unsigned a, b, c, d;
        return __get_cpuid(1, &a, &b, &c, &d) && (c & bit_PCLMUL);
    }();
    return v;
}
#endif
void clmul64h(uint64_t a, uint64_t b, uint64_t* lo, uint64_t* hi) {
    uint64_t l = 0, h = 0;
    for (int i = 0; i < 64; i++)
        if ((a >> i) & 1) {
            l ^= b << i;
            if (i) h ^= b >> (64 - i);
        }
    *lo = l;
    *hi = h;
}
H hm(H a, H b) {
#if defined(__x86_64__)
    if (has_clmul()) return hmul_clmul(a, b);
#endif
    uint64_t ll, lh, hl, hh, t0, t1, t2, t3;
    clmul64h(a.lo, b.lo, &ll, &t0);
    clmul64h(a.lo, b.hi, &lh, &t1);
    clmul64h(a.hi, b.lo, &hl, &t2);
    clmul64h(a.hi, b.hi, &hh, &t3);
    return hreduce(ll, t0 ^ lh ^ hl, hh ^ t1 ^ t2, t3);
}
H hinv(H x) {  // x^(2^128 - 2)
    H r = h1(), cur = hm(x, x);
    for (int i = 1; i < 128; i++) {
        r = hm(r, cur);
        cur = hm(cur, cur);
    }
    return r;
}
std::vector<H> heq(const H* r, size_t n) {  // build_eq, LSB-first
    std::vector<H> t((size_t)1 << n);
    t[0] = h1();
    for (size_t i = 0; i < n; i++) {
        const size_t len = (size_t)1 << i;
        const H rp = ha(h1(), r[i]);
        for (size_t x = 0; x < len; x++) {
            t[x + len] = hm(t[x], r[i]);
            t[x] = hm(t[x], rp);
        }
    }
    return t;
}
H heq_at(const H* a, const H* b, size_t n) {
    H acc = h1();
    for (size_t i = 0; i < n; i++) acc = hm(acc, ha(hm(a[i], b[i]), hm(ha(h1(), a[i]), ha(h1(), b[i]))));
    return acc;
}

// F256 = F128[u]/(u^2 + u + x^-1) (field/gf2_256.rs:1-160)
struct E {
    H c0, c1;
};
inline E ez() { return E{hz(), hz()}; }
inline E e1() { return E{h1(), hz()}; }
inline E ea(E a, E b) { return E{ha(a.c0, b.c0), ha(a.c1, b.c1)}; }
inline H hmulxinv(H z) {
    const uint64_t mask = 0 - (z.lo & 1);
    return H{((z.lo >> 1) | (z.hi << 63)) ^ (0x43 & mask), (z.hi >> 1) ^ ((1ull << 63) & mask)};
}
inline E em(E a, E b) {
    H p0 = hm(a.c0, b.c0), p1 = hm(a.c1, b.c1), p2 = hm(ha(a.c0, a.c1), ha(b.c0, b.c1));
    return E{ha(p0, hmulxinv(p1)), ha(p2, p0)};
}
inline E emb(E a, H b) { return E{hm(a.c0, b), hm(a.c1, b)}; }
std::vector<E> heq256(const E* r, size_t n) {
    std::vector<E> t((size_t)1 << n);
    t[0] = e1();
    for (size_t j = 0; j < n; j++) {
        const size_t len = (size_t)1 << j;
        const E rp = ea(e1(), r[j]);
        for (size_t i = 0; i < len; i++) {
            const E v = t[i];
            t[i + len] = em(v, r[j]);
            t[i] = em(v, rp);
        }
    }
    return t;
}

// x^(2^-1) = x^(2^127) through the square-root images of the bit basis (jagged.rs:2266-2300)
const H* sqrt_basis() {
    static const std::vector<H> t = [] {
        std::vector<H> v(128);
        for (int b = 0; b < 128; b++) {
            H y = hbit(b);
            for (int i = 0; i < 127; i++) y = hm(y, y);
            v[b] = y;
        }
        return v;
    }();
    return t.data();
}
// the same F2-linear map through 16 byte tables: T[k][v] = frob_inv(v x^(8k))
const H* frob_inv_tables() {
    static const std::vector<H> t = [] {
        const H* sb = sqrt_basis();
        std::vector<H> v(16 * 256);
        for (int k = 0; k < 16; k++)
            for (int n = 1; n < 256; n++) {
                const int low = __builtin_ctz((unsigned)n);
                v[k * 256 + n] = ha(v[k * 256 + (n & (n - 1))], sb[8 * k + low]);
            }
        return v;
    }();
    return t.data();
}
H frob_inv(H x) {
    const H* t = frob_inv_tables();
    H acc = hz();
    for (int k = 0; k < 8; k++) {
        acc = ha(acc, t[k * 256 + ((x.lo >> (8 * k)) & 255)]);
        acc = ha(acc, t[(k + 8) * 256 + ((x.hi >> (8 * k)) & 255)]);
    }
    return acc;
}

// ---------------------------------------------------------------------------
// bincode 1.3 (fixint, little-endian, u64 lengths)
// ---------------------------------------------------------------------------
struct Bin {
    std::vector<uint8_t> b;
    void u64(uint64_t v) {
        for (int i = 0; i < 8; i++) b.push_back((uint8_t)(v >> (8 * i)));
    }
    void f(H v) {
        u64(v.lo);
        u64(v.hi);
    }
    void e(E v) {
        f(v.c0);
        f(v.c1);
    }
    void raw(const uint8_t* p, size_t n) { b.insert(b.end(), p, p + n); }
    void vec_f(const H* v, size_t n) {
        u64(n);
        for (size_t i = 0; i < n; i++) f(v[i]);
    }
    void vec_u64(const uint64_t* v, size_t n) {
        u64(n);
        for (size_t i = 0; i < n; i++) u64(v[i]);
    }
    void vec_hash(const uint8_t* h, size_t n) {
        u64(n);
        raw(h, 32 * n);
    }
};

// ---------------------------------------------------------------------------
// Ligerito `fast` configurations (crates/flock-core/configs/ligerito/m{22..35}_fast.toml,
// looked up by ligerito.rs:542-620): per level rate, log message columns, log lanes,
// queries, query-PoW bits, claim-batch bits, consistency-batch bits, OOD samples.
// ---------------------------------------------------------------------------
struct Lvl {
    uint8_t rate, cols, lanes;
    uint16_t queries;
    uint8_t grind, claim, cons, ood;
};
struct Lig {
    uint8_t m, initial_k, n_levels;
    Lvl lv[7];
};
const Lig kLigFast[] = {
    {22, 6, 3, {{1, 9, 6, 279, 0, 6, 9, 1}, {2, 6, 4, 136, 0, 6, 9, 2}, {4, 3, 4, 68, 0, 7, 10, 2}}},
    {23, 6, 3, {{1, 10, 6, 279, 0, 6, 9, 1}, {2, 7, 4, 136, 0, 6, 9, 2}, {3, 4, 4, 91, 0, 7, 9, 2}}},
    {24, 6, 3, {{1, 11, 6, 279, 0, 6, 9, 1}, {2, 8, 4, 136, 0, 6, 9, 2}, {3, 5, 4, 91, 0, 7, 9, 2}}},
    {25, 6, 4, {{1, 12, 6, 279, 0, 6, 9, 1}, {2, 9, 4, 136, 0, 6, 9, 2}, {3, 6, 4, 91, 0, 7, 9, 2},
                {4, 3, 4, 68, 0, 7, 10, 2}}},
    {26, 6, 4, {{1, 13, 6, 279, 0, 6, 9, 1}, {2, 10, 4, 136, 0, 6, 9, 2}, {3, 7, 4, 91, 0, 7, 9, 2},
                {4, 4, 4, 68, 0, 7, 10, 2}}},
    {27, 6, 4, {{1, 14, 6, 279, 0, 6, 9, 1}, {2, 11, 4, 136, 0, 6, 9, 2}, {3, 8, 4, 91, 0, 7, 9, 2},
                {4, 5, 4, 68, 0, 7, 10, 2}}},
    {28, 4, 5, {{1, 17, 4, 279, 0, 6, 9, 1}, {2, 14, 4, 136, 0, 6, 9, 2}, {3, 11, 4, 91, 0, 7, 9, 2},
                {4, 8, 4, 68, 0, 7, 10, 2}, {5, 5, 4, 55, 0, 8, 10, 2}}},
    {29, 5, 5, {{1, 17, 5, 279, 0, 6, 9, 1}, {2, 14, 4, 136, 0, 6, 9, 2}, {3, 11, 4, 91, 0, 7, 9, 2},
                {4, 8, 4, 68, 0, 7, 10, 2}, {5, 5, 4, 55, 0, 8, 10, 2}}},
    {30, 6, 5, {{1, 17, 6, 279, 0, 6, 9, 1}, {2, 14, 4, 136, 0, 6, 9, 2}, {3, 11, 4, 91, 0, 7, 9, 2},
                {4, 8, 4, 68, 0, 7, 10, 2}, {5, 5, 4, 55, 0, 8, 10, 2}}},
    {31, 6, 6, {{1, 18, 6, 279, 0, 6, 9, 1}, {2, 15, 4, 136, 0, 6, 9, 2}, {3, 12, 4, 91, 0, 7, 9, 2},
                {4, 9, 4, 68, 0, 7, 10, 2}, {5, 6, 4, 55, 0, 8, 10, 2}, {6, 3, 4, 46, 0, 8, 11, 2}}},
    {32, 6, 6, {{1, 19, 6, 279, 0, 6, 9, 1}, {2, 16, 4, 136, 0, 6, 9, 2}, {3, 13, 4, 91, 0, 7, 9, 2},
                {4, 10, 4, 68, 0, 7, 10, 2}, {5, 7, 4, 55, 0, 8, 10, 2}, {6, 4, 4, 46, 0, 8, 11, 2}}},
    {33, 6, 6, {{1, 20, 6, 279, 0, 6, 9, 1}, {2, 17, 4, 136, 0, 6, 9, 2}, {3, 14, 4, 91, 0, 7, 9, 2},
                {4, 11, 4, 68, 0, 7, 10, 2}, {5, 8, 4, 55, 0, 8, 10, 2}, {6, 5, 4, 46, 0, 8, 11, 2}}},
    {34, 6, 7, {{1, 21, 6, 279, 0, 6, 9, 1}, {2, 18, 4, 136, 0, 6, 9, 2}, {3, 15, 4, 91, 0, 7, 9, 2},
                {4, 12, 4, 68, 0, 7, 10, 2}, {5, 9, 4, 55, 0, 8, 10, 2}, {6, 6, 4, 46, 0, 8, 11, 2},
                {7, 3, 4, 40, 0, 9, 11, 2}}},
    {35, 6, 7, {{1, 22, 6, 279, 0, 6, 9, 1}, {2, 19, 4, 136, 0, 6, 9, 2}, {3, 16, 4, 91, 0, 7, 9, 2},
                {4, 13, 4, 68, 0, 7, 10, 2}, {5, 10, 4, 55, 0, 8, 10, 2}, {6, 7, 4, 46, 0, 8, 11, 2},
                {7, 4, 4, 40, 0, 9, 11, 2}}},
};
const Lig* lig_for(uint32_t m) {
    for (const Lig& c : kLigFast)
        if (c.m == m) return &c;
    return nullptr;
}

// stratified schedule (pcs/stratified.rs:73-130)
struct Sched {
    int d = 0;
    std::vector<int> depth;
    int cap() const { return depth.empty() ? 0 : depth[0]; }
    size_t queries() const {
        size_t q = 0;
        for (int c : depth) q += (size_t)1 << c;
        return q;
    }
};
Sched sched_of(size_t q, int d) {
    Sched s;
    s.d = d;
    while (q) {
        int c = 63 - __builtin_clzll((unsigned long long)q);
        if (c > d) c = d;
        s.depth.push_back(c);
        q -= (size_t)1 << c;
    }
    return s;
}
// queries_from_words (ligerito.rs:4530-4547)
std::vector<uint64_t> queries_from(const Sched& s, const std::vector<H>& w) {
    std::vector<uint64_t> out;
    size_t k = 0;
    for (int c : s.depth) {
        const int lo = s.d - c;
        for (uint64_t st = 0; st < (1ull << c); st++, k++)
            out.push_back((st << lo) | (w[k].lo & ((1ull << lo) - 1)));
    }
    return out;
}
size_t ceil_log2(size_t n) {
    if (n <= 1) return 0;
    size_t b = 0;
    for (n -= 1; n; n >>= 1) b++;
    return b;
}
uint32_t bits_for_degree(size_t d) {
    uint32_t b = 0;
    for (; d; d >>= 1) b++;
    return b;
}

// ---------------------------------------------------------------------------
// transcript helpers (challenger.rs:57-216 default methods over the host FsChallenger)
// ---------------------------------------------------------------------------
H sample(flock_fs* fs) {
    H v;
    flock_fs_sample_f128(fs, &v);
    return v;
}
std::vector<H> sample_vec(flock_fs* fs, size_t n) {
    std::vector<H> v(n ? n : 1);
    flock_fs_sample_f128_vec(fs, v.data(), n);
    v.resize(n);
    return v;
}
E sample_e(flock_fs* fs) {
    std::vector<H> w = sample_vec(fs, 2);
    return E{w[0], w[1]};
}
void observe(flock_fs* fs, H v) { flock_fs_observe_f128(fs, v); }
void observe_e(flock_fs* fs, E v) {
    H w[2] = {v.c0, v.c1};
    flock_fs_observe_f128_slice(fs, w, 2);
}
void label(flock_fs* fs, const char* s) {
    flock_fs_observe_label(fs, reinterpret_cast<const uint8_t*>(s), std::strlen(s));
}
H grind_sample(flock_fs* fs, uint32_t bits, uint64_t* nonce) {
    *nonce = flock_fs_grind_pow(fs, bits);
    return sample(fs);
}
// fold_round_claim (jagged.rs:4592-4597): G(r) from (claim, G(1), G(inf))
H round_claim(H claim, H g1, H gi, H r) {
    const H g0 = ha(claim, g1);
    return ha(ha(g0, hm(ha(ha(g1, g0), gi), r)), hm(gi, hm(r, r)));
}

}  // namespace

// ---------------------------------------------------------------------------
// device kernels
// ---------------------------------------------------------------------------
namespace flock {
namespace {

struct e256d {
    f128 c0, c1;
};
__device__ __forceinline__ e256d eadd(e256d a, e256d b) {
    return e256d{f128_add(a.c0, b.c0), f128_add(a.c1, b.c1)};
}
__device__ __forceinline__ f128 mul_x_inv_d(f128 z) {
    const uint32_t mask = 0u - (z.x & 1u);
    return make_uint4(((z.x >> 1) | (z.y << 31)) ^ (0x43u & mask), (z.y >> 1) | (z.z << 31),
                      (z.z >> 1) | (z.w << 31), (z.w >> 1) ^ (0x80000000u & mask));
}
__device__ __forceinline__ e256d emul(e256d a, e256d b) {
    const f128 p0 = f128_mul(a.c0, b.c0), p1 = f128_mul(a.c1, b.c1);
    const f128 p2 = f128_mul(f128_add(a.c0, a.c1), f128_add(b.c0, b.c1));
    return e256d{f128_add(p0, mul_x_inv_d(p1)), f128_add(p2, p0)};
}
__device__ __forceinline__ e256d emulb(e256d a, f128 b) {
    return e256d{f128_mul(a.c0, b), f128_mul(a.c1, b)};
}

constexpr int kThreads = 256;

// XOR-reduce W words across the block and fold them into dst with atomicXor.
template <int W>
__device__ __forceinline__ void block_xor_atomic(uint32_t (&v)[W], uint32_t* dst) {
    __shared__ uint32_t sm[kThreads / 64][W];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < W; k++) v[k] ^= __shfl_xor(v[k], off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < W; k++) sm[wid][k] = v[k];
    __syncthreads();
    if (threadIdx.x < W) {
        uint32_t acc = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) acc ^= sm[w][threadIdx.x];
        if (acc) atomicXor(dst + threadIdx.x, acc);
    }
}
__device__ __forceinline__ void put4(uint32_t* v, f128 a) {
    v[0] ^= a.x;
    v[1] ^= a.y;
    v[2] ^= a.z;
    v[3] ^= a.w;
}

__global__ void k_zero32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Hand `words` (<= 256) accumulator words to the host's coherent pinned mirror `pub`, clear them
// for the next accumulation, and announce them by storing `seq` into `flag` at system scope. One
// wave: its system-scope fence orders every lane's stores before the flag.
__global__ void __launch_bounds__(64) k_acc_publish(uint32_t* acc, uint32_t words, uint32_t* pub,
                                                    uint32_t* flag, uint32_t seq) {
    for (uint32_t i = threadIdx.x; i < words; i += 64) {
        pub[i] = acc[i];
        acc[i] = 0u;
    }
    __threadfence_system();
    if (threadIdx.x == 0) __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// F2-linear map through 32 nibble tables: T[k * 16 + n] = Phi(n x^(4k)) (16 entries per row)
__device__ __forceinline__ f128 lin_apply_d(const f128* T, f128 x) {
    f128 acc = f128_zero();
    const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int j = 0; j < 8; j++) acc = f128_add(acc, T[(q * 8 + j) * 16 + ((w[q] >> (4 * j)) & 15)]);
    return acc;
}

constexpr int kMaxClaims = 4;
struct JagClaims {
    const f128* er[kMaxClaims];  // eq(z_row_c, .) over 2^n rows
    const f128* ec[kMaxClaims];  // eq(z_col_c, .) (scaled for the scalar form) over the columns
    const f128* tab[kMaxClaims]; // 512-word nibble tables of Phi_c (weight form only)
    uint32_t n;
};

__device__ __forceinline__ uint32_t col_of(const uint64_t* ps, uint32_t n_cols, uint64_t d) {
    // the column whose range [ps[c], ps[c+1]) holds d
    uint32_t lo = 0, hi = n_cols;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ps[mid] <= d)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// W[d] = sum_c Phi_c(er_c[row(d)] ec_c[col(d)]) for d < area, zero past it; acc gets the
// round-0 message (u0, u2) of <q, W> over LSB pairs (jagged.rs:1521-1650).
__global__ void __launch_bounds__(kThreads) k_jag_weight(const f128* q, f128* W, uint64_t N,
                                                         uint64_t area, const uint64_t* ps,
                                                         uint32_t n_cols, JagClaims jc,
                                                         uint32_t* acc) {
    __shared__ f128 T[kMaxClaims * 512];
    for (uint32_t i = threadIdx.x; i < jc.n * 512; i += blockDim.x) T[i] = jc.tab[i / 512][i % 512];
    __syncthreads();
    uint32_t v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < N / 2;
         u += (uint64_t)gridDim.x * blockDim.x) {
        f128 w2[2];
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const uint64_t d = 2 * u + e;
            f128 s = f128_zero();
            if (d < area) {
                const uint32_t c = col_of(ps, n_cols, d);
                const uint64_t row = d - ps[c];
                for (uint32_t k = 0; k < jc.n; k++)
                    s = f128_add(s, lin_apply_d(T + k * 512, f128_mul(jc.er[k][row], jc.ec[k][c])));
            }
            w2[e] = s;
            W[d] = s;
        }
        const f128 q0 = q[2 * u], q1 = q[2 * u + 1];
        put4(v, f128_mul(q0, w2[0]));
        put4(v + 4, f128_mul(f128_add(q0, q1), f128_add(w2[0], w2[1])));
    }
    block_xor_atomic<8>(v, acc);
}

// out[d] = sum_c er_c[row(d)] ec_c[col(d)] for d < area, zero past it (the combined jagged
// weight of the multipoint two-product sumcheck, jagged.rs:2729-2783).
__global__ void __launch_bounds__(kThreads) k_jag_scalar(f128* out, uint64_t N, uint64_t area,
                                                         const uint64_t* ps, uint32_t n_cols,
                                                         JagClaims jc) {
    for (uint64_t d = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; d < N;
         d += (uint64_t)gridDim.x * blockDim.x) {
        f128 s = f128_zero();
        if (d < area) {
            const uint32_t c = col_of(ps, n_cols, d);
            const uint64_t row = d - ps[c];
            for (uint32_t k = 0; k < jc.n; k++) s = f128_add(s, f128_mul(jc.er[k][row], jc.ec[k][c]));
        }
        out[d] = s;
    }
}

// Tiled forms of the two jagged tables: a tile is a run of one column's rows (col, d0, len), so
// ec_c[col] is fixed inside it and x -> Phi_c(x ec_c[col]) (weight form) or x -> x ec_c[col]
// (scalar form) is one F2-linear map per claim, held as 32 nibble tables in LDS. Entries past
// the jagged area are written by k_zero_f128.
struct JagTile {
    uint32_t col;
    uint32_t len;
    uint64_t d0;
};
constexpr uint32_t kJagTile = 8192;
__global__ void k_zero_f128(f128* p, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = f128_zero();
}
// WEIGHT: W[d] = sum_c Phi_c(er_c[row] ec_c[col]); else out[d] = sum_c er_c[row] ec_c[col]
template <bool WEIGHT>
__global__ void __launch_bounds__(kThreads) k_jag_tiles(f128* out, const JagTile* tiles,
                                                        const uint64_t* ps, JagClaims jc) {
    __shared__ f128 T[kMaxClaims * 512];
    __shared__ f128 img[kMaxClaims * 128];
    const JagTile tl = tiles[blockIdx.x];
    if constexpr (WEIGHT) {
        // basis images Phi_c(x^b ec) (Phi_c's tables staged in T first), then the nibble
        // tables of the composed map over them
        for (uint32_t i = threadIdx.x; i < jc.n * 512; i += blockDim.x) T[i] = jc.tab[i / 512][i % 512];
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < jc.n * 128; i += blockDim.x) {
            const uint32_t c = i / 128, b = i % 128;
            img[i] = lin_apply_d(T + c * 512, f128_shl_mod(jc.ec[c][tl.col], (int)b));
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < jc.n * 512; i += blockDim.x) {
            const uint32_t c = i / 512, k = (i % 512) / 16, nib = i % 16;
            f128 v = f128_zero();
#pragma unroll
            for (int j = 0; j < 4; j++)
                if ((nib >> j) & 1) v = f128_add(v, img[c * 128 + 4 * k + j]);
            T[i] = v;
        }
    } else {
        for (uint32_t r = threadIdx.x; r < 32 * jc.n; r += blockDim.x)
            shoup_build_row(T + (r / 32) * SHOUP_ENTRIES, jc.ec[r / 32][tl.col], (int)(r % 32));
    }
    __syncthreads();
    const uint64_t row0 = tl.d0 - ps[tl.col];
    for (uint32_t i = threadIdx.x; i < tl.len; i += blockDim.x) {
        f128 s = f128_zero();
        for (uint32_t c = 0; c < jc.n; c++) {
            const f128 e = jc.er[c][row0 + i];
            s = f128_add(s, WEIGHT ? lin_apply_d(T + c * 512, e) : shoup_mul(T + c * SHOUP_ENTRIES, e));
        }
        out[tl.d0 + i] = s;
    }
}

// v[i] = L(v[i]) for an F2-linear map L given by its nibble tables (512 words)
__global__ void __launch_bounds__(kThreads) k_lin_apply(f128* v, uint64_t N, const f128* tab) {
    __shared__ f128 T[512];
    for (uint32_t i = threadIdx.x; i < 512; i += blockDim.x) T[i] = tab[i];
    __syncthreads();
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x)
        v[i] = lin_apply_d(T, v[i]);
}

// The jagged branching program (jagged.rs:163-240, 690-720): one block per (claim, j) computes
// sum_runs w_run * ghat(z_row, x_j, t_c, t_n), x_j = rho^(2^-j) (multipoint_values,
// jagged.rs:2547-2585). eq4[l] = eq((z_row[l], x[l]), .) per layer l <= m; sparse[cd][s][a] =
// (eq4 index, next state).
struct BpArgs {
    const uint64_t* bc;   // run boundaries t_c
    const uint64_t* bn;   // run boundaries t_n
    uint32_t n_runs;
    const f128* w;        // per claim: n_runs weights
    const f128* zrow;     // per claim: n row coords
    const f128* xs;       // 128 points x_j, m coords each
    uint32_t n, m;
    int sparse[4][4][2][2];
};
__global__ void __launch_bounds__(kThreads) k_bp_values(BpArgs a, f128* out) {
    const uint32_t claim = blockIdx.x / 128, j = blockIdx.x % 128;
    // blockIdx.y picks this block's chunk of blockDim / 8 runs
    __shared__ f128 eq4[64][4];
    __shared__ f128 tab[4 * SHOUP_ENTRIES];
    for (uint32_t l = threadIdx.x; l <= a.m; l += blockDim.x) {
        const f128 r = l < a.n ? a.zrow[claim * a.n + l] : f128_zero();
        const f128 x = l < a.m ? a.xs[(size_t)j * a.m + l] : f128_zero();
        const f128 r1 = f128_add(f128_one(), r), x1 = f128_add(f128_one(), x);
        eq4[l][0] = f128_mul(r1, x1);
        eq4[l][1] = f128_mul(r, x1);
        eq4[l][2] = f128_mul(r1, x);
        eq4[l][3] = f128_mul(r, x);
    }
    __syncthreads();
    // eight lanes per run: lane (s, q) forms the q-th product of state s; the pair XORs its two
    // products and every lane of the group gathers the four new states (all lanes run every
    // layer, so the shuffles stay converged)
    uint32_t v[4] = {0, 0, 0, 0};
    const int lane = (int)(threadIdx.x & 63), gbase = lane & ~7, st = (lane & 7) >> 1, q = lane & 1;
    {
        const uint32_t b = blockIdx.y * (blockDim.x / 8) + threadIdx.x / 8;
        const bool live = b < a.n_runs;
        f128 S[4] = {f128_zero(), f128_zero(), f128_one(), f128_zero()};  // e_SUCCESS = state 2
        const uint64_t tc = live ? a.bc[b] : 0, tn = live ? a.bn[b] : 0;
        for (int l = (int)a.m; l >= 0; l--) {
            // Shoup tables of the layer's four eq4 values (shared by every run of the block)
            __syncthreads();
            if (threadIdx.x < 128)
                shoup_build_row(tab + (threadIdx.x >> 5) * SHOUP_ENTRIES, eq4[l][threadIdx.x >> 5],
                                (int)(threadIdx.x & 31));
            __syncthreads();
            const int cd = (int)(((tc >> l) & 1) + 2 * ((tn >> l) & 1));
            const int ei = a.sparse[cd][st][q][0], si = a.sparse[cd][st][q][1];
            const f128 sv = si == 0 ? S[0] : si == 1 ? S[1] : si == 2 ? S[2] : S[3];
            f128 p = shoup_mul(tab + ei * SHOUP_ENTRIES, sv);
            p.x ^= __shfl_xor(p.x, 1, 64);
            p.y ^= __shfl_xor(p.y, 1, 64);
            p.z ^= __shfl_xor(p.z, 1, 64);
            p.w ^= __shfl_xor(p.w, 1, 64);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int src = gbase + 2 * t;
                S[t] = make_uint4(__shfl(p.x, src, 64), __shfl(p.y, src, 64), __shfl(p.z, src, 64),
                                  __shfl(p.w, src, 64));
            }
        }
        if (live && (lane & 7) == 0)
            put4(v, f128_mul(a.w[(size_t)claim * a.n_runs + b], S[0]));  // read out e_INITIAL
    }
    block_xor_atomic<4>(v, reinterpret_cast<uint32_t*>(out + blockIdx.x));
}

__global__ void k_scale(f128* v, uint64_t N, f128 s) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x)
        v[i] = f128_mul(v[i], s);
}
// y += beta x
__global__ void k_axpy(f128* y, const f128* x, uint64_t N, f128 beta) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x)
        y[i] = f128_add(y[i], f128_mul(beta, x[i]));
}
__global__ void __launch_bounds__(kThreads) k_dot(const f128* a, const f128* b, uint64_t N,
                                                  uint32_t* acc) {
    uint32_t v[4] = {0, 0, 0, 0};
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x)
        put4(v, f128_mul(a[i], b[i]));
    block_xor_atomic<4>(v, acc);
}
// the blocked pair of output o: lo = 2 (o / d) d + o % d, hi = lo + d (extension.rs:145-224)
__device__ __forceinline__ uint64_t pair_lo(uint64_t o, uint64_t d) {
    return 2 * (o / d) * d + o % d;
}
// F128 round message (u0, u2) over blocked pairs
__global__ void __launch_bounds__(kThreads) k_msg_blocked(const f128* f, const f128* b, uint64_t n,
                                                          uint64_t d, uint32_t* acc) {
    uint32_t v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint64_t o = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; o < n / 2;
         o += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t lo = pair_lo(o, d), hi = lo + d;
        put4(v, f128_mul(f[lo], b[lo]));
        put4(v + 4, f128_mul(f128_add(f[lo], f[hi]), f128_add(b[lo], b[hi])));
    }
    block_xor_atomic<8>(v, acc);
}
// first F256 fold of base tables: out[o] = v[lo] + r (v[hi] + v[lo])
__global__ void k_fold_first(const f128* f, const f128* b, uint64_t n, uint64_t d, e256d r,
                             e256d* nf, e256d* nb) {
    for (uint64_t o = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; o < n / 2;
         o += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t lo = pair_lo(o, d), hi = lo + d;
        const f128 df = f128_add(f[hi], f[lo]), db = f128_add(b[hi], b[lo]);
        nf[o] = e256d{f128_add(f[lo], f128_mul(r.c0, df)), f128_mul(r.c1, df)};
        nb[o] = e256d{f128_add(b[lo], f128_mul(r.c0, db)), f128_mul(r.c1, db)};
    }
}
__global__ void k_fold_e(const e256d* f, const e256d* b, uint64_t n, uint64_t d, e256d r,
                         e256d* nf, e256d* nb) {
    for (uint64_t o = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; o < n / 2;
         o += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t lo = pair_lo(o, d), hi = lo + d;
        nf[o] = eadd(f[lo], emul(r, eadd(f[hi], f[lo])));
        nb[o] = eadd(b[lo], emul(r, eadd(b[hi], b[lo])));
    }
}
__global__ void __launch_bounds__(kThreads) k_msg_e(const e256d* f, const e256d* b, uint64_t n,
                                                    uint64_t d, uint32_t* acc) {
    uint32_t v[16];
    for (int i = 0; i < 16; i++) v[i] = 0;
    for (uint64_t o = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; o < n / 2;
         o += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t lo = pair_lo(o, d), hi = lo + d;
        const e256d u0 = emul(f[lo], b[lo]);
        const e256d u2 = emul(eadd(f[lo], f[hi]), eadd(b[lo], b[hi]));
        put4(v, u0.c0);
        put4(v + 4, u0.c1);
        put4(v + 8, u2.c0);
        put4(v + 12, u2.c1);
    }
    block_xor_atomic<16>(v, acc);
}
// the code switch (extension.rs:17-45, 1525-1533): f -> (c0, c1) pairs, b -> (b, u b) pairs;
// u (b0 + b1 u) = x^-1 b1 + (b0 + b1) u
__global__ void k_switch(const e256d* f, const e256d* b, uint64_t n, e256d* nf, e256d* nb) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const e256d x = f[i], y = b[i];
        nf[2 * i] = e256d{x.c0, f128_zero()};
        nf[2 * i + 1] = e256d{x.c1, f128_zero()};
        nb[2 * i] = y;
        nb[2 * i + 1] = e256d{mul_x_inv_d(y.c1), f128_add(y.c0, y.c1)};
    }
}
__global__ void k_base_table(const e256d* f, uint64_t n, f128* out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = f[i].c0;
}
// introduce an F128 basis e (OOD eq table): y = sum f.c0 e; message over LSB pairs of (f, e)
__global__ void __launch_bounds__(kThreads) k_intro_base(const e256d* f, const f128* e, uint64_t n,
                                                         uint32_t* acc /* 4 y + 16 msg */) {
    extern __shared__ f128 ws[];
    f128* wl = comb_lane(ws);
    // y = sum f.c0 e; u0 = sum f0 e0 (per component), u2 = sum (f0 + f1)(e0 + e1)
    f256u Y, A0, A1, B0, B1;
    f256u_zero(Y);
    f256u_zero(A0);
    f256u_zero(A1);
    f256u_zero(B0);
    f256u_zero(B1);
    for (uint64_t o = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; o < n / 2;
         o += (uint64_t)gridDim.x * blockDim.x) {
        const e256d f0 = f[2 * o], f1 = f[2 * o + 1];
        const f128 e0 = e[2 * o], e1 = e[2 * o + 1], es = f128_add(e0, e1);
        uint32_t bh = comb_build(wl, e0);
        comb_apply(Y, wl, bh, f0.c0);
        comb_apply(A0, wl, bh, f0.c0);
        comb_apply(A1, wl, bh, f0.c1);
        bh = comb_build(wl, e1);
        comb_apply(Y, wl, bh, f1.c0);
        bh = comb_build(wl, es);
        comb_apply(B0, wl, bh, f128_add(f0.c0, f1.c0));
        comb_apply(B1, wl, bh, f128_add(f0.c1, f1.c1));
    }
    uint32_t v[20] = {0};
    put4(v, f256u_reduce(Y));
    put4(v + 4, f256u_reduce(A0));
    put4(v + 8, f256u_reduce(A1));
    put4(v + 12, f256u_reduce(B0));
    put4(v + 16, f256u_reduce(B1));
    block_xor_atomic<20>(v, acc);
}
__global__ void k_glue_base(e256d* b, const f128* e, uint64_t n, f128 beta) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        b[i].c0 = f128_add(b[i].c0, f128_mul(e[i], beta));
}
// introduce a presplit basis B (2^k words) on the split table f (2^(k+1)): the basis pairs
// are (B_x, u B_x) = ((B_x, 0), (0, B_x)) (extension.rs:1570-1580)
__global__ void __launch_bounds__(kThreads) k_intro_presplit(const e256d* f, const f128* B,
                                                             uint64_t nb, uint32_t* acc) {
    uint32_t v[16];
    for (int i = 0; i < 16; i++) v[i] = 0;
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < nb;
         x += (uint64_t)gridDim.x * blockDim.x) {
        const e256d f0 = f[2 * x], f1 = f[2 * x + 1];
        const f128 bx = B[x];
        const e256d u0 = emulb(f0, bx);
        const e256d u2 = emul(eadd(f0, f1), e256d{bx, bx});
        put4(v, u0.c0);
        put4(v + 4, u0.c1);
        put4(v + 8, u2.c0);
        put4(v + 12, u2.c1);
    }
    block_xor_atomic<16>(v, acc);
}
__global__ void k_glue_presplit(e256d* b, const f128* B, uint64_t nb, f128 beta) {
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < nb;
         x += (uint64_t)gridDim.x * blockDim.x) {
        const f128 t = f128_mul(B[x], beta);
        b[2 * x].c0 = f128_add(b[2 * x].c0, t);
        b[2 * x + 1].c1 = f128_add(b[2 * x + 1].c1, t);
    }
}
// rows of a row-major matrix at the given positions, packed
__global__ void k_gather_rows(const f128* mat, uint64_t row_len, const uint64_t* idx, uint32_t nq,
                              f128* out) {
    const uint64_t total = (uint64_t)nq * row_len;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = mat[idx[i / row_len] * row_len + i % row_len];
}
// 32-byte tree nodes at the given node indices
__global__ void k_gather_nodes(const uint4* tree, const uint64_t* idx, uint32_t n, uint4* out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 2 * n; i += gridDim.x * blockDim.x)
        out[i] = tree[2 * idx[i / 2] + (i & 1)];
}
// induced basis (ligerito.rs:2676-2770 with empty rows): out[jl + 2^h jh] =
// sum_i lo[i][jl] hi[i][jh], lo/hi the per-query tensors over the low h and high bits
__global__ void k_induced(const f128* lo, const f128* hi, uint32_t nq, uint32_t h, uint32_t dims,
                          f128* out) {
    const uint64_t n = 1ull << dims, nl = 1ull << h, nh = n >> h;
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t jl = j & (nl - 1), jh = j >> h;
        f128 s = f128_zero();
        for (uint32_t i = 0; i < nq; i++) s = f128_add(s, f128_mul(lo[i * nl + jl], hi[i * nh + jh]));
        out[j] = s;
    }
}
// s_hat_v[b] = sum_k eq[k] z_vec[128 k + b] (pcs/ring_switch.rs:1278-1330)
// split over the k range: each block takes kShvK consecutive k for all 128 b and XORs its
// partial sums into out (zeroed beforehand)
constexpr uint32_t kShvK = 4;
__global__ void __launch_bounds__(kThreads) k_shv_zvec(const f128* zv, const f128* eq,
                                                       uint64_t n_tail, f128* out) {
    const uint32_t b = threadIdx.x & 127, half = threadIdx.x >> 7;
    f128 s = f128_zero();
    const uint64_t k0 = (uint64_t)blockIdx.x * kShvK;
    for (uint64_t k = k0 + half; k < k0 + kShvK && k < n_tail; k += 2)
        s = f128_add(s, f128_mul(eq[k], zv[128 * k + b]));
    __shared__ f128 part[128];
    if (half) part[b] = s;
    __syncthreads();
    if (!half) {
        s = f128_add(s, part[b]);
        uint32_t* o = reinterpret_cast<uint32_t*>(out + b);
        if (s.x) atomicXor(o, s.x);
        if (s.y) atomicXor(o + 1, s.y);
        if (s.z) atomicXor(o + 2, s.z);
        if (s.w) atomicXor(o + 3, s.w);
    }
}

// ---- fixed multipliers through LDS Shoup tables ----

// the block builds the Shoup tables of m[0..nt) at tab + i * SHOUP_ENTRIES (one row per thread)
template <int NT>
__device__ __forceinline__ void build_tables(f128* tab, const f128 (&m)[NT]) {
    for (int r = threadIdx.x; r < 32 * NT; r += blockDim.x) {
        const int t = r >> 5;
        f128 v = m[0];
#pragma unroll
        for (int i = 1; i < NT; i++)
            if (t == i) v = m[i];
        shoup_build_row(tab + t * SHOUP_ENTRIES, v, r & 31);
    }
    __syncthreads();
}
// x r for r in F256 with the tables of (r0, r1, r0 + r1) at tab: p0 = x0 r0, p1 = x1 r1,
// p2 = (x0 + x1)(r0 + r1); x r = (p0 + x^-1 p1, p2 + p0)
__device__ __forceinline__ e256d emul_tab(const f128* tab, e256d x) {
    const f128 p0 = shoup_mul(tab, x.c0), p1 = shoup_mul(tab + SHOUP_ENTRIES, x.c1);
    const f128 p2 = shoup_mul(tab + 2 * SHOUP_ENTRIES, f128_add(x.c0, x.c1));
    return e256d{f128_add(p0, mul_x_inv_d(p1)), f128_add(p2, p0)};
}

__global__ void __launch_bounds__(kThreads) k_scale_tab(f128* v, uint64_t N, f128 s) {
    __shared__ f128 tab[SHOUP_ENTRIES];
    const f128 m[1] = {s};
    build_tables<1>(tab, m);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x)
        v[i] = shoup_mul(tab, v[i]);
}
// y += beta x
__global__ void __launch_bounds__(kThreads) k_axpy_tab(f128* y, const f128* x, uint64_t N, f128 beta) {
    __shared__ f128 tab[SHOUP_ENTRIES];
    const f128 m[1] = {beta};
    build_tables<1>(tab, m);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x)
        y[i] = f128_add(y[i], shoup_mul(tab, x[i]));
}
__global__ void __launch_bounds__(kThreads) k_glue_base_tab(e256d* b, const f128* e, uint64_t n, f128 beta) {
    __shared__ f128 tab[SHOUP_ENTRIES];
    const f128 m[1] = {beta};
    build_tables<1>(tab, m);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        b[i].c0 = f128_add(b[i].c0, shoup_mul(tab, e[i]));
}
__global__ void __launch_bounds__(kThreads) k_glue_presplit_tab(e256d* b, const f128* B, uint64_t nb,
                                                                f128 beta) {
    __shared__ f128 tab[SHOUP_ENTRIES];
    const f128 m[1] = {beta};
    build_tables<1>(tab, m);
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < nb;
         x += (uint64_t)gridDim.x * blockDim.x) {
        const f128 t = shoup_mul(tab, B[x]);
        b[2 * x].c0 = f128_add(b[2 * x].c0, t);
        b[2 * x + 1].c1 = f128_add(b[2 * x + 1].c1, t);
    }
}

// ---- products of two varying values: per-lane comb tables, unreduced accumulation ----

// acc ^= a b, unreduced (the lane's comb column is rebuilt for b)
__device__ __forceinline__ void comb_acc(f256u& acc, f128* wl, f128 a, f128 b) {
    const uint32_t bh = comb_build(wl, b);
    comb_apply(acc, wl, bh, a);
}
// the three Karatsuba sums of an F256 product a b: S0 += a0 b0, S1 += a1 b1,
// S2 += (a0 + a1)(b0 + b1)
__device__ __forceinline__ void emul_acc(f256u (&S)[3], f128* wl, e256d a, e256d b) {
    comb_acc(S[0], wl, a.c0, b.c0);
    comb_acc(S[1], wl, a.c1, b.c1);
    comb_acc(S[2], wl, f128_add(a.c0, a.c1), f128_add(b.c0, b.c1));
}
// sum of the products from the three sums: (S0 + x^-1 S1, S2 + S0), each reduced once
__device__ __forceinline__ e256d emul_sum(const f256u (&S)[3]) {
    const f128 p0 = f256u_reduce(S[0]), p1 = f256u_reduce(S[1]), p2 = f256u_reduce(S[2]);
    return e256d{f128_add(p0, mul_x_inv_d(p1)), f128_add(p2, p0)};
}
__device__ __forceinline__ void f256u_zero3(f256u (&S)[3]) {
    f256u_zero(S[0]);
    f256u_zero(S[1]);
    f256u_zero(S[2]);
}

// one fold at r over blocked pairs at distance d (n -> n/2), then, with `msg`, the next
// round's message over blocked pairs at distance d2 of the folded tables: acc[0..16) ^= (u0, u2)
// with u0 = sum g_lo h_lo, u2 = sum (g_lo + g_hi)(h_lo + h_hi) (extension.rs:145-224). BASE:
// the inputs are F128 tables (the first fold, x -> (x_lo + r0 dx, r1 dx), dx = x_hi + x_lo).
template <bool BASE>
__device__ __forceinline__ e256d fold_at(const void* src, uint64_t o, uint64_t d, const f128* tab) {
    const uint64_t lo = pair_lo(o, d), hi = lo + d;
    if constexpr (BASE) {
        const f128* v = static_cast<const f128*>(src);
        const f128 dx = f128_add(v[hi], v[lo]);
        return e256d{f128_add(v[lo], shoup_mul(tab, dx)), shoup_mul(tab + SHOUP_ENTRIES, dx)};
    } else {
        const e256d* v = static_cast<const e256d*>(src);
        return eadd(v[lo], emul_tab(tab, eadd(v[hi], v[lo])));
    }
}
// MSG = false: the fold alone (a separate kernel, so its register and LDS footprint stays small)
template <bool BASE, bool MSG>
__global__ void __launch_bounds__(kThreads) k_fold_msg(const void* f, const void* b, uint64_t n,
                                                       uint64_t d, e256d r, uint64_t d2,
                                                       e256d* nf, e256d* nb, uint32_t* acc) {
    __shared__ f128 tab[3 * SHOUP_ENTRIES];
    extern __shared__ f128 ws[];
    const f128 m[3] = {r.c0, r.c1, f128_add(r.c0, r.c1)};
    build_tables<3>(tab, m);
    const uint64_t half = n / 2;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    if (!MSG || half < 2) {
        for (uint64_t o = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; o < half; o += step) {
            nf[o] = fold_at<BASE>(f, o, d, tab);
            nb[o] = fold_at<BASE>(b, o, d, tab);
        }
        return;
    }
    f128* wl = comb_lane(ws);
    f256u U0[3], U2[3];
    f256u_zero3(U0);
    f256u_zero3(U2);
    for (uint64_t o = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; o < half / 2; o += step) {
        const uint64_t p0 = pair_lo(o, d2), p1 = p0 + d2;
        const e256d g0 = fold_at<BASE>(f, p0, d, tab), g1 = fold_at<BASE>(f, p1, d, tab);
        const e256d h0 = fold_at<BASE>(b, p0, d, tab), h1 = fold_at<BASE>(b, p1, d, tab);
        nf[p0] = g0;
        nf[p1] = g1;
        nb[p0] = h0;
        nb[p1] = h1;
        emul_acc(U0, wl, g0, h0);
        emul_acc(U2, wl, eadd(g0, g1), eadd(h0, h1));
    }
    const e256d u0 = emul_sum(U0), u2 = emul_sum(U2);
    uint32_t v[16] = {0};
    put4(v, u0.c0);
    put4(v + 4, u0.c1);
    put4(v + 8, u2.c0);
    put4(v + 12, u2.c1);
    block_xor_atomic<16>(v, acc);
}

// F256 message over blocked pairs at d2 (the unfused form for large tables)
__global__ void __launch_bounds__(kThreads) k_msg_e_comb(const e256d* f, const e256d* b, uint64_t n,
                                                         uint64_t d2, uint32_t* acc) {
    extern __shared__ f128 ws[];
    f128* wl = comb_lane(ws);
    f256u U0[3], U2[3];
    f256u_zero3(U0);
    f256u_zero3(U2);
    for (uint64_t o = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; o < n / 2;
         o += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t p0 = pair_lo(o, d2), p1 = p0 + d2;
        const e256d g0 = f[p0], g1 = f[p1], h0 = b[p0], h1 = b[p1];
        emul_acc(U0, wl, g0, h0);
        emul_acc(U2, wl, eadd(g0, g1), eadd(h0, h1));
    }
    const e256d u0 = emul_sum(U0), u2 = emul_sum(U2);
    uint32_t v[16] = {0};
    put4(v, u0.c0);
    put4(v + 4, u0.c1);
    put4(v + 8, u2.c0);
    put4(v + 12, u2.c1);
    block_xor_atomic<16>(v, acc);
}

// F128 message over blocked pairs (u0 = sum f_lo b_lo, u2 = sum (f_lo + f_hi)(b_lo + b_hi))
__global__ void __launch_bounds__(kThreads) k_msg_blocked_comb(const f128* f, const f128* b, uint64_t n,
                                                               uint64_t d, uint32_t* acc) {
    extern __shared__ f128 ws[];
    f128* wl = comb_lane(ws);
    f256u u0, u2;
    f256u_zero(u0);
    f256u_zero(u2);
    for (uint64_t o = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; o < n / 2;
         o += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t lo = pair_lo(o, d), hi = lo + d;
        const f128 f0 = f[lo], b0 = b[lo];
        comb_acc(u0, wl, f0, b0);
        comb_acc(u2, wl, f128_add(f0, f[hi]), f128_add(b0, b[hi]));
    }
    uint32_t v[8] = {0};
    put4(v, f256u_reduce(u0));
    put4(v + 4, f256u_reduce(u2));
    block_xor_atomic<8>(v, acc);
}
// <a, b>
__global__ void __launch_bounds__(kThreads) k_dot_comb(const f128* a, const f128* b, uint64_t N,
                                                       uint32_t* acc) {
    extern __shared__ f128 ws[];
    f128* wl = comb_lane(ws);
    f256u s;
    f256u_zero(s);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x)
        comb_acc(s, wl, a[i], b[i]);
    uint32_t v[4] = {0};
    put4(v, f256u_reduce(s));
    block_xor_atomic<4>(v, acc);
}

// Induced basis (ligerito.rs:2676-2770 with empty rows) as out = sum_i lo_i (x) hi_i:
// lo_i[j] = aw_i prod_{k < h, bit k of j} s_ik over the low h coordinates, hi_i[j] =
// prod_{k >= h, bit k - h of j} s_ik over the rest (one block per query i), with
// s_ik = s_k(q_i) / s_k(v_k) (eval_sk_at_vks, ligerito.rs:2402-2420; zero where s_k(v_k) = 0).
// The subspace polynomials s_k are F2-linear, so s_ik is the XOR of row k of skb
// (skb[k * 64 + b] = s_k(x^b) / s_k(v_k)) over the set bits b of q_i: eight lanes per k, one
// byte of q_i each, joined by shuffles. Both tables are then built by doubling in LDS, level k
// doubling the table by its factor s_ik (one product per entry, through the level's Shoup table:
// every product of a level shares its multiplier).
constexpr int kSkRows = 64;  // rows of skb: dims <= kSkRows
__global__ void __launch_bounds__(kThreads) k_induced_tables(const uint64_t* qs, const f128* skb,
                                                             const f128* aw, uint32_t dims, uint32_t h,
                                                             f128* lo, f128* hi) {
    __shared__ f128 tab[2 * SHOUP_ENTRIES];  // the level's lo and hi multipliers
    __shared__ f128 si[kSkRows];
    extern __shared__ f128 tt[];             // lo (2^h entries), then hi (2^(dims - h))
    const uint32_t i = blockIdx.x;
    const uint64_t nl = 1ull << h, nh = 1ull << (dims - h);
    const uint64_t q = qs[i];
    // groups of eight lanes are whole (blockDim is a multiple of 8): all active or none
    for (uint32_t e = threadIdx.x; e < 8 * dims; e += blockDim.x) {
        const uint32_t k = e >> 3, p = e & 7, byte = (uint32_t)(q >> (8 * p)) & 0xffu;
        const f128* row = skb + (size_t)k * 64 + 8 * p;
        f128 v = f128_zero();
#pragma unroll
        for (int j = 0; j < 8; j++)
            if ((byte >> j) & 1u) v = f128_add(v, row[j]);
#pragma unroll
        for (int m = 1; m < 8; m <<= 1)
            v = make_uint4(v.x ^ __shfl_xor(v.x, m, 8), v.y ^ __shfl_xor(v.y, m, 8),
                           v.z ^ __shfl_xor(v.z, m, 8), v.w ^ __shfl_xor(v.w, m, 8));
        if (p == 0) si[k] = v;
    }
    f128* tl = tt;
    f128* th = tt + nl;
    if (threadIdx.x == 0) {
        tl[0] = aw[i];
        th[0] = f128_one();
    }
    __syncthreads();  // si is written
    const uint32_t levels = h > dims - h ? h : dims - h;
    for (uint32_t k = 0; k < levels; k++) {
        if (threadIdx.x < 64) {
            const uint32_t t = threadIdx.x >> 5;
            if (t ? k < dims - h : k < h)
                shoup_build_row(tab + t * SHOUP_ENTRIES, si[t ? h + k : k], threadIdx.x & 31);
        }
        __syncthreads();  // the tables and the previous level's entries are written
        const uint64_t len = 1ull << k;
        for (uint64_t j = threadIdx.x; j < 2 * len; j += blockDim.x) {
            if (j < len) {
                if (k < h) tl[len + j] = shoup_mul(tab, tl[j]);
            } else if (k < dims - h) {
                th[j] = shoup_mul(tab + SHOUP_ENTRIES, th[j - len]);
            }
        }
        __syncthreads();  // the tables are consumed
    }
    for (uint64_t j = threadIdx.x; j < nl; j += blockDim.x) lo[i * nl + j] = tl[j];
    for (uint64_t j = threadIdx.x; j < nh; j += blockDim.x) hi[i * nh + j] = th[j];
}
// out[jl + 2^h jh] = sum_{i in this block's chunk} lo_i[jl] hi_i[jh]: kIndTL rows jl per block
// (their Shoup tables in LDS, rebuilt per i), kIndTH columns jh per thread; chunk z of the
// queries writes its partial sums at out + z * 2^dims (k_xor_chunks adds them up)
constexpr int kIndTL = 4;  // jl rows per block
constexpr int kIndTH = 2;  // jh columns per thread (blockDim * kIndTH per block)
__global__ void __launch_bounds__(kThreads) k_induced_gemm(const f128* lo, const f128* hi, uint32_t nq,
                                                           uint32_t h, uint32_t dims, uint32_t chunk,
                                                           f128* out) {
    __shared__ f128 tab[kIndTL * SHOUP_ENTRIES];
    const uint64_t nl = 1ull << h, nh = 1ull << (dims - h);
    const uint64_t jl0 = (uint64_t)blockIdx.x * kIndTL;
    const uint64_t jh0 = (uint64_t)blockIdx.y * blockDim.x * kIndTH + threadIdx.x;
    const uint32_t i0 = blockIdx.z * chunk, i1 = min(nq, i0 + chunk);
    f128 acc[kIndTH][kIndTL];
#pragma unroll
    for (int u = 0; u < kIndTH; u++)
#pragma unroll
        for (int t = 0; t < kIndTL; t++) acc[u][t] = f128_zero();
    for (uint32_t i = i0; i < i1; i++) {
        __syncthreads();  // the previous tables are consumed
        if (threadIdx.x < 32 * kIndTL) {
            const int t = threadIdx.x >> 5;
            const f128 m = jl0 + t < nl ? lo[i * nl + jl0 + t] : f128_zero();
            shoup_build_row(tab + t * SHOUP_ENTRIES, m, threadIdx.x & 31);
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kIndTH; u++) {
            const uint64_t jh = jh0 + (uint64_t)u * blockDim.x;
            if (jh < nh) {
                const f128 x = hi[i * nh + jh];
#pragma unroll
                for (int t = 0; t < kIndTL; t++)
                    acc[u][t] = f128_add(acc[u][t], shoup_mul(tab + t * SHOUP_ENTRIES, x));
            }
        }
    }
#pragma unroll
    for (int u = 0; u < kIndTH; u++) {
        const uint64_t jh = jh0 + (uint64_t)u * blockDim.x;
        if (jh >= nh) break;
#pragma unroll
        for (int t = 0; t < kIndTL; t++) {
            if (jl0 + t >= nl) break;
            out[(blockIdx.z << dims) + (jl0 + t) + nl * jh] = acc[u][t];
        }
    }
}

// Two rounds per pass of the (G(0), G(inf)) product sumcheck of <f, b> over LSB pairs (the
// reference's lookahead, cuda-ghash/sumcheck_ab.cuh:125-239). The tables are folded by the
// `nfold` (0..2) pending challenges r1, r2 (v' = v0 + r (v0 + v1)) into nf, nb of length
// L = n >> nfold; then the kernel sums round r's pair on the folded table and, when L >= 4,
// round r+1's pair as quadratics in its own challenge rho, by their values at rho = 0, 1 and
// their leading coefficients. For a four-entry group g0..g3 (c0..c3 of b):
//   G0 += g0 c0 + g2 c2, Ginf += (g0+g1)(c0+c1) + (g2+g3)(c2+c3);
//   next G0:   at 0 g0 c0, at 1 g1 c1, leading (g0+g1)(c0+c1);
//   next Ginf: at 0 (g0+g2)(c0+c2), at 1 (g1+g3)(c1+c3), leading (g0+..+g3)(c0+..+c3).
// acc words (32): [g0c0, g2c2, (g0+g1)(c0+c1), (g2+g3)(c2+c3), g1c1, M0, M1, Minf].
__device__ __forceinline__ f128 shfl_f128(f128 v, uint32_t src) {
    return make_uint4(__shfl(v.x, src, 64), __shfl(v.y, src, 64), __shfl(v.z, src, 64), __shfl(v.w, src, 64));
}
__device__ __forceinline__ f128 fold_by(const f128* T, f128 v0, f128 v1) {
    return f128_add(v0, shoup_mul(T, f128_add(v0, v1)));
}
// One thread per folded entry (the groups' four lanes exchange their entries by shuffles and
// share the eight products), so the small tables this runs on see short dependency chains.
__global__ void __launch_bounds__(kThreads) k_sc_fold_la(const f128* f, const f128* b, uint64_t n,
                                                         int nfold, f128 r1, f128 r2, f128* nf,
                                                         f128* nb, uint32_t* acc) {
    __shared__ f128 tab[2 * SHOUP_ENTRIES];
    extern __shared__ f128 ws[];
    f128* wl = comb_lane(ws);
    if (threadIdx.x < 64)
        shoup_build_row(tab + (threadIdx.x >> 5) * SHOUP_ENTRIES, threadIdx.x < 32 ? r1 : r2,
                        threadIdx.x & 31);
    __syncthreads();
    const f128* T1 = tab;
    const f128* T2 = tab + SHOUP_ENTRIES;
    const uint64_t L = n >> nfold;
    const int per = L >= 4 ? 4 : (int)L;  // entries per group
    const int t = threadIdx.x & 3;        // entry in the group (groups never straddle a wave)
    f128 s0 = f128_zero(), s1 = f128_zero();  // this lane's two product sums
    // L is a power of two, so every lane of a group iterates alike
    for (uint64_t base = blockIdx.x * (uint64_t)blockDim.x + (threadIdx.x & ~3u); base < 4 * (L / per);
         base += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = (base / 4) * per + t;  // this lane's folded entry
        f128 g = f128_zero(), c = f128_zero();
        if (t < per) {
            if (nfold == 0) {
                g = f[k];
                c = b[k];
            } else if (nfold == 1) {
                g = fold_by(T1, f[2 * k], f[2 * k + 1]);
                c = fold_by(T1, b[2 * k], b[2 * k + 1]);
            } else {
                g = fold_by(T2, fold_by(T1, f[4 * k], f[4 * k + 1]), fold_by(T1, f[4 * k + 2], f[4 * k + 3]));
                c = fold_by(T2, fold_by(T1, b[4 * k], b[4 * k + 1]), fold_by(T1, b[4 * k + 2], b[4 * k + 3]));
            }
            if (nfold) {
                nf[k] = g;
                nb[k] = c;
            }
        }
        f128 G[4], C[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            G[u] = shfl_f128(g, (threadIdx.x & ~3u) + u);
            C[u] = shfl_f128(c, (threadIdx.x & ~3u) + u);
        }
        if (per == 4) {
            // lane 0: g0c0, (g0+g1)(c0+c1); lane 1: g2c2, (g2+g3)(c2+c3);
            // lane 2: g1c1, (g0+g2)(c0+c2); lane 3: (g1+g3)(c1+c3), (g0+..+g3)(c0+..+c3)
            f128 x0, y0, x1, y1;
            if (t == 0) {
                x0 = G[0]; y0 = C[0];
                x1 = f128_add(G[0], G[1]); y1 = f128_add(C[0], C[1]);
            } else if (t == 1) {
                x0 = G[2]; y0 = C[2];
                x1 = f128_add(G[2], G[3]); y1 = f128_add(C[2], C[3]);
            } else if (t == 2) {
                x0 = G[1]; y0 = C[1];
                x1 = f128_add(G[0], G[2]); y1 = f128_add(C[0], C[2]);
            } else {
                x0 = f128_add(G[1], G[3]); y0 = f128_add(C[1], C[3]);
                x1 = f128_add(f128_add(G[0], G[1]), f128_add(G[2], G[3]));
                y1 = f128_add(f128_add(C[0], C[1]), f128_add(C[2], C[3]));
            }
            s0 = f128_add(s0, comb_mul(wl, x0, y0));
            s1 = f128_add(s1, comb_mul(wl, x1, y1));
        } else if (per == 2 && t == 0) {
            s0 = f128_add(s0, comb_mul(wl, G[0], C[0]));
            s1 = f128_add(s1, comb_mul(wl, f128_add(G[0], G[1]), f128_add(C[0], C[1])));
        }
    }
    // slots: lane 0 -> (0, 2), lane 1 -> (1, 3), lane 2 -> (4, 5), lane 3 -> (6, 7)
    uint32_t v[32];
#pragma unroll
    for (int k = 0; k < 32; k++) v[k] = 0u;
    const int a0 = t == 0 ? 0 : t == 1 ? 1 : t == 2 ? 4 : 6;
    const int a1 = t == 0 ? 2 : t == 1 ? 3 : t == 2 ? 5 : 7;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        if (k == a0) put4(v + 4 * k, s0);
        if (k == a1) put4(v + 4 * k, s1);
    }
    block_xor_atomic<32>(v, acc);
}

// <q, eq(z)> and the eq-basis table of the opening without materialising an eq table: with
// x = xl + 2^nlo xh, eq(z)[x] = lo[xl] hi[xh] (the two halves, k_eq_direct).
// k_ood_dot: block (bl, bh) adds hi[xh] q[x] over its kOodXh rows xh (one Shoup table of
// hi[xh] at a time) for each of its columns xl, multiplies each column's sum by lo[xl] (one
// comb product per column and block) and XORs the block's total into acc.
constexpr int kOodXh = 64;  // rows xh per block
constexpr int kOodXl = 2;   // columns xl per thread
__global__ void __launch_bounds__(kThreads) k_ood_dot(const f128* q, int nlo, int nhi, const f128* lo,
                                                      const f128* hi, uint32_t* acc) {
    __shared__ f128 tab[SHOUP_ENTRIES];
    extern __shared__ f128 ws[];
    f128* wl = comb_lane(ws);
    const uint64_t nl = 1ull << nlo, nh = 1ull << nhi;
    const uint64_t xl0 = (uint64_t)blockIdx.x * (kThreads * kOodXl);
    const uint64_t xh0 = (uint64_t)blockIdx.y * kOodXh;
    f128 col[kOodXl];
#pragma unroll
    for (int u = 0; u < kOodXl; u++) col[u] = f128_zero();
    for (int r = 0; r < kOodXh && xh0 + r < nh; r++) {
        const uint64_t xh = xh0 + r;
        __syncthreads();  // the previous row's table is consumed
        if (threadIdx.x < 32) shoup_build_row(tab, hi[xh], threadIdx.x);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kOodXl; u++) {
            const uint64_t xl = xl0 + (uint64_t)u * kThreads + threadIdx.x;
            if (xl < nl) col[u] = f128_add(col[u], shoup_mul(tab, q[(xh << nlo) + xl]));
        }
    }
    f256u s;
    f256u_zero(s);
#pragma unroll
    for (int u = 0; u < kOodXl; u++) {
        const uint64_t xl = xl0 + (uint64_t)u * kThreads + threadIdx.x;
        if (xl < nl) comb_acc(s, wl, col[u], lo[xl]);
    }
    uint32_t v[4] = {0};
    put4(v, f256u_reduce(s));
    block_xor_atomic<4>(v, acc);
}
// b[x] = sum_t (c_t hi_t[xh]) lo_t[xl]: the sum of up to kB0Terms scaled eq tables given by
// their halves, one block per row xh with one Shoup table per term
constexpr int kB0Terms = 4;
struct B0Terms {
    const f128* lo[kB0Terms];
    const f128* hi[kB0Terms];
    f128 c[kB0Terms];
    int n;
};
constexpr int kB0Rows = 1;  // rows xh per block
__global__ void __launch_bounds__(kThreads) k_b0_build(B0Terms T, int nlo, int nhi, f128* b) {
    extern __shared__ f128 tab[];  // T.n tables
    __shared__ f128 mult[kB0Terms];
    const uint64_t nl = 1ull << nlo, nh = 1ull << nhi;
    for (int r = 0; r < kB0Rows; r++) {
        const uint64_t xh = (uint64_t)blockIdx.x * kB0Rows + r;
        if (xh >= nh) break;
        __syncthreads();  // the previous row's tables are consumed
        if (threadIdx.x < (uint32_t)T.n) mult[threadIdx.x] = f128_mul(T.c[threadIdx.x], T.hi[threadIdx.x][xh]);
        __syncthreads();
        if (threadIdx.x < 32u * T.n)
            shoup_build_row(tab + (threadIdx.x >> 5) * SHOUP_ENTRIES, mult[threadIdx.x >> 5], threadIdx.x & 31);
        __syncthreads();
        for (uint64_t xl = threadIdx.x; xl < nl; xl += blockDim.x) {
            f128 v = f128_zero();
            for (int t = 0; t < T.n; t++) v = f128_add(v, shoup_mul(tab + t * SHOUP_ENTRIES, T.lo[t][xl]));
            b[(xh << nlo) + xl] = v;
        }
    }
}

// out[i] = sum_z part[z * n + i], z < gz
__global__ void __launch_bounds__(kThreads) k_xor_chunks(const f128* part, uint32_t gz, uint64_t n,
                                                         f128* out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        f128 v = part[i];
#pragma unroll 8
        for (uint32_t z = 1; z < gz; z++) v = f128_add(v, part[z * n + i]);
        out[i] = v;
    }
}

}  // namespace
}  // namespace flock

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
namespace {
using flock::e256d;

struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

// One set of reduction accumulators: the device words the kernels XOR into and the host's
// pinned mirror they are published to.
struct AccSet {
    uint32_t* acc = nullptr;   // device accumulators (256 words), zero between uses
    uint32_t* hacc = nullptr;  // coherent pinned mirror: 256 words, then the flag word
    uint32_t* dacc = nullptr;  // its device view
    uint32_t seq = 0;          // last sequence number announced through the flag
};

struct flock_open_ws {
    std::vector<DBuf> bufs;
    AccSet accs[2];  // one per host thread of the opening (the inner opening; the assist)
    double stage_ms[6] = {};        // host wall time of the last opening's stages
    std::chrono::steady_clock::time_point mark;
    void stamp(int i) {
        const auto t = std::chrono::steady_clock::now();
        if (i >= 0) stage_ms[i] = std::chrono::duration<double, std::milli>(t - mark).count();
        mark = t;
    }
    // pinned host staging of the inner opening's thread (grow-only): [0] the level openings'
    // rows and paths, written there by the gather kernels, [1] the induced basis's inputs going out, [2] the openings'
    // row and node indices going out
    // the induced basis's table of subspace polynomials (slot 54), built once
    std::mutex sk_mu;
    bool sk_ready = false;
    void* pin[3] = {nullptr, nullptr, nullptr};
    size_t pin_bytes[3] = {0, 0, 0};
    void* pinned(int i, size_t bytes) {
        if (pin_bytes[i] >= bytes) return pin[i];
        if (pin[i]) (void)hipHostFree(pin[i]);
        pin[i] = nullptr;
        pin_bytes[i] = 0;
        // [0] is written by kernels directly (coherent: visible to the host once they complete)
        if (hipHostMalloc(&pin[i], bytes, i == 0 ? hipHostMallocCoherent : 0) != hipSuccess) {
            (void)hipGetLastError();
            pin[i] = nullptr;
            return nullptr;
        }
        pin_bytes[i] = bytes;
        return pin[i];
    }
    // the second context the multipoint block runs on beside the inner opening (lazily made on
    // the prover's device)
    flock_hip_ctx* aux = nullptr;
    int aux_dev = -1;
    flock_hip_ctx* aux_ctx(int device) {
        if (aux && aux_dev == device) return aux;
        if (aux) flock_hip_ctx_destroy(aux);
        aux = nullptr;
        if (device < 0 || flock_hip_ctx_create(device, &aux) != 0) return nullptr;
        aux_dev = device;
        return aux;
    }
    // Slots are fixed per use and bufs never grows past kSlots, so the two threads of an opening
    // (which use disjoint slots) never touch the same entry or resize the vector.
    static constexpr size_t kSlots = 64;
    void* get(size_t slot, size_t bytes) {
        if (slot >= kSlots) return nullptr;
        if (bufs.size() < kSlots) bufs.resize(kSlots);
        DBuf& b = bufs[slot];
        if (b.bytes >= bytes) return b.p;
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
        if (hipMalloc(&b.p, bytes ? bytes : 16) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        b.bytes = bytes;
        return b.p;
    }
};

namespace {

using flock::kThreads;

// kernels whose dynamic LDS is the per-lane comb workspace
hipError_t comb_attrs() {
    static const hipError_t e = [] {
        const void* fns[] = {reinterpret_cast<const void*>(&flock::k_fold_msg<true, true>),
                             reinterpret_cast<const void*>(&flock::k_fold_msg<false, true>),
                             reinterpret_cast<const void*>(&flock::k_msg_e_comb),
                             reinterpret_cast<const void*>(&flock::k_intro_base),
                             reinterpret_cast<const void*>(&flock::k_msg_blocked_comb),
                             reinterpret_cast<const void*>(&flock::k_dot_comb),
                             reinterpret_cast<const void*>(&flock::k_sc_fold_la),
                             reinterpret_cast<const void*>(&flock::k_induced_tables),
                             reinterpret_cast<const void*>(&flock::k_ood_dot)};
        for (const void* f : fns)
            if (hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   kCombWsBytes);
                r != hipSuccess)
                return r;
        return hipSuccess;
    }();
    return e;
}

struct Ctx {
    flock_hip_ctx* c;
    flock_open_ws* ws;
    hipStream_t st;
    AccSet* A;  // this thread's accumulators
    int err = 0;
    std::string what;

    uint32_t grid(uint64_t n) const {
        uint64_t g = (n + kThreads - 1) / kThreads;
        if (g > 2048) g = 2048;
        return (uint32_t)(g ? g : 1);
    }
    bool fail(int code, const std::string& m) {
        if (!err) {
            err = code;
            what = m;
        }
        return false;
    }
    bool hip(hipError_t e, const char* where) {
        if (e != hipSuccess) return fail(FLOCK_HIP_E_HIP, std::string(where) + ": " + hipGetErrorString(e));
        return true;
    }
    template <class T>
    T* buf(size_t slot, size_t count) {
        void* p = ws->get(slot, count * sizeof(T));
        if (!p) fail(FLOCK_HIP_E_NOMEM, "opening workspace");
        return static_cast<T*>(p);
    }
    // Every accumulation kernel XORs into A->acc, which read_acc leaves zero again; an entry
    // point clears all of it once before its first accumulation (an earlier call may have
    // stopped between a kernel and its read).
    bool clear_acc() {
        hipLaunchKernelGGL(flock::k_zero32, dim3(1), dim3(256), 0, st, A->acc, 256u);
        return hip(hipGetLastError(), "zero");
    }
    // Read back `words` accumulator words (and clear them): a one-wave kernel queued behind the
    // accumulation copies them into the pinned mirror and then raises the flag word, which the
    // host polls instead of synchronising the stream. Bounded: past 10 s, or once the stream
    // reports an error, the wait ends with an error.
    bool read_acc(uint32_t words) {
        if (++A->seq == 0) ++A->seq;
        const uint32_t seq = A->seq;
        hipLaunchKernelGGL(flock::k_acc_publish, dim3(1), dim3(64), 0, st, A->acc, words, A->dacc,
                           A->dacc + 256, seq);
        if (!hip(hipGetLastError(), "acc publish")) return false;
        const uint32_t* flag = A->hacc + 256;
        const auto t0 = std::chrono::steady_clock::now();
        for (uint64_t it = 1; __atomic_load_n(flag, __ATOMIC_ACQUIRE) != seq; it++) {
            if ((it & 4095) == 0) {
                const hipError_t q = hipStreamQuery(st);
                if (q != hipSuccess && q != hipErrorNotReady) return hip(q, "accumulation kernel");
                if (q == hipSuccess && __atomic_load_n(flag, __ATOMIC_ACQUIRE) != seq)
                    return fail(FLOCK_HIP_E_INTERNAL, "accumulators not published");
                if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(10))
                    return fail(FLOCK_HIP_E_INTERNAL, "accumulators: no result within 10 s");
            }
            _mm_pause();
        }
        return true;
    }
    H acc_f(int k) const {
        const uint32_t* v = A->hacc + 4 * k;
        return H{(uint64_t)v[1] << 32 | v[0], (uint64_t)v[3] << 32 | v[2]};
    }
    bool h2d(void* d, const void* h, size_t n) {
        return hip(hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, st), "h2d");
    }
    bool d2h(void* h, const void* d, size_t n) {
        if (!hip(hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, st), "d2h")) return false;
        return hip(hipStreamSynchronize(st), "sync");
    }
    bool chk(const char* where) { return hip(hipGetLastError(), where); }
    // the two halves of the eq table of host coordinates (flock::build_eq_halves_async)
    bool eq_halves(const H* r, size_t n, flock_f128* lo, flock_f128* hi) {
        if (int rc = flock::build_eq_halves_async(c, r, n, lo, hi)) return fail(rc, flock_hip_last_error());
        return true;
    }
    // device eq table of host coordinates
    bool eq_dev(const H* r, size_t n, flock_f128* out) {
        if (int rc = flock::build_eq_async(c, r, n, out)) return fail(rc, flock_hip_last_error());
        return true;
    }
};

inline uint4 U4(H v) {
    return make_uint4((uint32_t)v.lo, (uint32_t)(v.lo >> 32), (uint32_t)v.hi, (uint32_t)(v.hi >> 32));
}
inline e256d E4(E v) { return e256d{U4(v.c0), U4(v.c1)}; }

// 512-word nibble tables of an F2-linear map with basis images img[128]
std::vector<H> lin_tables(const H* img) {
    std::vector<H> t(512);
    for (int k = 0; k < 32; k++) {
        t[k * 16] = hz();
        for (int n = 1; n < 16; n++) {
            const int low = __builtin_ctz((unsigned)n);
            t[k * 16 + n] = ha(t[k * 16 + (n & (n - 1))], img[4 * k + low]);
        }
    }
    return t;
}

// jagged layout (pcs/jagged.rs:87-140, 569-584)
struct Jag {
    int n = 0, k = 0, m = 0;
    std::vector<uint64_t> ps;        // 2^k + 1
    std::vector<uint64_t> bc, bn;    // deduped boundary runs
    std::vector<uint32_t> brun;
    uint64_t area() const { return ps.back(); }
};
Jag jag_make(const uint64_t* heights, size_t nh, int n, int m) {
    Jag J;
    J.n = n;
    J.m = m;
    J.k = __builtin_ctzll(nh);
    J.ps.assign(nh + 1, 0);
    for (size_t y = 0; y < nh; y++) J.ps[y + 1] = J.ps[y] + heights[y];
    for (size_t y = 0; y < nh; y++) {
        const uint64_t c = J.ps[y], d = J.ps[y + 1];
        if (!J.bc.empty() && J.bc.back() == c && J.bn.back() == d)
            J.brun.back()++;
        else {
            J.bc.push_back(c);
            J.bn.push_back(d);
            J.brun.push_back(1);
        }
    }
    return J;
}
std::vector<H> run_weights(const Jag& J, const H* z_col) {
    const std::vector<H> ec = heq(z_col, (size_t)J.k);
    std::vector<H> w(J.bc.size());
    size_t y = 0;
    for (size_t b = 0; b < J.bc.size(); b++) {
        H acc = hz();
        for (uint32_t r = 0; r < J.brun[b]; r++) acc = ha(acc, ec[y++]);
        w[b] = acc;
    }
    return w;
}

// branching-program transitions (jagged.rs:163-190, 700-716)
int bp_transition(int row, int index, int curr, int next, int state) {
    const int carry = state & 1, cmp = (state >> 1) & 1, sum = row + carry + curr;
    if (index != (sum & 1)) return -1;
    const int ncmp = (index == next) ? cmp : next;
    return (sum >> 1) + (ncmp << 1);
}
struct Sparse {
    int t[4][4][2][2];
    Sparse() {
        for (int cd = 0; cd < 4; cd++)
            for (int s = 0; s < 4; s++)
                for (int a = 0; a < 2; a++) {
                    const int c = cd & 1, d = cd >> 1, b = (a + (s & 1) + c) & 1;
                    t[cd][s][a][0] = a + 2 * b;
                    t[cd][s][a][1] = bp_transition(a, b, c, d, s);
                }
    }
};
const Sparse& sparse_tr() {
    static const Sparse s;
    return s;
}
inline int col_cd(uint64_t tc, uint64_t tn, int l) { return (int)(((tc >> l) & 1) + 2 * ((tn >> l) & 1)); }
void eq4h(H a, H b, H out[4]) {
    const H a1 = ha(h1(), a), b1 = ha(h1(), b);
    out[0] = hm(a1, b1);
    out[1] = hm(a, b1);
    out[2] = hm(a1, b);
    out[3] = hm(a, b);
}
void bp_step_h(const H e[4], int cd, const H src[4], H dst[4]) {
    const Sparse& S = sparse_tr();
    for (int s = 0; s < 4; s++) {
        H acc = hz();
        for (int a = 0; a < 2; a++) acc = ha(acc, hm(e[S.t[cd][s][a][0]], src[S.t[cd][s][a][1]]));
        dst[s] = acc;
    }
}

// Frobenius assist (pcs/jagged.rs:1712-2025) on the host: statements (z_row, run weights),
// sumcheck over the 2(m+1) boundary bits c_0, d_0, c_1, d_1, ...
void frobenius_assist(flock_fs* fs, const Jag& J, const std::vector<std::vector<H>>& zr,
                      const std::vector<std::vector<H>>& ww, const std::vector<H>& rho, uint32_t bits,
                      Bin& out) {
    const int m = J.m;
    const size_t nb = J.bc.size(), ns = zr.size(), sl = (size_t)m + 2;
    const Sparse& SP = sparse_tr();
    std::vector<H> E((size_t)ns * (m + 1) * 4);
    for (size_t s = 0; s < ns; s++)
        for (int l = 0; l <= m; l++)
            eq4h((size_t)l < (size_t)J.n ? zr[s][l] : hz(), l < m ? rho[l] : hz(),
                 &E[(s * (m + 1) + l) * 4]);
    std::vector<H> S(ns * nb * sl * 4);
    for (size_t s = 0; s < ns; s++)
        for (size_t b = 0; b < nb; b++) {
            H* Sb = &S[(s * nb + b) * sl * 4];
            for (int i = 0; i < 4; i++) Sb[(m + 1) * 4 + i] = hz();
            Sb[(m + 1) * 4 + 2] = h1();
            for (int l = m; l >= 0; l--)
                bp_step_h(&E[(s * (m + 1) + l) * 4], col_cd(J.bc[b], J.bn[b], l), Sb + (l + 1) * 4,
                          Sb + l * 4);
        }
    H v = hz();
    for (size_t s = 0; s < ns; s++)
        for (size_t b = 0; b < nb; b++) v = ha(v, hm(ww[s][b], S[(s * nb + b) * sl * 4 + 0]));
    label(fs, "flock-frobenius-assist-v0");
    observe(fs, v);
    out.f(v);
    std::vector<H> p(ns * nb), pre(ns * 4, hz()), rounds;
    std::vector<uint64_t> nonces;
    for (size_t s = 0; s < ns; s++) {
        for (size_t b = 0; b < nb; b++) p[s * nb + b] = ww[s][b];
        pre[s * 4 + 0] = h1();
    }
    H ch4[4];
    std::vector<H> U(ns * 16), B(ns * 16), UD(ns * 8);
    for (int l = 0; l <= m; l++) {
        for (size_t s = 0; s < ns; s++) {
            H* ps = &p[s * nb];
            if (l > 0)
                for (size_t b = 0; b < nb; b++) ps[b] = hm(ps[b], ch4[col_cd(J.bc[b], J.bn[b], l - 1)]);
            H* Bs = &B[s * 16];
            for (int i = 0; i < 16; i++) Bs[i] = hz();
            for (size_t b = 0; b < nb; b++) {
                const int cd = col_cd(J.bc[b], J.bn[b], l);
                const H* sv = &S[((s * nb + b) * sl + l + 1) * 4];
                for (int i = 0; i < 4; i++) Bs[cd * 4 + i] = ha(Bs[cd * 4 + i], hm(ps[b], sv[i]));
            }
            const H* e = &E[(s * (m + 1) + l) * 4];
            H* Us = &U[s * 16];
            for (int i = 0; i < 16; i++) Us[i] = hz();
            for (int cd = 0; cd < 4; cd++)
                for (int st = 0; st < 4; st++)
                    for (int a = 0; a < 2; a++) {
                        const int ix = SP.t[cd][st][a][0], o = SP.t[cd][st][a][1];
                        Us[cd * 4 + o] = ha(Us[cd * 4 + o], hm(pre[s * 4 + st], e[ix]));
                    }
        }
        H g1 = hz(), gi = hz();
        for (size_t s = 0; s < ns; s++) {
            const H* Us = &U[s * 16];
            const H* Bs = &B[s * 16];
            for (int i = 0; i < 4; i++) {
                g1 = ha(g1, ha(hm(Us[4 + i], Bs[4 + i]), hm(Us[12 + i], Bs[12 + i])));
                gi = ha(gi, ha(hm(ha(Us[i], Us[4 + i]), ha(Bs[i], Bs[4 + i])),
                               hm(ha(Us[8 + i], Us[12 + i]), ha(Bs[8 + i], Bs[12 + i]))));
            }
        }
        observe(fs, g1);
        observe(fs, gi);
        uint64_t nonce = 0;
        const H rc = bits ? grind_sample(fs, bits, &nonce) : sample(fs);
        if (bits) nonces.push_back(nonce);
        rounds.push_back(g1);
        rounds.push_back(gi);
        const H rc1 = ha(h1(), rc);
        g1 = hz();
        gi = hz();
        for (size_t s = 0; s < ns; s++) {
            const H* Us = &U[s * 16];
            const H* Bs = &B[s * 16];
            for (int i = 0; i < 4; i++) {
                const H ud0 = ha(hm(rc1, Us[i]), hm(rc, Us[4 + i]));
                const H ud1 = ha(hm(rc1, Us[8 + i]), hm(rc, Us[12 + i]));
                const H d0 = ha(hm(rc1, Bs[i]), hm(rc, Bs[4 + i]));
                const H d1 = ha(hm(rc1, Bs[8 + i]), hm(rc, Bs[12 + i]));
                g1 = ha(g1, hm(ud1, d1));
                gi = ha(gi, hm(ha(ud0, ud1), ha(d0, d1)));
                UD[s * 8 + i] = ud0;
                UD[s * 8 + 4 + i] = ud1;
            }
        }
        observe(fs, g1);
        observe(fs, gi);
        const H rd = bits ? grind_sample(fs, bits, &nonce) : sample(fs);
        if (bits) nonces.push_back(nonce);
        rounds.push_back(g1);
        rounds.push_back(gi);
        const H rd1 = ha(h1(), rd);
        for (size_t s = 0; s < ns; s++)
            for (int i = 0; i < 4; i++)
                pre[s * 4 + i] = ha(hm(rd1, UD[s * 8 + i]), hm(rd, UD[s * 8 + 4 + i]));
        ch4[0] = hm(rc1, rd1);
        ch4[1] = hm(rc, rd1);
        ch4[2] = hm(rc1, rd);
        ch4[3] = hm(rc, rd);
    }
    out.u64(rounds.size() / 2);
    for (const H& x : rounds) out.f(x);
    out.vec_u64(nonces.data(), nonces.size());
}

// Product-sumcheck round on two device tables through the public kernels, re-expressed in the
// jagged convention: the kernels give (u0, u2) = (G(0), G(inf)); G(1) = claim + G(0).
struct PairSumcheck {
    Ctx* X;
    flock_f128* a;
    flock_f128* b;
    flock_f128* ha_;  // half buffers (ping-pong)
    flock_f128* hb_;
    flock_f128* qa_;
    flock_f128* qb_;
};

// Runs the rounds of the (G(1), G(inf)) product sumcheck of <a, b> = claim over LSB pairs
// from the round whose pair (G(0), G(inf)) is m2, on tables (ca, cb) of length cur, to the
// end. Each round observes (claim + G(0), G(inf)), grinds `bits` (none when 0), samples its
// challenge and updates the claim (jagged.rs:4592-4597). Tables longer than 2^sc_la_max_log
// (a launch tunable, 2^18 by default) entries
// fold one round per launch (flock_hip_sumcheck_fold_msg); shorter ones two rounds per launch
// (k_sc_fold_la: the second round's pair comes from the lookahead at the first's challenge),
// which halves the host round trips of the small rounds. With final_fold the last challenges
// are folded in too and *fin holds a's folded value. s1/s2 are ping-pong tables of cur / 2.
bool prod_rounds(Ctx& X, flock_fs* fs, const flock_f128* ca, const flock_f128* cb, uint64_t cur,
                 const flock_f128 m2[2], H& claim, uint32_t bits, bool final_fold, flock_f128* s1a,
                 flock_f128* s1b, flock_f128* s2a, flock_f128* s2b, std::vector<H>& msgs,
                 std::vector<uint64_t>& nonces, std::vector<H>& chal, H* fin) {
    H pend[2] = {hz(), hz()};
    int np = 0, flip = 0;
    bool have_la = false;
    H la[6];
    auto round = [&](H u0, H u2) {
        const H g1 = ha(claim, u0), gi = u2;
        observe(fs, g1);
        observe(fs, gi);
        uint64_t nonce = 0;
        const H ch = bits ? grind_sample(fs, bits, &nonce) : sample(fs);
        if (bits) nonces.push_back(nonce);
        msgs.push_back(g1);
        msgs.push_back(gi);
        chal.push_back(ch);
        claim = round_claim(claim, g1, gi, ch);
        pend[np++] = ch;
    };
    // a quadratic from its values at 0 and 1 and its leading coefficient, at r
    auto quad = [](const H* p, H r) { return ha(ha(p[0], hm(ha(ha(p[0], p[1]), p[2]), r)), hm(p[2], hm(r, r))); };
    auto next = [&]() -> flock_f128* {
        flock_f128* na = flip ? s2a : s1a;
        flip ^= 1;
        return na;
    };
    auto launch_la = [&](flock_f128* na, flock_f128* nb) {
        hipLaunchKernelGGL(flock::k_sc_fold_la, dim3(X.grid(std::max<uint64_t>(cur >> np, 4))),
                           dim3(kThreads), kCombWsBytes, X.st, reinterpret_cast<const uint4*>(ca),
                           reinterpret_cast<const uint4*>(cb), cur, np, U4(pend[0]), U4(pend[1]),
                           reinterpret_cast<uint4*>(na), reinterpret_cast<uint4*>(nb), X.A->acc);
        return X.chk("fold la") && X.read_acc(32);
    };
    H u0 = m2[0], u2 = m2[1];
    while (true) {
        round(u0, u2);
        const uint64_t after = cur >> np;  // length once the pending challenges are folded in
        if (after == 1) break;             // that was the last round
        if (have_la) {
            u0 = quad(la, pend[0]);
            u2 = quad(la + 3, pend[0]);
            have_la = false;
            continue;
        }
        flock_f128* na = next();
        flock_f128* nb = na == s1a ? s1b : s2b;
        if (cur > (1ull << flock::tunable(flock::kTuneScLaMaxLog))) {  // one pending challenge here
            flock_f128 m[2];
            if (int rc = flock_hip_sumcheck_fold_msg(X.c, ca, cb, cur, pend[0], na, nb, m))
                return X.fail(rc, flock_hip_last_error());
            u0 = m[0];
            u2 = m[1];
        } else {
            if (!launch_la(na, nb)) return false;
            u0 = ha(X.acc_f(0), X.acc_f(1));
            u2 = ha(X.acc_f(2), X.acc_f(3));
            if (after >= 4) {
                la[0] = X.acc_f(0);
                la[1] = X.acc_f(4);
                la[2] = X.acc_f(2);
                la[3] = X.acc_f(5);
                la[4] = X.acc_f(6);
                la[5] = X.acc_f(7);
                have_la = true;
            }
        }
        ca = na;
        cb = nb;
        cur = after;
        np = 0;
    }
    if (final_fold) {
        flock_f128* na = next();
        flock_f128* nb = na == s1a ? s1b : s2b;
        if (!launch_la(na, nb)) return false;
        if (fin && !X.d2h(fin, na, 16)) return false;
    }
    return true;
}

// Runs n rounds of the (G(1), G(inf)) sumcheck of sum_x a(x) b(x) = claim over LSB pairs,
// each round: observe the pair, grind `bits`, sample. The final fold is skipped when
// `final_fold` is false; on return *fa holds a's final folded value if folded.
bool pair_sumcheck(Ctx& X, flock_f128* a, flock_f128* b, flock_f128* s1a, flock_f128* s1b,
                   flock_f128* s2a, flock_f128* s2b, uint64_t N, uint32_t n_rounds, H claim,
                   flock_fs* fs, uint32_t bits, bool final_fold, std::vector<H>& msgs,
                   std::vector<uint64_t>& nonces, std::vector<H>& chal, H* fa) {
    if (n_rounds == 0 || (1ull << n_rounds) != N) return X.fail(FLOCK_HIP_E_INVALID, "pair sumcheck: n != 2^rounds");
    flock_f128 m2[2];
    if (int rc = flock_hip_sumcheck_round_msg(X.c, a, b, N, m2)) return X.fail(rc, flock_hip_last_error());
    return prod_rounds(X, fs, a, b, N, m2, claim, bits, final_fold, s1a, s1b, s2a, s2b, msgs, nonces, chal, fa);
}

// ---- the eq-basis Ligerito opening (pcs.rs:214-560; extension.rs:1636-2290) ----
struct LevelW {
    flock_f128* cw = nullptr;  // block_len x lanes, row-major
    uint8_t* tree = nullptr;
    uint64_t block_len = 0, lanes = 0;
};

bool ligero_commit(Ctx& X, const flock_f128* d_poly, int log_cols, int log_lanes, int log_rate,
                   uint32_t kind, size_t slot, LevelW* w, int cap, std::vector<uint8_t>* cap_out) {
    flock_pcs_params p = {(uint32_t)(log_cols + log_lanes + 7), (uint32_t)log_rate,
                          (uint32_t)log_lanes, 0, kind};
    w->lanes = 1ull << log_lanes;
    w->block_len = 1ull << (log_cols + log_rate);
    w->cw = X.buf<flock_f128>(slot, w->block_len * w->lanes);
    w->tree = X.buf<uint8_t>(slot + 1, 32 * (2 * w->block_len - 1));
    if (X.err) return false;
    cap_out->assign(32ull << cap, 0);
    // nothing reads a level tree above its cap (the openings' paths end there): the tree stops at
    // the cap layer
    flock::ctx_set_tree_to_cap(X.c, true);
    const int rc = flock_hip_commit(X.c, d_poly, w->lanes << log_cols, &p, w->cw, w->tree, (uint32_t)cap,
                                    cap_out->data());
    flock::ctx_set_tree_to_cap(X.c, false);
    if (rc) return X.fail(rc, flock_hip_last_error());
    return true;
}

// rows at the queries and their capped paths (merkle.rs:663-682), appended to out as
// RecursiveProof {opened_rows, merkle_proof}; the rows also come back in rows_h
bool open_level(Ctx& X, const flock_f128* mat, uint64_t row_len, const uint8_t* tree,
                uint64_t block_len, const std::vector<uint64_t>& qs, int cap, Bin& out,
                std::vector<H>* rows_h, bool write_rows_first = true) {
    const uint32_t nq = (uint32_t)qs.size();
    uint64_t* d_idx = X.buf<uint64_t>(40, nq);
    const int d = __builtin_ctzll(block_len);
    const int depth = d - cap;
    std::vector<uint64_t> nodes;
    nodes.reserve((size_t)nq * depth);
    for (uint64_t q : qs) {
        uint64_t start = 0, len = block_len, idx = q;
        while (len > (1ull << cap)) {
            nodes.push_back(start + (idx ^ 1));
            start += len;
            len >>= 1;
            idx >>= 1;
        }
    }
    uint64_t* d_nidx = X.buf<uint64_t>(42, nodes.size() + 1);
    // rows, then the path nodes, written by the gather kernels straight into coherent pinned host
    // memory: the host waits for the kernels, not for a copy behind them
    const size_t row_bytes = 16 * (size_t)nq * row_len, node_bytes = 32 * nodes.size();
    uint8_t* h_out = static_cast<uint8_t*>(X.ws->pinned(0, row_bytes + node_bytes + 32));
    // the indices go out through pinned staging (rewritten only by the next level's call, after
    // this call's gathers have synchronised the stream)
    uint64_t* h_idx = static_cast<uint64_t*>(X.ws->pinned(2, 8 * (nq + nodes.size())));
    if (X.err) return false;
    if (!h_out || !h_idx) return X.fail(FLOCK_HIP_E_NOMEM, "pinned staging");
    void* d_out = nullptr;
    if (!X.hip(hipHostGetDevicePointer(&d_out, h_out, 0), "pinned device pointer")) return false;
    flock_f128* d_rows = static_cast<flock_f128*>(d_out);
    uint8_t* d_nodes = static_cast<uint8_t*>(d_out) + row_bytes;
    std::memcpy(h_idx, qs.data(), 8 * nq);
    std::memcpy(h_idx + nq, nodes.data(), 8 * nodes.size());
    if (!X.h2d(d_idx, h_idx, 8 * nq)) return false;
    hipLaunchKernelGGL(flock::k_gather_rows, dim3(X.grid((uint64_t)nq * row_len)), dim3(kThreads), 0,
                       X.st, reinterpret_cast<const uint4*>(mat), row_len, d_idx, nq,
                       reinterpret_cast<uint4*>(d_rows));
    if (!X.chk("gather rows")) return false;
    if (!nodes.empty()) {
        if (!X.h2d(d_nidx, h_idx + nq, 8 * nodes.size())) return false;
        hipLaunchKernelGGL(flock::k_gather_nodes, dim3(X.grid(2 * nodes.size())), dim3(kThreads), 0,
                           X.st, reinterpret_cast<const uint4*>(tree), d_nidx, (uint32_t)nodes.size(),
                           reinterpret_cast<uint4*>(d_nodes));
        if (!X.chk("gather nodes")) return false;
    }
    if (!X.hip(hipStreamSynchronize(X.st), "sync")) return false;
    rows_h->resize((size_t)nq * row_len);
    std::memcpy(rows_h->data(), h_out, row_bytes);
    (void)write_rows_first;
    out.u64(nq);
    for (uint32_t i = 0; i < nq; i++) out.vec_f(rows_h->data() + (size_t)i * row_len, row_len);
    out.vec_hash(h_out + row_bytes, nodes.size());
    return true;
}

// induced basis over 2^dims coefficients (ligerito.rs:2676-2770 with empty rows)
// skb[k * 64 + b] = s_k(x^b) / s_k(v_k) for k < kSkRows, b < 64 (zero where s_k(v_k) = 0), with
// s_0(x) = x, s_k(x) = s_{k-1}(x)^2 + s_{k-1}(v_{k-1}) s_{k-1}(x), v_k = x^k (eval_sk_at_vks,
// ligerito.rs:2402-2420). It depends on neither the level nor the proof: a workspace builds it
// once (slot 54).
const flock_f128* sk_basis(Ctx& X) {
    std::lock_guard<std::mutex> lk(X.ws->sk_mu);
    flock_f128* d = X.buf<flock_f128>(54, (size_t)flock::kSkRows * 64);
    if (!d || X.ws->sk_ready) return d;
    constexpr int K = flock::kSkRows;
    // s_k(v_k) by the reference's layered recursion
    std::vector<H> skv(K + 1), layer(K);
    skv[0] = h1();
    for (int i = 1; i <= K; i++) layer[i - 1] = hbit(i);
    for (int i = 0, cur = K; i < K; i++, cur--)
        for (int j = 0; j < cur; j++) {
            const H sv = layer[j], nx = ha(hm(sv, sv), hm(skv[i], sv));
            if (j == 0)
                skv[i + 1] = nx;
            else
                layer[j - 1] = nx;
        }
    // their inverses with one field inversion (prefix products)
    std::vector<H> inv(K, hz()), pre(K + 1, h1());
    for (int i = 0; i < K; i++) pre[i + 1] = hzero(skv[i]) ? pre[i] : hm(pre[i], skv[i]);
    H acc = hinv(pre[K]);
    for (int i = K - 1; i >= 0; i--) {
        if (hzero(skv[i])) continue;
        inv[i] = hm(acc, pre[i]);
        acc = hm(acc, skv[i]);
    }
    std::vector<H> tab((size_t)K * 64);
    for (int b = 0; b < 64; b++) {
        H cur = hbit(b);
        for (int k = 0; k < K; k++) {
            tab[(size_t)k * 64 + b] = hm(cur, inv[k]);
            cur = ha(hm(cur, cur), hm(skv[k], cur));
        }
    }
    if (!X.h2d(d, tab.data(), 16 * tab.size()) || !X.hip(hipStreamSynchronize(X.st), "sync")) return nullptr;
    X.ws->sk_ready = true;
    return d;
}

bool induced_basis(Ctx& X, int dims, const std::vector<uint64_t>& qs, const std::vector<H>& alpha,
                   flock_f128* out) {
    if (dims < 0 || dims > flock::kSkRows) return X.fail(FLOCK_HIP_E_INVALID, "induced basis: dims");
    const flock_f128* d_skb = sk_basis(X);
    if (!d_skb) return X.err ? false : X.fail(FLOCK_HIP_E_NOMEM, "induced basis table");
    const std::vector<H> aw = heq(alpha.data(), alpha.size());
    const uint32_t nq = (uint32_t)qs.size();
    const int h = dims / 2;
    const uint64_t nl = 1ull << h, nh = 1ull << (dims - h);
    flock_f128* d_aw = X.buf<flock_f128>(49, nq);
    flock_f128* d_lo = X.buf<flock_f128>(44, (size_t)nq * nl);
    flock_f128* d_hi = X.buf<flock_f128>(45, (size_t)nq * nh);
    uint64_t* d_q = X.buf<uint64_t>(51, nq);
    if (X.err) return false;
    // the inputs go out through pinned staging, so nothing waits for the copies: the staging is
    // rewritten only by the next level's call, after that level's row copy-back has synchronised
    // the stream
    const size_t up_bytes = 8 * (size_t)nq + 16 * (size_t)nq;
    uint8_t* up = static_cast<uint8_t*>(X.ws->pinned(1, up_bytes));
    if (!up) return X.fail(FLOCK_HIP_E_NOMEM, "pinned staging");
    std::memcpy(up, qs.data(), 8 * nq);
    std::memcpy(up + 8 * nq, aw.data(), 16 * nq);
    if (!X.h2d(d_q, up, 8 * nq) || !X.h2d(d_aw, up + 8 * nq, 16 * nq)) return false;
    hipLaunchKernelGGL(flock::k_induced_tables, dim3(nq), dim3(kThreads), 16 * (nl + nh), X.st, d_q,
                       reinterpret_cast<const uint4*>(d_skb), reinterpret_cast<const uint4*>(d_aw),
                       (uint32_t)dims, (uint32_t)h, reinterpret_cast<uint4*>(d_lo),
                       reinterpret_cast<uint4*>(d_hi));
    if (!X.chk("induced tables")) return false;
    // jl rows in blocks of kIndTL, jh columns in blocks of kThreads * kIndTH
    const uint32_t gx = (uint32_t)((nl + flock::kIndTL - 1) / flock::kIndTL);
    const uint32_t gy = (uint32_t)((nh + kThreads * flock::kIndTH - 1) / (kThreads * flock::kIndTH));
    // split the query sum into chunks until the grid holds about 1024 workgroups, keeping at
    // least 4 queries per chunk (the chunk partials cost a write and a read of 2^dims words each)
    uint32_t gz = std::max<uint32_t>(
        1, std::min<uint32_t>((1024 + gx * gy - 1) / std::max<uint32_t>(1, gx * gy), nq / 4));
    const uint32_t chunk = (nq + gz - 1) / gz;
    gz = (nq + chunk - 1) / chunk;
    flock_f128* part = gz > 1 ? X.buf<flock_f128>(53, (size_t)gz << dims) : out;
    if (X.err) return false;
    hipLaunchKernelGGL(flock::k_induced_gemm, dim3(gx, gy, gz), dim3(kThreads), 0, X.st,
                       reinterpret_cast<const uint4*>(d_lo), reinterpret_cast<const uint4*>(d_hi), nq, (uint32_t)h,
                       (uint32_t)dims, chunk, reinterpret_cast<uint4*>(part));
    if (gz > 1)
        hipLaunchKernelGGL(flock::k_xor_chunks, dim3(X.grid(1ull << dims)), dim3(kThreads), 0, X.st,
                           reinterpret_cast<const uint4*>(part), gz, 1ull << dims, reinterpret_cast<uint4*>(out));
    return X.chk("induced basis");
}

struct Ladder {
    Ctx& X;
    flock_fs* fs;
    e256d* f;   // current table
    e256d* b;   // combined basis
    e256d* f2;  // ping-pong
    e256d* b2;
    uint64_t n = 0;
    Bin msgs;
    size_t n_msgs = 0;

    void push(E u0, E u2) {
        msgs.e(u0);
        msgs.e(u2);
        n_msgs++;
        observe_e(fs, u0);
        observe_e(fs, u2);
    }
    void swap() {
        std::swap(f, f2);
        std::swap(b, b2);
    }
    bool msg(uint64_t d, E* u0, E* u2) {
        const uint64_t dd = (d == 1 || n == d) ? 1 : d;
        hipLaunchKernelGGL(flock::k_msg_e_comb, dim3(X.grid(n / 2)), dim3(kThreads), kCombWsBytes, X.st, f, b, n,
                           dd, X.A->acc);
        if (!X.chk("msg") || !X.read_acc(16)) return false;
        *u0 = E{X.acc_f(0), X.acc_f(1)};
        *u2 = E{X.acc_f(2), X.acc_f(3)};
        return true;
    }
    bool fold(uint64_t d, E r) {
        hipLaunchKernelGGL((flock::k_fold_msg<false, false>), dim3(X.grid(n / 2)), dim3(kThreads), 0, X.st,
                           static_cast<const void*>(f), static_cast<const void*>(b), n, d, E4(r),
                           (uint64_t)1, f2, b2, X.A->acc);
        if (!X.chk("fold")) return false;
        swap();
        n /= 2;
        return true;
    }
    // the fold and the next round's message (over pairs at d, or LSB pairs once a single block
    // of d is left); BASE: the inputs are the F128 tables (fb, bb) of length n. Small tables take
    // one fused pass; large ones a fold pass and a message pass, each at a higher occupancy
    // than the fused kernel's.
    template <bool BASE>
    bool fold_msg(uint64_t d, E r, E* u0, E* u2, const void* fb = nullptr, const void* bb = nullptr) {
        const uint64_t nn = n / 2, d2 = (d == 1 || nn == d) ? 1 : d;
        const void* src_f = BASE ? fb : static_cast<const void*>(f);
        const void* src_b = BASE ? bb : static_cast<const void*>(b);
        e256d* dst_f = BASE ? f : f2;
        e256d* dst_b = BASE ? b : b2;
        if (nn >= (1ull << flock::tunable(flock::kTuneOpenSplitFoldLog))) {  // fold and message split
            hipLaunchKernelGGL((flock::k_fold_msg<BASE, false>), dim3(X.grid(nn)), dim3(kThreads), 0, X.st, src_f,
                               src_b, n, d, E4(r), d2, dst_f, dst_b, X.A->acc);
            hipLaunchKernelGGL(flock::k_msg_e_comb, dim3(X.grid(nn / 2)), dim3(kThreads), kCombWsBytes, X.st,
                               dst_f, dst_b, nn, d2, X.A->acc);
        } else {
            hipLaunchKernelGGL((flock::k_fold_msg<BASE, true>), dim3(X.grid(nn / 2)), dim3(kThreads), kCombWsBytes,
                               X.st, src_f, src_b, n, d, E4(r), d2, dst_f, dst_b, X.A->acc);
        }
        if (!X.chk("fold msg") || !X.read_acc(16)) return false;
        if (!BASE) swap();
        n = nn;
        *u0 = E{X.acc_f(0), X.acc_f(1)};
        *u2 = E{X.acc_f(2), X.acc_f(3)};
        return true;
    }
    bool sw(E* u0, E* u2) {  // code switch + its LSB message
        hipLaunchKernelGGL(flock::k_switch, dim3(X.grid(n)), dim3(kThreads), 0, X.st, f, b, n, f2, b2);
        if (!X.chk("switch")) return false;
        swap();
        n *= 2;
        return msg(1, u0, u2);
    }
};

// FLOCK_OPEN_TRACE=1: wall time of the inner opening's steps, summed per step name, on stderr
struct StepTrace {
    bool on = std::getenv("FLOCK_OPEN_TRACE") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    std::vector<std::pair<const char*, double>> acc;
    void mark(const char* name) {
        if (!on) return;
        const auto now = std::chrono::steady_clock::now();
        const double ms = std::chrono::duration<double, std::milli>(now - t).count();
        t = now;
        for (auto& a : acc)
            if (a.first == name) {
                a.second += ms;
                return;
            }
        acc.emplace_back(name, ms);
    }
    ~StepTrace() {
        if (!on) return;
        for (auto& a : acc) std::fprintf(stderr, "open-trace %-14s %8.3f ms\n", a.first, a.second);
    }
};

// A one-shot gate between the opening's two host threads.
struct Gate {
    std::mutex m;
    std::condition_variable cv;
    bool open = false;
    void release() {
        {
            std::lock_guard<std::mutex> lk(m);
            open = true;
        }
        cv.notify_all();
    }
    void wait() {
        std::unique_lock<std::mutex> lk(m);
        cv.wait(lk, [&] { return open; });
    }
};

// `lanes_done` (optional) is released once the lane folds (the opening's large, GPU-bound
// passes) are enqueued and their last message has come back.
bool inner_open(Ctx& X, flock_fs* fs, const flock_f128* q, const flock_pcs_params* P,
                const flock_f128* cw0, const uint8_t* tree0, const std::vector<H>& rho, H q_eval,
                Bin& out, Gate* lanes_done = nullptr) {
    const Lig* C = lig_for(P->m);
    if (!C || C->initial_k != P->log_batch_size)
        return X.fail(FLOCK_HIP_E_INVALID, "no Ligerito config for this m / log_batch_size");
    const int log_n = (int)P->m - 7, ik = C->initial_k, R = C->n_levels - 1;
    const uint64_t N = 1ull << log_n;
    const uint64_t t = P->num_lanes ? P->num_lanes : (1ull << ik);
    const int n1 = log_n - ik;
    const uint64_t D = t < (1ull << ik) ? (1ull << n1) : 1;
    const uint64_t bl0 = 1ull << (n1 + C->lv[0].rate);
    std::vector<Sched> sch;
    for (int l = 0; l <= R; l++) sch.push_back(sched_of(C->lv[l].queries, C->lv[l].cols + C->lv[l].rate));
    const uint32_t kind = P->merkle_hash;
    StepTrace T;

    // compute_combined_basis_and_target: one packed-direct EqPoint claim (pcs.rs:423-560)
    label(fs, "flock-pcs-open-batch-v0");
    label(fs, "flock-pcs-packed-direct-v0");
    observe(fs, q_eval);
    const H gamma = sample_vec(fs, 1)[0];
    H target = hm(gamma, q_eval);
    // b0 = gamma eq(rho) + sum_o beta_o eq(z_o), written once after the OOD samples: each eq
    // table stays as its two halves (eq(r)[x] = lo[xl] hi[xh]), <q, eq(z_o)> is summed from q
    // and the halves, and k_b0_build writes b0 from all the halves at once
    const int nlo = log_n / 2, nhi = log_n - nlo;
    const uint64_t nl = 1ull << nlo, nh = 1ull << nhi;
    const int n_ood = C->lv[0].ood;
    if (n_ood + 1 > flock::kB0Terms) return X.fail(FLOCK_HIP_E_INVALID, "too many L0 OOD samples");
    flock_f128* b0 = X.buf<flock_f128>(10, N);
    flock_f128* halves = X.buf<flock_f128>(11, (size_t)(1 + n_ood) * (nl + nh));
    const uint32_t nbh = (uint32_t)((nh + flock::kOodXh - 1) / flock::kOodXh);
    if (X.err) return false;
    flock::B0Terms terms{};
    auto term = [&](int t, const H* r, H coef) {
        terms.lo[t] = reinterpret_cast<const uint4*>(halves + (size_t)t * (nl + nh));
        terms.hi[t] = reinterpret_cast<const uint4*>(halves + (size_t)t * (nl + nh) + nl);
        terms.c[t] = U4(coef);
        terms.n = t + 1;
        return X.eq_halves(r, (size_t)log_n, halves + (size_t)t * (nl + nh), halves + (size_t)t * (nl + nh) + nl);
    };
    if (!term(0, rho.data(), gamma)) return false;
    label(fs, "flock-ligerito-basis-f256-split-v0");
    observe(fs, target);
    const int c0 = sch[0].cap();
    std::vector<uint8_t> icap(32ull << c0);
    if (!X.d2h(icap.data(), tree0 + 32 * (2 * bl0 - 2 * (1ull << c0)), icap.size())) return false;
    flock_fs_observe_bytes(fs, icap.data(), icap.size());
    std::vector<H> ood;
    std::vector<uint64_t> claim_n, cons_n, grind_n;
    for (int o = 0; o < n_ood; o++) {  // L0 OOD samples
        const std::vector<H> z = sample_vec(fs, (size_t)log_n);
        if (!term(1 + o, z.data(), hz())) return false;
        hipLaunchKernelGGL(flock::k_ood_dot,
                           dim3((uint32_t)((nl + kThreads * flock::kOodXl - 1) / (kThreads * flock::kOodXl)), nbh),
                           dim3(kThreads), kCombWsBytes, X.st, reinterpret_cast<const uint4*>(q), nlo, nhi,
                           terms.lo[1 + o], terms.hi[1 + o], X.A->acc);
        if (!X.chk("ood dot") || !X.read_acc(4)) return false;
        const H y = X.acc_f(0);
        observe(fs, y);
        ood.push_back(y);
        uint64_t nonce;
        const H beta = grind_sample(fs, C->lv[0].claim, &nonce);
        claim_n.push_back(nonce);
        target = ha(target, hm(beta, y));
        terms.c[1 + o] = U4(beta);
    }
    hipLaunchKernelGGL(flock::k_b0_build, dim3((uint32_t)((nh + flock::kB0Rows - 1) / flock::kB0Rows)),
                       dim3(kThreads), (size_t)terms.n * SHOUP_ENTRIES * 16, X.st, terms, nlo, nhi,
                       reinterpret_cast<uint4*>(b0));
    if (!X.chk("basis table")) return false;
    T.mark("ood0");
    // first message over blocked pairs (F128)
    hipLaunchKernelGGL(flock::k_msg_blocked_comb, dim3(X.grid(N / 2)), dim3(kThreads), kCombWsBytes, X.st,
                       reinterpret_cast<const uint4*>(q), reinterpret_cast<const uint4*>(b0), N, D, X.A->acc);
    if (!X.chk("first msg") || !X.read_acc(8)) return false;
    Ladder Lg{X, fs, X.buf<e256d>(12, N / 2), X.buf<e256d>(13, N / 2), X.buf<e256d>(14, N / 2),
              X.buf<e256d>(15, N / 2)};
    if (X.err) return false;
    Lg.push(E{X.acc_f(0), hz()}, E{X.acc_f(1), hz()});
    // the initial_k lane folds; the last one ends in the code switch
    for (int j = 0; j < ik; j++) {
        const E r = sample_e(fs);
        E u0, u2;
        if (j == 0) {  // initial_k >= 2: a message always follows the base fold
            Lg.n = N;
            if (!Lg.fold_msg<true>(D, r, &u0, &u2, q, b0)) return false;
        } else if (j + 1 < ik) {
            if (!Lg.fold_msg<false>(D, r, &u0, &u2)) return false;
        } else if (!Lg.fold(D, r) || !Lg.sw(&u0, &u2)) {
            return false;
        }
        Lg.push(u0, u2);
    }
    // the split table's buffers must hold the doubled length: the first fold wrote N/2, the
    // switch's output is 2 * 2^n1 <= N/2 for ik >= 2
    T.mark("lane folds");
    if (lanes_done) lanes_done->release();
    int split_dim = n1 + 1;
    flock_f128* base = X.buf<flock_f128>(16, 1ull << split_dim);
    if (X.err) return false;
    auto base_of = [&]() {
        hipLaunchKernelGGL(flock::k_base_table, dim3(X.grid(Lg.n)), dim3(kThreads), 0, X.st, Lg.f, Lg.n,
                           reinterpret_cast<uint4*>(base));
        return X.chk("base table");
    };
    int cur_slot = 20;
    LevelW prev;
    std::vector<uint8_t> capb;
    Bin caps;
    size_t n_caps = 0;
    if (!base_of() ||
        !ligero_commit(X, base, split_dim - C->lv[1].lanes, C->lv[1].lanes, C->lv[1].rate, kind, cur_slot,
                       &prev, sch[1].cap(), &capb))
        return false;
    flock_fs_observe_bytes(fs, capb.data(), capb.size());
    caps.vec_hash(capb.data(), capb.size() / 32);
    n_caps++;
    T.mark("commit");
    flock_f128* ezl = X.buf<flock_f128>(17, 1ull << split_dim);
    flock_f128* basis = X.buf<flock_f128>(18, 1ull << split_dim);
    if (X.err) return false;
    auto level_ood = [&](int level) -> bool {
        for (int o = 0; o < C->lv[level].ood; o++) {
            const std::vector<H> z = sample_vec(fs, (size_t)split_dim);
            if (!X.eq_dev(z.data(), (size_t)split_dim, ezl)) return false;
            hipLaunchKernelGGL(flock::k_intro_base, dim3(X.grid(Lg.n / 2)), dim3(kThreads), kCombWsBytes, X.st, Lg.f,
                               reinterpret_cast<const uint4*>(ezl), Lg.n, X.A->acc);
            if (!X.chk("intro ood") || !X.read_acc(20)) return false;
            const H y = X.acc_f(0);
            const E u0{X.acc_f(1), X.acc_f(2)}, u2{X.acc_f(3), X.acc_f(4)};
            Lg.msgs.e(u0);
            Lg.msgs.e(u2);
            Lg.n_msgs++;
            observe(fs, y);
            ood.push_back(y);
            observe_e(fs, u0);
            observe_e(fs, u2);
            uint64_t nonce;
            const H beta = grind_sample(fs, C->lv[level].claim, &nonce);
            claim_n.push_back(nonce);
            hipLaunchKernelGGL(flock::k_glue_base_tab, dim3(X.grid(Lg.n)), dim3(kThreads), 0, X.st, Lg.b,
                               reinterpret_cast<const uint4*>(ezl), Lg.n, U4(beta));
            if (!X.chk("glue")) return false;
        }
        return true;
    };
    auto presplit = [&](int dims, int next_level) -> bool {
        hipLaunchKernelGGL(flock::k_intro_presplit, dim3(X.grid(1ull << dims)), dim3(kThreads), 0, X.st, Lg.f,
                           reinterpret_cast<const uint4*>(basis), 1ull << dims, X.A->acc);
        if (!X.chk("intro presplit") || !X.read_acc(16)) return false;
        const E u0{X.acc_f(0), X.acc_f(1)}, u2{X.acc_f(2), X.acc_f(3)};
        Lg.push(u0, u2);
        uint64_t nonce;
        const H beta = grind_sample(fs, C->lv[next_level].claim, &nonce);
        claim_n.push_back(nonce);
        hipLaunchKernelGGL(flock::k_glue_presplit_tab, dim3(X.grid(1ull << dims)), dim3(kThreads), 0, X.st, Lg.b,
                           reinterpret_cast<const uint4*>(basis), 1ull << dims, U4(beta));
        return X.chk("glue presplit");
    };
    auto queries = [&](int level, uint64_t block_len, std::vector<uint64_t>* qs, std::vector<H>* alpha) {
        const size_t nq = C->lv[level].queries;
        grind_n.push_back(flock_fs_grind_pow(fs, C->lv[level].grind));
        const std::vector<H> w = sample_vec(fs, nq);
        (void)block_len;
        *qs = queries_from(sch[level], w);
        cons_n.push_back(flock_fs_grind_pow(fs, C->lv[level].cons));
        *alpha = sample_vec(fs, ceil_log2(nq));
    };
    if (!level_ood(1)) return false;
    T.mark("ood");
    Bin l0p;
    {
        std::vector<uint64_t> qs;
        std::vector<H> alpha, rows;
        queries(0, bl0, &qs, &alpha);
        T.mark("queries");
        if (!open_level(X, cw0, t, tree0, bl0, qs, c0, l0p, &rows)) return false;
        T.mark("open level");
        if (!induced_basis(X, n1, qs, alpha, basis)) return false;
        T.mark("induced");
        if (!presplit(n1, 1)) return false;
        T.mark("presplit");
    }
    Bin recp, finp;
    size_t n_rec = 0;
    for (int i = 0; i < R; i++) {
        const int k = C->lv[i + 1].lanes;
        for (int j = 0; j < k; j++) {
            const E r = sample_e(fs);
            const bool switching = (j + 1 == k) && (i + 1 != R);
            E u0, u2;
            if (switching) {
                if (!Lg.fold(1, r) || !Lg.sw(&u0, &u2)) return false;
            } else if (!Lg.fold_msg<false>(1, r, &u0, &u2)) {
                return false;
            }
            Lg.push(u0, u2);
        }
        T.mark("level folds");
        const int ext_dim = split_dim - k, level = i + 1;
        if (i + 1 == R) {
            std::vector<e256d> fv(Lg.n);
            if (!X.d2h(fv.data(), Lg.f, sizeof(e256d) * Lg.n)) return false;
            std::vector<H> yr;
            for (const e256d& v : fv) {
                yr.push_back(H{(uint64_t)v.c0.y << 32 | v.c0.x, (uint64_t)v.c0.w << 32 | v.c0.z});
                yr.push_back(H{(uint64_t)v.c1.y << 32 | v.c1.x, (uint64_t)v.c1.w << 32 | v.c1.z});
            }
            for (const H& y : yr) observe(fs, y);
            std::vector<uint64_t> qs;
            std::vector<H> alpha, rows;
            queries(level, prev.block_len, &qs, &alpha);
            uint64_t nonce;
            (void)grind_sample(fs, C->lv[level].claim, &nonce);
            claim_n.push_back(nonce);
            finp.vec_f(yr.data(), yr.size());
            if (!open_level(X, prev.cw, prev.lanes, prev.tree, prev.block_len, qs, sch[level].cap(), finp,
                            &rows))
                return false;
            break;
        }
        split_dim = ext_dim + 1;
        const int next_level = i + 2;
        LevelW next;
        cur_slot = cur_slot == 20 ? 22 : 20;
        if (!base_of() ||
            !ligero_commit(X, base, split_dim - C->lv[next_level].lanes, C->lv[next_level].lanes,
                           C->lv[next_level].rate, kind, cur_slot, &next, sch[next_level].cap(), &capb))
            return false;
        flock_fs_observe_bytes(fs, capb.data(), capb.size());
        caps.vec_hash(capb.data(), capb.size() / 32);
        n_caps++;
        T.mark("commit");
        if (!level_ood(next_level)) return false;
        T.mark("ood");
        {
            std::vector<uint64_t> qs;
            std::vector<H> alpha, rows;
            queries(level, prev.block_len, &qs, &alpha);
            T.mark("queries");
            if (!open_level(X, prev.cw, prev.lanes, prev.tree, prev.block_len, qs, sch[level].cap(), recp,
                            &rows))
                return false;
            n_rec++;
            T.mark("open level");
            if (!induced_basis(X, ext_dim, qs, alpha, basis)) return false;
            T.mark("induced");
            if (!presplit(ext_dim, next_level)) return false;
            T.mark("presplit");
        }
        prev = next;
    }
    // BatchOpeningProofLigerito {ring_switches: [], batching_nonces: [], ligerito}
    out.u64(0);
    out.u64(0);
    // LigeritoProof (ligerito.rs:2177-2218)
    out.vec_hash(icap.data(), icap.size() / 32);
    out.raw(l0p.b.data(), l0p.b.size());
    out.u64(n_caps);
    out.raw(caps.b.data(), caps.b.size());
    out.u64(n_rec);
    out.raw(recp.b.data(), recp.b.size());
    out.raw(finp.b.data(), finp.b.size());
    out.u64(0);  // sumcheck_transcript (F128)
    out.u64(Lg.n_msgs);
    out.raw(Lg.msgs.b.data(), Lg.msgs.b.size());
    out.vec_u64(grind_n.data(), grind_n.size());
    out.vec_f(ood.data(), ood.size());
    out.u64(0);  // fold_grinding_nonces
    out.vec_u64(claim_n.data(), claim_n.size());
    out.vec_u64(cons_n.data(), cons_n.size());
    return true;
}

}  // namespace

extern "C" {

int flock_open_ws_create(flock_open_ws** out) {
    if (!out) return flock::set_error(FLOCK_HIP_E_INVALID, "null output");
    flock_open_ws* w = new (std::nothrow) flock_open_ws();
    if (!w) return flock::set_error(FLOCK_HIP_E_NOMEM, "workspace");
    for (AccSet& a : w->accs) {
        if (hipMalloc(&a.acc, 4 * 256) != hipSuccess ||
            hipHostMalloc((void**)&a.hacc, 4 * 260, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
            hipHostGetDevicePointer((void**)&a.dacc, a.hacc, 0) != hipSuccess) {
            (void)hipGetLastError();
            for (AccSet& b : w->accs) {
                if (b.acc) (void)hipFree(b.acc);
                if (b.hacc) (void)hipHostFree(b.hacc);
            }
            delete w;
            return flock::set_error(FLOCK_HIP_E_NOMEM, "workspace accumulators");
        }
        // the flag word starts below every sequence number
        std::memset(a.hacc, 0, 4 * 260);
    }
    w->bufs.resize(flock_open_ws::kSlots);
    *out = w;
    return 0;
}

void flock_open_ws_destroy(flock_open_ws* w) {
    if (!w) return;
    if (w->aux) flock_hip_ctx_destroy(w->aux);
    for (DBuf& b : w->bufs)
        if (b.p) (void)hipFree(b.p);
    for (AccSet& a : w->accs) {
        if (a.acc) (void)hipFree(a.acc);
        if (a.hacc) (void)hipHostFree(a.hacc);
    }
    for (void* p : w->pin)
        if (p) (void)hipHostFree(p);
    delete w;
}

uint64_t flock_open_ws_bytes(const flock_open_ws* w) {
    uint64_t s = 0;
    if (w)
        for (const DBuf& b : w->bufs) s += b.bytes;
    return s;
}

int flock_lig_initial_k(uint32_t m) {
    const Lig* c = lig_for(m);
    return c ? c->initial_k : -1;
}

int flock_l0_cap_depth(uint32_t m) {
    const Lig* c = lig_for(m);
    if (!c) return -1;
    return sched_of(c->lv[0].queries, c->lv[0].cols + c->lv[0].rate).cap();
}

int flock_hip_s_hat_v_from_z_vec(flock_hip_ctx* ctx, flock_open_ws* ws, const flock_f128* d_z_vec,
                                 size_t z_vec_len, const flock_f128* h_tail, size_t tail_len,
                                 flock_f128 h_out[128]) {
    if (!ctx || !ws || !d_z_vec || !h_out || z_vec_len != (128ull << tail_len))
        return flock::set_error(FLOCK_HIP_E_INVALID, "s_hat_v_from_z_vec: bad arguments");
    Ctx X{ctx, ws, static_cast<hipStream_t>(flock_hip_ctx_stream(ctx)), &ws->accs[0]};
    flock_f128* eq = X.buf<flock_f128>(46, 1ull << tail_len);
    flock_f128* out = X.buf<flock_f128>(47, 128);
    if (X.err) return flock::set_error(X.err, X.what.c_str());
    if (!X.eq_dev(h_tail, tail_len, eq)) return flock::set_error(X.err, X.what.c_str());
    hipLaunchKernelGGL(flock::k_zero32, dim3(2), dim3(256), 0, X.st, reinterpret_cast<uint32_t*>(out), 512u);
    const uint64_t n_tail = 1ull << tail_len;
    hipLaunchKernelGGL(flock::k_shv_zvec, dim3((uint32_t)((n_tail + flock::kShvK - 1) / flock::kShvK)),
                       dim3(kThreads), 0, X.st, reinterpret_cast<const uint4*>(d_z_vec),
                       reinterpret_cast<const uint4*>(eq), n_tail, reinterpret_cast<uint4*>(out));
    if (!X.chk("s_hat_v") || !X.d2h(h_out, out, 16 * 128)) return flock::set_error(X.err, X.what.c_str());
    return 0;
}

int flock_hip_open_merged(flock_hip_ctx* ctx, flock_open_ws* ws, flock_fs* fs, const flock_f128* d_q,
                          const flock_pcs_params* P, const flock_f128* d_codeword, const uint8_t* d_tree,
                          size_t n_claims, const flock_f128* h_x_outers, size_t x_len,
                          const flock_f128* h_s_hat_v, const uint64_t* h_heights, size_t n_heights,
                          uint32_t n_log, uint8_t** out_bytes, size_t* out_len, flock_f128* h_q_eval) {
    const flock::ProofScope in_flight;
    if (!ctx || !ws || !fs || !d_q || !P || !d_codeword || !d_tree || !h_x_outers || !h_s_hat_v ||
        !h_heights || !out_bytes || !out_len || n_claims < 1 || n_claims > flock::kMaxClaims ||
        n_heights == 0 || (n_heights & (n_heights - 1)) || P->m < 14 || P->m > 40)
        return flock::set_error(FLOCK_HIP_E_INVALID, "open_merged: bad arguments");
    *out_bytes = nullptr;
    *out_len = 0;
    Ctx X{ctx, ws, static_cast<hipStream_t>(flock_hip_ctx_stream(ctx)), &ws->accs[0]};
    if (!X.hip(comb_attrs(), "kernel attributes") || !X.clear_acc())
        return flock::set_error(X.err, X.what.c_str());
    const int dense_log = (int)P->m - 7;
    const uint64_t N = 1ull << dense_log;
    const Jag J = jag_make(h_heights, n_heights, (int)n_log, dense_log);
    if (x_len != 1 + n_log + (size_t)J.k || J.area() > N)
        return flock::set_error(FLOCK_HIP_E_INVALID, "open_merged: point / heights mismatch");
    const uint64_t area = J.area();
    const uint32_t n_cols = (uint32_t)n_heights;
    Bin out;
    ws->stamp(-1);
    label(fs, "flock-merged-open-v1");
    // ring switch per claim (ring_switch.rs:2723-2760), 3 PoW bits each
    std::vector<std::vector<H>> eqr(n_claims);
    std::vector<H> rs_claim(n_claims);
    std::vector<uint64_t> rs_nonce(n_claims);
    for (size_t c = 0; c < n_claims; c++) {
        const H* sv = h_s_hat_v + 128 * c;
        label(fs, "flock-ring-switch-v0");
        flock_fs_observe_f128_slice(fs, sv, 128);
        rs_nonce[c] = flock_fs_grind_pow(fs, 3);
        const std::vector<H> rdp = sample_vec(fs, 7);
        eqr[c] = heq(rdp.data(), 7);
        H acc = hz();
        for (int bt = 0; bt < 128; bt++) {
            H u = hz();
            for (int i = 0; i < 128; i++)
                if (hgetbit(sv[i], bt)) {
                    if (i < 64)
                        u.lo |= 1ull << i;
                    else
                        u.hi |= 1ull << (i - 64);
                }
            acc = ha(acc, hm(u, eqr[c][bt]));
        }
        rs_claim[c] = acc;
    }
    const uint32_t bb = n_claims > 1 ? 6 : 0;
    // claim_batch_bits_for (pcs.rs:193-199): no PoW, and nothing absorbed, for one claim
    const uint64_t batch_nonce = bb ? flock_fs_grind_pow(fs, bb) : 0;
    const std::vector<H> gam = sample_vec(fs, n_claims);
    H target = hz();
    std::vector<std::vector<H>> tabs(n_claims);
    for (size_t c = 0; c < n_claims; c++) {
        target = ha(target, hm(gam[c], rs_claim[c]));
        H ge[128];
        for (int bt = 0; bt < 128; bt++) ge[bt] = hm(gam[c], eqr[c][bt]);
        tabs[c] = lin_tables(ge);
    }
    // device tables: prefix sums, per-claim eq_row / eq_col / Phi tables
    uint64_t* d_ps = X.buf<uint64_t>(0, J.ps.size());
    flock_f128* d_tab = X.buf<flock_f128>(1, 512 * n_claims);
    flock_f128* d_er = X.buf<flock_f128>(2, (n_claims << n_log));
    flock_f128* d_ec = X.buf<flock_f128>(3, n_claims * (size_t)n_cols);
    flock_f128* W = X.buf<flock_f128>(4, N);
    flock_f128* A = X.buf<flock_f128>(5, N);
    flock_f128* s1a = X.buf<flock_f128>(6, N / 2);
    flock_f128* s1b = X.buf<flock_f128>(7, N / 2);
    flock_f128* s2a = X.buf<flock_f128>(8, N / 4 + 1);
    flock_f128* s2b = X.buf<flock_f128>(9, N / 4 + 1);
    if (X.err) return flock::set_error(X.err, X.what.c_str());
    std::vector<H> ec_all(n_claims * (size_t)n_cols), tab_all(512 * n_claims);
    flock::JagClaims jc{};
    jc.n = (uint32_t)n_claims;
    for (size_t c = 0; c < n_claims; c++) {
        const H* x = h_x_outers + c * x_len;
        if (!X.eq_dev(x + 1, n_log, d_er + (c << n_log))) return flock::set_error(X.err, X.what.c_str());
        const std::vector<H> ec = heq(x + 1 + n_log, (size_t)J.k);
        std::copy(ec.begin(), ec.end(), ec_all.begin() + c * n_cols);
        std::copy(tabs[c].begin(), tabs[c].end(), tab_all.begin() + 512 * c);
        jc.er[c] = reinterpret_cast<const uint4*>(d_er + (c << n_log));
        jc.ec[c] = reinterpret_cast<const uint4*>(d_ec + c * n_cols);
        jc.tab[c] = reinterpret_cast<const uint4*>(d_tab + 512 * c);
    }
    if (!X.h2d(d_ps, J.ps.data(), 8 * J.ps.size()) || !X.h2d(d_ec, ec_all.data(), 16 * ec_all.size()) ||
        !X.h2d(d_tab, tab_all.data(), 16 * tab_all.size()))
        return flock::set_error(X.err, X.what.c_str());
    // tiles of at most kJagTile rows of one column; W past the area is zero
    std::vector<flock::JagTile> tiles;
    for (uint32_t c = 0; c < n_cols; c++)
        for (uint64_t off = 0; off < J.ps[c + 1] - J.ps[c]; off += flock::kJagTile)
            tiles.push_back({c, (uint32_t)std::min<uint64_t>(flock::kJagTile, J.ps[c + 1] - J.ps[c] - off),
                             J.ps[c] + off});
    flock::JagTile* d_tiles = X.buf<flock::JagTile>(50, tiles.size() + 1);
    if (X.err || (!tiles.empty() && !X.h2d(d_tiles, tiles.data(), sizeof(flock::JagTile) * tiles.size())))
        return flock::set_error(X.err, X.what.c_str());
    auto jag_table = [&](Ctx& Y, bool weight, flock_f128* dst) -> bool {
        if (area < N)
            hipLaunchKernelGGL(flock::k_zero_f128, dim3(Y.grid(N - area)), dim3(kThreads), 0, Y.st,
                               reinterpret_cast<uint4*>(dst + area), N - area);
        if (!tiles.empty()) {
            if (weight)
                hipLaunchKernelGGL(flock::k_jag_tiles<true>, dim3((uint32_t)tiles.size()), dim3(kThreads), 0, Y.st,
                                   reinterpret_cast<uint4*>(dst), d_tiles, d_ps, jc);
            else
                hipLaunchKernelGGL(flock::k_jag_tiles<false>, dim3((uint32_t)tiles.size()), dim3(kThreads), 0,
                                   Y.st, reinterpret_cast<uint4*>(dst), d_tiles, d_ps, jc);
        }
        return Y.chk("jagged table");
    };
    // W and the round-0 message of <q, W> over LSB pairs
    if (!jag_table(X, true, W)) return flock::set_error(X.err, X.what.c_str());
    hipLaunchKernelGGL(flock::k_msg_blocked_comb, dim3(X.grid(N / 2)), dim3(kThreads), kCombWsBytes, X.st,
                       reinterpret_cast<const uint4*>(d_q), reinterpret_cast<const uint4*>(W), N, (uint64_t)1,
                       X.A->acc);
    if (!X.chk("jagged weight") || !X.read_acc(8)) return flock::set_error(X.err, X.what.c_str());
    ws->stamp(0);
    // the merged sumcheck <q, W> = target (pcs.rs:1640-1700), merged_round_bits = 2
    std::vector<H> mrounds, rho;
    std::vector<uint64_t> mnon;
    H q_eval = hz();
    {
        // round 0's message is the fused prime; later rounds come from the fold kernels, the
        // first of which reads q in place
        const flock_f128 m2[2] = {X.acc_f(0), X.acc_f(1)};
        H claim = target;
        if (!prod_rounds(X, fs, d_q, W, N, m2, claim, 2, true, s1a, s1b, s2a, s2b, mrounds, mnon, rho,
                         &q_eval))
            return flock::set_error(X.err, X.what.c_str());
    }
    if (h_q_eval) *h_q_eval = q_eval;
    ws->stamp(1);
    // fork the assist's transcript (challenger.rs:167-191)
    const H seed0 = sample(fs), seed1 = sample(fs);
    static const uint8_t kFork[] = "flock-par-assist-v1";
    flock_fs* ch = flock_fs_new(kFork, sizeof(kFork) - 1);
    if (!ch) return flock::set_error(FLOCK_HIP_E_NOMEM, "transcript");
    observe(ch, seed0);
    observe(ch, seed1);
    Bin mp;  // MultipointTwistedProof
    const int m = dense_log;
    // The multipoint evaluation, its twisted sumcheck and the Frobenius assist draw only on the
    // forked transcript ch, and the inner Ligerito opening only on fs (ch is merged back after
    // both), so the two run at once: this block on a second host thread with a context (stream,
    // reduction scratch) of its own, the opening on this one. Their workspace slots are disjoint.
    double t_mp = 0.0, t_fa = 0.0;
    auto multipoint = [&](Ctx& X) -> bool {
        bool ok = true;
        const auto t0 = std::chrono::steady_clock::now();
        // rho^(2^-j), j < 128
        std::vector<H> rp(128 * (size_t)m);
        for (int i = 0; i < m; i++) rp[i] = rho[i];
        for (int j = 1; j < 128; j++)
            for (int i = 0; i < m; i++) rp[(size_t)j * m + i] = frob_inv(rp[(size_t)(j - 1) * m + i]);
        const size_t nr = J.bc.size();
        std::vector<H> wall(n_claims * nr), zrow(n_claims * (size_t)n_log + 1);
        for (size_t c = 0; c < n_claims; c++) {
            const H* x = h_x_outers + c * x_len;
            const std::vector<H> w = run_weights(J, x + 1 + n_log);
            std::copy(w.begin(), w.end(), wall.begin() + c * nr);
            std::copy(x + 1, x + 1 + n_log, zrow.begin() + c * n_log);
        }
        uint64_t* d_bc = X.buf<uint64_t>(30, nr);
        uint64_t* d_bn = X.buf<uint64_t>(31, nr);
        flock_f128* d_w = X.buf<flock_f128>(32, wall.size());
        flock_f128* d_zr = X.buf<flock_f128>(33, zrow.size());
        flock_f128* d_xs = X.buf<flock_f128>(34, rp.size());
        flock_f128* d_vals = X.buf<flock_f128>(35, 128 * n_claims);
        std::vector<H> vals(128 * n_claims);
        if (X.err || !X.h2d(d_bc, J.bc.data(), 8 * nr) || !X.h2d(d_bn, J.bn.data(), 8 * nr) ||
            !X.h2d(d_w, wall.data(), 16 * wall.size()) || !X.h2d(d_zr, zrow.data(), 16 * zrow.size()) ||
            !X.h2d(d_xs, rp.data(), 16 * rp.size()) ||
            !X.hip(hipMemsetAsync(d_vals, 0, 16 * 128 * n_claims, X.st), "memset"))
            ok = false;
        if (ok) {
            flock::BpArgs ba{};
            ba.bc = d_bc;
            ba.bn = d_bn;
            ba.n_runs = (uint32_t)nr;
            ba.w = reinterpret_cast<const uint4*>(d_w);
            ba.zrow = reinterpret_cast<const uint4*>(d_zr);
            ba.xs = reinterpret_cast<const uint4*>(d_xs);
            ba.n = n_log;
            ba.m = (uint32_t)m;
            std::memcpy(ba.sparse, sparse_tr().t, sizeof(ba.sparse));
            hipLaunchKernelGGL(flock::k_bp_values, dim3(128 * (uint32_t)n_claims, (uint32_t)((nr + 31) / 32)),
                               dim3(kThreads), 0, X.st, ba,
                               reinterpret_cast<uint4*>(d_vals));
            ok = X.chk("bp values") && X.d2h(vals.data(), d_vals, 16 * vals.size());
        }
        if (ok) {
            label(ch, "flock-multipoint-twisted-v1");
            for (const H& v : vals) observe(ch, v);
            uint32_t gbits = bits_for_degree(128 * n_claims - 1);
            if (gbits < 1) gbits = 1;
            uint64_t gnonce;
            const H gmm = grind_sample(ch, gbits, &gnonce);
            std::vector<H> gpow(128 * n_claims);
            gpow[0] = h1();
            for (size_t i = 1; i < gpow.size(); i++) gpow[i] = hm(gpow[i - 1], gmm);
            // the claim sum_{c,j} gamma^(128c+j) A_cj = sum_d abar_d g_d
            H claim2 = hz();
            for (size_t i = 0; i < vals.size(); i++) claim2 = ha(claim2, hm(gpow[i], vals[i]));
            // g = L(eq(rho, .)), L(x) = sum_{j<128} gamma^j x^(2^-j)
            H img[128];
            for (int b = 0; b < 128; b++) {
                H acc = hz(), y = hbit(b);
                for (int j = 0; j < 128; j++) {
                    acc = ha(acc, hm(gpow[j], y));
                    y = frob_inv(y);
                }
                img[b] = acc;
            }
            const std::vector<H> lt = lin_tables(img);
            flock_f128* d_lt = X.buf<flock_f128>(36, 512);
            if (X.err || !X.h2d(d_lt, lt.data(), 16 * 512) || !X.eq_dev(rho.data(), (size_t)m, A)) ok = false;
            if (ok) {
                hipLaunchKernelGGL(flock::k_lin_apply, dim3(X.grid(N)), dim3(kThreads), 0, X.st,
                                   reinterpret_cast<uint4*>(A), N, reinterpret_cast<const uint4*>(d_lt));
                ok = X.chk("lin apply");
            }
            // abar = sum_c gamma^(128 c) eq_row_c (x) eq_col_c on the jagged area
            std::vector<H> ecs(n_claims * (size_t)n_cols);
            for (size_t c = 0; c < n_claims && ok; c++)
                for (uint32_t y = 0; y < n_cols; y++) ecs[c * n_cols + y] = hm(gpow[128 * c], ec_all[c * n_cols + y]);
            if (ok && X.h2d(d_ec, ecs.data(), 16 * ecs.size())) {
                ok = jag_table(X, false, W);
            } else {
                ok = false;
            }
            std::vector<H> trounds, point;
            std::vector<uint64_t> tnon;
            if (ok)
                ok = pair_sumcheck(X, W, A, s1a, s1b, s2a, s2b, N, (uint32_t)m, claim2, ch, 2, false,
                                   trounds, tnon, point, nullptr);
            // anchor: the untwisted assist at the endpoint (jagged.rs:3240-3285)
            if (ok) {
                H g_at = hz();
                for (int j = 0; j < 128; j++) g_at = ha(g_at, hm(gpow[j], heq_at(&rp[(size_t)j * m], point.data(), m)));
                std::vector<std::vector<H>> zr, ww;
                for (size_t c = 0; c < n_claims; c++) {
                    const H* x = h_x_outers + c * x_len;
                    const H co = hm(gpow[128 * c], g_at);
                    if (hzero(co)) continue;
                    std::vector<H> w = run_weights(J, x + 1 + n_log);
                    for (H& v : w) v = hm(v, co);
                    std::vector<H> z(x + 1, x + 1 + n_log);
                    size_t s = 0;
                    for (; s < zr.size(); s++)
                        if (std::memcmp(zr[s].data(), z.data(), 16 * n_log) == 0) break;
                    if (s < zr.size()) {
                        for (size_t b = 0; b < w.size(); b++) ww[s][b] = ha(ww[s][b], w[b]);
                    } else {
                        zr.push_back(z);
                        ww.push_back(w);
                    }
                }
                mp.u64(n_claims);
                for (size_t c = 0; c < n_claims; c++) mp.vec_f(&vals[128 * c], 128);
                mp.u64(0);  // group_values
                mp.u64(gnonce);
                mp.u64(trounds.size() / 2);
                for (const H& x : trounds) mp.f(x);
                mp.vec_u64(tnon.data(), tnon.size());
                const auto t1 = std::chrono::steady_clock::now();
                t_mp = std::chrono::duration<double, std::milli>(t1 - t0).count();
                frobenius_assist(ch, J, zr, ww, point, 2, mp);
                t_fa = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
            }
        }
        // the stream is idle when the thread ends
        return X.hip(hipStreamSynchronize(X.st), "assist sync") && ok;
    };
    bool ok = true;
    flock_hip_ctx* actx = ws->aux_ctx(flock::ctx_device(ctx));
    if (!actx) {
        flock_fs_free(ch);
        return flock::set_error(FLOCK_HIP_E_HIP, "open_merged: second context");
    }
    Ctx X2{actx, ws, static_cast<hipStream_t>(flock_hip_ctx_stream(actx)), &ws->accs[1]};
    bool ok_mp = false;
    std::thread th;
    // On large stacks the multipoint block starts once the inner opening's lane folds are
    // through: those are large GPU-bound passes, while the rest of the inner opening (its levels)
    // is mostly host round trips that leave the GPU idle, which the multipoint block's kernels
    // then fill. Small stacks are latency-bound throughout, and both start at once.
    Gate lanes_done;
    if (N < (1ull << 20)) lanes_done.release();
    try {
        th = std::thread([&] {
            lanes_done.wait();
            if (X2.hip(hipSetDevice(flock::ctx_device(actx)), "assist device"))
                ok_mp = X2.clear_acc() && multipoint(X2);
        });
    } catch (...) {
        ok_mp = X2.clear_acc() && multipoint(X2);  // no thread: one after the other
    }
    Bin inner;
    ok = inner_open(X, fs, d_q, P, d_codeword, d_tree, rho, q_eval, inner, th.joinable() ? &lanes_done : nullptr);
    lanes_done.release();  // also when the inner opening stopped early
    ws->stamp(4);
    if (th.joinable()) th.join();
    ws->stage_ms[2] = t_mp;
    ws->stage_ms[3] = t_fa;
    if (ok && !ok_mp) {
        X.err = X2.err ? X2.err : FLOCK_HIP_E_INTERNAL;
        X.what = X2.what;
        ok = false;
    }
    // merge the child (challenger.rs:193-203)
    const H d0 = sample(ch), d1 = sample(ch);
    observe(fs, d0);
    observe(fs, d1);
    flock_fs_free(ch);
    if (!ok) return flock::set_error(X.err ? X.err : FLOCK_HIP_E_INTERNAL, X.what.c_str());
    // MergedOpenProof (pcs.rs:1362-1377)
    out.u64(n_claims);
    for (size_t c = 0; c < n_claims; c++) {
        out.vec_f(h_s_hat_v + 128 * c, 128);
        out.u64(rs_nonce[c]);
    }
    out.vec_u64(&batch_nonce, bb ? 1 : 0);
    out.u64(mrounds.size() / 2);
    for (const H& x : mrounds) out.f(x);
    out.vec_u64(mnon.data(), mnon.size());
    out.f(q_eval);
    out.raw(mp.b.data(), mp.b.size());
    out.raw(inner.b.data(), inner.b.size());
    uint8_t* p = static_cast<uint8_t*>(std::malloc(out.b.size()));
    if (!p) return flock::set_error(FLOCK_HIP_E_NOMEM, "proof bytes");
    std::memcpy(p, out.b.data(), out.b.size());
    *out_bytes = p;
    *out_len = out.b.size();
    ws->stamp(5);
    return 0;
}

// ---------------------------------------------------------------------------
// The merged opening's verifier (pcs.rs:1971-2300): ring switch, merged sumcheck replay,
// multipoint twisted evaluation with its Frobenius anchor (pcs/jagged.rs:2034-2230,
// 4335-4585) on the forked transcript, and the succinct Ligerito verifier of q^(rho)
// (pcs.rs:1129-1320, pcs/ligerito/extension.rs:2409-2880). Host only.
// ---------------------------------------------------------------------------
namespace {

// bincode 1.3 reader; an overrun or a length larger than the bytes left marks it bad
struct Rd {
    const uint8_t* p;
    size_t n, off = 0;
    bool ok = true;
    bool need(size_t k) {
        if (!ok || k > n - off) ok = false;
        return ok;
    }
    uint64_t u64() {
        if (!need(8)) return 0;
        uint64_t v = 0;
        for (int i = 0; i < 8; i++) v |= (uint64_t)p[off + i] << (8 * i);
        off += 8;
        return v;
    }
    H f() {
        H v;
        v.lo = u64();
        v.hi = u64();
        return v;
    }
    E e() {
        E v;
        v.c0 = f();
        v.c1 = f();
        return v;
    }
    size_t len(size_t elem) {
        const uint64_t k = u64();
        if (ok && k > (n - off) / (elem ? elem : 1)) ok = false;
        return ok ? (size_t)k : 0;
    }
    std::vector<H> vec_f() {
        std::vector<H> v(len(16));
        for (H& x : v) x = f();
        return v;
    }
    std::vector<uint64_t> vec_u64() {
        std::vector<uint64_t> v(len(8));
        for (uint64_t& x : v) x = u64();
        return v;
    }
    std::vector<uint8_t> vec_hash() {
        const size_t k = len(32);
        std::vector<uint8_t> v;
        if (need(32 * k)) {
            v.assign(p + off, p + off + 32 * k);
            off += 32 * k;
        }
        return v;
    }
    std::vector<H> pairs() {  // Vec<(F128, F128)> flattened
        std::vector<H> v(2 * len(32));
        for (H& x : v) x = f();
        return v;
    }
};

// RecursiveProof {opened_rows: Vec<Vec<F128>>, merkle_proof: Vec<Hash>} (ligerito.rs:2160-2176)
struct Opened {
    std::vector<std::vector<H>> rows;
    std::vector<uint8_t> paths;
};
Opened rd_opened(Rd& r) {
    Opened o;
    o.rows.resize(r.len(8));
    for (auto& row : o.rows) row = r.vec_f();
    o.paths = r.vec_hash();
    return o;
}

// LigeritoProof (ligerito.rs:2177-2218) of the basis / F256 route
struct LigProof {
    std::vector<uint8_t> initial_cap;
    Opened initial;
    std::vector<std::vector<uint8_t>> caps;
    std::vector<Opened> recursive;
    std::vector<H> yr;
    Opened fin;
    size_t n_sc128 = 0;
    std::vector<E> msgs;  // (u0, u2) per message
    std::vector<uint64_t> grind_n, fold_n, claim_n, cons_n;
    std::vector<H> ood;
};
LigProof rd_lig(Rd& r) {
    LigProof L;
    L.initial_cap = r.vec_hash();
    L.initial = rd_opened(r);
    L.caps.resize(r.len(8));
    for (auto& c : L.caps) c = r.vec_hash();
    L.recursive.resize(r.len(16));
    for (auto& o : L.recursive) o = rd_opened(r);
    L.yr = r.vec_f();
    L.fin = rd_opened(r);
    L.n_sc128 = r.len(32);
    for (size_t i = 0; i < L.n_sc128; i++) r.f(), r.f();
    L.msgs.resize(2 * r.len(64));
    for (E& x : L.msgs) x = r.e();
    L.grind_n = r.vec_u64();
    L.ood = r.vec_f();
    L.fold_n = r.vec_u64();
    L.claim_n = r.vec_u64();
    L.cons_n = r.vec_u64();
    return L;
}

// RoundQuad256 (extension.rs:71-97): q(X) = c + b X + a X^2 from (u0, u2) and the claim
struct Quad {
    E c, b, a;
    static Quad from(E u0, E u2, E claim) { return Quad{u0, ea(claim, u2), u2}; }
    E eval(E r) const { return ea(ea(c, em(r, b)), em(em(r, r), a)); }
    Quad fold(const Quad& o, H alpha) const {
        return Quad{ea(c, emb(o.c, alpha)), ea(b, emb(o.b, alpha)), ea(a, emb(o.a, alpha))};
    }
};
inline E eh(H x) { return E{x, hz()}; }
inline bool eeq(E a, E b) { return std::memcmp(&a, &b, sizeof(E)) == 0; }
inline bool heqv(H a, H b) { return a.lo == b.lo && a.hi == b.hi; }

// verify_pow then sample (challenger.rs verify_pow_and_sample_f128{,_vec})
bool vps(flock_fs* fs, uint64_t nonce, uint32_t bits, size_t n, std::vector<H>* out) {
    if (!flock_fs_verify_pow(fs, nonce, bits)) return false;
    *out = sample_vec(fs, n);
    return true;
}
bool vps1(flock_fs* fs, uint64_t nonce, uint32_t bits, H* out) {
    if (!flock_fs_verify_pow(fs, nonce, bits)) return false;
    *out = sample(fs);
    return true;
}

// verify_level_opens (ligerito.rs:5152-5220) with verify_merkle_proof_capped (merkle.rs:
// 698-732): SHA-256 leaves over the rows' bytes, capped paths in sample order
bool level_opens_ok(const std::vector<uint8_t>& cap, uint64_t block_len,
                    const std::vector<uint64_t>& qs, const Opened& o, size_t lanes, const Sched& s) {
    const int d = __builtin_ctzll(block_len);
    if (qs.size() != o.rows.size() || s.d != d || s.queries() != qs.size()) return false;
    const int c1 = s.cap();
    if (cap.size() != (32ull << c1) || o.paths.size() != 32 * qs.size() * (size_t)(d - c1))
        return false;
    size_t off = 0;
    for (size_t i = 0; i < qs.size(); i++) {
        if (o.rows[i].size() != lanes || qs[i] >= block_len) return false;
        uint8_t acc[32], pair[64];
        flock::sha256_host(o.rows[i].data(), 16 * lanes, acc);
        uint64_t idx = qs[i];
        for (int l = 0; l < d - c1; l++, off += 32) {
            if (idx & 1) {
                std::memcpy(pair, &o.paths[off], 32);
                std::memcpy(pair + 32, acc, 32);
            } else {
                std::memcpy(pair, acc, 32);
                std::memcpy(pair + 32, &o.paths[off], 32);
            }
            flock::sha256_host(pair, 64, acc);
            idx >>= 1;
        }
        if (std::memcmp(acc, &cap[32 * idx], 32) != 0) return false;
    }
    return true;
}

// induce_enforced_sum (extension.rs:2885-2910)
E enforced_sum(const std::vector<std::vector<H>>& rows, const std::vector<E>& lane_ch,
               const std::vector<H>& alpha, bool* ok) {
    const std::vector<E> lw = heq256(lane_ch.data(), lane_ch.size());
    const std::vector<H> rw = heq(alpha.data(), alpha.size());
    E acc = ez();
    for (size_t q = 0; q < rows.size(); q++) {
        if (rows[q].size() > lw.size() || q >= rw.size()) {
            *ok = false;
            return acc;
        }
        E s = ez();
        for (size_t i = 0; i < rows[q].size(); i++) s = ea(s, emb(lw[i], rows[q][i]));
        acc = ea(acc, emb(s, rw[q]));
    }
    return acc;
}

const E kU{hz(), h1()};  // u, the F256 generator over F128
// coordinate_fold_factor (extension.rs:2291-2293): 1 + c (1 + u)
inline E coord_factor(E c) { return ea(e1(), em(c, E{h1(), h1()})); }

// eval_sk_at_vks (ligerito.rs:2402-2420)
std::vector<H> sk_at_vk(int log_n) {
    std::vector<H> skv(log_n + 1, hz()), layer(log_n > 0 ? log_n : 1);
    skv[0] = h1();
    for (int i = 1; i <= log_n; i++) layer[i - 1] = H{1ull << i, 0};
    for (int i = 0, cur = log_n; i < log_n; i++, cur--)
        for (int j = 0; j < cur; j++) {
            const H nx = ha(hm(layer[j], layer[j]), hm(skv[i], layer[j]));
            if (j == 0)
                skv[i + 1] = nx;
            else
                layer[j - 1] = nx;
        }
    return skv;
}

// induced_basis_at_residual (extension.rs:2295-2343)
std::vector<E> induced_at_residual(int log_cols, const std::vector<uint64_t>& qs,
                                   const std::vector<H>& alpha, const std::vector<E>& fixed,
                                   int res_log) {
    const std::vector<H> skv = sk_at_vk(log_cols);
    std::vector<H> inv(skv.size());
    for (size_t i = 0; i < skv.size(); i++) inv[i] = hzero(skv[i]) ? hz() : hinv(skv[i]);
    const std::vector<H> wts = heq(alpha.data(), alpha.size());
    std::vector<E> out((size_t)1 << res_log, ez());
    for (size_t qi = 0; qi < qs.size(); qi++) {
        std::vector<H> w(log_cols, hz());
        if (log_cols > 0) {
            w[0] = H{qs[qi], 0};
            for (int k = 1; k < log_cols; k++) w[k] = ha(hm(w[k - 1], w[k - 1]), hm(skv[k - 1], w[k - 1]));
            for (int k = 0; k < log_cols; k++) w[k] = hm(w[k], inv[k]);
        }
        E pre = e1();
        for (size_t j = 0; j < fixed.size(); j++) pre = em(pre, ea(e1(), emb(fixed[j], ha(h1(), w[j]))));
        const E pw = emb(pre, wts[qi]);
        for (size_t y = 0; y < out.size(); y++) {
            E t = pw;
            for (int j = 0; j < res_log; j++)
                if ((y >> j) & 1) t = emb(t, w[fixed.size() + j]);
            out[y] = ea(out[y], t);
        }
    }
    return out;
}

// eq_residual (extension.rs:2386-2404)
std::vector<E> eq_residual(const std::vector<H>& pt, const std::vector<E>& fixed, int res_log, E scale) {
    E pre = scale;
    for (size_t j = 0; j < fixed.size(); j++) pre = em(pre, ea(ea(e1(), eh(pt[j])), fixed[j]));
    std::vector<E> out((size_t)1 << res_log);
    for (size_t y = 0; y < out.size(); y++) {
        E t = pre;
        for (int j = 0; j < res_log; j++) {
            const H z = pt[fixed.size() + j];
            t = emb(t, ((y >> j) & 1) ? z : ha(h1(), z));
        }
        out[y] = t;
    }
    return out;
}

// the original challenges from start_level on (residual_original_challenges, :2360-2378)
std::vector<E> originals(const std::vector<E>& initial, const std::vector<std::vector<E>>& levels,
                         size_t start) {
    std::vector<E> out;
    if (start == 0) out = initial;
    for (size_t l = start; l < levels.size(); l++) out.insert(out.end(), levels[l].begin() + 1, levels[l].end());
    return out;
}
E coord_product(const std::vector<std::vector<E>>& levels, size_t start) {
    E acc = e1();
    for (size_t l = start; l < levels.size(); l++) acc = em(acc, coord_factor(levels[l][0]));
    return acc;
}

// recursive_verifier_with_basis_succinct (extension.rs:2409-2880) for one packed-direct
// claim (point rho, batched by gamma), the shape verify_opening_batch_ligerito_mixed_with_
// grinding drives for the merged opening's inner claim (pcs.rs:1150-1320)
bool lig_verify(flock_fs* fs, const Lig* C, const LigProof& P, int log_n, H target,
                const std::vector<uint8_t>& cap, size_t l0_lanes, const std::vector<H>& rho,
                H gamma, std::string* why) {
    const int ik = C->initial_k, R = C->n_levels - 1;
    auto fail = [&](const char* w) {
        *why = w;
        return false;
    };
    if (R < 1 || P.initial_cap != cap || P.n_sc128 != 0 || P.msgs.empty() || !P.fold_n.empty())
        return fail("Ligerito proof shape");
    for (int i = 0; i < R; i++)
        if (C->lv[i + 1].lanes < 2) return fail("Ligerito config");
    std::vector<Sched> sch;
    for (int l = 0; l <= R; l++) sch.push_back(sched_of(C->lv[l].queries, C->lv[l].cols + C->lv[l].rate));
    label(fs, "flock-ligerito-basis-f256-split-v0");
    observe(fs, target);
    flock_fs_observe_bytes(fs, P.initial_cap.data(), P.initial_cap.size());
    const int log_cols_0 = log_n - ik;
    const uint64_t block_len_0 = 1ull << (log_cols_0 + C->lv[0].rate);
    size_t tx = 0, oi = 0, cn = 0, kn = 0, qn = 0;
    E claim = eh(target);
    struct Ood {
        std::vector<H> point;
        H beta;
        int split;  // -1: L0; s: consumed by recursive level s
    };
    struct Cons {
        int log_cols;
        std::vector<uint64_t> qs;
        std::vector<H> alpha;
        H beta;
        size_t start;
    };
    std::vector<Ood> oods;
    std::vector<Cons> cons;
    const bool lane_major = l0_lanes < (1ull << ik);
    auto next_msg = [&](E* u0, E* u2) {
        if (2 * tx + 1 >= P.msgs.size()) return false;
        *u0 = P.msgs[2 * tx];
        *u2 = P.msgs[2 * tx + 1];
        tx++;
        observe_e(fs, *u0);
        observe_e(fs, *u2);
        return true;
    };
    for (int o = 0; o < C->lv[0].ood; o++) {
        std::vector<H> pt = sample_vec(fs, (size_t)log_n);
        if (lane_major) std::rotate(pt.begin(), pt.begin() + (log_n - ik), pt.end());
        if (oi >= P.ood.size() || cn >= P.claim_n.size()) return fail("OOD values");
        const H y = P.ood[oi++];
        observe(fs, y);
        H beta;
        if (!vps1(fs, P.claim_n[cn++], C->lv[0].claim, &beta)) return fail("claim-batch proof of work");
        claim = ea(claim, eh(hm(beta, y)));
        oods.push_back({pt, beta, -1});
    }
    E u0, u2;
    if (!next_msg(&u0, &u2)) return fail("sumcheck messages");
    Quad quad = Quad::from(u0, u2, claim);
    std::vector<E> initial;
    for (int j = 0; j < ik; j++) {
        const E ch = sample_e(fs);
        claim = quad.eval(ch);
        initial.push_back(ch);
        if (!next_msg(&u0, &u2)) return fail("sumcheck messages");
        quad = Quad::from(u0, u2, claim);
    }
    int split_dim = log_n - ik + 1;
    if (P.caps.empty()) return fail("recursive caps");
    flock_fs_observe_bytes(fs, P.caps[0].data(), P.caps[0].size());
    auto level_ood = [&](int lvl, int split) -> bool {
        for (int o = 0; o < C->lv[lvl].ood; o++) {
            std::vector<H> pt = sample_vec(fs, (size_t)split_dim);
            if (oi >= P.ood.size() || cn >= P.claim_n.size()) return fail("OOD values");
            const H y = P.ood[oi++];
            observe(fs, y);
            E a0, a2;
            if (!next_msg(&a0, &a2)) return fail("sumcheck messages");
            const Quad intro = Quad::from(a0, a2, eh(y));
            H beta;
            if (!vps1(fs, P.claim_n[cn++], C->lv[lvl].claim, &beta))
                return fail("claim-batch proof of work");
            quad = quad.fold(intro, beta);
            claim = ea(claim, eh(hm(beta, y)));
            oods.push_back({pt, beta, split});
        }
        return true;
    };
    auto queries = [&](int lvl, uint64_t block_len, std::vector<uint64_t>* qs, std::vector<H>* alpha) {
        const size_t nq = C->lv[lvl].queries;
        std::vector<H> w;
        if (qn >= P.grind_n.size() || !vps(fs, P.grind_n[qn++], C->lv[lvl].grind, nq, &w))
            return fail("query proof of work");
        (void)block_len;
        *qs = queries_from(sch[lvl], w);
        if (kn >= P.cons_n.size() || !vps(fs, P.cons_n[kn++], C->lv[lvl].cons, ceil_log2(nq), alpha))
            return fail("consistency proof of work");
        return true;
    };
    if (!level_ood(1, 0)) return false;
    std::vector<uint64_t> qs;
    std::vector<H> alpha;
    if (!queries(0, block_len_0, &qs, &alpha)) return false;
    if (!level_opens_ok(P.initial_cap, block_len_0, qs, P.initial, l0_lanes, sch[0]))
        return fail("L0 Merkle openings");
    bool ok = true;
    E enforced = enforced_sum(P.initial.rows, initial, alpha, &ok);
    if (!ok) return fail("L0 opened rows");
    if (!next_msg(&u0, &u2)) return fail("sumcheck messages");
    {
        const Quad intro = Quad::from(u0, u2, enforced);
        H beta;
        if (cn >= P.claim_n.size() || !vps1(fs, P.claim_n[cn++], C->lv[1].claim, &beta))
            return fail("claim-batch proof of work");
        quad = quad.fold(intro, beta);
        claim = ea(claim, emb(enforced, beta));
        cons.push_back({log_n - ik, qs, alpha, beta, 0});
    }
    std::vector<std::vector<E>> levels;
    const std::vector<uint8_t>* prev_cap = &P.caps[0];
    int prev_lanes = C->lv[1].lanes;
    if (split_dim < prev_lanes) return fail("Ligerito config");
    int prev_cols = split_dim - prev_lanes, prev_rate = C->lv[1].rate;
    size_t cap_i = 1, rec_i = 0;
    for (int i = 0; i < R; i++) {
        const int k = C->lv[i + 1].lanes;
        if (split_dim < k) return fail("Ligerito config");
        std::vector<E> chs;
        for (int j = 0; j < k; j++) {
            const E ch = sample_e(fs);
            claim = quad.eval(ch);
            chs.push_back(ch);
            if (!next_msg(&u0, &u2)) return fail("sumcheck messages");
            quad = Quad::from(u0, u2, claim);
        }
        levels.push_back(chs);
        const int ext = split_dim - k, level = i + 1;
        const uint64_t block_len = 1ull << (prev_cols + prev_rate);
        if (i + 1 == R) {
            if (tx != P.msgs.size() / 2 || P.yr.size() != (2ull << ext)) return fail("final shape");
            for (const H& y : P.yr) observe(fs, y);
            if (!queries(level, block_len, &qs, &alpha)) return false;
            if (!level_opens_ok(*prev_cap, block_len, qs, P.fin, 1ull << prev_lanes, sch[level]))
                return fail("final Merkle openings");
            enforced = enforced_sum(P.fin.rows, levels[i], alpha, &ok);
            if (!ok) return fail("final opened rows");
            H beta;
            if (cn >= P.claim_n.size() || !vps1(fs, P.claim_n[cn++], C->lv[level].claim, &beta))
                return fail("claim-batch proof of work");
            claim = ea(claim, emb(enforced, beta));
            cons.push_back({ext, qs, alpha, beta, (size_t)R});
            const std::vector<E> orig = originals(initial, levels, 0);
            if ((int)orig.size() + ext != log_n) return fail("residual dimension");
            // the basis at the residual: the packed-direct claim's eq (pcs.rs:1214-1304), the
            // OOD eqs, the consistency bases, each with its coordinate fold factors
            std::vector<H> pt(rho);
            if (lane_major) std::rotate(pt.begin(), pt.begin() + (pt.size() - ik), pt.end());
            std::vector<E> residual = eq_residual(pt, orig, ext, emb(e1(), gamma));
            const E s0 = coord_product(levels, 0);
            for (E& v : residual) v = em(v, s0);
            for (const Ood& o : oods) {
                std::vector<E> fixed;
                E cs;
                if (o.split < 0) {
                    fixed = orig;
                    cs = coord_product(levels, 0);
                } else {
                    fixed = levels[o.split];
                    for (size_t l = o.split + 1; l < levels.size(); l++)
                        fixed.insert(fixed.end(), levels[l].begin() + 1, levels[l].end());
                    cs = coord_product(levels, (size_t)o.split + 1);
                }
                if (fixed.size() + ext != o.point.size()) return fail("OOD point dimension");
                const std::vector<E> v = eq_residual(o.point, fixed, ext, emb(cs, o.beta));
                for (size_t y = 0; y < residual.size(); y++) residual[y] = ea(residual[y], v[y]);
            }
            for (const Cons& c : cons) {
                const std::vector<E> fixed = originals({}, levels, c.start);
                if ((int)fixed.size() + ext != c.log_cols) return fail("consistency dimension");
                const std::vector<E> v = induced_at_residual(c.log_cols, c.qs, c.alpha, fixed, ext);
                const E sc = emb(coord_product(levels, c.start), c.beta);
                for (size_t y = 0; y < residual.size(); y++) residual[y] = ea(residual[y], em(v[y], sc));
            }
            // split_basis + split_inner_product against yr (extension.rs:32-55)
            E rc = ez();
            for (size_t y = 0; y < residual.size(); y++) {
                rc = ea(rc, emb(residual[y], P.yr[2 * y]));
                rc = ea(rc, emb(em(kU, residual[y]), P.yr[2 * y + 1]));
            }
            if (!eeq(rc, claim)) return fail("residual claim");
            if (oi != P.ood.size() || qn != P.grind_n.size() || cn != P.claim_n.size() ||
                kn != P.cons_n.size() || cap_i != P.caps.size() || rec_i != P.recursive.size())
                return fail("unused proof entries");
            return true;
        }
        split_dim = ext + 1;
        if (cap_i >= P.caps.size()) return fail("recursive caps");
        const std::vector<uint8_t>* cap_n = &P.caps[cap_i++];
        flock_fs_observe_bytes(fs, cap_n->data(), cap_n->size());
        if (!level_ood(level + 1, i + 1)) return false;
        if (!queries(level, block_len, &qs, &alpha)) return false;
        if (rec_i >= P.recursive.size()) return fail("recursive openings");
        const Opened& op = P.recursive[rec_i++];
        if (!level_opens_ok(*prev_cap, block_len, qs, op, 1ull << prev_lanes, sch[level]))
            return fail("recursive Merkle openings");
        enforced = enforced_sum(op.rows, levels[i], alpha, &ok);
        if (!ok) return fail("recursive opened rows");
        if (!next_msg(&u0, &u2)) return fail("sumcheck messages");
        const Quad intro = Quad::from(u0, u2, enforced);
        H beta;
        if (cn >= P.claim_n.size() || !vps1(fs, P.claim_n[cn++], C->lv[level + 1].claim, &beta))
            return fail("claim-batch proof of work");
        quad = quad.fold(intro, beta);
        claim = ea(claim, emb(enforced, beta));
        cons.push_back({ext, qs, alpha, beta, (size_t)i + 1});
        prev_cap = cap_n;
        prev_lanes = C->lv[i + 2].lanes;
        if (split_dim < prev_lanes) return fail("Ligerito config");
        prev_cols = split_dim - prev_lanes;
        prev_rate = C->lv[i + 2].rate;
    }
    return fail("Ligerito rounds");
}

// the trace-dual basis of the bit basis, raised to every 2^j: D[j][b] = delta_b^(2^j), with
// bit_b(x) = Tr(delta_b x); then an F2-linear map with basis images g has the linearized
// coefficients c_j = sum_b g[b] D[j][b] (ring_switch linearized_coefficients)
const std::vector<H>& dual_powers() {
    static const std::vector<H> D = [] {
        // Tr(x^e) for e < 255
        std::vector<int> tr(255);
        for (int e = 0; e < 255; e++) {
            H y = h1();
            for (int i = 0; i < e; i++) y = hm(y, H{2, 0});
            H acc = hz(), s = y;
            for (int k = 0; k < 128; k++) {
                acc = ha(acc, s);
                s = hm(s, s);
            }
            tr[e] = (int)(acc.lo & 1);
        }
        // invert T[a][b] = Tr(x^(a+b)) over F2 (rows as 128-bit masks)
        std::vector<H> T(128), I(128);
        for (int a = 0; a < 128; a++) {
            T[a] = hz();
            for (int b = 0; b < 128; b++)
                if (tr[a + b]) T[a] = ha(T[a], hbit(b));
            I[a] = hbit(a);
        }
        for (int col = 0; col < 128; col++) {
            int piv = col;
            while (piv < 128 && !hgetbit(T[piv], col)) piv++;
            if (piv == 128) return std::vector<H>();  // singular (cannot happen: Tr is nondegenerate)
            std::swap(T[col], T[piv]);
            std::swap(I[col], I[piv]);
            for (int r = 0; r < 128; r++)
                if (r != col && hgetbit(T[r], col)) {
                    T[r] = ha(T[r], T[col]);
                    I[r] = ha(I[r], I[col]);
                }
        }
        // delta_a = sum_b Tinv[a][b] x^b = I[a] as an element; then its 2^j powers
        std::vector<H> out(128 * 128);
        for (int b = 0; b < 128; b++) {
            H v = I[b];
            for (int j = 0; j < 128; j++) {
                out[(size_t)j * 128 + b] = v;
                v = hm(v, v);
            }
        }
        return out;
    }();
    return D;
}

}  // namespace

int flock_open_merged_verify(flock_fs* fs, const flock_pcs_params* P, const uint8_t* cap, size_t cap_len,
                             size_t n_claims, const flock_f128* h_values, const flock_f128* h_z_skips,
                             const flock_f128* h_x_outers, size_t x_len, const uint64_t* h_heights,
                             size_t n_heights, uint32_t n_log, const uint8_t* proof, size_t proof_len,
                             size_t* consumed) {
    if (!fs || !P || !cap || !h_values || !h_z_skips || !h_x_outers || !h_heights || !proof || n_claims < 1 ||
        n_claims > flock::kMaxClaims || n_heights == 0 || (n_heights & (n_heights - 1)) || P->m < 14 ||
        P->m > 40 || x_len < 1)
        return flock::set_error(FLOCK_HIP_E_INVALID, "open_merged_verify: bad arguments");
    if (P->merkle_hash != 0)
        return flock::set_error(FLOCK_HIP_E_INVALID, "open_merged_verify: SHA-256 Merkle commitments only");
    const Lig* C = lig_for(P->m);
    if (!C || C->initial_k != P->log_batch_size || P->log_inv_rate != 1)
        return flock::set_error(FLOCK_HIP_E_INVALID, "open_merged_verify: no Ligerito config for this m");
    const int dense_log = (int)P->m - 7;
    const Jag J = jag_make(h_heights, n_heights, (int)n_log, dense_log);
    if (x_len != 1 + n_log + (size_t)J.k || J.area() > (1ull << dense_log))
        return flock::set_error(FLOCK_HIP_E_INVALID, "open_merged_verify: point / heights mismatch");
    const size_t lanes0 = P->num_lanes ? P->num_lanes : (1ull << P->log_batch_size);
    auto reject = [](const std::string& w) {
        return flock::set_error(FLOCK_HIP_E_REJECT, ("proof rejected: " + w).c_str());
    };
    // MergedOpenProof (pcs.rs:1362-1377)
    Rd r{proof, proof_len};
    const size_t n_rs = r.len(16 * 128 + 16);
    std::vector<std::vector<H>> shv(n_rs);
    std::vector<uint64_t> rs_nonce(n_rs);
    for (size_t c = 0; c < n_rs; c++) {
        shv[c] = r.vec_f();
        rs_nonce[c] = r.u64();
    }
    const std::vector<uint64_t> batch_n = r.vec_u64();
    const std::vector<H> mrounds = r.pairs();
    const std::vector<uint64_t> mnon = r.vec_u64();
    const H q_eval = r.f();
    std::vector<std::vector<H>> vals(r.len(8));
    for (auto& v : vals) v = r.vec_f();
    const std::vector<H> gvals = r.vec_f();
    const uint64_t gnonce = r.u64();
    const std::vector<H> trounds = r.pairs();
    const std::vector<uint64_t> tnon = r.vec_u64();
    const H av = r.f();
    const std::vector<H> arounds = r.pairs();
    const std::vector<uint64_t> anon = r.vec_u64();
    const size_t inner_rs = r.len(16), inner_bn = r.len(8);
    LigProof L = rd_lig(r);
    if (!r.ok) return reject("malformed merged opening bytes");
    const int m = dense_log;
    if (n_rs != n_claims || inner_rs != 0 || inner_bn != 0 || mrounds.size() != 2ull * m ||
        mnon.size() != (size_t)m || vals.size() != n_claims || !gvals.empty() ||
        trounds.size() != 2ull * m || tnon.size() != (size_t)m || arounds.size() != 4ull * (m + 1) ||
        anon.size() != 2ull * (m + 1))
        return reject("merged opening shape");
    for (size_t c = 0; c < n_claims; c++)
        if (shv[c].size() != 128 || vals[c].size() != 128) return reject("merged opening shape");
    const uint32_t bb = n_claims > 1 ? 6 : 0;
    if (batch_n.size() != (bb ? 1u : 0u)) return reject("batching nonces");
    label(fs, "flock-merged-open-v1");
    // 1. ring switch per claim (ring_switch.rs:2966-3005), 3 PoW bits
    std::vector<std::vector<H>> eqr(n_claims);
    std::vector<H> sc_claim(n_claims);
    for (size_t c = 0; c < n_claims; c++) {
        const H* x = h_x_outers + c * x_len;
        label(fs, "flock-ring-switch-v0");
        flock_fs_observe_f128_slice(fs, shv[c].data(), 128);
        H lam[64];
        flock::lagrange_phi8_host(0, h_z_skips[c], lam);
        const H e0 = ha(h1(), x[0]);
        H chk = hz();
        for (int i = 0; i < 128; i++) chk = ha(chk, hm(hm(lam[i & 63], (i >> 6) ? x[0] : e0), shv[c][i]));
        if (!heqv(chk, h_values[c])) return reject("ring-switch claim");
        std::vector<H> rdp;
        if (!vps(fs, rs_nonce[c], 3, 7, &rdp)) return reject("ring-switch proof of work");
        eqr[c] = heq(rdp.data(), 7);
        H acc = hz();
        for (int bt = 0; bt < 128; bt++) {
            H u = hz();
            for (int i = 0; i < 128; i++)
                if (hgetbit(shv[c][i], bt)) u = ha(u, hbit(i));
            acc = ha(acc, hm(u, eqr[c][bt]));
        }
        sc_claim[c] = acc;
    }
    // 2. batching coefficients (no PoW, nothing absorbed, for one claim)
    std::vector<H> gam;
    if (bb) {
        if (!vps(fs, batch_n[0], bb, n_claims, &gam)) return reject("batching proof of work");
    } else {
        gam = sample_vec(fs, n_claims);
    }
    H running = hz();
    for (size_t c = 0; c < n_claims; c++) running = ha(running, hm(gam[c], sc_claim[c]));
    // 3. the dense merged sumcheck replay, 2 PoW bits a round
    std::vector<H> rho;
    for (int i = 0; i < m; i++) {
        observe(fs, mrounds[2 * i]);
        observe(fs, mrounds[2 * i + 1]);
        H ch;
        if (!vps1(fs, mnon[i], 2, &ch)) return reject("merged sumcheck proof of work");
        running = round_claim(running, mrounds[2 * i], mrounds[2 * i + 1], ch);
        rho.push_back(ch);
    }
    // 4. the multipoint twisted evaluation on the forked transcript (challenger.rs:167-191)
    const H seed0 = sample(fs), seed1 = sample(fs);
    static const uint8_t kFork[] = "flock-par-assist-v1";
    flock_fs* ch = flock_fs_new(kFork, sizeof(kFork) - 1);
    if (!ch) return flock::set_error(FLOCK_HIP_E_NOMEM, "transcript");
    observe(ch, seed0);
    observe(ch, seed1);
    std::string why;
    H V = hz();
    {
        label(ch, "flock-multipoint-twisted-v1");
        for (const auto& v : vals)
            for (const H& x : v) observe(ch, x);
        uint32_t gbits = bits_for_degree(128 * n_claims - 1);
        if (gbits < 1) gbits = 1;
        H gmm;
        std::vector<H> point;
        H mrun = hz(), g_at = hz();
        std::vector<H> gpow(128 * n_claims);
        if (!vps1(ch, gnonce, gbits, &gmm)) {
            why = "multipoint gamma proof of work";
        } else {
            gpow[0] = h1();
            for (size_t i = 1; i < gpow.size(); i++) gpow[i] = hm(gpow[i - 1], gmm);
            for (size_t c = 0; c < n_claims; c++)
                for (int j = 0; j < 128; j++) mrun = ha(mrun, hm(gpow[128 * c + j], vals[c][j]));
            for (int i = 0; i < m && why.empty(); i++) {
                observe(ch, trounds[2 * i]);
                observe(ch, trounds[2 * i + 1]);
                H rr;
                if (!vps1(ch, tnon[i], 2, &rr)) why = "multipoint round proof of work";
                mrun = round_claim(mrun, trounds[2 * i], trounds[2 * i + 1], rr);
                point.push_back(rr);
            }
        }
        if (why.empty()) {
            // g_at = sum_j gamma^j eq(rho^(2^-j), point) (twisted_eq_at)
            std::vector<H> rj(rho);
            for (int j = 0; j < 128; j++) {
                g_at = ha(g_at, hm(gpow[j], heq_at(rj.data(), point.data(), m)));
                for (H& x : rj) x = frob_inv(x);
            }
            // the anchor: the Frobenius assist (untwisted) at the endpoint
            label(ch, "flock-frobenius-assist-v0");
            observe(ch, av);
            H aclaim = av;
            std::vector<H> sigma;
            for (int i = 0; i < 2 * (m + 1) && why.empty(); i++) {
                observe(ch, arounds[2 * i]);
                observe(ch, arounds[2 * i + 1]);
                H rr;
                if (!vps1(ch, anon[i], 2, &rr)) why = "assist round proof of work";
                aclaim = round_claim(aclaim, arounds[2 * i], arounds[2 * i + 1], rr);
                sigma.push_back(rr);
            }
            if (why.empty()) {
                // per layer l, e_l[cd] = eq((c, d), (rc_l, rd_l))
                std::vector<H> el(4 * (size_t)(m + 1));
                for (int l = 0; l <= m; l++) {
                    const H rc = sigma[2 * l], rd = sigma[2 * l + 1];
                    const H rc1 = ha(h1(), rc), rd1 = ha(h1(), rd);
                    el[4 * l + 0] = hm(rc1, rd1);
                    el[4 * l + 1] = hm(rc, rd1);
                    el[4 * l + 2] = hm(rc1, rd);
                    el[4 * l + 3] = hm(rc, rd);
                }
                // eq over each boundary run's (c, d) bits
                std::vector<H> ecd(J.bc.size());
                for (size_t b = 0; b < J.bc.size(); b++) {
                    H acc = h1();
                    for (int l = 0; l <= m; l++) acc = hm(acc, el[4 * l + col_cd(J.bc[b], J.bn[b], l)]);
                    ecd[b] = acc;
                }
                H expect = hz();
                for (size_t c = 0; c < n_claims; c++) {
                    const H co = hm(gpow[128 * c], g_at);
                    if (hzero(co)) continue;
                    const H* x = h_x_outers + c * x_len;
                    const std::vector<H> w = run_weights(J, x + 1 + n_log);
                    H wsum = hz();
                    for (size_t b = 0; b < w.size(); b++) wsum = ha(wsum, hm(w[b], ecd[b]));
                    // the branching program at sigma, from the success state down (jagged.rs:2160-2190)
                    H g[4] = {hz(), hz(), h1(), hz()};
                    for (int l = m; l >= 0; l--) {
                        H e4[4];
                        eq4h((uint32_t)l < n_log ? x[1 + l] : hz(), l < m ? point[l] : hz(), e4);
                        H prev[4] = {hz(), hz(), hz(), hz()};
                        for (int cd = 0; cd < 4; cd++) {
                            H t[4];
                            bp_step_h(e4, cd, g, t);
                            for (int s = 0; s < 4; s++) prev[s] = ha(prev[s], hm(el[4 * l + cd], t[s]));
                        }
                        for (int s = 0; s < 4; s++) g[s] = prev[s];
                    }
                    expect = ha(expect, hm(hm(co, wsum), g[0]));
                }
                if (!heqv(aclaim, expect)) why = "Frobenius assist";
                else if (!heqv(mrun, av)) why = "multipoint endpoint";
            }
        }
        if (why.empty()) {
            // V = sum_{c,j} coeff_{c,j} A_{c,j}^(2^j), coeff = the linearized coefficients of
            // Phi_c: bit b -> gamma_c eq(r''_c)[b]
            const std::vector<H>& D = dual_powers();
            if (D.empty()) why = "trace-dual basis";
            for (size_t c = 0; c < n_claims && why.empty(); c++) {
                H ge[128];
                for (int b = 0; b < 128; b++) ge[b] = hm(gam[c], eqr[c][b]);
                for (int j = 0; j < 128; j++) {
                    H cj = hz();
                    for (int b = 0; b < 128; b++) cj = ha(cj, hm(ge[b], D[(size_t)j * 128 + b]));
                    H x = vals[c][j];
                    for (int s = 0; s < j; s++) x = hm(x, x);
                    V = ha(V, hm(cj, x));
                }
            }
        }
    }
    if (!why.empty()) {
        flock_fs_free(ch);
        return reject(why);
    }
    // 5. the inner eq-basis Ligerito opening of q^(rho) on the parent transcript
    label(fs, "flock-pcs-open-batch-v0");
    label(fs, "flock-pcs-packed-direct-v0");
    observe(fs, q_eval);
    const H gamma = sample_vec(fs, 1)[0];
    const std::vector<uint8_t> capv(cap, cap + cap_len);
    const bool lig_ok = lig_verify(fs, C, L, dense_log, hm(gamma, q_eval), capv, lanes0, rho, gamma, &why);
    // merge the child (challenger.rs:193-203)
    const H d0 = sample(ch), d1 = sample(ch);
    flock_fs_free(ch);
    if (!lig_ok) return reject(why);
    observe(fs, d0);
    observe(fs, d1);
    if (!heqv(running, hm(q_eval, V))) return reject("merged sumcheck endpoint");
    if (consumed) *consumed = r.off;
    return 0;
}

int flock_open_ws_stage_ms(const flock_open_ws* ws, double* out, size_t n) {
    if (!ws || !out) return FLOCK_HIP_E_INVALID;
    for (size_t i = 0; i < n && i < 6; i++) out[i] = ws->stage_ms[i];
    return 0;
}

}  // extern "C"
