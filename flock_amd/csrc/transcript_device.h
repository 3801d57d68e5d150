// Device-resident Fiat-Shamir transcript: the FsChallenger of the host (SHA-256 discipline,
// crates/flock-core/src/challenger.rs:609-1044) kept in device memory for the sumcheck rounds,
// so a round's message, proof of work and challenge never leave the GPU.
//
// Every op of the rounds absorbs whole 16-byte units: a header [op][kind][0 x6][len u64 LE]
// (:708-714), an F128 as 16 little-endian bytes, or 8 nonce bytes zero-padded to 16 (:720-726).
// The running SHA-256 therefore always holds 0, 16, 32 or 48 pending bytes, kept as
// big-endian message words. A squeeze hashes (state || ctr_le) on a copy (:745-763); a
// PoW nonce is the smallest n with SHA-256(digest || n_le) starting with `bits` zero bits,
// MSB first (:884-958, 1096-1135), then absorbed as bytes.
//
// One round step, run by the workgroup that folds a round kernel's final message pair:
//   observe_f128(m0); observe_f128(m1); [grind_pow(bits)]; rho = sample_f128()
// It logs (m0, m1, nonce, rho) at log[round] and bumps round; the next round's kernel reads
// rho from that log entry. The host uploads the state before the rounds and replays the log
// afterwards through its own transcript, checking every nonce and challenge.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sha256_device.h"

namespace flock {

struct DevRound {
    uint4 m0, m1, rho;
    uint64_t nonce;
    uint32_t err;  // 1: no nonce found within the search bound
    uint32_t pad;
};

struct DevTranscript {
    uint32_t h[8];    // chaining value of the running SHA-256
    uint32_t buf[16]; // pending message words (big-endian), blen / 4 of them valid
    uint64_t total;   // bytes absorbed
    uint32_t blen;    // pending bytes: 0, 16, 32 or 48
    uint32_t round;   // rounds logged so far
    uint32_t grind_bits;
    uint32_t max_rounds;
    uint32_t pad[2];
};
static_assert(sizeof(DevRound) == 64 && sizeof(DevTranscript) == 128, "layout shared with host");

// Where a round kernel's final message pair goes: `out` (may be null); with `flag`, the value
// `seq` is stored there (system scope) once the pair is visible to the host, which polls it
// instead of synchronising the stream; with `fs`, the device transcript's round step runs on it.
struct RoundOut {
    uint4* out;
    uint32_t* flag;
    uint32_t seq;
    DevTranscript* fs;
};
// A plain destination for a pair.
__host__ __device__ inline RoundOut round_out(uint4* out) {
    return RoundOut{out, nullptr, 0u, nullptr};
}

// nonces tested per search at most: 2^16 iterations of one workgroup
constexpr uint32_t kFsGrindMaxIters = 1u << 16;
constexpr uint32_t kFsMaxGrindBits = 12;

enum : uint32_t { kFsOpObserve = 3, kFsOpSqueeze = 4, kFsOpBytes = 5 };
enum : uint32_t { kFsKindNone = 0, kFsKindScalar = 1 };

// The transcript held in registers by the thread that runs the step.
struct FsLocal {
    uint32_t h[8];
    uint32_t buf[16];
    uint64_t total;
    uint32_t blen;
};

__device__ __forceinline__ void fs_put4(uint32_t buf[16], uint32_t slot, uint32_t w0, uint32_t w1,
                                        uint32_t w2, uint32_t w3) {
    // static indices only (slot = blen / 16), so buf stays in registers
    switch (slot) {
        case 0: buf[0] = w0; buf[1] = w1; buf[2] = w2; buf[3] = w3; break;
        case 1: buf[4] = w0; buf[5] = w1; buf[6] = w2; buf[7] = w3; break;
        case 2: buf[8] = w0; buf[9] = w1; buf[10] = w2; buf[11] = w3; break;
        default: buf[12] = w0; buf[13] = w1; buf[14] = w2; buf[15] = w3; break;
    }
}

__device__ __forceinline__ void fs_absorb16(FsLocal& s, uint32_t w0, uint32_t w1, uint32_t w2,
                                            uint32_t w3) {
    fs_put4(s.buf, s.blen >> 4, w0, w1, w2, w3);
    s.blen += 16;
    s.total += 16;
    if (s.blen == 64) {
        uint32_t w[16];
#pragma unroll
        for (int i = 0; i < 16; i++) w[i] = s.buf[i];
        sha256_compress(s.h, w);
        s.blen = 0;
    }
}

__device__ __forceinline__ void fs_header(FsLocal& s, uint32_t op, uint32_t kind, uint64_t len) {
    fs_absorb16(s, op << 24 | kind << 16, 0u, bswap32((uint32_t)len), bswap32((uint32_t)(len >> 32)));
}

__device__ __forceinline__ void fs_absorb_f128(FsLocal& s, uint4 v) {
    fs_absorb16(s, bswap32(v.x), bswap32(v.y), bswap32(v.z), bswap32(v.w));
}

// SHA-256 finalisation of (state || tail) where tail is `tail_words` big-endian words
// (0 or 2) appended at the pending position; writes the 8 digest words.
__device__ __forceinline__ void fs_finish(const FsLocal& s, uint32_t tail_words, uint32_t t0,
                                          uint32_t t1, uint32_t out[8]) {
    uint32_t h[8], w[16];
#pragma unroll
    for (int i = 0; i < 8; i++) h[i] = s.h[i];
    const uint32_t at = s.blen >> 2;  // first free word
#pragma unroll
    for (int i = 0; i < 16; i++) {
        uint32_t v = (uint32_t)i < at ? s.buf[i] : 0u;
        if (tail_words == 2 && (uint32_t)i == at) v = t0;
        if (tail_words == 2 && (uint32_t)i == at + 1) v = t1;
        if ((uint32_t)i == at + tail_words) v = 0x80000000u;
        w[i] = v;
    }
    const uint64_t bits = (s.total + 4ull * tail_words) * 8ull;
    if (at + tail_words + 1 <= 14) {
        w[14] = (uint32_t)(bits >> 32);
        w[15] = (uint32_t)bits;
        sha256_compress(h, w);
    } else {
        sha256_compress(h, w);
#pragma unroll
        for (int i = 0; i < 14; i++) w[i] = 0u;
        w[14] = (uint32_t)(bits >> 32);
        w[15] = (uint32_t)bits;
        sha256_compress(h, w);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) out[i] = h[i];
}

// SHA-256(digest || nonce_le) has `bits` (1..32) leading zero bits
__device__ __forceinline__ bool fs_pow_ok(const uint32_t d[8], uint64_t nonce, uint32_t bits) {
    uint32_t h[8], w[16];
    sha256_iv(h);
#pragma unroll
    for (int i = 0; i < 8; i++) w[i] = d[i];
    w[8] = bswap32((uint32_t)nonce);
    w[9] = bswap32((uint32_t)(nonce >> 32));
    w[10] = 0x80000000u;
#pragma unroll
    for (int i = 11; i < 15; i++) w[i] = 0u;
    w[15] = 40u * 8u;
    sha256_compress(h, w);
    return bits >= 32 ? h[0] == 0u : (h[0] >> (32 - bits)) == 0u;
}

// LDS the round step works in: the transcript state, the PoW digest and the best nonce.
struct FsShared {
    FsLocal st;
    uint32_t digest[8];
    unsigned long long found;
};

// The round step; called by every thread of one workgroup with the message pair valid in
// thread 0, working in the workgroup's LDS `sh`.
__device__ __noinline__ void fs_round_step(DevTranscript* t, uint4 m0, uint4 m1, FsShared* sh) {
    FsLocal& st = sh->st;
    uint32_t* digest = sh->digest;
    unsigned long long& found = sh->found;
    const uint32_t bits = t->grind_bits;
    if (threadIdx.x == 0) {
        FsLocal s;
#pragma unroll
        for (int i = 0; i < 8; i++) s.h[i] = t->h[i];
#pragma unroll
        for (int i = 0; i < 16; i++) s.buf[i] = t->buf[i];
        s.total = t->total;
        s.blen = t->blen;
        fs_header(s, kFsOpObserve, kFsKindScalar, 1);
        fs_absorb_f128(s, m0);
        fs_header(s, kFsOpObserve, kFsKindScalar, 1);
        fs_absorb_f128(s, m1);
        if (bits) fs_finish(s, 0, 0u, 0u, digest);
        st = s;
        found = ~0ull;
    }
    __syncthreads();
    uint32_t err = 0;
    if (bits) {
        for (uint32_t it = 0;; it++) {
            const uint64_t n = (uint64_t)it * blockDim.x + threadIdx.x;
            if (fs_pow_ok(digest, n, bits)) atomicMin(&found, (unsigned long long)n);
            __syncthreads();
            const unsigned long long f = found;
            __syncthreads();  // every thread has read `found` before the next iteration
            if (f != ~0ull) break;
            if (it + 1 == kFsGrindMaxIters) {
                err = 1;
                break;
            }
        }
    }
    if (threadIdx.x == 0) {
        FsLocal s = st;
        const uint64_t nonce = found;
        if (bits) {
            fs_header(s, kFsOpBytes, kFsKindNone, 8);
            fs_absorb16(s, bswap32((uint32_t)nonce), bswap32((uint32_t)(nonce >> 32)), 0u, 0u);
        }
        fs_header(s, kFsOpSqueeze, kFsKindScalar, 1);
        uint32_t d[8];
        fs_finish(s, 2, 0u, 0u, d);  // SHA256(state || 0u64 LE): the first squeeze block
        const uint4 rho = make_uint4(bswap32(d[0]), bswap32(d[1]), bswap32(d[2]), bswap32(d[3]));
        const uint32_t r = t->round;
        if (r < t->max_rounds) {
            DevRound* log = reinterpret_cast<DevRound*>(t + 1);
            log[r].m0 = m0;
            log[r].m1 = m1;
            log[r].rho = rho;
            log[r].nonce = bits ? nonce : 0ull;
            log[r].err = err;
        }
#pragma unroll
        for (int i = 0; i < 8; i++) t->h[i] = s.h[i];
#pragma unroll
        for (int i = 0; i < 16; i++) t->buf[i] = s.buf[i];
        t->total = s.total;
        t->blen = s.blen;
        t->round = r + 1;
    }
}

}  // namespace flock
