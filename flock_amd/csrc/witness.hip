// Witness generation and data movement.
//
// 1. The SHA-256 compression R1CS witness in the batch-major layout
//    (crates/flock-prover/src/r1cs_hashes/sha2.rs:896-1128, 1743-1898; common.rs:560-960).
//    One thread builds one compression instance. Its (z, a, b) rows are 2^15 bits, of which
//    the first 25,470 are used; they are produced region by region (H_in, M, the schedule
//    carries, then per round the Ch/Maj AND words, the round's carry-save block and the
//    periodic E_NEW/A_NEW words, then the output carries and H_out), each region strictly in
//    bit order. A region is a bit stream whose finished 64-bit words go straight to HBM at the
//    batch-major address: u64 word w of row o lives in F128 ((w >> 1) << n_log) + o, half
//    w & 1, so the 64 lanes of a wave (64 consecutive rows) store one contiguous 1 KiB span.
//    The four words that straddle two regions are OR-collected in LDS and stored at the end.
//    Rows past the declared count are written as zeros (dummy rows).
// 2. The lincheck byte stripe from z: an 8-row x 64-bit block transposes into 64 bytes.
// 3. The union dense-stack gather (compact_witness, crates/flock-core/src/union.rs:418-490).
//
// The padded union buffer is batch-major: chunk-column c of slot t holds its 2^nu rows at
// words [start_t + (c << nu), +2^nu). The committed stack q takes the first n_t rows of each
// used column, columns of a slot back to back, slots in order, then zeros up to the
// committed length. Every word moves once: an HBM copy, 32 bytes per kept word.
#include "f128_device.h"
#include "flock_internal.h"

namespace flock {

// ---- SHA-256 layout (sha2.rs:60-240) ----
namespace sha2 {
__constant__ uint32_t kK[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

constexpr uint32_t kHostK[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

constexpr uint32_t ctz32(uint32_t x) {
    uint32_t n = 0;
    while (!(x & 1u)) {
        x >>= 1;
        n++;
    }
    return n;
}
constexpr uint32_t kCpa = 31, kRip = 30, kSchedAdd = 92;
constexpr uint32_t kHBase = 0, kHOutBase = 256, kMBase = 512, kChBase = 1024, kMajBase = 3072;
constexpr uint32_t kRoundBase0 = 5120;
constexpr uint32_t round_bits(int r) { return 184 + kCpa - (ctz32(kHostK[r]) + 1); }
constexpr uint32_t sched_base() {
    uint32_t acc = kRoundBase0;
    for (int r = 0; r < 64; r++) acc += round_bits(r);
    return acc;
}
constexpr uint32_t kSchedBase = sched_base();                // 18,757
constexpr uint32_t kEBase = kSchedBase + 48 * kSchedAdd;     // 23,173
constexpr uint32_t kABase = kEBase + 32 * 32;                // 24,197
constexpr uint32_t kOutBase = kABase + 32 * 32;              // 25,221
constexpr uint32_t kZConst = kOutBase + 8 * kCpa;            // 25,469
constexpr uint32_t kUseful = kZConst + 1;                    // 25,470
constexpr uint32_t kUsefulWords = (kUseful + 63) / 64;       // 398
static_assert(kSchedBase == 18757 && kUseful == 25470, "SHA-256 layout");
// words shared by two regions (a region boundary inside the word)
constexpr uint32_t kShared[4] = {kSchedBase / 64, kEBase / 64, kABase / 64, kOutBase / 64};
static_assert(kSchedBase % 64 && kEBase % 64 && kABase % 64 && kOutBase % 64, "shared words");

__device__ __forceinline__ uint32_t ror(uint32_t x, int n) { return __builtin_rotateright32(x, n); }
__device__ __forceinline__ uint32_t bs0(uint32_t x) { return ror(x, 2) ^ ror(x, 13) ^ ror(x, 22); }
__device__ __forceinline__ uint32_t bs1(uint32_t x) { return ror(x, 6) ^ ror(x, 11) ^ ror(x, 25); }
__device__ __forceinline__ uint32_t ss0(uint32_t x) { return ror(x, 7) ^ ror(x, 18) ^ (x >> 3); }
__device__ __forceinline__ uint32_t ss1(uint32_t x) { return ror(x, 17) ^ ror(x, 19) ^ (x >> 10); }
}  // namespace sha2

constexpr int kWitThreads = 128;

// Output of one witness row: a finished 64-bit word goes to its batch-major address, except
// the words listed by the policy (SHARED: words that straddle two regions written by
// different streams), which are OR-collected in LDS and stored at the end of the row.
template <class SHARED>
struct WitOut {
    uint64_t* z;
    uint64_t* a;
    uint64_t* b;
    uint64_t row;       // this thread's row o
    uint32_t n_log;
    bool live;          // a declared row (else zeros are written)
    uint64_t* lds;      // SHARED::kN words x 3 buffers, lane-minor: lds[(k*3+j)*blockDim + tid]
    __device__ __forceinline__ void emit(uint32_t w, uint64_t vz, uint64_t va, uint64_t vb) const {
        if (!live) vz = va = vb = 0;
        if constexpr (SHARED::kN > 0) {
            int k = -1;
#pragma unroll
            for (int i = 0; i < SHARED::kN; i++)
                if (w == SHARED::word(i)) k = i;
            if (k >= 0) {
                lds[(k * 3 + 0) * blockDim.x + threadIdx.x] |= vz;
                lds[(k * 3 + 1) * blockDim.x + threadIdx.x] |= va;
                lds[(k * 3 + 2) * blockDim.x + threadIdx.x] |= vb;
                return;
            }
        }
        const uint64_t at = 2 * ((((uint64_t)w >> 1) << n_log) + row) + (w & 1u);
        z[at] = vz;
        a[at] = va;
        b[at] = vb;
    }
};

struct Sha2Shared {
    static constexpr int kN = 4;
    __device__ static constexpr uint32_t word(int i) { return sha2::kShared[i]; }
};
struct NoShared {
    static constexpr int kN = 0;
    __device__ static constexpr uint32_t word(int) { return 0; }
};

// A region written in bit order: partial 64-bit words of z, a, b.
struct BitStream {
    uint32_t pos;
    uint64_t z, a, b;
    __device__ __forceinline__ explicit BitStream(uint32_t base) : pos(base), z(0), a(0), b(0) {}
    // append n <= 32 bits (values already masked to n bits)
    template <class O>
    __device__ __forceinline__ void push(const O& o, uint32_t vz, uint32_t va, uint32_t vb,
                                         uint32_t n) {
        const uint32_t sh = pos & 63u;
        z |= (uint64_t)vz << sh;
        a |= (uint64_t)va << sh;
        b |= (uint64_t)vb << sh;
        if (sh + n >= 64u) {
            o.emit(pos >> 6, z, a, b);
            const uint32_t used = 64u - sh;  // 1..32
            z = (uint64_t)vz >> used;
            a = (uint64_t)va >> used;
            b = (uint64_t)vb >> used;
        }
        pos += n;
    }
    template <class O>
    __device__ __forceinline__ void push3(const O& o, const uint32_t t[3], uint32_t n) {
        push(o, t[2], t[0], t[1], n);  // (z, a, b) = (product, left, right)
    }
    template <class O>
    __device__ __forceinline__ void finish(const O& o) {
        if (pos & 63u) o.emit(pos >> 6, z, a, b);
    }
};

#define LO31 0x7FFFFFFFu
#define LO30 0x3FFFFFFFu

__device__ __forceinline__ uint32_t add_parts(uint32_t x, uint32_t y, uint32_t t[3]) {
    const uint32_t s = x + y, cin = s ^ x ^ y;
    t[0] = (x ^ cin) & LO31;
    t[1] = (y ^ cin) & LO31;
    t[2] = t[0] & t[1];
    return s;
}

__device__ __forceinline__ uint32_t add3_parts(uint32_t x, uint32_t y, uint32_t m, uint32_t maj[3],
                                               uint32_t rip[3]) {
    const uint32_t xm = x ^ m, ym = y ^ m, w = xm & ym, p = x ^ y ^ m, bw = (w ^ m) << 1;
    const uint32_t s = p + bw, cin = s ^ p ^ bw, rl = p ^ cin, rr = bw ^ cin;
    maj[0] = xm & LO31;
    maj[1] = ym & LO31;
    maj[2] = w & LO31;
    rip[0] = (rl >> 1) & LO30;
    rip[1] = (rr >> 1) & LO30;
    rip[2] = ((rl & rr) >> 1) & LO30;
    return s;
}

__device__ __forceinline__ uint32_t add4_parts(uint32_t x, uint32_t y, uint32_t z, uint32_t w,
                                               uint32_t m1[3], uint32_t m2[3], uint32_t rip[3]) {
    const uint32_t xz = x ^ z, yz = y ^ z, w1 = xz & yz, p1 = x ^ y ^ z, b1 = (w1 ^ z) << 1;
    const uint32_t pw = p1 ^ w, bw = b1 ^ w, w2 = pw & bw, p2 = p1 ^ b1 ^ w, b2 = (w2 ^ w) << 1;
    const uint32_t s = p2 + b2, cin = s ^ p2 ^ b2, rl = p2 ^ cin, rr = b2 ^ cin;
    m1[0] = xz & LO31;
    m1[1] = yz & LO31;
    m1[2] = w1 & LO31;
    m2[0] = pw & LO31;
    m2[1] = bw & LO31;
    m2[2] = w2 & LO31;
    rip[0] = (rl >> 1) & LO30;
    rip[1] = (rr >> 1) & LO30;
    rip[2] = ((rl & rr) >> 1) & LO30;
    return s;
}

// One thread per row o < 2^n_log; inputs[24 o ..] = (h_in[8], m[16]) for o < n_decl.
__global__ __launch_bounds__(kWitThreads) void k_sha2_witness(const uint32_t* __restrict__ inputs,
                                                               uint64_t n_decl, uint32_t n_log,
                                                               uint64_t* __restrict__ z,
                                                               uint64_t* __restrict__ a,
                                                               uint64_t* __restrict__ b) {
    __shared__ uint32_t wsched[64 * kWitThreads];
    __shared__ uint64_t shared_words[4 * 3 * kWitThreads];
    const int tid = threadIdx.x;
    const uint64_t row = (uint64_t)blockIdx.x * kWitThreads + tid;
    if (row >= (1ull << n_log)) return;  // the block is wider than 2^n_log (no barriers below)
    WitOut<Sha2Shared> o{z, a, b, row, n_log, row < n_decl, shared_words};
#pragma unroll
    for (int i = 0; i < 12; i++) shared_words[i * kWitThreads + tid] = 0;
    uint32_t h_in[8];
    const uint32_t* in = inputs + 24 * (o.live ? row : 0);
#pragma unroll
    for (int i = 0; i < 8; i++) h_in[i] = o.live ? in[i] : 0u;
    for (int i = 0; i < 16; i++) wsched[i * kWitThreads + tid] = o.live ? in[8 + i] : 0u;

    // H_in, M: (z, a, b) = (v, v, 1); aligned 32-bit words
    {
        BitStream s(sha2::kHBase);
        for (int w = 0; w < 8; w++) s.push(o, h_in[w], h_in[w], 0xFFFFFFFFu, 32);
        s.finish(o);
        BitStream m(sha2::kMBase);
        for (int i = 0; i < 16; i++) {
            const uint32_t v = wsched[i * kWitThreads + tid];
            m.push(o, v, v, 0xFFFFFFFFu, 32);
        }
        m.finish(o);
    }
    // message schedule: fused 4-operand trees, 92 carry bits per step
    {
        BitStream s(sha2::kSchedBase);
        for (int t = 16; t < 64; t++) {
            uint32_t m1[3], m2[3], rp[3];
            const uint32_t wt = add4_parts(sha2::ss1(wsched[(t - 2) * kWitThreads + tid]),
                                           sha2::ss0(wsched[(t - 15) * kWitThreads + tid]),
                                           wsched[(t - 7) * kWitThreads + tid],
                                           wsched[(t - 16) * kWitThreads + tid], m1, m2, rp);
            s.push3(o, m1, sha2::kCpa);
            s.push3(o, m2, sha2::kCpa);
            s.push3(o, rp, sha2::kRip);
            wsched[t * kWitThreads + tid] = wt;
        }
        s.finish(o);
    }
    // rounds
    uint32_t A = h_in[0], B = h_in[1], C = h_in[2], D = h_in[3], E = h_in[4], F = h_in[5],
             G = h_in[6], H = h_in[7];
    {
        BitStream sch(sha2::kChBase), smj(sha2::kMajBase), srd(sha2::kRoundBase0);
        BitStream se(sha2::kEBase), sa(sha2::kABase);
        for (int r = 0; r < 64; r++) {
            const uint32_t fg = F ^ G, ch = E & fg;
            sch.push(o, ch, E, fg, 32);
            const uint32_t ba = B ^ A, ca = C ^ A, mj = ba & ca;
            smj.push(o, mj, ba, ca, 32);
            const uint32_t ch_out = ch ^ G, maj_out = mj ^ A;
            const uint32_t k = sha2::kK[r];
            const uint32_t kt = (uint32_t)__builtin_ctz(k) + 1;
            uint32_t kp[3];
            const uint32_t hk = add_parts(k, H, kp);
            const uint32_t kmask = (1u << (31u - kt)) - 1u;
            const uint32_t kadd[3] = {(kp[0] >> kt) & kmask, (kp[1] >> kt) & kmask,
                                      (kp[2] >> kt) & kmask};
            uint32_t m1[3], m2[3], rp[3];
            const uint32_t t1 =
                add4_parts(hk, sha2::bs1(E), ch_out, wsched[r * kWitThreads + tid], m1, m2, rp);
            uint32_t am[3], ar[3];
            const uint32_t a_new = add3_parts(t1, sha2::bs0(A), maj_out, am, ar);
            uint32_t ep[3];
            const uint32_t e_new = add_parts(D, t1, ep);
            srd.push3(o, m1, sha2::kCpa);
            srd.push3(o, m2, sha2::kCpa);
            srd.push3(o, rp, sha2::kRip);
            srd.push3(o, am, sha2::kCpa);
            srd.push3(o, ar, sha2::kRip);
            srd.push3(o, ep, sha2::kCpa);
            srd.push3(o, kadd, 31u - kt);
            if (r & 1) {
                sa.push(o, a_new, a_new, 0xFFFFFFFFu, 32);
                se.push(o, e_new, e_new, 0xFFFFFFFFu, 32);
            }
            H = G;
            G = F;
            F = E;
            E = e_new;
            D = C;
            C = B;
            B = A;
            A = a_new;
        }
        sch.finish(o);
        smj.finish(o);
        srd.finish(o);
        se.finish(o);
        sa.finish(o);
    }
    // output feed-forward, then the constant wire
    {
        const uint32_t fin[8] = {A, B, C, D, E, F, G, H};
        BitStream sc(sha2::kOutBase), sh(sha2::kHOutBase);
        for (int w = 0; w < 8; w++) {
            uint32_t t[3];
            const uint32_t s = add_parts(fin[w], h_in[w], t);
            sc.push3(o, t, sha2::kCpa);
            sh.push(o, s, s, 0xFFFFFFFFu, 32);
        }
        sc.push(o, 1u, 1u, 1u, 1);  // Z_CONST: (1, 1, 1)
        sc.finish(o);
        sh.finish(o);
    }
    // the straddling words, OR-collected from both regions
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t w = sha2::kShared[k];
        const uint64_t at = 2 * ((((uint64_t)w >> 1) << n_log) + row) + (w & 1u);
        z[at] = shared_words[(k * 3 + 0) * kWitThreads + tid];
        a[at] = shared_words[(k * 3 + 1) * kWitThreads + tid];
        b[at] = shared_words[(k * 3 + 2) * kWitThreads + tid];
    }
}

// 8x8 bit transpose of a u64 (bit 8r + c -> bit 8c + r), three delta swaps.
__device__ __forceinline__ uint64_t transpose8(uint64_t x) {
    uint64_t t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
    x = x ^ t ^ (t << 7);
    t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
    x = x ^ t ^ (t << 14);
    t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
    return x ^ t ^ (t << 28);
}

// stripe[g*K + 64 w + j] bit r = bit j of u64 word w of row 8g + r (stripe_from_rows,
// common.rs:587-599); one thread per (group g, word w < K/64); words past useful_words are 0.
__global__ __launch_bounds__(256) void k_stripe_from_z(const uint64_t* __restrict__ z,
                                                        uint32_t n_log, uint32_t words_per_row,
                                                        uint32_t useful_words,
                                                        uint64_t* __restrict__ stripe) {
    const uint64_t n_groups = (1ull << n_log) / 8;
    const uint64_t total = n_groups * words_per_row;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t g = e / words_per_row;
        const uint32_t w = (uint32_t)(e - g * words_per_row);
        uint64_t x[8];
        if (w < useful_words) {
#pragma unroll
            for (int r = 0; r < 8; r++)
                x[r] = z[2 * ((((uint64_t)w >> 1) << n_log) + 8 * g + r) + (w & 1u)];
        } else {
#pragma unroll
            for (int r = 0; r < 8; r++) x[r] = 0;
        }
        uint64_t* out = stripe + (g * words_per_row * 64 + (uint64_t)w * 64) / 8;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            // M: byte r = byte c of x[r]
            uint64_t m = 0;
#pragma unroll
            for (int r = 0; r < 8; r++) m |= ((x[r] >> (8 * c)) & 0xFFull) << (8 * r);
            out[c] = transpose8(m);  // bytes 8c..8c+7: byte t bit r = bit 8c+t of x[r]
        }
    }
}

hipError_t launch_sha2_witness(const uint32_t* inputs, uint64_t n_decl, uint32_t n_log,
                               uint4* z, uint4* a, uint4* b, uint8_t* stripe, hipStream_t s) {
    const uint64_t n = 1ull << n_log;
    const uint32_t useful_chunks = (sha2::kUseful + 127) / 128;  // 199
    const uint64_t words = 256ull << n_log;                      // F128 per buffer
    // chunk-columns past the useful ones are zero
    for (uint4* buf : {z, a, b}) {
        hipError_t e = launch_zero_words(buf + ((uint64_t)useful_chunks << n_log),
                                         words - ((uint64_t)useful_chunks << n_log), s);
        if (e != hipSuccess) return e;
    }
    const unsigned grid = (unsigned)((n + kWitThreads - 1) / kWitThreads);
    k_sha2_witness<<<grid, kWitThreads, 0, s>>>(inputs, n_decl, n_log,
                                                reinterpret_cast<uint64_t*>(z),
                                                reinterpret_cast<uint64_t*>(a),
                                                reinterpret_cast<uint64_t*>(b));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !stripe) return e;
    const uint64_t total = (n / 8) * 512;
    uint64_t g = (total + 255) / 256;
    if (g > 16384) g = 16384;
    k_stripe_from_z<<<(unsigned)g, 256, 0, s>>>(reinterpret_cast<const uint64_t*>(z), n_log, 512,
                                                sha2::kUsefulWords,
                                                reinterpret_cast<uint64_t*>(stripe));
    return hipGetLastError();
}

// ---- BLAKE3 compression (blake3.rs:122-340, 1283-1452) ----
// Row layout (2^14 bits): cv [0,256), out_lo [256,512), m [512,1024), counter_lo | counter_hi |
// block_len | flags [1024,1152), out_hi [1152,1408), then the 56 G blocks (183/183/182/182 bits
// for round 1's column G's, 184 otherwise) and the constant at 11,706. Every region starts on
// a 64-bit boundary, so each is its own stream and no word is shared.
namespace b3 {
constexpr uint32_t kIV[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                             0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
constexpr int kPerm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
constexpr int kLanes[8][4] = {{0, 4, 8, 12}, {1, 5, 9, 13}, {2, 6, 10, 14}, {3, 7, 11, 15},
                              {0, 5, 10, 15}, {1, 6, 11, 12}, {2, 7, 8, 13}, {3, 4, 9, 14}};
struct MsgTable {
    uint8_t v[7][8][2];  // (mx, my) message word of round r, G g (per_round_msg_idx)
};
constexpr MsgTable make_msg_table() {
    MsgTable t{};
    int perm[16] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15};
    for (int r = 0; r < 7; r++) {
        for (int g = 0; g < 8; g++) {
            t.v[r][g][0] = (uint8_t)perm[2 * g];
            t.v[r][g][1] = (uint8_t)perm[2 * g + 1];
        }
        int next[16] = {};
        for (int i = 0; i < 16; i++) next[i] = perm[kPerm[i]];
        for (int i = 0; i < 16; i++) perm[i] = next[i];
    }
    return t;
}
__constant__ MsgTable kMsg = make_msg_table();
constexpr uint32_t kGsBase = 1408, kZConst = 11706, kUseful = 11707;
constexpr uint32_t kUsefulWords = (kUseful + 63) / 64;  // 183
}  // namespace b3

constexpr int kB3Threads = 128;

// One thread per row o < 2^n_log; inputs[28 o ..] = (cv[8], m[16], counter_lo, counter_hi,
// block_len, flags) for o < n_decl; rows past n_decl are zero (the partial drivers' dummies).
__global__ __launch_bounds__(kB3Threads) void k_blake3_witness(const uint32_t* __restrict__ inputs,
                                                                uint64_t n_decl, uint32_t n_log,
                                                                uint64_t* __restrict__ z,
                                                                uint64_t* __restrict__ a,
                                                                uint64_t* __restrict__ b) {
    __shared__ uint32_t msg[16 * kB3Threads];
    const int tid = threadIdx.x;
    const uint64_t row = (uint64_t)blockIdx.x * kB3Threads + tid;
    if (row >= (1ull << n_log)) return;  // no barriers below
    WitOut<NoShared> o{z, a, b, row, n_log, row < n_decl, nullptr};
    const uint32_t* in = inputs + 28 * (o.live ? row : 0);
    uint32_t s[16];
#pragma unroll
    for (int i = 0; i < 8; i++) s[i] = o.live ? in[i] : 0u;
    for (int i = 0; i < 16; i++) msg[i * kB3Threads + tid] = o.live ? in[8 + i] : 0u;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        s[8 + i] = b3::kIV[i];
        s[12 + i] = o.live ? in[24 + i] : 0u;
    }
    {  // cv, m, (counter_lo, counter_hi, block_len, flags): (z, a, b) = (v, v, 1)
        BitStream cv(0);
#pragma unroll
        for (int w = 0; w < 8; w++) cv.push(o, s[w], s[w], 0xFFFFFFFFu, 32);
        cv.finish(o);
        BitStream mp(512);
        for (int i = 0; i < 16; i++) {
            const uint32_t v = msg[i * kB3Threads + tid];
            mp.push(o, v, v, 0xFFFFFFFFu, 32);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) mp.push(o, s[12 + i], s[12 + i], 0xFFFFFFFFu, 32);
        mp.finish(o);
    }
    BitStream gs(b3::kGsBase);
#pragma unroll 1
    for (int r = 0; r < 7; r++) {
#pragma unroll
        for (int gi = 0; gi < 8; gi++) {
            const int la = b3::kLanes[gi][0], lb = b3::kLanes[gi][1], lc = b3::kLanes[gi][2],
                      ld = b3::kLanes[gi][3];
            const uint32_t mx = msg[b3::kMsg.v[r][gi][0] * kB3Threads + tid];
            const uint32_t my = msg[b3::kMsg.v[r][gi][1] * kB3Threads + tid];
            uint32_t maj1[3], rip1[3], t1[3], maj2[3], rip2[3], t2[3];
            const uint32_t a1 = add3_parts(s[la], s[lb], mx, maj1, rip1);
            gs.push3(o, maj1, 31);
            gs.push3(o, rip1, 30);
            const uint32_t d1 = sha2::ror(s[ld] ^ a1, 16);
            uint32_t c1;
            if (r == 0 && gi < 4) {
                // c = IV[gi], a constant: rows for bits t..30 only (g_c1_rows, blake3.rs:154-160)
                c1 = add_parts(b3::kIV[gi], d1, t1);
                const uint32_t n = gi < 2 ? 30u : 29u, t = 31u - n, mask = (1u << n) - 1u;
                const uint32_t tt[3] = {(t1[0] >> t) & mask, (t1[1] >> t) & mask,
                                        (t1[2] >> t) & mask};
                gs.push3(o, tt, n);
            } else {
                c1 = add_parts(s[lc], d1, t1);
                gs.push3(o, t1, 31);
            }
            const uint32_t b1 = sha2::ror(s[lb] ^ c1, 12);
            const uint32_t a2 = add3_parts(a1, b1, my, maj2, rip2);
            const uint32_t d2 = sha2::ror(d1 ^ a2, 8);
            const uint32_t c2 = add_parts(c1, d2, t2);
            gs.push3(o, maj2, 31);
            gs.push3(o, rip2, 30);
            gs.push3(o, t2, 31);
            s[la] = a2;
            s[lb] = sha2::ror(b1 ^ c2, 7);
            s[lc] = c2;
            s[ld] = d2;
        }
    }
    gs.push(o, 1u, 1u, 1u, 1);  // the constant wire, (1, 1, 1)
    gs.finish(o);
    o.emit(b3::kUsefulWords, 0, 0, 0);  // the upper half of the last useful chunk
    // finalization rows: out_lo = s[w] ^ s[w+8], out_hi = s[w+8] ^ cv[w]
    BitStream lo(256), hi(1152);
#pragma unroll
    for (int w = 0; w < 8; w++) {
        const uint32_t v = s[w] ^ s[w + 8];
        lo.push(o, v, v, 0xFFFFFFFFu, 32);
        const uint32_t cvw = o.live ? in[w] : 0u;
        const uint32_t h = s[w + 8] ^ cvw;
        hi.push(o, h, h, 0xFFFFFFFFu, 32);
    }
    lo.finish(o);
    hi.finish(o);
}

hipError_t launch_blake3_witness(const uint32_t* inputs, uint64_t n_decl, uint32_t n_log,
                                 uint4* z, uint4* a, uint4* b, uint8_t* stripe, hipStream_t s) {
    const uint64_t n = 1ull << n_log;
    const uint32_t useful_chunks = (b3::kUseful + 127) / 128;  // 92
    const uint64_t words = 128ull << n_log;                    // F128 per buffer
    for (uint4* buf : {z, a, b}) {
        hipError_t e = launch_zero_words(buf + ((uint64_t)useful_chunks << n_log),
                                         words - ((uint64_t)useful_chunks << n_log), s);
        if (e != hipSuccess) return e;
    }
    const unsigned grid = (unsigned)((n + kB3Threads - 1) / kB3Threads);
    k_blake3_witness<<<grid, kB3Threads, 0, s>>>(inputs, n_decl, n_log,
                                                 reinterpret_cast<uint64_t*>(z),
                                                 reinterpret_cast<uint64_t*>(a),
                                                 reinterpret_cast<uint64_t*>(b));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !stripe) return e;
    const uint64_t total = (n / 8) * 256;
    uint64_t g = (total + 255) / 256;
    if (g > 16384) g = 16384;
    k_stripe_from_z<<<(unsigned)g, 256, 0, s>>>(reinterpret_cast<const uint64_t*>(z), n_log, 256,
                                                b3::kUsefulWords,
                                                reinterpret_cast<uint64_t*>(stripe));
    return hipGetLastError();
}

// ---- Keccak-f[1600] (keccak.rs:98-330, 436-643) ----
// Row layout (u64 words of a 2^16-bit row): state_0 lanes 0..24, state_24 lanes 32..56, the
// constant at word 64 (bit 0), t_r lanes at 65 + 25 r + lane; 42,560 useful bits. Rows are
// whole lanes, so every store is one aligned u64 at its batch-major address.
namespace keccak {
__constant__ uint64_t kRC[24] = {
    0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
    0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
    0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
    0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
    0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
    0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};
// rho offset of lane (x, y), indexed [x][y]
constexpr int kRho[5][5] = {{0, 36, 3, 41, 18}, {1, 44, 10, 45, 2}, {62, 6, 43, 15, 61},
                            {28, 55, 25, 21, 56}, {27, 20, 39, 8, 14}};
constexpr uint32_t kUseful = 42560;
constexpr uint32_t kUsefulWords = kUseful / 64;  // 665
__device__ __forceinline__ uint64_t rotl(uint64_t v, int r) {
    return r ? (v << r) | (v >> (64 - r)) : v;
}
}  // namespace keccak

constexpr int kKecThreads = 128;

// One thread per row o < 2^n_log; inputs[25 o ..] = the initial lanes for o < n_decl, the zero
// state otherwise (the padding invocations of generate_witness_batch_major, keccak.rs:1228-1285).
__global__ __launch_bounds__(kKecThreads) void k_keccak_witness(const uint64_t* __restrict__ inputs,
                                                                 uint64_t n_decl, uint32_t n_log,
                                                                 uint64_t* __restrict__ z,
                                                                 uint64_t* __restrict__ a,
                                                                 uint64_t* __restrict__ b) {
    const uint64_t row = (uint64_t)blockIdx.x * kKecThreads + threadIdx.x;
    if (row >= (1ull << n_log)) return;
    auto at = [&](uint32_t w) { return 2 * ((((uint64_t)w >> 1) << n_log) + row) + (w & 1u); };
    auto put = [&](uint32_t w, uint64_t vz, uint64_t va, uint64_t vb) {
        const uint64_t i = at(w);
        z[i] = vz;
        a[i] = va;
        b[i] = vb;
    };
    uint64_t s[25];
    const bool live = row < n_decl;
#pragma unroll
    for (int l = 0; l < 25; l++) s[l] = live ? inputs[25 * row + l] : 0ull;
#pragma unroll
    for (int l = 0; l < 25; l++) put(l, s[l], s[l], ~0ull);
    for (uint32_t w = 25; w < 32; w++) put(w, 0, 0, 0);
    for (uint32_t w = 57; w < 64; w++) put(w, 0, 0, 0);
    put(64, 1, 1, 1);
    put(keccak::kUsefulWords, 0, 0, 0);  // the upper half of the last useful chunk
#pragma unroll 1
    for (int r = 0; r < 24; r++) {
        uint64_t c[5], d[5], B[25];
#pragma unroll
        for (int x = 0; x < 5; x++) c[x] = s[x] ^ s[x + 5] ^ s[x + 10] ^ s[x + 15] ^ s[x + 20];
#pragma unroll
        for (int x = 0; x < 5; x++) d[x] = c[(x + 4) % 5] ^ keccak::rotl(c[(x + 1) % 5], 1);
#pragma unroll
        for (int y = 0; y < 5; y++)
#pragma unroll
            for (int x = 0; x < 5; x++) {
                const int sx = (x + 3 * y) % 5, sy = x;
                B[x + 5 * y] = keccak::rotl(s[sx + 5 * sy] ^ d[sx], keccak::kRho[sx][sy]);
            }
        const uint32_t tb = 65 + 25 * r;
#pragma unroll
        for (int y = 0; y < 5; y++)
#pragma unroll
            for (int x = 0; x < 5; x++) {
                const uint64_t b1 = B[(x + 1) % 5 + 5 * y], b2 = B[(x + 2) % 5 + 5 * y];
                const uint64_t t = ~b1 & b2;
                put(tb + x + 5 * y, t, ~b1, b2);
                s[x + 5 * y] = B[x + 5 * y] ^ t;
            }
        s[0] ^= keccak::kRC[r];
    }
#pragma unroll
    for (int l = 0; l < 25; l++) put(32 + l, s[l], s[l], ~0ull);
}

hipError_t launch_keccak_witness(const uint64_t* inputs, uint64_t n_decl, uint32_t n_log,
                                 uint4* z, uint4* a, uint4* b, uint8_t* stripe, hipStream_t s) {
    const uint64_t n = 1ull << n_log;
    const uint32_t useful_chunks = (keccak::kUseful + 127) / 128;  // 333
    const uint64_t words = 512ull << n_log;                        // F128 per buffer
    for (uint4* buf : {z, a, b}) {
        hipError_t e = launch_zero_words(buf + ((uint64_t)useful_chunks << n_log),
                                         words - ((uint64_t)useful_chunks << n_log), s);
        if (e != hipSuccess) return e;
    }
    const unsigned grid = (unsigned)((n + kKecThreads - 1) / kKecThreads);
    k_keccak_witness<<<grid, kKecThreads, 0, s>>>(inputs, n_decl, n_log,
                                                  reinterpret_cast<uint64_t*>(z),
                                                  reinterpret_cast<uint64_t*>(a),
                                                  reinterpret_cast<uint64_t*>(b));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !stripe) return e;
    const uint64_t total = (n / 8) * 1024;
    uint64_t g = (total + 255) / 256;
    if (g > 16384) g = 16384;
    k_stripe_from_z<<<(unsigned)g, 256, 0, s>>>(reinterpret_cast<const uint64_t*>(z), n_log, 1024,
                                                keccak::kUsefulWords,
                                                reinterpret_cast<uint64_t*>(stripe));
    return hipGetLastError();
}

// One slot: q[c * n_t + i] = z[start + (c << nu) + i] for c < cols, i < n_t.
__global__ __launch_bounds__(256) void k_compact_slot(const f128* __restrict__ z, uint64_t start,
                                                       uint32_t nu, uint64_t cols, uint64_t n_t,
                                                       f128* __restrict__ q) {
    const uint64_t total = cols * n_t;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t c = e / n_t, i = e - c * n_t;
        q[e] = z[start + (c << nu) + i];
    }
}

__global__ __launch_bounds__(256) void k_zero_words(f128* __restrict__ q, uint64_t n) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (uint64_t)gridDim.x * blockDim.x)
        q[e] = f128_zero();
}

static inline unsigned copy_grid(uint64_t n) {
    uint64_t g = (n + 255) / 256;
    if (g < 1) g = 1;
    if (g > 16384) g = 16384;
    return (unsigned)g;
}

hipError_t launch_compact_slot(const uint4* z, uint64_t start, uint32_t nu, uint64_t cols,
                               uint64_t n_t, uint4* q, hipStream_t s) {
    if (cols == 0 || n_t == 0) return hipSuccess;
    k_compact_slot<<<copy_grid(cols * n_t), 256, 0, s>>>(z, start, nu, cols, n_t, q);
    return hipGetLastError();
}

hipError_t launch_zero_words(uint4* q, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_zero_words<<<copy_grid(n), 256, 0, s>>>(q, n);
    return hipGetLastError();
}

}  // namespace flock
