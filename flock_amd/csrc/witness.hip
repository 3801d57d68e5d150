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

struct WitOut {
    uint64_t* z;
    uint64_t* a;
    uint64_t* b;
    uint64_t row;       // this thread's row o
    uint32_t n_log;
    bool live;          // a declared row (else zeros are written)
    uint64_t* lds;      // 4 shared words x 3 buffers, lane-minor: lds[(k*3+j)*kWitThreads + tid]
    __device__ __forceinline__ void emit(uint32_t w, uint64_t vz, uint64_t va, uint64_t vb) const {
        if (!live) vz = va = vb = 0;
        int k = -1;
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (w == sha2::kShared[i]) k = i;
        if (k >= 0) {
            lds[(k * 3 + 0) * kWitThreads + threadIdx.x] |= vz;
            lds[(k * 3 + 1) * kWitThreads + threadIdx.x] |= va;
            lds[(k * 3 + 2) * kWitThreads + threadIdx.x] |= vb;
            return;
        }
        const uint64_t at = 2 * ((((uint64_t)w >> 1) << n_log) + row) + (w & 1u);
        z[at] = vz;
        a[at] = va;
        b[at] = vb;
    }
};

// A region written in bit order: partial 64-bit words of z, a, b.
struct BitStream {
    uint32_t pos;
    uint64_t z, a, b;
    __device__ __forceinline__ explicit BitStream(uint32_t base) : pos(base), z(0), a(0), b(0) {}
    // append n <= 32 bits (values already masked to n bits)
    __device__ __forceinline__ void push(const WitOut& o, uint32_t vz, uint32_t va, uint32_t vb,
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
    __device__ __forceinline__ void push3(const WitOut& o, const uint32_t t[3], uint32_t n) {
        push(o, t[2], t[0], t[1], n);  // (z, a, b) = (product, left, right)
    }
    __device__ __forceinline__ void finish(const WitOut& o) {
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
    WitOut o{z, a, b, row, n_log, row < n_decl, shared_words};
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
