// Interleaved LCH additive NTT (neighbors-last) and the commit's replicate fills.
//
// Layout (crates/flock-core/src/ntt/additive_ntt_f128.rs:168-186): position-major SoA,
// data[pos * t + lane], 2^k positions x t lanes, F128 = 16 B. Butterfly at layer l,
// block b (additive_ntt_f128.rs:303-316):  u += v * tw(l, b);  v += u,  pairs
// (pos, pos + 2^(k-l-1)). tw(l, b) = XOR of evals[k-l-1][1 + j] over the set bits j of
// b (:141-144).
//
// One pass fuses K consecutive layers [l0, l0+K): a "group" is the 2^K positions
// hb*2^(k-l0) + i*S + lowp (S = 2^(k-l0-K), i < 2^K). All groups with the same high
// block hb share the same 2^K - 1 twiddles, so a workgroup serves one hb: it derives
// those twiddles, builds one 8 KiB Shoup table per twiddle in LDS, and then each thread
// owns one column c = lowp*t + lane, keeps its 2^K elements in registers and runs the K
// butterfly layers there. Columns are contiguous in memory, so every load/store of a
// wave is one coalesced 1 KiB segment.
#include "f128_device.h"
#include "flock_internal.h"

namespace flock {

// ------------------------------------------------------------------ fills --

// cw[(rep*D + p)*t + l] = q[l*D + p]  (replicate_lane_major_fill, pcs/commit.rs:516-536)
__global__ __launch_bounds__(256) void k_fill_lane_major(const f128* __restrict__ q,
                                                          f128* __restrict__ cw, uint64_t d,
                                                          uint32_t t, uint32_t reps) {
    __shared__ f128 tile[32][33];
    const uint64_t p0 = (uint64_t)blockIdx.x * 32;
    const uint32_t l0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
    for (int ly = 0; ly < 32; ly += 8) {
        const uint32_t l = l0 + ly + ty;
        const uint64_t p = p0 + tx;
        if (l < t && p < d) tile[ly + ty][tx] = q[(uint64_t)l * d + p];
    }
    __syncthreads();
    const uint64_t msg_len = d * t;
#pragma unroll
    for (int py = 0; py < 32; py += 8) {
        const uint64_t p = p0 + py + ty;
        const uint32_t l = l0 + tx;
        if (l < t && p < d) {
            const f128 v = tile[tx][py + ty];
            for (uint32_t rep = 0; rep < reps; rep++) cw[rep * msg_len + p * t + l] = v;
        }
    }
}

// cw[i] = msg[i % msg_len]  (replicate_message_fill, pcs/commit.rs:543-568)
__global__ __launch_bounds__(256) void k_fill_replicate(const f128* __restrict__ msg,
                                                         f128* __restrict__ cw, uint64_t msg_len,
                                                         uint32_t reps) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < msg_len;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const f128 v = msg[i];
        for (uint32_t rep = 0; rep < reps; rep++) cw[rep * msg_len + i] = v;
    }
}

hipError_t launch_fill_lane_major(const uint4* q, uint4* cw, uint64_t d, uint32_t t,
                                  uint32_t reps, hipStream_t s) {
    dim3 grid((unsigned)((d + 31) / 32), (t + 31) / 32);
    k_fill_lane_major<<<grid, 256, 0, s>>>(q, cw, d, t, reps);
    return hipGetLastError();
}

hipError_t launch_fill_replicate(const uint4* msg, uint4* cw, uint64_t msg_len, uint32_t reps,
                                 hipStream_t s) {
    uint64_t g = (msg_len + 255) / 256;
    if (g > 8192) g = 8192;
    k_fill_replicate<<<(unsigned)g, 256, 0, s>>>(msg, cw, msg_len, reps);
    return hipGetLastError();
}

// ------------------------------------------------------------------- NTT --

constexpr int kNttThreads = 512;
constexpr int kMaxNttK = 40;       // largest transform log-size the library accepts

// One fused pass over layers [l0, l0 + K). src == data for an in-place pass; for the
// first pass of a commit src is the message itself: the state after the log_inv_rate
// copy layers is the message replicated, so position P reads message position
// P mod 2^(k-l0) and the group's source offset drops the high block (base) term. With
// lm_d != 0 that message is lane-major (lane l at src[l * lm_d ..]), which fuses
// replicate_lane_major_fill (pcs/commit.rs:516-536) into the pass.
template <int K>
__global__ __launch_bounds__(kNttThreads) void k_ntt_pass(const f128* __restrict__ src,
                                                           f128* __restrict__ data, uint32_t t,
                                                           uint32_t live, int k, int l0,
                                                           const f128* __restrict__ evals,
                                                           uint64_t cols_per_hb,
                                                           uint32_t wgs_per_hb, int from_msg,
                                                           uint64_t lm_d) {
    constexpr int NT = (1 << K) - 1;  // twiddles (and Shoup tables) per group
    // one LDS array, tables first: they sit at LDS address 0, so a lookup's address is its
    // masked nibble plus an immediate
    __shared__ f128 lds[NT * SHOUP_ENTRIES + NT];
    f128* tables = lds;
    f128* tw = lds + NT * SHOUP_ENTRIES;

    const uint64_t hb = blockIdx.x / wgs_per_hb;
    const uint32_t chunk = blockIdx.x - (uint32_t)(hb * wgs_per_hb);

    // Twiddle index tix = 2^j - 1 + u serves layer l0 + j, block (hb << j) | u.
    if (threadIdx.x < NT) {
        const int tix = threadIdx.x;
        const int j = 31 - __clz(tix + 1);
        const uint64_t u = (uint64_t)(tix + 1 - (1 << j));
        const int layer = l0 + j;
        const uint64_t block = (hb << j) | u;
        const f128* row = evals + (size_t)(k - layer - 1) * k + 1;
        f128 acc = f128_zero();
        for (int bit = 0; bit < layer; bit++)
            if ((block >> bit) & 1) acc = f128_add(acc, row[bit]);
        tw[tix] = acc;
    }
    __syncthreads();
    for (int r = threadIdx.x; r < NT * 32; r += blockDim.x)
        shoup_build_row(tables + (r >> 5) * SHOUP_ENTRIES, tw[r >> 5], r & 31);
    __syncthreads();

    const uint64_t S = 1ull << (k - l0 - K);
    const uint64_t stride = S * t;                         // elements between group members
    const uint64_t base = (hb << (k - l0)) * (uint64_t)t;  // first element of the group
    const uint64_t sbase = from_msg ? 0 : base;
    for (uint64_t c = (uint64_t)chunk * blockDim.x + threadIdx.x; c < cols_per_hb;
         c += (uint64_t)wgs_per_hb * blockDim.x) {
        const uint32_t lane = (uint32_t)(c % t);
        if (lane >= live) continue;  // dead lanes are zero and stay zero
        f128 e[1 << K];
        if (lm_d) {
            // lane-major source (the first pass of commit_lane_major): message position
            // i * S + lowp of this lane is q[lane * D + i * S + lowp]
            const f128* ql = src + (uint64_t)lane * lm_d + c / t;
#pragma unroll
            for (int i = 0; i < (1 << K); i++) e[i] = ql[i * S];
        } else {
#pragma unroll
            for (int i = 0; i < (1 << K); i++) e[i] = src[sbase + i * stride + c];
        }
#pragma unroll
        for (int j = 0; j < K; j++) {
            const int h = 1 << (K - 1 - j);
#pragma unroll
            for (int i = 0; i < (1 << K); i++) {
                if (i & h) continue;
                const int tix = (1 << j) - 1 + (i >> (K - j));
                e[i] = f128_add(e[i], shoup_mul(tables + tix * SHOUP_ENTRIES, e[i | h]));
                e[i | h] = f128_add(e[i | h], e[i]);
            }
        }
#pragma unroll
        for (int i = 0; i < (1 << K); i++) data[base + i * stride + c] = e[i];
    }
}

// Deep passes (at most 256 columns per high block): a 32-row Shoup table per twiddle would be
// used only a few hundred times and a 512-thread group would mostly idle. Here a group of
// H * cols threads (rounded up to whole waves) takes H = min(256 / cols, 5) consecutive high
// blocks (50 columns: 250 of 256 lanes busy; 74 KB of LDS for K = 3), and builds, per
// (block, twiddle),
// the 8-row (2 KiB) table of shoup8_mul; each thread runs the K layers of one (block, column)
// with those tables. The tables come first in the group's LDS (address 0), then the twiddle
// rows and the twiddles.
constexpr int kDeepThreads = 256;
constexpr int kDeepMaxHb = 5;       // high blocks per group (tables: 5 x 7 x 2 KiB = 70 KiB)
constexpr int kDeepTabEntries = 128;  // 8 rows of 16

template <int K>
__host__ __device__ constexpr size_t deep_lds_bytes(uint32_t hbs) {
    return ((size_t)hbs * ((1 << K) - 1) * (kDeepTabEntries + 1) + (size_t)K * kMaxNttK) *
           sizeof(f128);
}

template <int K>
__global__ __launch_bounds__(kDeepThreads) void k_ntt_pass_deep(
    const f128* __restrict__ src, f128* __restrict__ data, uint32_t t, uint32_t live, int k,
    int l0, const f128* __restrict__ evals, uint32_t cols_per_hb, uint32_t hbs_per_wg,
    uint64_t n_hb, int from_msg) {
    constexpr int NT = (1 << K) - 1;
    extern __shared__ f128 lds[];
    f128* tabs = lds;                                   // hbs_per_wg * NT tables
    f128* rows = tabs + (size_t)hbs_per_wg * NT * kDeepTabEntries;  // [K][kMaxNttK]
    f128* tws = rows + K * kMaxNttK;                    // hbs_per_wg * NT twiddles
    for (int i = threadIdx.x; i < K * kMaxNttK; i += blockDim.x) {
        const int j = i / kMaxNttK, bit = i - j * kMaxNttK;
        const int layer = l0 + j;
        rows[i] = bit < layer ? evals[(size_t)(k - layer - 1) * k + 1 + bit] : f128_zero();
    }
    __syncthreads();
    const uint64_t hb0 = (uint64_t)blockIdx.x * hbs_per_wg;
    // twiddle (h, tix): tix = 2^j - 1 + u serves layer l0 + j, block (hb << j) | u
    for (uint32_t i = threadIdx.x; i < hbs_per_wg * NT; i += blockDim.x) {
        const uint32_t h = i / NT, tix = i - h * NT;
        const uint64_t hb = hb0 + h;
        f128 tw = f128_zero();
        if (hb < n_hb) {
            const int j = 31 - __clz(tix + 1);
            uint64_t blk = (hb << j) | (uint64_t)(tix + 1 - (1u << j));
            while (blk) {
                tw = f128_add(tw, rows[j * kMaxNttK + __builtin_ctzll(blk)]);
                blk &= blk - 1;
            }
        }
        tws[i] = tw;
    }
    __syncthreads();
    for (uint32_t r = threadIdx.x; r < hbs_per_wg * NT * 8; r += blockDim.x)
        shoup_build_row(tabs + (size_t)(r >> 3) * kDeepTabEntries, tws[r >> 3], (int)(r & 7));
    __syncthreads();
    const uint32_t h = threadIdx.x / cols_per_hb;
    const uint32_t c = threadIdx.x - h * cols_per_hb;
    const uint64_t hb = hb0 + h;
    if (h >= hbs_per_wg || hb >= n_hb || (c % t) >= live) return;  // no barriers below
    const uint64_t S = 1ull << (k - l0 - K);
    const uint64_t stride = S * t;
    const uint64_t base = (hb << (k - l0)) * (uint64_t)t;
    const uint64_t sbase = from_msg ? 0 : base;
    f128 e[1 << K];
#pragma unroll
    for (int i = 0; i < (1 << K); i++) e[i] = src[sbase + i * stride + c];
    const f128* mytabs = tabs + (size_t)NT * kDeepTabEntries * h;
#pragma unroll
    for (int j = 0; j < K; j++) {
        const int hh = 1 << (K - 1 - j);
#pragma unroll
        for (int i = 0; i < (1 << K); i++) {
            if (i & hh) continue;
            const int tix = (1 << j) - 1 + (i >> (K - j));
            e[i] = f128_add(e[i], shoup8_mul(mytabs + tix * kDeepTabEntries, e[i | hh]));
            e[i | hh] = f128_add(e[i | hh], e[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < (1 << K); i++) data[base + i * stride + c] = e[i];
}

template <int K>
static hipError_t launch_pass(const f128* src, f128* data, uint32_t t, uint32_t live, int k,
                              int l0, const f128* evals, bool from_msg, hipStream_t s,
                              uint64_t lm_d = 0) {
    constexpr int NT = (1 << K) - 1;
    const uint64_t n_hb = 1ull << l0;
    const uint64_t cols = (1ull << (k - l0 - K)) * t;
    if (cols <= (uint64_t)tunable(kTuneNttDeepCols)) {
        if (lm_d) return hipErrorInvalidValue;  // the caller fills instead
        uint32_t per = (uint32_t)(kDeepThreads / cols);  // high blocks per group
        if (per > (uint32_t)kDeepMaxHb) per = kDeepMaxHb;
        const uint64_t grid = (n_hb + per - 1) / per;
        if (grid > 0x7fffffffull) return hipErrorInvalidConfiguration;
        const size_t dyn = deep_lds_bytes<K>(per);
        static const hipError_t deep_attr = hipFuncSetAttribute(
            (const void*)k_ntt_pass_deep<K>, hipFuncAttributeMaxDynamicSharedMemorySize,
            (int)deep_lds_bytes<K>(kDeepMaxHb));
        if (deep_attr != hipSuccess) return deep_attr;
        const unsigned threads = (unsigned)((per * cols + 63) / 64 * 64);
        k_ntt_pass_deep<K><<<(unsigned)grid, threads, dyn, s>>>(
            src, data, t, live, k, l0, evals, (uint32_t)cols, per, n_hb, from_msg ? 1 : 0);
        return hipGetLastError();
    }
    // up to ntt_cols_per_thread (2) columns per thread so the twiddle-table build is amortised
    const uint64_t cpt = (uint64_t)tunable(kTuneNttColsPerThread);
    uint64_t wgs = (cols + cpt * kNttThreads - 1) / (cpt * kNttThreads);
    if (wgs < 1) wgs = 1;
    const uint64_t grid = n_hb * wgs;
    if (grid > 0x7fffffffull) return hipErrorInvalidConfiguration;
    k_ntt_pass<K><<<(unsigned)grid, kNttThreads, 0, s>>>(src, data, t, live, k, l0, evals, cols,
                                                           (uint32_t)wgs, from_msg ? 1 : 0, lm_d);
    return hipGetLastError();
}

// Fused depth of each pass: 3 layers (7 Shoup tables = 56 KiB of LDS, two groups per CU);
// a remainder of 1 or 2 layers is covered by one or two 2-layer passes.
static int plan_passes(int layers, int* ks) {
    int n = 0;
    if (layers == 1) {
        ks[n++] = 1;
        return n;
    }
    int threes = layers / 3, rem = layers % 3;
    if (rem == 1) {
        threes -= 1;
        rem = 4;
    }
    for (int i = 0; i < threes; i++) ks[n++] = 3;
    while (rem > 0) {
        ks[n++] = 2;
        rem -= 2;
    }
    return n;
}

hipError_t launch_ntt_pass(uint4* data, uint32_t t, uint32_t live, int k, int l0, int layers,
                           const uint4* evals, hipStream_t s) {
    switch (layers) {
        case 3: return launch_pass<3>(data, data, t, live, k, l0, evals, false, s);
        case 2: return launch_pass<2>(data, data, t, live, k, l0, evals, false, s);
        case 1: return launch_pass<1>(data, data, t, live, k, l0, evals, false, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ntt_forward(uint4* data, uint32_t t, uint32_t live, int k, int start_layer,
                              const uint4* evals, const uint4* msg, hipStream_t s,
                              uint64_t lm_d) {
    int ks[64];
    const int n = plan_passes(k - start_layer, ks);
    if (lm_d && n > 0 && (1ull << (k - start_layer - ks[0])) * t <= (uint64_t)tunable(kTuneNttDeepCols)) {
        // the first pass is a deep one: fill the replicated lane-major message, then transform
        // in place
        if (hipError_t e = launch_fill_lane_major(msg, data, lm_d, t, 1u << start_layer, s))
            return e;
        msg = nullptr;
        lm_d = 0;
    }
    int l = start_layer;
    for (int p = 0; p < n; p++) {
        const bool from_msg = (p == 0) && msg != nullptr;
        const f128* src = from_msg ? msg : data;
        // a message-sourced pass writes every lane (dead lanes copy their zeros)
        const uint32_t lv = from_msg ? t : live;
        const uint64_t lmd = from_msg ? lm_d : 0;
        hipError_t e;
        switch (ks[p]) {
            case 3: e = launch_pass<3>(src, data, t, lv, k, l, evals, from_msg, s, lmd); break;
            case 2: e = launch_pass<2>(src, data, t, lv, k, l, evals, from_msg, s, lmd); break;
            default: e = launch_pass<1>(src, data, t, lv, k, l, evals, from_msg, s, lmd); break;
        }
        if (e != hipSuccess) return e;
        l += ks[p];
    }
    return hipSuccess;
}

}  // namespace flock
