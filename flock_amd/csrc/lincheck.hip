// Lincheck kernels (crates/flock-core/src/lincheck.rs, lincheck/union.rs).
//
// The column-domain vectors have K = 2^k_log words (K = 2^15 for the SHA-256 circuit):
//   comb[c]  = alpha * (A^T eq_inner)[c] + (B^T eq_inner)[c]       CscCircuit, lincheck.rs:432-497
//   z_vec[i] = sum_o z[i, o] eq(x_outer, o)                         partial fold, :804-990
// followed by a product sumcheck that binds the TOP remaining index bit each round
// (sumcheck_round_eval_par / sumcheck_bind_both_and_eval_next / sumcheck_bind_top_in_place_par,
// :1479-1612) down to 64 words.
//
// Partial fold layout (the lincheck stripe, lincheck.rs:1128-1199): byte z[s*K + i] holds
// the outer rows 8s..8s+7 of column i. A workgroup takes a tile of 1024 columns (four per
// thread, one u32 load per stripe) and a group of stripes; per stripe it keeps two nibble
// tables T_lo[n] = sum of eq_outer[8s + b] over the bits b of n (b < 4) and T_hi for bits
// 4..7. One table is 16 x 16 B = 256 B, all 64 LDS banks once, so the 64 lanes of a wave
// read it without bank conflicts. Each byte costs two lookups and two XORs.
#include "f128_device.h"
#include "flock_internal.h"

namespace flock {

// ---- CSC fold (CscCircuit::fold_split / fold_alpha_batched, lincheck.rs:432-497) ----
// One wave per column. Row indexes must be < n_cols and col_ptr entries <= nnz; a
// violation sets *err and the offending entries are skipped.
// XOR of eq over the rows rows[k0..k1) of one column, lanes striding by 64. Four independent
// index loads (each a coalesced 256-byte segment) and four gathers are in flight per lane per
// step, so a column's walk is not a chain of dependent round trips.
__device__ __forceinline__ f128 csc_column_sum(const uint32_t* __restrict__ rows, uint64_t k0,
                                               uint64_t k1, int lane, uint64_t n_cols,
                                               const f128* __restrict__ eq, bool& bad) {
    f128 s0 = f128_zero(), s1 = f128_zero();
    uint64_t k = k0 + lane;
    for (; k + 192 < k1; k += 256) {
        const uint32_t r0 = rows[k], r1 = rows[k + 64], r2 = rows[k + 128], r3 = rows[k + 192];
        const bool ok = r0 < n_cols && r1 < n_cols && r2 < n_cols && r3 < n_cols;
        if (ok) {
            s0 = f128_add3(s0, eq[r0], eq[r1]);
            s1 = f128_add3(s1, eq[r2], eq[r3]);
        } else {
            bad = true;
        }
    }
    for (; k < k1; k += 64) {
        const uint32_t r = rows[k];
        if (r < n_cols) s0 = f128_add(s0, eq[r]);
        else bad = true;
    }
    return f128_add(s0, s1);
}

__global__ __launch_bounds__(256) void k_lc_csc_fold(
    uint64_t n_cols, const uint32_t* __restrict__ a_ptr, const uint32_t* __restrict__ a_rows,
    uint64_t nnz_a, const uint32_t* __restrict__ b_ptr, const uint32_t* __restrict__ b_rows,
    uint64_t nnz_b, const f128* __restrict__ eq, f128 alpha, f128* __restrict__ sa_out,
    f128* __restrict__ sb_out, f128* __restrict__ comb_out, int* __restrict__ err) {
    __shared__ f128 table[SHOUP_ENTRIES];
    if (threadIdx.x < 32) shoup_build_row(table, alpha, threadIdx.x);
    __syncthreads();
    // one wave per column: the lanes stride over its nonzeros (coalesced row indices, eq
    // gathers from a table that stays in L2), then XOR-reduce across the wave
    const int lane = threadIdx.x & 63;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t c = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < n_cols;
         c += nwaves) {
        f128 sa = f128_zero(), sb = f128_zero();
        bool bad = false;
        uint64_t k0 = a_ptr[c], k1 = a_ptr[c + 1];
        if (k1 > nnz_a) {
            k1 = nnz_a;
            bad = true;
        }
        sa = csc_column_sum(a_rows, k0, k1, lane, n_cols, eq, bad);
        k0 = b_ptr[c];
        k1 = b_ptr[c + 1];
        if (k1 > nnz_b) {
            k1 = nnz_b;
            bad = true;
        }
        sb = csc_column_sum(b_rows, k0, k1, lane, n_cols, eq, bad);
        sa = wave_xor_f128(sa);
        sb = wave_xor_f128(sb);
        const bool any_bad = __any(bad);
        if (lane == 0) {
            if (any_bad) atomicOr(err, 1);
            if (sa_out) sa_out[c] = sa;
            if (sb_out) sb_out[c] = sb;
            if (comb_out) comb_out[c] = f128_add(shoup_mul(table, sa), sb);
        }
    }
}

// comb[pin] += beta (the constant-wire pin, lincheck/union.rs:318-335).
__global__ void k_lc_add_at(f128* __restrict__ v, uint64_t idx, f128 beta) {
    if (threadIdx.x == 0 && blockIdx.x == 0) v[idx] = f128_add(v[idx], beta);
}

// v[c] = w * v[c]: a union slot's comb scaled by its prefix weight (eq_prefix_weight,
// lincheck/union.rs:90-101).
__global__ __launch_bounds__(256) void k_lc_scale(f128* __restrict__ v, f128 w, uint64_t n) {
    __shared__ f128 table[SHOUP_ENTRIES];
    if (threadIdx.x < 32) shoup_build_row(table, w, threadIdx.x);
    __syncthreads();
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (uint64_t)gridDim.x * blockDim.x)
        v[c] = shoup_mul(table, v[c]);
}

__global__ __launch_bounds__(256) void k_fill_i32_zero(int* p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// ---- stripe partial fold ----
constexpr int kPfCols = 1024;   // columns per workgroup (4 per thread)
constexpr int kPfMaxSpg = 64;   // stripes per group: 64 x 512 B of tables = 32 KiB

// Grid (tiles, gy): workgroup (x, y) folds stripe groups y, y + gy, ... (spg stripes each)
// into out[y*K + ..] for its 1024 columns.
__global__ __launch_bounds__(256) void k_lc_partial_fold(const uint8_t* __restrict__ z, uint64_t K,
                                                          uint64_t useful_bits,
                                                          uint64_t n_stripes, uint32_t spg,
                                                          uint64_t n_groups,
                                                          const f128* __restrict__ eq_outer,
                                                          f128* __restrict__ out) {
    __shared__ f128 tabs[kPfMaxSpg * 32];
    const uint64_t i0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    const bool live = i0 < K && i0 < useful_bits;
    f128 a0 = f128_zero(), a1 = f128_zero(), a2 = f128_zero(), a3 = f128_zero();
    for (uint64_t g = blockIdx.y; g < n_groups; g += gridDim.y) {
        const uint64_t s0 = g * spg;
        uint64_t s1 = s0 + spg;
        if (s1 > n_stripes) s1 = n_stripes;
        const int ns = (int)(s1 - s0);
        __syncthreads();  // the previous group's tables are no longer read
        for (int e = threadIdx.x; e < ns * 32; e += blockDim.x) {
            const int sl = e >> 5, h = (e >> 4) & 1, n = e & 15;
            const f128* q = eq_outer + 8 * (s0 + sl) + 4 * h;
            f128 acc = f128_zero();
#pragma unroll
            for (int b = 0; b < 4; b++)
                if ((n >> b) & 1) acc = f128_add(acc, q[b]);
            tabs[e] = acc;
        }
        __syncthreads();
        if (live) {
            const uint8_t* zp = z + s0 * K + i0;
            for (int sl = 0; sl < ns; sl++, zp += K) {
                const uint32_t w = *reinterpret_cast<const uint32_t*>(zp);
                const f128* T = tabs + sl * 32;
                a0 = f128_add(a0, f128_add(T[w & 15u], T[16 + ((w >> 4) & 15u)]));
                a1 = f128_add(a1, f128_add(T[(w >> 8) & 15u], T[16 + ((w >> 12) & 15u)]));
                a2 = f128_add(a2, f128_add(T[(w >> 16) & 15u], T[16 + ((w >> 20) & 15u)]));
                a3 = f128_add(a3, f128_add(T[(w >> 24) & 15u], T[16 + (w >> 28)]));
            }
        }
    }
    if (i0 >= K) return;
    if (i0 + 1 >= useful_bits) a1 = f128_zero();
    if (i0 + 2 >= useful_bits) a2 = f128_zero();
    if (i0 + 3 >= useful_bits) a3 = f128_zero();
    f128* o = out + (uint64_t)blockIdx.y * K + i0;
    o[0] = a0;
    o[1] = a1;
    o[2] = a2;
    o[3] = a3;
}

// out[i] = XOR over g < groups of partials[g*K + i]
__global__ __launch_bounds__(256) void k_lc_xor_rows(const f128* __restrict__ partials, int groups,
                                                      uint64_t K, f128* __restrict__ out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < K;
         i += (uint64_t)gridDim.x * blockDim.x) {
        f128 acc = f128_zero();
        for (int g = 0; g < groups; g++) acc = f128_add(acc, partials[(uint64_t)g * K + i]);
        out[i] = acc;
    }
}

__global__ __launch_bounds__(256) void k_fill_f128_zero(f128* __restrict__ out, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = f128_zero();
}

// ---- top-bit product sumcheck ----
// (e1, einf) = (sum c_hi z_hi, sum (c_hi + c_lo)(z_hi + z_lo)) over i < half.
__global__ __launch_bounds__(256) void k_lc_eval(const f128* __restrict__ c,
                                                  const f128* __restrict__ z, uint64_t half,
                                                  f128* __restrict__ partials,
                                                  RoundOut ro) {
    extern __shared__ f128 ws[];
    f128* wl = comb_lane(ws);
    f256u e1, ei;
    f256u_zero(e1);
    f256u_zero(ei);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < half;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const f128 cl = c[i], ch = c[half + i], zl = z[i], zh = z[half + i];
        comb_apply(e1, wl, comb_build(wl, zh), ch);
        comb_apply(ei, wl, comb_build(wl, f128_add(zh, zl)), f128_add(ch, cl));
    }
    grid_xor_pair(f256u_reduce(e1), f256u_reduce(ei), partials, ro);
}

// Bind the top bit of c and z at r in place (len -> len/2) and evaluate the next round on the
// bound tables. Thread i reads quarters i, q+i, 2q+i, 3q+i and writes i and q+i only.
__global__ __launch_bounds__(256) void k_lc_bind_eval(f128* __restrict__ c, f128* __restrict__ z,
                                                       uint64_t q, f128 r,
                                                       const f128* __restrict__ rin,
                                                       f128* __restrict__ partials,
                                                       RoundOut ro) {
    __shared__ f128 table[SHOUP_ENTRIES];
    extern __shared__ f128 ws[];
    f128* wl = comb_lane(ws);
    if (rin) r = *rin;
    if (threadIdx.x < 32) shoup_build_row(table, r, threadIdx.x);
    __syncthreads();
    f256u e1, ei;
    f256u_zero(e1);
    f256u_zero(ei);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < q;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const f128 c0 = c[i], c1 = c[q + i], c2 = c[2 * q + i], c3 = c[3 * q + i];
        const f128 z0 = z[i], z1 = z[q + i], z2 = z[2 * q + i], z3 = z[3 * q + i];
        const f128 lo = f128_add(c0, shoup_mul(table, f128_add(c2, c0)));
        const f128 hi = f128_add(c1, shoup_mul(table, f128_add(c3, c1)));
        const f128 zlo = f128_add(z0, shoup_mul(table, f128_add(z2, z0)));
        const f128 zhi = f128_add(z1, shoup_mul(table, f128_add(z3, z1)));
        c[i] = lo;
        c[q + i] = hi;
        z[i] = zlo;
        z[q + i] = zhi;
        comb_apply(e1, wl, comb_build(wl, zhi), hi);
        comb_apply(ei, wl, comb_build(wl, f128_add(zhi, zlo)), f128_add(hi, lo));
    }
    grid_xor_pair(f256u_reduce(e1), f256u_reduce(ei), partials, ro);
}

// Final binding: v[i] += r (v[half + i] + v[i]) for c and z, in place.
__global__ __launch_bounds__(256) void k_lc_bind(f128* __restrict__ c, f128* __restrict__ z,
                                                  uint64_t half, f128 r,
                                                  const f128* __restrict__ rin) {
    __shared__ f128 table[SHOUP_ENTRIES];
    if (rin) r = *rin;
    if (threadIdx.x < 32) shoup_build_row(table, r, threadIdx.x);
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < half;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const f128 c0 = c[i], z0 = z[i];
        c[i] = f128_add(c0, shoup_mul(table, f128_add(c[half + i], c0)));
        z[i] = f128_add(z0, shoup_mul(table, f128_add(z[half + i], z0)));
    }
}

// (sum w a, sum w b)
__global__ __launch_bounds__(256) void k_dot2(const f128* __restrict__ w, const f128* __restrict__ a,
                                               const f128* __restrict__ b, uint64_t n,
                                               f128* __restrict__ partials,
        f128* __restrict__ out) {
    extern __shared__ f128 ws[];
    f128* wl = comb_lane(ws);
    f256u sa, sb;
    f256u_zero(sa);
    f256u_zero(sb);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t wh = comb_build(wl, w[i]);
        comb_apply(sa, wl, wh, a[i]);
        comb_apply(sb, wl, wh, b[i]);
    }
    grid_xor_pair(f256u_reduce(sa), f256u_reduce(sb), partials, round_out(out));
}

// ---- Keccak-f[1600] circuit marginals (KeccakLincheckCircuit, r1cs_hashes/keccak.rs:692-835) ----
// sa = A_0^T eq and sb = B_0^T eq over the 2^16 columns without forming A_0, B_0. state_r
// (r >= 1) is the linear form phi(state_{r-1}) + t_{r-1} + RC_{r-1} with phi = pi o rho o
// theta, so the coefficient vector on state_r is carried from r = 24 down to 0,
//   K_{r-1} = phi^T(K_r) + chi[r-1],
// and K_r lands on the t_{r-1} columns (K_0 on state_0). chi[r][s] sums eq over the t rows of
// round r whose chi operand (A: B[x+1], B: B[x+2]) reads state_r bit s; the constant column
// collects the NOT of the A operand, the propagated round constants and the B = [1] sides.
// Workgroup 0 walks the A side (K_24 = the pin rows' eq), workgroup 1 the B side (K_24 = 0);
// 1024 threads hold the 1600 state coefficients in LDS.
//   inv[s][i]: the 11 outputs s_out of phi whose preimage holds s (phi^T gathers);
//   trow[side][s][i]: column offset within a round's t block of the row of chi[.][s], term i;
//   rcm[r][w]: bit s of RC_r, the constant vector of state_r, r <= 24 (50 words per r).
constexpr int kKcStateBits = 1600;
constexpr uint32_t kKcZc = 4096, kKcS24 = 2048, kKcT = 4160;

__device__ __forceinline__ uint32_t kc_lane_pos(int s) { return 64u * (uint32_t)(s % 25) + (uint32_t)(s / 25); }

__global__ __launch_bounds__(1024) void k_kc_walk(const f128* __restrict__ eq,
                                                   const uint16_t* __restrict__ inv,
                                                   const uint16_t* __restrict__ trow,
                                                   const uint32_t* __restrict__ rcm,
                                                   f128* __restrict__ sa, f128* __restrict__ sb) {
    __shared__ f128 kv[2][kKcStateBits];  // 50 KiB ping-pong
    __shared__ f128 red[16];
    const int side = blockIdx.x;  // 0: A, 1: B
    f128* out = side ? sb : sa;
    const uint16_t* tr = trow + (size_t)side * kKcStateBits * 11;
    const int t = threadIdx.x;
    f128 zc = f128_zero();  // this thread's share of the constant column
    for (int s = t; s < kKcStateBits; s += blockDim.x) {
        const f128 pin = eq[kKcS24 + kc_lane_pos(s)];
        if (side == 0) {
            kv[0][s] = pin;                              // K_24
            out[kKcT + 23u * 1600u + kc_lane_pos(s)] = pin;  // t_23 takes K_24
            if ((rcm[24 * 50 + (s >> 5)] >> (s & 31)) & 1u) zc = f128_add(zc, pin);  // RC_24 . pin
            for (int r = 0; r < 24; r++)  // NOT of every chi A operand
                zc = f128_add(zc, eq[kKcT + 1600u * r + kc_lane_pos(s)]);
        } else {
            kv[0][s] = f128_zero();
            zc = f128_add3(zc, pin, eq[kc_lane_pos(s)]);  // pin and state_0 rows: B = [1]
        }
    }
    if (t == 0) zc = f128_add(zc, eq[kKcZc]);  // the constant row
    __syncthreads();
    int cur = 0;
    for (int r = 24; r >= 1; r--) {
        for (int s = t; s < kKcStateBits; s += blockDim.x) {
            const uint16_t* iv = inv + s * 11;
            const uint16_t* tv = tr + s * 11;
            const f128* erow = eq + kKcT + 1600u * (r - 1);
            f128 v = f128_zero(), c = f128_zero();
#pragma unroll
            for (int i = 0; i < 11; i++) {
                v = f128_add(v, kv[cur][iv[i]]);
                c = f128_add(c, erow[tv[i]]);
            }
            if ((rcm[(r - 1) * 50 + (s >> 5)] >> (s & 31)) & 1u) zc = f128_add(zc, c);
            const f128 k = f128_add(v, c);  // K_{r-1}
            kv[cur ^ 1][s] = k;
            if (r >= 2) {
                out[kKcT + 1600u * (r - 2) + kc_lane_pos(s)] = k;
            } else {
                const uint32_t p = kc_lane_pos(s);
                out[p] = side == 0 ? f128_add(k, eq[p]) : k;  // state_0 self-loop on A
            }
        }
        __syncthreads();
        cur ^= 1;
    }
    zc = wave_xor_f128(zc);
    if ((t & 63) == 0) red[t >> 6] = zc;
    __syncthreads();
    if (t == 0) {
        f128 v = f128_zero();
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) v = f128_add(v, red[w]);
        out[kKcZc] = v;
    }
}

// comb = alpha * sa + sb
__global__ __launch_bounds__(256) void k_lc_combine(const f128* __restrict__ sa,
                                                     const f128* __restrict__ sb, f128 alpha,
                                                     uint64_t n, f128* __restrict__ comb) {
    __shared__ f128 table[SHOUP_ENTRIES];
    if (threadIdx.x < 32) shoup_build_row(table, alpha, threadIdx.x);
    __syncthreads();
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (uint64_t)gridDim.x * blockDim.x)
        comb[c] = f128_add(shoup_mul(table, sa[c]), sb[c]);
}

static hipError_t lc_attrs() {
    static hipError_t st = [] {
        hipError_t e = hipFuncSetAttribute((const void*)k_lc_eval,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           kCombWsBytes);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)k_lc_bind_eval,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kCombWsBytes);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)k_dot2,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kCombWsBytes);
        return e;
    }();
    return st;
}

static inline unsigned grid_for(uint64_t work, uint64_t cap) {
    uint64_t g = (work + 255) / 256;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

hipError_t launch_lc_csc_fold(uint64_t n_cols, const uint32_t* a_ptr, const uint32_t* a_rows,
                              uint64_t nnz_a, const uint32_t* b_ptr, const uint32_t* b_rows,
                              uint64_t nnz_b, const uint4* eq, uint4 alpha, uint4* sa, uint4* sb,
                              uint4* comb, int* err, hipStream_t s) {
    k_fill_i32_zero<<<1, 64, 0, s>>>(err);
    // four columns (waves) per workgroup
    k_lc_csc_fold<<<grid_for(n_cols * 64, 65535), 256, 0, s>>>(n_cols, a_ptr, a_rows, nnz_a, b_ptr,
                                                          b_rows, nnz_b, eq, alpha, sa, sb, comb,
                                                          err);
    return hipGetLastError();
}

hipError_t launch_kc_fold(const uint4* eq, const uint16_t* inv, const uint16_t* trow,
                          const uint32_t* rcm, uint4 alpha, uint4* sa, uint4* sb, uint4* comb,
                          hipStream_t s) {
    constexpr uint64_t K = 1ull << 16;
    k_fill_f128_zero<<<grid_for(K, 4096), 256, 0, s>>>(sa, K);
    k_fill_f128_zero<<<grid_for(K, 4096), 256, 0, s>>>(sb, K);
    k_kc_walk<<<2, 1024, 0, s>>>(eq, inv, trow, rcm, sa, sb);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !comb) return e;
    k_lc_combine<<<grid_for(K, 1024), 256, 0, s>>>(sa, sb, alpha, K, comb);
    return hipGetLastError();
}

hipError_t launch_lc_add_at(uint4* v, uint64_t idx, uint4 beta, hipStream_t s) {
    k_lc_add_at<<<1, 64, 0, s>>>(v, idx, beta);
    return hipGetLastError();
}

hipError_t launch_lc_scale(uint4* v, uint64_t n, uint4 w, hipStream_t s) {
    k_lc_scale<<<grid_for(n, 1024), 256, 0, s>>>(v, w, n);
    return hipGetLastError();
}

// Stripes per group (<= 64, so the tables fit 32 KiB) and the grid's y extent (<= 64 partial
// rows of K words): about 1024 workgroups in all.
uint32_t lc_partial_fold_rows(uint64_t K) {
    const uint64_t tiles = (K + kPfCols - 1) / kPfCols;
    uint64_t want = (1024 + tiles - 1) / tiles;
    return (uint32_t)(want > 64 ? 64 : want);
}

void lc_partial_fold_plan(uint64_t K, uint64_t n_stripes, uint32_t* spg, uint64_t* n_groups,
                          uint32_t* gy) {
    const uint64_t want = lc_partial_fold_rows(K);
    uint64_t per = (n_stripes + want - 1) / want;
    if (per < 8) per = 8;
    if (per > (uint64_t)kPfMaxSpg) per = kPfMaxSpg;
    *spg = (uint32_t)per;
    *n_groups = (n_stripes + per - 1) / per;
    *gy = (uint32_t)(*n_groups < want ? *n_groups : want);
}

hipError_t launch_lc_partial_fold(const uint8_t* z, uint64_t K, uint64_t useful_bits,
                                  uint64_t n_stripes, const uint4* eq_outer, uint4* partials,
                                  uint4* out, hipStream_t s) {
    if (n_stripes == 0) {
        k_fill_f128_zero<<<grid_for(K, 4096), 256, 0, s>>>(out, K);
        return hipGetLastError();
    }
    uint32_t spg = 0, gy = 0;
    uint64_t n_groups = 0;
    lc_partial_fold_plan(K, n_stripes, &spg, &n_groups, &gy);
    const uint64_t tiles = (K + kPfCols - 1) / kPfCols;
    dim3 grid((unsigned)tiles, gy);
    k_lc_partial_fold<<<grid, 256, 0, s>>>(z, K, useful_bits, n_stripes, spg, n_groups, eq_outer,
                                           gy == 1 ? out : partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || gy == 1) return e;
    k_lc_xor_rows<<<grid_for(K, 4096), 256, 0, s>>>(partials, (int)gy, K, out);
    return hipGetLastError();
}

hipError_t launch_lc_eval(const uint4* c, const uint4* z, uint64_t len, uint4* partials,
                          RoundOut ro, hipStream_t s) {
    if (hipError_t ae = lc_attrs(); ae != hipSuccess) return ae;
    const unsigned g = grid_for(len / 2, kMaxReduceBlocks);
    k_lc_eval<<<g, 256, kCombWsBytes, s>>>(c, z, len / 2, partials, ro);
    return hipGetLastError();
}

hipError_t launch_lc_bind_eval(uint4* c, uint4* z, uint64_t len, uint4 r, uint4* partials,
                               RoundOut ro, hipStream_t s, const uint4* rin) {
    if (hipError_t ae = lc_attrs(); ae != hipSuccess) return ae;
    const unsigned g = grid_for(len / 4, kMaxReduceBlocks);
    k_lc_bind_eval<<<g, 256, kCombWsBytes, s>>>(c, z, len / 4, r, rin, partials, ro);
    return hipGetLastError();
}

hipError_t launch_lc_bind(uint4* c, uint4* z, uint64_t len, uint4 r, hipStream_t s,
                          const uint4* rin) {
    k_lc_bind<<<grid_for(len / 2, 8192), 256, 0, s>>>(c, z, len / 2, r, rin);
    return hipGetLastError();
}

hipError_t launch_dot2(const uint4* w, const uint4* a, const uint4* b, uint64_t n,
                       uint4* partials, uint4* out, hipStream_t s) {
    if (hipError_t ae = lc_attrs(); ae != hipSuccess) return ae;
    const unsigned g = grid_for(n, kMaxReduceBlocks);
    k_dot2<<<g, 256, kCombWsBytes, s>>>(w, a, b, n, partials, out);
    return hipGetLastError();
}

}  // namespace flock
