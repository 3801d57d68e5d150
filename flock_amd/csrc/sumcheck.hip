// Sumcheck round kernels: the product sumcheck of the Ligerito opening and the
// eq-weighted zerocheck tail.
//
// Product sumcheck (crates/flock-core/src/pcs/ligerito.rs):
//   round_msg_lsb  (:3204-3240)  u0 = sum f0*b0, u2 = sum (f0+f1)(b0+b1)
//   fold_and_msg_lsb (:3372-3414) v' = v0 + r(v0+v1), then the next message on v'
// Zerocheck tail (crates/flock-core/src/zerocheck/multilinear.rs):
//   round_pair_naive (:331-355)  (G(1), G(inf)) = (sum eq*a1*b1, sum eq*(a0+a1)(b0+b1))
//   fold_in_place_pair (:1148-1163)
//
// Sums are exact in GF(2^128) and order-free, so every kernel accumulates
// unreduced 256-bit products per thread (one reduction per thread), XOR-reduces across
// the wave with shuffles and across the block in LDS, and writes one partial pair per
// block; k_xor_reduce folds the partials. Fixed multipliers (the fold challenge) go
// through an LDS Shoup table; products of two varying values use f128_mul.
//
// The eq table of round i+1 is the pairwise XOR of round i's: eq(r', (0,x)) +
// eq(r', (1,x)) = ((1+r0) + r0) eq(r'', x) = eq(r'', x), so each message kernel also
// emits the next round's table (n/4 -> n/8 entries) instead of rebuilding it.
#include <cstring>
#include <mutex>

#include "../../include/flock_hip.h"
#include "f128_device.h"
#include "flock_internal.h"

namespace flock {

static_assert(kMaxReduceBlocks == kGridReduceMax, "grid_xor_pair's counter sits after the partials");

__global__ __launch_bounds__(256) void k_sc_round_msg(const f128* __restrict__ f,
                                                       const f128* __restrict__ b, uint64_t n,
                                                       f128* __restrict__ partials,
                                                       RoundOut ro) {
    extern __shared__ f128 ws[];
    f128* wl = comb_lane(ws);
    f256u u0, u2;
    f256u_zero(u0);
    f256u_zero(u2);
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n / 2;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const f128 f0 = f[2 * j], f1 = f[2 * j + 1], b0 = b[2 * j], b1 = b[2 * j + 1];
        comb_apply(u0, wl, comb_build(wl, b0), f0);
        comb_apply(u2, wl, comb_build(wl, f128_add(b0, b1)), f128_add(f0, f1));
    }
    grid_xor_pair(f256u_reduce(u0), f256u_reduce(u2), partials, ro);
}

__global__ __launch_bounds__(256) void k_sc_fold_msg(const f128* __restrict__ f,
                                                      const f128* __restrict__ b, uint64_t n,
                                                      f128 r, const f128* __restrict__ rin,
                                                      f128* __restrict__ nf,
                                                      f128* __restrict__ nb,
                                                      f128* __restrict__ partials,
                                                      RoundOut ro) {
    __shared__ f128 table[SHOUP_ENTRIES];
    extern __shared__ f128 ws[];
    f128* wl = comb_lane(ws);
    if (rin) r = *rin;
    if (threadIdx.x < 32) shoup_build_row(table, r, threadIdx.x);
    __syncthreads();
    const uint64_t half = n / 2;
    f256u u0, u2;
    f256u_zero(u0);
    f256u_zero(u2);
    if (half == 1) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            nf[0] = f128_add(f[0], shoup_mul(table, f128_add(f[0], f[1])));
            nb[0] = f128_add(b[0], shoup_mul(table, f128_add(b[0], b[1])));
        }
    } else {
        for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < half / 2;
             k += (uint64_t)gridDim.x * blockDim.x) {
            const f128 f0 = f[4 * k], f1 = f[4 * k + 1], f2 = f[4 * k + 2], f3 = f[4 * k + 3];
            const f128 b0 = b[4 * k], b1 = b[4 * k + 1], b2 = b[4 * k + 2], b3 = b[4 * k + 3];
            const f128 g0 = f128_add(f0, shoup_mul(table, f128_add(f0, f1)));
            const f128 g1 = f128_add(f2, shoup_mul(table, f128_add(f2, f3)));
            const f128 c0 = f128_add(b0, shoup_mul(table, f128_add(b0, b1)));
            const f128 c1 = f128_add(b2, shoup_mul(table, f128_add(b2, b3)));
            nf[2 * k] = g0;
            nf[2 * k + 1] = g1;
            nb[2 * k] = c0;
            nb[2 * k + 1] = c1;
            comb_apply(u0, wl, comb_build(wl, c0), g0);
            comb_apply(u2, wl, comb_build(wl, f128_add(c0, c1)), f128_add(g0, g1));
        }
    }
    grid_xor_pair(f256u_reduce(u0), f256u_reduce(u2), partials, ro);
}

__global__ __launch_bounds__(256) void k_zc_round(const f128* __restrict__ a,
                                                   const f128* __restrict__ b, uint64_t n,
                                                   const f128* __restrict__ e,
                                                   f128* __restrict__ e_next,
                                                   f128* __restrict__ partials,
                                                   RoundOut ro) {
    extern __shared__ f128 ws[];
    f128* wl = comb_lane(ws);
    f256u g1, gi;
    f256u_zero(g1);
    f256u_zero(gi);
    const uint64_t half = n / 2;
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < half;
         x += (uint64_t)gridDim.x * blockDim.x) {
        const f128 a0 = a[2 * x], a1 = a[2 * x + 1], b0 = b[2 * x], b1 = b[2 * x + 1];
        const f128 ex = e[x];
        const f128 p1 = comb_mul(wl, a1, b1);
        const f128 pi = comb_mul(wl, f128_add(a0, a1), f128_add(b0, b1));
        const uint32_t eh = comb_build(wl, ex);
        comb_apply(g1, wl, eh, p1);
        comb_apply(gi, wl, eh, pi);
        if (e_next && !(x & 1) && x + 1 < half) e_next[x >> 1] = f128_add(ex, e[x + 1]);
    }
    grid_xor_pair(f256u_reduce(g1), f256u_reduce(gi), partials, ro);
}

__global__ __launch_bounds__(256) void k_zc_fold_round(const f128* __restrict__ a,
                                                        const f128* __restrict__ b, uint64_t n,
                                                        f128 rho, const f128* __restrict__ rin,
                                                        f128* __restrict__ ao,
                                                        f128* __restrict__ bo,
                                                        const f128* __restrict__ e,
                                                        f128* __restrict__ e_next,
                                                        f128* __restrict__ partials,
                                                        RoundOut ro) {
    __shared__ f128 table[SHOUP_ENTRIES];
    extern __shared__ f128 ws[];
    f128* wl = comb_lane(ws);
    if (rin) rho = *rin;
    if (threadIdx.x < 32) shoup_build_row(table, rho, threadIdx.x);
    __syncthreads();
    f256u g1, gi;
    f256u_zero(g1);
    f256u_zero(gi);
    const uint64_t quarter = n / 4;
    for (uint64_t y = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; y < quarter;
         y += (uint64_t)gridDim.x * blockDim.x) {
        const f128 a0 = a[4 * y], a1 = a[4 * y + 1], a2 = a[4 * y + 2], a3 = a[4 * y + 3];
        const f128 b0 = b[4 * y], b1 = b[4 * y + 1], b2 = b[4 * y + 2], b3 = b[4 * y + 3];
        const f128 p0 = f128_add(a0, shoup_mul(table, f128_add(a0, a1)));
        const f128 p1 = f128_add(a2, shoup_mul(table, f128_add(a2, a3)));
        const f128 q0 = f128_add(b0, shoup_mul(table, f128_add(b0, b1)));
        const f128 q1 = f128_add(b2, shoup_mul(table, f128_add(b2, b3)));
        ao[2 * y] = p0;
        ao[2 * y + 1] = p1;
        bo[2 * y] = q0;
        bo[2 * y + 1] = q1;
        const f128 ey = e[y];
        const f128 m1 = comb_mul(wl, p1, q1);
        const f128 mi = comb_mul(wl, f128_add(p0, p1), f128_add(q0, q1));
        const uint32_t eh = comb_build(wl, ey);
        comb_apply(g1, wl, eh, m1);
        comb_apply(gi, wl, eh, mi);
        if (e_next && !(y & 1) && y + 1 < quarter) e_next[y >> 1] = f128_add(ey, e[y + 1]);
    }
    grid_xor_pair(f256u_reduce(g1), f256u_reduce(gi), partials, ro);
}

__global__ __launch_bounds__(256) void k_fold_pair(const f128* __restrict__ a,
                                                    const f128* __restrict__ b, uint64_t n, f128 c,
                                                    const f128* __restrict__ cin,
                                                    f128* __restrict__ ao, f128* __restrict__ bo) {
    __shared__ f128 table[SHOUP_ENTRIES];
    if (cin) c = *cin;
    if (threadIdx.x < 32) shoup_build_row(table, c, threadIdx.x);
    __syncthreads();
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n / 2;
         x += (uint64_t)gridDim.x * blockDim.x) {
        const f128 a0 = a[2 * x], a1 = a[2 * x + 1], b0 = b[2 * x], b1 = b[2 * x + 1];
        ao[x] = f128_add(a0, shoup_mul(table, f128_add(a0, a1)));
        bo[x] = f128_add(b0, shoup_mul(table, f128_add(b0, b1)));
    }
}

// ---------------------------------------------------------------------------
// Split-eq zerocheck rounds. For a round whose eq table has nv >= 16 variables,
// eq(x) = e_lo[x & 255] * e_hi[x >> 8]. Thread t of a 256-thread workgroup keeps x_lo = t and
// the workgroup walks x_hi together, so e_hi[x_hi] is one multiplier per step: its Shoup table
// is built once per step by the whole workgroup, and each thread accumulates
// sum_{x_hi} e_hi[x_hi] * P(x) with table lookups, then multiplies by e_lo[t] once at the end.
// Per pair this replaces a per-lane comb table of eq(x) and two comb products with two Shoup
// products. The split tables come from k_eq_windows / k_eq_levels.
// ---------------------------------------------------------------------------

// The Shoup table of t (SHOUP_ENTRIES words) by all 256 threads of the workgroup: row = t / 8,
// two entries per thread.
__device__ __forceinline__ void shoup_build_block(f128* table, f128 t) {
    const int row = threadIdx.x >> 3, e = (threadIdx.x & 7) * 2;
    const f128 e1 = f128_shl_mod(t, 4 * row);
    const f128 e2 = f128_mul_x(e1);
    const f128 e4 = f128_mul_x(e2);
    const f128 e8 = f128_mul_x(e4);
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int n = e + k;
        f128 v = f128_zero();
        if (n & 1) v = f128_add(v, e1);
        if (n & 2) v = f128_add(v, e2);
        if (n & 4) v = f128_add(v, e4);
        if (n & 8) v = f128_add(v, e8);
        shoup_put(table, 16 * row + n, v);
    }
}

// Message of (a, b) (n = 2 * half words each): (sum eq*a1*b1, sum eq*(a0+a1)(b0+b1)),
// half = 256 * n_hi.
__global__ __launch_bounds__(256) void k_zc_round_split(const f128* __restrict__ a,
                                                         const f128* __restrict__ b,
                                                         uint64_t half,
                                                         const f128* __restrict__ e_lo,
                                                         const f128* __restrict__ e_hi,
                                                         f128* __restrict__ partials,
                                                         RoundOut ro) {
    // one 8 KiB table, so that the workgroup (with the comb workspace) stays under 80 KiB of
    // LDS, two per CU; two barriers per step
    __shared__ f128 tab[SHOUP_ENTRIES];
    extern __shared__ f128 ws[];
    f128* wl = comb_lane(ws);
    f128 s1 = f128_zero(), si = f128_zero();
    const uint64_t n_hi = half >> 8;
    for (uint64_t xh = blockIdx.x; xh < n_hi; xh += gridDim.x) {
        __syncthreads();  // the previous step's lookups into tab are done
        shoup_build_block(tab, e_hi[xh]);
        __syncthreads();
        const uint64_t x = (xh << 8) + threadIdx.x;
        const f128 a0 = a[2 * x], a1 = a[2 * x + 1], b0 = b[2 * x], b1 = b[2 * x + 1];
        const f128 p1 = comb_mul(wl, a1, b1);
        const f128 pi = comb_mul(wl, f128_add(a0, a1), f128_add(b0, b1));
        s1 = f128_add(s1, shoup_mul(tab, p1));
        si = f128_add(si, shoup_mul(tab, pi));
    }
    const uint32_t eh = comb_build(wl, e_lo[threadIdx.x]);
    f256u g1, gi;
    f256u_zero(g1);
    f256u_zero(gi);
    comb_apply(g1, wl, eh, s1);
    comb_apply(gi, wl, eh, si);
    grid_xor_pair(f256u_reduce(g1), f256u_reduce(gi), partials, ro);
}

// Fused fold at rho (n -> n/2) and the split-eq message on the folded tables; quarter = n / 4
// = 256 * n_hi output pairs. Static LDS is the rho table and the e_hi table only (16 KiB), and
// the final reduction runs in the comb workspace once it is free: with the 64 KiB workspace the
// workgroup takes exactly 80 KiB, two per CU.
__global__ __launch_bounds__(256) void k_zc_fold_round_split(const f128* __restrict__ a,
                                                              const f128* __restrict__ b,
                                                              uint64_t quarter, f128 rho,
                                                              const f128* __restrict__ rin,
                                                              f128* __restrict__ ao,
                                                              f128* __restrict__ bo,
                                                              const f128* __restrict__ e_lo,
                                                              const f128* __restrict__ e_hi,
                                                              f128* __restrict__ partials,
                                                              RoundOut ro) {
    __shared__ f128 rt[SHOUP_ENTRIES];
    __shared__ f128 tab[SHOUP_ENTRIES];
    extern __shared__ f128 ws[];
    f128* wl = comb_lane(ws);
    if (rin) rho = *rin;
    if (threadIdx.x < 32) shoup_build_row(rt, rho, threadIdx.x);
    f128 s1 = f128_zero(), si = f128_zero();
    const uint64_t n_hi = quarter >> 8;
    for (uint64_t yh = blockIdx.x; yh < n_hi; yh += gridDim.x) {
        __syncthreads();  // the previous step's lookups into tab are done (and rt is built)
        shoup_build_block(tab, e_hi[yh]);
        __syncthreads();
        const uint64_t y = (yh << 8) + threadIdx.x;
        const f128 a0 = a[4 * y], a1 = a[4 * y + 1], a2 = a[4 * y + 2], a3 = a[4 * y + 3];
        const f128 b0 = b[4 * y], b1 = b[4 * y + 1], b2 = b[4 * y + 2], b3 = b[4 * y + 3];
        const f128 p0 = f128_add(a0, shoup_mul(rt, f128_add(a0, a1)));
        const f128 p1 = f128_add(a2, shoup_mul(rt, f128_add(a2, a3)));
        const f128 q0 = f128_add(b0, shoup_mul(rt, f128_add(b0, b1)));
        const f128 q1 = f128_add(b2, shoup_mul(rt, f128_add(b2, b3)));
        ao[2 * y] = p0;
        ao[2 * y + 1] = p1;
        bo[2 * y] = q0;
        bo[2 * y + 1] = q1;
        const f128 m1 = comb_mul(wl, p1, q1);
        const f128 mi = comb_mul(wl, f128_add(p0, p1), f128_add(q0, q1));
        s1 = f128_add(s1, shoup_mul(tab, m1));
        si = f128_add(si, shoup_mul(tab, mi));
    }
    const uint32_t eh = comb_build(wl, e_lo[threadIdx.x]);
    f256u g1, gi;
    f256u_zero(g1);
    f256u_zero(gi);
    comb_apply(g1, wl, eh, s1);
    comb_apply(gi, wl, eh, si);
    const f128 r1 = f256u_reduce(g1), ri = f256u_reduce(gi);
    __syncthreads();  // every wave is done with the comb workspace
    grid_xor_pair_in(r1, ri, partials, ro, reinterpret_cast<PairShared*>(ws));
}

// out[j * 256 + t] = eq(r[j .. j+8), t) for j < count: one workgroup per window.
__global__ __launch_bounds__(256) void k_eq_windows(const f128* __restrict__ r, uint32_t count,
                                                     f128* __restrict__ out) {
    __shared__ f128 tabs[8][16];
    __shared__ uint32_t his[8];
    const uint32_t j = blockIdx.x;
    if (threadIdx.x < 8) {
        tabs[threadIdx.x][0] = f128_zero();
        his[threadIdx.x] = comb_build<1>(tabs[threadIdx.x], r[j + threadIdx.x]);
    }
    __syncthreads();
    const uint32_t t = threadIdx.x;
    f128 acc = f128_one();
#pragma unroll 1
    for (int i = 0; i < 8; i++) {
        f256u p;
        f256u_zero(p);
        comb_apply<1>(p, tabs[i], his[i], acc);
        const f128 ar = f256u_reduce(p);  // acc * r_i
        acc = ((t >> i) & 1) ? ar : f128_add(acc, ar);
    }
    out[(uint64_t)j * 256 + t] = acc;
}

// Levels 1..levels-1 of an eq table over nv variables: level j is the table without its j
// lowest variables, entry y = XOR of e0[y * 2^j .. (y + 1) * 2^j) (the lowest variables'
// eq factors sum to one). Level j starts at out + sum_{1 <= i < j} 2^(nv - i).
__global__ __launch_bounds__(256) void k_eq_levels(const f128* __restrict__ e0, uint32_t nv,
                                                    uint32_t levels, f128* __restrict__ out) {
    uint64_t total = 0;
    for (uint32_t j = 1; j < levels; j++) total += 1ull << (nv - j);
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t j = 1;
        uint64_t off = 0;
        while (g >= off + (1ull << (nv - j))) {
            off += 1ull << (nv - j);
            j++;
        }
        const uint64_t y = g - off;
        f128 v = f128_zero();
        for (uint64_t z = 0; z < (1ull << j); z++) v = f128_add(v, e0[(y << j) + z]);
        out[g] = v;
    }
}

// eq(r[0..nv), x) for every x as a product of nv factors (nv <= kEqDirectMax). Every factor
// multiply is by a fixed r_i: acc * (1 + r_i) = acc + acc * r_i, with acc * r_i through a
// 16-entry comb table of r_i shared by the workgroup, so no thread runs a chain of generic
// products.
constexpr int kEqDirectMax = 20;

__global__ __launch_bounds__(256) void k_eq_direct(const f128* __restrict__ r, int nlo, int nhi,
                                                    uint32_t lo_blocks, f128* __restrict__ lo,
                                                    f128* __restrict__ hi) {
    // workgroups [0, lo_blocks) build eq(r[0..nlo)) into lo, the rest eq(r[nlo..nlo+nhi)) into hi
    const bool is_lo = blockIdx.x < lo_blocks;
    const int nv = is_lo ? nlo : nhi;
    const f128* rv = is_lo ? r : r + nlo;
    f128* out = is_lo ? lo : hi;
    __shared__ f128 tabs[kEqDirectMax][16];
    __shared__ uint32_t his[kEqDirectMax];
    for (int i = threadIdx.x; i < nv; i += blockDim.x) {
        tabs[i][0] = f128_zero();
        his[i] = comb_build<1>(tabs[i], rv[i]);
    }
    __syncthreads();
    const uint64_t x = (uint64_t)(is_lo ? blockIdx.x : blockIdx.x - lo_blocks) * blockDim.x +
                       threadIdx.x;
    if (x >= (1ull << nv)) return;
    f128 acc = f128_one();
    for (int i = 0; i < nv; i++) {
        f256u p;
        f256u_zero(p);
        comb_apply<1>(p, tabs[i], his[i], acc);
        const f128 ar = f256u_reduce(p);  // acc * r_i
        acc = ((x >> i) & 1) ? ar : f128_add(acc, ar);
    }
    out[x] = acc;
}

// out[x_lo + 2^nlo * x_hi] = lo[x_lo] * hi[x_hi]; one group per x_hi (Shoup table of hi[x_hi]).
__global__ __launch_bounds__(256) void k_eq_outer(const f128* __restrict__ lo, int nlo,
                                                   const f128* __restrict__ hi,
                                                   f128* __restrict__ out) {
    __shared__ f128 table[SHOUP_ENTRIES];
    const uint64_t xh = blockIdx.x;
    if (threadIdx.x < 32) shoup_build_row(table, hi[xh], threadIdx.x);
    __syncthreads();
    const uint64_t nl = 1ull << nlo;
    for (uint64_t xl = threadIdx.x; xl < nl; xl += blockDim.x)
        out[(xh << nlo) + xl] = shoup_mul(table, lo[xl]);
}

// XOR-reduce W words over the grid and hand them to the host: each workgroup folds its words
// into lo.acc with atomicXor and draws a ticket from lo.acc[W]; the workgroup that draws the
// last one moves the sums to lo.out (coherent pinned memory), leaves lo.acc and the ticket at
// zero, and stores lo.seq into lo.flag at system scope, which the host polls.
template <int W>
__device__ __forceinline__ void grid_publish(uint32_t (&v)[W], const LaOut& lo) {
    __shared__ uint32_t sm[4][W];
    __shared__ int last;
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
        uint32_t x = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) x ^= sm[w][threadIdx.x];
        if (x) atomicXor(lo.acc + threadIdx.x, x);
        __threadfence();
    }
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(lo.acc + W, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    if (threadIdx.x < W) {
        lo.out[threadIdx.x] = atomicExch(lo.acc + threadIdx.x, 0u);
        __threadfence_system();
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicExch(lo.acc + W, 0u);
        __hip_atomic_store(lo.flag, lo.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Two rounds per pass of the eq-weighted zerocheck tail (the reference's lookahead,
// cuda-ghash/zerocheck_tail.cuh:429-552; multilinear.rs:331-355 per round). The tables a, b
// (length n) are folded by the nfold (1 or 2) pending challenges r1, r2 into ao, bo of length
// L = n >> nfold. With L >= 2 the kernel sums round j's pair over the folded tables,
//   G(1) = sum_x e[x] a1 b1,  G(inf) = sum_x e[x] (a0+a1)(b0+b1)   (e: L/2 entries),
// and with L >= 4 round j+1's pair as quadratics in its challenge rho, by their values at
// rho = 0, 1 and leading coefficients, weighted by e'[y] = e[2y] + e[2y+1] (round j+1's eq).
// For a four-entry group g0..g3 (c0..c3 of b), e0 = e[2y], e1 = e[2y+1]:
//   G(1) += e0 g1c1 + e1 g3c3,  G(inf) += e0 (g0+g1)(c0+c1) + e1 (g2+g3)(c2+c3);
//   next G(1):   e' g2c2, e' g3c3, e' (g2+g3)(c2+c3)     (at 0, at 1, leading);
//   next G(inf): e' (g0+g2)(c0+c2), e' (g1+g3)(c1+c3), e' (g0+..+g3)(c0+..+c3).
// One thread per folded entry; the four lanes of a group exchange entries by shuffles and
// share the products. With L >= 8 it also writes e2 (L/8 entries), round j+2's eq table:
// e2[z] = e[4z] + e[4z+1] + e[4z+2] + e[4z+3]. Output words: [G(1), G(inf), next G(1) at
// 0, 1, leading, next G(inf) at 0, 1, leading].
__device__ __forceinline__ f128 shfl_f128(f128 v, uint32_t src) {
    return make_uint4(__shfl(v.x, src, 64), __shfl(v.y, src, 64), __shfl(v.z, src, 64),
                      __shfl(v.w, src, 64));
}
__device__ __forceinline__ f128 fold_by(const f128* T, f128 v0, f128 v1) {
    return f128_add(v0, shoup_mul(T, f128_add(v0, v1)));
}
__global__ __launch_bounds__(256) void k_zc_fold_la(const f128* __restrict__ a,
                                                     const f128* __restrict__ b, uint64_t n,
                                                     int nfold, f128 r1, f128 r2,
                                                     f128* __restrict__ ao, f128* __restrict__ bo,
                                                     const f128* __restrict__ e,
                                                     f128* __restrict__ e2, LaOut lo) {
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
    const int per = L >= 4 ? 4 : (int)L;
    const int t = threadIdx.x & 3;
    f256u acc[4];  // lane 0: G1, Ginf; lane 1: G1, Ginf, P1, Pinf; lane 2: P0, Q0; lane 3: Q1, Qinf
#pragma unroll
    for (int k = 0; k < 4; k++) f256u_zero(acc[k]);
    // L is a power of two, so the four lanes of a group always iterate together
    for (uint64_t base = blockIdx.x * (uint64_t)blockDim.x + (threadIdx.x & ~3u); base < 4 * (L / per);
         base += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t grp = base / 4;
        const uint64_t k = grp * per + t;
        f128 g = f128_zero(), c = f128_zero();
        if (t < per) {
            if (nfold == 1) {
                g = fold_by(T1, a[2 * k], a[2 * k + 1]);
                c = fold_by(T1, b[2 * k], b[2 * k + 1]);
            } else {
                g = fold_by(T2, fold_by(T1, a[4 * k], a[4 * k + 1]), fold_by(T1, a[4 * k + 2], a[4 * k + 3]));
                c = fold_by(T2, fold_by(T1, b[4 * k], b[4 * k + 1]), fold_by(T1, b[4 * k + 2], b[4 * k + 3]));
            }
            ao[k] = g;
            bo[k] = c;
        }
        if (L >= 8 && t == 0 && (grp & 1) == 0)  // round j+2's eq: 4 entries of e per entry
            e2[grp >> 1] = f128_add(f128_add(e[2 * grp], e[2 * grp + 1]), f128_add(e[2 * grp + 2], e[2 * grp + 3]));
        if (per < 2) continue;
        f128 G[4], C[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            G[u] = shfl_f128(g, (threadIdx.x & ~3u) + u);
            C[u] = shfl_f128(c, (threadIdx.x & ~3u) + u);
        }
        if (per == 2) {
            if (t == 0) {  // one pair: G(1) = e0 g1c1, G(inf) = e0 (g0+g1)(c0+c1)
                const f128 x0 = comb_mul(wl, G[1], C[1]);
                const f128 x1 = comb_mul(wl, f128_add(G[0], G[1]), f128_add(C[0], C[1]));
                const uint32_t h = comb_build(wl, e[0]);
                comb_apply(acc[0], wl, h, x0);
                comb_apply(acc[1], wl, h, x1);
            }
            continue;
        }
        const f128 e0 = e[2 * grp], e1 = e[2 * grp + 1];
        f128 x0, y0, x1, y1;
        if (t == 0) {
            x0 = G[1]; y0 = C[1];
            x1 = f128_add(G[0], G[1]); y1 = f128_add(C[0], C[1]);
        } else if (t == 1) {
            x0 = G[3]; y0 = C[3];
            x1 = f128_add(G[2], G[3]); y1 = f128_add(C[2], C[3]);
        } else if (t == 2) {
            x0 = G[2]; y0 = C[2];
            x1 = f128_add(G[0], G[2]); y1 = f128_add(C[0], C[2]);
        } else {
            x0 = f128_add(G[1], G[3]); y0 = f128_add(C[1], C[3]);
            x1 = f128_add(f128_add(G[0], G[1]), f128_add(G[2], G[3]));
            y1 = f128_add(f128_add(C[0], C[1]), f128_add(C[2], C[3]));
        }
        const f128 p0 = comb_mul(wl, x0, y0), p1 = comb_mul(wl, x1, y1);
        // weights: lane 0 by e0; lane 1 by e1 (this round) and e' (the next); lanes 2, 3 by e'
        const f128 w0 = t == 0 ? e0 : t == 1 ? e1 : f128_add(e0, e1);
        uint32_t h = comb_build(wl, w0);
        comb_apply(acc[0], wl, h, p0);
        comb_apply(acc[1], wl, h, p1);
        if (t == 1) {
            h = comb_build(wl, f128_add(e0, e1));
            comb_apply(acc[2], wl, h, p0);
            comb_apply(acc[3], wl, h, p1);
        }
    }
    // output slots of each lane's sums
    uint32_t v[32];
#pragma unroll
    for (int k = 0; k < 32; k++) v[k] = 0u;
    const int sl[4][4] = {{0, 1, -1, -1}, {0, 1, 3, 4}, {2, 5, -1, -1}, {6, 7, -1, -1}};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const f128 r = f256u_reduce(acc[q]);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int want = t == 0 ? sl[0][q] : t == 1 ? sl[1][q] : t == 2 ? sl[2][q] : sl[3][q];
            if (k == want) {
                v[4 * k] ^= r.x;
                v[4 * k + 1] ^= r.y;
                v[4 * k + 2] ^= r.z;
                v[4 * k + 3] ^= r.w;
            }
        }
    }
    grid_publish<32>(v, lo);
}

static hipError_t set_ws_attr(const void* fn) {
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, kCombWsBytes);
}

static hipError_t init_attrs() {
    static hipError_t st = [] {
        hipError_t e = set_ws_attr((const void*)k_sc_round_msg);
        if (e == hipSuccess) e = set_ws_attr((const void*)k_sc_fold_msg);
        if (e == hipSuccess) e = set_ws_attr((const void*)k_zc_round);
        if (e == hipSuccess) e = set_ws_attr((const void*)k_zc_fold_round);
        if (e == hipSuccess) e = set_ws_attr((const void*)k_zc_fold_la);
        return e;
    }();
    return st;
}

static hipError_t split_attrs() {
    static hipError_t st = [] {
        hipError_t e = set_ws_attr((const void*)k_zc_round_split);
        if (e == hipSuccess) e = set_ws_attr((const void*)k_zc_fold_round_split);
        return e;
    }();
    return st;
}

static inline unsigned red_grid(uint64_t work) {
    uint64_t g = (work + 255) / 256;
    if (g < 1) g = 1;
    if (g > (uint64_t)kMaxReduceBlocks) g = kMaxReduceBlocks;
    return (unsigned)g;
}

hipError_t launch_sc_round_msg(const uint4* f, const uint4* b, uint64_t n, uint4* partials,
                               RoundOut ro, hipStream_t s) {
    if (hipError_t ae = init_attrs(); ae != hipSuccess) return ae;
    const unsigned g = red_grid(n / 2);
    k_sc_round_msg<<<g, 256, kCombWsBytes, s>>>(f, b, n, partials, ro);
    return hipGetLastError();
}

hipError_t launch_sc_fold_msg(const uint4* f, const uint4* b, uint64_t n, uint4 r, uint4* nf,
                              uint4* nb, uint4* partials, RoundOut ro, hipStream_t s,
                              const uint4* rin) {
    if (hipError_t ae = init_attrs(); ae != hipSuccess) return ae;
    const unsigned g = red_grid(n / 4);
    k_sc_fold_msg<<<g, 256, kCombWsBytes, s>>>(f, b, n, r, rin, nf, nb, partials, ro);
    return hipGetLastError();
}

hipError_t launch_zc_round(const uint4* a, const uint4* b, uint64_t n, const uint4* e,
                           uint4* e_next, uint4* partials, RoundOut ro, hipStream_t s) {
    if (hipError_t ae = init_attrs(); ae != hipSuccess) return ae;
    const unsigned g = red_grid(n / 2);
    k_zc_round<<<g, 256, kCombWsBytes, s>>>(a, b, n, e, e_next, partials, ro);
    return hipGetLastError();
}

hipError_t launch_zc_fold_round(const uint4* a, const uint4* b, uint64_t n, uint4 rho,
                                uint4* ao, uint4* bo, const uint4* e, uint4* e_next,
                                uint4* partials, RoundOut ro, hipStream_t s,
                                const uint4* rin) {
    if (hipError_t ae = init_attrs(); ae != hipSuccess) return ae;
    const unsigned g = red_grid(n / 4);
    k_zc_fold_round<<<g, 256, kCombWsBytes, s>>>(a, b, n, rho, rin, ao, bo, e, e_next, partials,
                                                 ro);
    return hipGetLastError();
}

hipError_t launch_zc_fold_la(const uint4* a, const uint4* b, uint64_t n, int nfold, uint4 r1,
                             uint4 r2, uint4* ao, uint4* bo, const uint4* e, uint4* e2, LaOut lo,
                             hipStream_t s) {
    if (nfold < 1 || nfold > 2 || (n >> nfold) == 0) return hipErrorInvalidValue;
    if (hipError_t ae = init_attrs(); ae != hipSuccess) return ae;
    const uint64_t L = n >> nfold;
    const unsigned g = red_grid(L < 4 ? 4 : L);
    k_zc_fold_la<<<g, 256, kCombWsBytes, s>>>(a, b, n, nfold, r1, r2, ao, bo, e, e2, lo);
    return hipGetLastError();
}

hipError_t launch_zc_round_split(const uint4* a, const uint4* b, uint64_t n, const uint4* e_lo,
                                 const uint4* e_hi, uint4* partials, RoundOut ro, hipStream_t s) {
    if (n < 512 || (n & 511)) return hipErrorInvalidValue;
    if (hipError_t ae = split_attrs(); ae != hipSuccess) return ae;
    // one workgroup per x_hi step, at most kMaxReduceBlocks (all resident)
    const uint64_t steps = (n / 2) >> 8;
    const unsigned g = (unsigned)(steps < (uint64_t)kMaxReduceBlocks ? steps : kMaxReduceBlocks);
    k_zc_round_split<<<g, 256, kCombWsBytes, s>>>(a, b, n / 2, e_lo, e_hi, partials, ro);
    return hipGetLastError();
}

hipError_t launch_zc_fold_round_split(const uint4* a, const uint4* b, uint64_t n, uint4 rho,
                                      uint4* ao, uint4* bo, const uint4* e_lo, const uint4* e_hi,
                                      uint4* partials, RoundOut ro, hipStream_t s,
                                      const uint4* rin) {
    if (n < 1024 || (n & 1023)) return hipErrorInvalidValue;
    if (hipError_t ae = split_attrs(); ae != hipSuccess) return ae;
    const uint64_t steps = (n / 4) >> 8;
    const unsigned g = (unsigned)(steps < (uint64_t)kMaxReduceBlocks ? steps : kMaxReduceBlocks);
    k_zc_fold_round_split<<<g, 256, kCombWsBytes, s>>>(a, b, n / 4, rho, rin, ao, bo, e_lo, e_hi,
                                                       partials, ro);
    return hipGetLastError();
}

hipError_t launch_eq_windows(const uint4* r, uint32_t count, uint4* out, hipStream_t s) {
    if (count == 0) return hipSuccess;
    k_eq_windows<<<count, 256, 0, s>>>(r, count, out);
    return hipGetLastError();
}

hipError_t launch_eq_levels(const uint4* e0, uint32_t nv, uint32_t levels, uint4* out,
                            hipStream_t s) {
    if (levels < 2) return hipSuccess;
    if (levels > nv + 1) return hipErrorInvalidValue;
    uint64_t total = 0;
    for (uint32_t j = 1; j < levels; j++) total += 1ull << (nv - j);
    uint64_t g = (total + 255) / 256;
    if (g > 4096) g = 4096;
    k_eq_levels<<<(unsigned)g, 256, 0, s>>>(e0, nv, levels, out);
    return hipGetLastError();
}

hipError_t launch_fold_pair(const uint4* a, const uint4* b, uint64_t n, uint4 c, uint4* ao,
                            uint4* bo, hipStream_t s, const uint4* cin) {
    uint64_t g = (n / 2 + 255) / 256;
    if (g < 1) g = 1;
    if (g > 8192) g = 8192;
    k_fold_pair<<<(unsigned)g, 256, 0, s>>>(a, b, n, c, cin, ao, bo);
    return hipGetLastError();
}

hipError_t launch_outer_product(const uint4* lo, int nlo, const uint4* hi, int nhi, uint4* out,
                               hipStream_t s) {
    k_eq_outer<<<(unsigned)(1ull << nhi), 256, 0, s>>>(lo, nlo, hi, out);
    return hipGetLastError();
}

// Coordinates carried in the launch's arguments (640 B of the 4 KiB argument space), so that a
// host array reaches the device with no copy and no wait for the stream.
struct EqCoords {
    uint4 r[kEqArgMax];
};
__global__ void k_eq_coords(EqCoords c, uint32_t n, uint4* __restrict__ dst) {
    if (threadIdx.x < n) dst[threadIdx.x] = c.r[threadIdx.x];
}

hipError_t launch_eq_coords(const uint4* h_r, uint32_t n, uint4* dst, hipStream_t s) {
    if (n > (uint32_t)kEqArgMax) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    EqCoords c{};
    for (uint32_t i = 0; i < n; i++) c.r[i] = h_r[i];
    k_eq_coords<<<1, 64, 0, s>>>(c, n, dst);
    return hipGetLastError();
}

hipError_t launch_eq_halves(const uint4* r, int nv, uint4* lo, uint4* hi, hipStream_t s) {
    const int nlo = nv / 2, nhi = nv - nlo;
    if (nhi > kEqDirectMax) return hipErrorInvalidValue;
    const uint32_t lo_blocks = (uint32_t)(((1ull << nlo) + 255) / 256);
    const uint32_t hi_blocks = (uint32_t)(((1ull << nhi) + 255) / 256);
    k_eq_direct<<<lo_blocks + hi_blocks, 256, 0, s>>>(r, nlo, nhi, lo_blocks, lo, hi);
    return hipGetLastError();
}

hipError_t launch_build_eq(const uint4* r, int nv, uint4* out, uint4* scratch, hipStream_t s) {
    // k_eq_direct with zero variables writes the single entry 1.
    const int nlo = nv / 2, nhi = nv - nlo;
    if (nhi > kEqDirectMax) return hipErrorInvalidValue;
    uint4* lo = scratch;
    uint4* hi = scratch + (1ull << nlo);
    const uint32_t lo_blocks = (uint32_t)(((1ull << nlo) + 255) / 256);
    const uint32_t hi_blocks = (uint32_t)(((1ull << nhi) + 255) / 256);
    k_eq_direct<<<lo_blocks + hi_blocks, 256, 0, s>>>(r, nlo, nhi, lo_blocks, lo, hi);
    k_eq_outer<<<(unsigned)(1ull << nhi), 256, 0, s>>>(lo, nlo, hi, out);
    return hipGetLastError();
}

}  // namespace flock

// ---- fold_and_msg_lsb without allocation ---------------------------------------------------
//
// The grid reduction's partial pairs and its ticket counter live in `scratch` below, a
// file-scope device array of kMaxReduceBlocks * 2 + 1 words (16,400 bytes) that each device
// holds for the life of the process. Host code reaches it only through the address
// hipGetSymbolAddress returns on that device (kept per device); the array's name is a host
// stand-in, never a device address. Every call on a device shares that device's array, so
// calls that overlap on one device race.

static __device__ __align__(256) uint4 scratch[flock::kMaxReduceBlocks * 2 + 1];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint4* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device.
int scratch_on_current_device(uint4** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        return flock::set_error(FLOCK_HIP_E_NODEVICE, "fold_msg_static: no current device");
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return flock::set_error(FLOCK_HIP_E_HIP, "fold_msg_static: hipGetSymbolAddress failed");
        }
        g_scratch_ptr[dev] = static_cast<uint4*>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// the ticket counter must read zero when a launch starts
__global__ void k_clear_ticket(uint4* partials) {
    partials[2 * flock::kMaxReduceBlocks] = make_uint4(0u, 0u, 0u, 0u);
}

}  // namespace

extern "C" int flock_hip_sumcheck_fold_msg_static(void* stream, const flock_f128* d_f,
                                                  const flock_f128* d_b, size_t n, flock_f128 r,
                                                  flock_f128* d_nf, flock_f128* d_nb,
                                                  flock_f128* d_msg) {
    if (n < 2 || (n & (n - 1)) || n > (1ull << 40))
        return flock::set_error(FLOCK_HIP_E_INVALID,
                                "fold_msg_static: n must be a power of two in [2, 2^40]");
    if (!d_f || !d_b || !d_nf || !d_nb || !d_msg)
        return flock::set_error(FLOCK_HIP_E_INVALID, "fold_msg_static: null buffer");
    uint4* partials = nullptr;
    if (int rc = scratch_on_current_device(&partials)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    k_clear_ticket<<<1, 1, 0, s>>>(partials);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        uint4 rr;
        std::memcpy(&rr, &r, sizeof(rr));
        e = flock::launch_sc_fold_msg(reinterpret_cast<const uint4*>(d_f),
                                      reinterpret_cast<const uint4*>(d_b), n, rr,
                                      reinterpret_cast<uint4*>(d_nf), reinterpret_cast<uint4*>(d_nb),
                                      partials, flock::round_out(reinterpret_cast<uint4*>(d_msg)), s);
    }
    if (e != hipSuccess) return flock::set_error(FLOCK_HIP_E_HIP, hipGetErrorString(e));
    return 0;
}
