// Zerocheck front: the univariate-skip round 1 and the round-2 fold at z
// (crates/flock-core/src/zerocheck/{univariate_skip,multilinear}.rs).
//
// Round 1 (round1_naive, univariate_skip.rs:68-124) on LSB-first packed bits, k_skip = 6:
//   row x = bits x*64 .. x*64+63 of a, b, c.  A_x, B_x, C_x are the GF(2^8) polynomials of
//   degree < 64 through those bits on S = {0..63}, read on Lambda = {64..127}:
//     p_ab[i] = sum_x eq(r[6..], x) phi8(A_x(64+i) B_x(64+i)),
//     p_c[i]  = sum_x eq(r[6..], x) phi8(C_x(64+i)).
// phi8 is F2-linear, so with Y_x = (A_x*B_x | C_x) as 128 bytes = 1024 bits,
//     p[j/8] = sum_b phi8(2^b) S[(j/8)*8 + b],   S[j] = XOR of eq(x) over rows with bit j of Y_x.
// The kernel therefore never multiplies by eq: it builds Y_x per row (the S->Lambda extension
// is F2-linear: 16 nibble lookups of 64-byte vectors from an LDS table; the byte products use
// log/exp tables), and every thread owns 4 of the 1024 bit accumulators, XORing eq(x) in
// when its bit is set. Block partials are XOR-reduced and combined with phi8(2^b) at the end.
//
// Round 2 (fold_at_z_naive, multilinear.rs:290-310): a_z[x] = sum_s bit(a, x*64+s) L_s(z),
// L_s the Lagrange basis on phi8(0..63): a byte table T[j][v] = sum of L_{8j+b}(z) over the
// set bits b of v (the UniSkipFoldTable layout, multilinear.rs:414-457) held in LDS.
#include "f128_device.h"
#include "flock_internal.h"

namespace flock {

constexpr int kR1Threads = 256;
constexpr int kR1Acc = 1024;  // 64 AB bytes + 64 C bytes, 8 bits each

// ext: 16 nibble groups x 16 values x 64 bytes (as 16 uint4 per entry).
// logexp: log[256] then exp[512] (bytes).
__global__ __launch_bounds__(kR1Threads) void k_zc_round1(
    const uint8_t* __restrict__ a, const uint8_t* __restrict__ b, const uint8_t* __restrict__ c,
    const f128* __restrict__ eq, uint64_t rows, uint32_t rows_per_group,
    const uint4* __restrict__ ext_g, const uint8_t* __restrict__ logexp_g,
    f128* __restrict__ partials) {
    // ext in LDS is k-major: ext[(g*4 + k)*16 + n] = 16-byte slice k of entry (g, n), so the
    // 16 entries of one slice are 256 B (all banks once) and random nibbles never conflict.
    __shared__ uint4 ext[16 * 16 * 4];       // 16 KiB
    __shared__ uint8_t lg[256];
    __shared__ uint8_t ex[512];
    __shared__ uint32_t ys[kR1Threads][32];  // Y of one batch, transposed words (32 KiB)
    __shared__ f128 eqs[kR1Threads];
    __shared__ f128 tabs[kR1Threads / 8 * 32];  // nibble sum tables per 8-row group (16 KiB)
    for (int i = threadIdx.x; i < 16 * 16 * 4; i += blockDim.x) {
        const int g = i >> 6, n = (i >> 2) & 15, k = i & 3;  // source index (g*16 + n)*4 + k
        ext[(g * 4 + k) * 16 + n] = ext_g[i];
    }
    for (int i = threadIdx.x; i < 256; i += blockDim.x) lg[i] = logexp_g[i];
    for (int i = threadIdx.x; i < 512; i += blockDim.x) ex[i] = logexp_g[256 + i];
    __syncthreads();

    f128 acc[4];
#pragma unroll
    for (int q = 0; q < 4; q++) acc[q] = f128_zero();
    const int t = threadIdx.x;
    const uint64_t row0 = (uint64_t)blockIdx.x * rows_per_group;
    uint64_t row_end = row0 + rows_per_group;
    if (row_end > rows) row_end = rows;

    for (uint64_t base = row0; base < row_end; base += kR1Threads) {
        const uint64_t x = base + t;
        const bool live = x < row_end;
        // ---- Y for row x ----
        uint32_t ya[16], yb[16], yc[16];
#pragma unroll
        for (int w = 0; w < 16; w++) ya[w] = yb[w] = yc[w] = 0u;
        if (live) {
            const uint2 ra = reinterpret_cast<const uint2*>(a)[x];
            const uint2 rb = reinterpret_cast<const uint2*>(b)[x];
            const uint2 rc = reinterpret_cast<const uint2*>(c)[x];
#pragma unroll 2  // fully unrolled, the 16 groups of loads spill
            for (int g = 0; g < 16; g++) {
                const uint32_t wa = g < 8 ? ra.x : ra.y, wb = g < 8 ? rb.x : rb.y,
                               wc = g < 8 ? rc.x : rc.y;
                const int sh = 4 * (g & 7);
                const uint4* ea = ext + g * 64 + ((wa >> sh) & 15u);
                const uint4* eb = ext + g * 64 + ((wb >> sh) & 15u);
                const uint4* ec = ext + g * 64 + ((wc >> sh) & 15u);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint4 va = ea[16 * k], vb = eb[16 * k], vc = ec[16 * k];
                    ya[4 * k] ^= va.x; ya[4 * k + 1] ^= va.y; ya[4 * k + 2] ^= va.z; ya[4 * k + 3] ^= va.w;
                    yb[4 * k] ^= vb.x; yb[4 * k + 1] ^= vb.y; yb[4 * k + 2] ^= vb.z; yb[4 * k + 3] ^= vb.w;
                    yc[4 * k] ^= vc.x; yc[4 * k + 1] ^= vc.y; yc[4 * k + 2] ^= vc.z; yc[4 * k + 3] ^= vc.w;
                }
            }
            // byte products A*B in GF(2^8) via log/exp
#pragma unroll
            for (int w = 0; w < 16; w++) {
                uint32_t r = 0u;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t u = (ya[w] >> (8 * k)) & 0xffu, v = (yb[w] >> (8 * k)) & 0xffu;
                    const uint32_t p = (u && v) ? (uint32_t)ex[lg[u] + lg[v]] : 0u;
                    r |= p << (8 * k);
                }
                ya[w] = r;
            }
            eqs[t] = eq[x];
        } else {
            eqs[t] = f128_zero();
        }
        // Y bytes: j-byte index jb in [0,128): AB bytes 0..63, C bytes 64..127.
        // Stored transposed: word d of row holds bytes d, d+32, d+64, d+96.
#pragma unroll
        for (int d = 0; d < 32; d++) {
            const uint32_t b0 = (ya[d >> 2] >> (8 * (d & 3))) & 0xffu;         // byte d (AB)
            const uint32_t b1 = (ya[(d + 32) >> 2] >> (8 * (d & 3))) & 0xffu;  // byte d+32 (AB)
            const uint32_t b2 = (yc[d >> 2] >> (8 * (d & 3))) & 0xffu;         // byte d+64 (C)
            const uint32_t b3 = (yc[(d + 32) >> 2] >> (8 * (d & 3))) & 0xffu;  // byte d+96 (C)
            ys[t][d] = b0 | (b1 << 8) | (b2 << 16) | (b3 << 24);
        }
        __syncthreads();
        // ---- accumulate (method of four Russians over 8-row groups) ----
        // tabs[32 g + 16 h + n] = XOR of eqs[8g + 4h + i] over the bits i of n. Rows past the
        // batch's end have zero Y and zero eq, so they add nothing.
        for (int e = t; e < kR1Threads / 8 * 32; e += kR1Threads) {
            const f128* q = eqs + 8 * (e >> 5) + 4 * ((e >> 4) & 1);
            const int n = e & 15;
            f128 te = f128_zero();
#pragma unroll
            for (int i = 0; i < 4; i++)
                if ((n >> i) & 1) te = f128_add(te, q[i]);
            tabs[e] = te;
        }
        __syncthreads();
        // thread t owns bits j = t + 256 q: byte (t>>3) + 32 q, bit t&7. For a group, byte q of
        // v collects that bit of the group's 8 rows (row i -> bit i); its nibbles index tabs.
        const int nrows = (int)((row_end - base) < (uint64_t)kR1Threads ? (row_end - base)
                                                                        : (uint64_t)kR1Threads);
        const int d = t >> 3, bt = t & 7;
        for (int g = 0; g < (nrows + 7) >> 3; g++) {
            uint32_t v = 0;
#pragma unroll
            for (int i = 0; i < 8; i++) v |= ((ys[8 * g + i][d] >> bt) & 0x01010101u) << i;
            const f128* T = tabs + 32 * g;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t pat = (v >> (8 * q)) & 0xffu;
                acc[q] = f128_add(acc[q], f128_add(T[pat & 15u], T[16 + (pat >> 4)]));
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 4; q++) partials[(size_t)blockIdx.x * kR1Acc + t + kR1Threads * q] = acc[q];
}

// S[j] = XOR over groups of partials; then p[k] = sum_b phi8(2^b) S[8k + b] for k < 128
// (k < 64: p_ab, k >= 64: p_c). One thread per accumulator column, then 128 combiners.
__global__ __launch_bounds__(kR1Acc) void k_zc_round1_finish(const f128* __restrict__ partials,
                                                              int groups,
                                                              const f128* __restrict__ phi_pow,
                                                              f128* __restrict__ out) {
    __shared__ f128 s[kR1Acc];
    const int j = threadIdx.x;
    f128 acc = f128_zero();
    for (int g = 0; g < groups; g++) acc = f128_add(acc, partials[(size_t)g * kR1Acc + j]);
    // accumulator column j = t + 256 q holds byte (t>>3) + 32 q, bit t & 7
    const int t = j & 255, q = j >> 8;
    const int byte = (t >> 3) + 32 * q, bit = t & 7;
    s[byte * 8 + bit] = acc;
    __syncthreads();
    if (j < 128) {
        f128 r = f128_zero();
        for (int bb = 0; bb < 8; bb++) r = f128_add(r, f128_mul(phi_pow[bb], s[j * 8 + bb]));
        out[j] = r;
    }
}

// a_z[x] = sum of table[j][byte j of row x] (8 lookups), for a and b at once.
__global__ __launch_bounds__(256) void k_zc_fold_at_z(const uint8_t* __restrict__ a,
                                                       const uint8_t* __restrict__ b,
                                                       uint64_t rows,
                                                       const f128* __restrict__ table_g,
                                                       f128* __restrict__ az,
                                                       f128* __restrict__ bz) {
    __shared__ f128 table[8 * 256];  // 32 KiB
    for (int i = threadIdx.x; i < 8 * 256; i += blockDim.x) table[i] = table_g[i];
    __syncthreads();
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < rows;
         x += (uint64_t)gridDim.x * blockDim.x) {
        const uint2 ra = reinterpret_cast<const uint2*>(a)[x];
        const uint2 rb = reinterpret_cast<const uint2*>(b)[x];
        f128 sa = f128_zero(), sb = f128_zero();
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const uint32_t wa = j < 4 ? ra.x : ra.y, wb = j < 4 ? rb.x : rb.y;
            const int sh = 8 * (j & 3);
            sa = f128_add(sa, table[j * 256 + ((wa >> sh) & 0xffu)]);
            sb = f128_add(sb, table[j * 256 + ((wb >> sh) & 0xffu)]);
        }
        az[x] = sa;
        bz[x] = sb;
    }
}

// XOR-fold of an eq table: out[z] = in[2z] ^ in[2z+1].
__global__ __launch_bounds__(256) void k_eq_fold(const f128* __restrict__ in, uint64_t n_out,
                                                  f128* __restrict__ out) {
    for (uint64_t z = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; z < n_out;
         z += (uint64_t)gridDim.x * blockDim.x)
        out[z] = f128_add(in[2 * z], in[2 * z + 1]);
}

// fold_1b_rows (crates/flock-core/src/pcs/ring_switch.rs:182-230, scalar form :410-470):
// out[j] = sum over rows y of t[y] * bit j of the 128-bit row y, j < 128.  Thread t of 256
// owns column j = t & 127 for the rows of parity t >> 7 in each 256-row batch staged in LDS;
// a 64-lane wave reads one row word per step (LDS broadcast).
__global__ __launch_bounds__(256) void k_fold_1b_rows(const uint4* __restrict__ rows_g,
                                                       const f128* __restrict__ t_g,
                                                       uint64_t rows, uint32_t rows_per_group,
                                                       f128* __restrict__ partials) {
    __shared__ uint4 rw[256];
    __shared__ f128 tw[256];
    const int t = threadIdx.x;
    const int j = t & 127, par = t >> 7, word = j >> 5, bit = j & 31;
    f128 acc = f128_zero();
    const uint64_t row0 = (uint64_t)blockIdx.x * rows_per_group;
    uint64_t row_end = row0 + rows_per_group;
    if (row_end > rows) row_end = rows;
    for (uint64_t base = row0; base < row_end; base += 256) {
        const uint64_t y = base + t;
        if (y < row_end) {
            rw[t] = rows_g[y];
            tw[t] = t_g[y];
        } else {
            rw[t] = make_uint4(0u, 0u, 0u, 0u);
            tw[t] = f128_zero();
        }
        __syncthreads();
        const int n = (int)((row_end - base) < 256u ? (row_end - base) : 256u);
        for (int rr = par; rr < n; rr += 2) {
            const uint4 r = rw[rr];
            const uint32_t w = word == 0 ? r.x : word == 1 ? r.y : word == 2 ? r.z : r.w;
            const uint32_t msk = 0u - ((w >> bit) & 1u);
            const f128 e = tw[rr];
            acc.x ^= e.x & msk;
            acc.y ^= e.y & msk;
            acc.z ^= e.z & msk;
            acc.w ^= e.w & msk;
        }
        __syncthreads();
    }
    partials[(size_t)blockIdx.x * 256 + t] = acc;
}

__global__ __launch_bounds__(128) void k_fold_1b_rows_finish(const f128* __restrict__ partials,
                                                              int groups,
                                                              f128* __restrict__ out) {
    const int j = threadIdx.x;
    f128 acc = f128_zero();
    for (int g = 0; g < groups; g++) {
        acc = f128_add(acc, partials[(size_t)g * 256 + j]);
        acc = f128_add(acc, partials[(size_t)g * 256 + 128 + j]);
    }
    out[j] = acc;
}

uint32_t zc_round1_groups(uint64_t rows, uint32_t* rows_per_group) {
    // at most 512 groups; each group walks whole 256-row batches
    uint64_t per = (rows + 511) / 512;
    per = ((per + kR1Threads - 1) / kR1Threads) * kR1Threads;
    if (per < kR1Threads) per = kR1Threads;
    *rows_per_group = (uint32_t)per;
    return (uint32_t)((rows + per - 1) / per);
}

hipError_t launch_zc_round1(const uint8_t* a, const uint8_t* b, const uint8_t* c, const uint4* eq,
                            uint64_t rows, const uint4* ext, const uint8_t* logexp,
                            const uint4* phi_pow, uint4* partials, uint4* out, hipStream_t s) {
    uint32_t per = 0;
    const uint32_t groups = zc_round1_groups(rows, &per);
    k_zc_round1<<<groups, kR1Threads, 0, s>>>(a, b, c, eq, rows, per, ext, logexp, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_zc_round1_finish<<<1, kR1Acc, 0, s>>>(partials, (int)groups, phi_pow, out);
    return hipGetLastError();
}

hipError_t launch_fold_1b_rows(const uint4* rows_g, const uint4* t, uint64_t rows,
                              uint4* partials, uint4* out, hipStream_t s) {
    uint32_t per = 0;
    const uint32_t groups = zc_round1_groups(rows, &per);  // <= 512 groups of 256-row batches
    k_fold_1b_rows<<<groups, 256, 0, s>>>(rows_g, t, rows, per, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_fold_1b_rows_finish<<<1, 128, 0, s>>>(partials, (int)groups, out);
    return hipGetLastError();
}

hipError_t launch_zc_fold_at_z(const uint8_t* a, const uint8_t* b, uint64_t rows,
                               const uint4* table, uint4* az, uint4* bz, hipStream_t s) {
    uint64_t g = (rows + 255) / 256;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    k_zc_fold_at_z<<<(unsigned)g, 256, 0, s>>>(a, b, rows, table, az, bz);
    return hipGetLastError();
}

hipError_t launch_eq_fold(const uint4* in, uint64_t n_out, uint4* out, hipStream_t s) {
    uint64_t g = (n_out + 255) / 256;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    k_eq_fold<<<(unsigned)g, 256, 0, s>>>(in, n_out, out);
    return hipGetLastError();
}

}  // namespace flock
