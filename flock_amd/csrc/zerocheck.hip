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

// red[j] = XOR over blocks of the partials (k_col_xor_reduce); then
// p[k] = sum_b phi8(2^b) S[8k + b] for k < 128 (k < 64: p_ab, k >= 64: p_c). One thread per
// accumulator column, then 128 combiners.
__global__ __launch_bounds__(kR1Acc) void k_zc_round1_finish(const f128* __restrict__ red,
                                                              const f128* __restrict__ phi_pow,
                                                              f128* __restrict__ out) {
    __shared__ f128 s[kR1Acc];
    const int j = threadIdx.x;
    const f128 acc = red[j];
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

// ---------------------------------------------------------------------------
// Method of four Russians over 8-item groups, shared by the round-1 group kernel and
// fold_1b_rows. A batch of up to 256 items is staged in LDS: item n's W words at
// ys[n * STRIDE], its weight at eqs[n]. For group g (items 8g..8g+7),
// tabs[32 g + 16 h + v] = XOR of eqs[8g + 4h + i] over the set bits i of v, so that
// sum_i eqs[8g + i] * bit_i(p) = tabs[32g + (p & 15)] ^ tabs[32g + 16 + (p >> 4)].
// Thread t owns slot t / (8W) and bits 32w + 8k + bt (k < 4) with w = (t % 8W) >> 3,
// bt = t & 7: one LDS word per item gives the pattern of all four. Slots take the groups
// g = slot (mod 256 / 8W). The partial of bit j in slot s is at s * 32W + j.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void m4rm_tabs(const f128* eqs, f128* tabs) {
    for (int e = threadIdx.x; e < 32 * 32; e += blockDim.x) {
        const f128* q = eqs + 8 * (e >> 5) + 4 * ((e >> 4) & 1);
        const int n = e & 15;
        f128 te = f128_zero();
#pragma unroll
        for (int i = 0; i < 4; i++)
            if ((n >> i) & 1) te = f128_add(te, q[i]);
        tabs[e] = te;
    }
}

template <int W, int STRIDE>
__device__ __forceinline__ void m4rm_accumulate(const uint32_t* ys, const f128* tabs, int ngroups,
                                                f128 acc[4]) {
    constexpr int kSlots = 256 / (8 * W);
    const int t = threadIdx.x;
    const int slot = t / (8 * W), w = (t % (8 * W)) >> 3, bt = t & 7;
    for (int g = slot; g < ngroups; g += kSlots) {
        uint32_t v = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) v |= ((ys[(8 * g + i) * STRIDE + w] >> bt) & 0x01010101u) << i;
        const f128* T = tabs + 32 * g;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t pat = (v >> (8 * k)) & 0xffu;
            acc[k] = f128_add3(acc[k], T[pat & 15u], T[16 + (pat >> 4)]);
        }
    }
}

template <int W>
__device__ __forceinline__ void m4rm_store(const f128 acc[4], f128* partials_block) {
    const int t = threadIdx.x;
    const int slot = t / (8 * W), w = (t % (8 * W)) >> 3, bt = t & 7;
#pragma unroll
    for (int k = 0; k < 4; k++) partials_block[slot * 32 * W + 32 * w + 8 * k + bt] = acc[k];
}

// ---------------------------------------------------------------------------
// Round 1 with the protocol's inner constants (the zerocheck driver's path).
//
// The driver fixes r[6..9) = phi8(0xF7), phi8(0x53), phi8(0xB5)
// (zerocheck/univariate_skip_optimized.rs:81, zerocheck.rs:644-669). phi8 is a field
// embedding GF(2^8) -> GF(2^128), so eq(r[6..m), x) = phi8(f(x mod 8)) * eq(r[9..m), x / 8)
// with f(l) = prod_t (bit t of l ? s_t : 1 + s_t) in GF(2^8), and with G = x / 8
//   p_ab[i] = sum_G eq(r[9..m), G) phi8(Q_i(G)),
//   Q_i(G)  = sum_{l < 8} f(l) A_{8G+l}(lambda_i) B_{8G+l}(lambda_i)     (in GF(2^8)).
// The bit accumulation then runs over 2^(m-9) group vectors instead of 2^(m-6) rows.
// Q is bit-sliced: word w = 2b + h holds bit b of Q_i for i = 32h .. 32h+31. The extension
// table holds bit-sliced entries, the byte products are a bit-sliced GF(2^8) multiply
// modulo x^8 + x^4 + x^3 + x + 1, and f(l) is a fixed GF(2)-linear map of the 8 planes.
// Bit j of the accumulation is (b, i) with j = 64 b + i.
// ---------------------------------------------------------------------------
constexpr int kQW = 16;        // words of a group vector (512 bits)
constexpr int kQStride = 20;   // LDS words per staged group vector (b128 stores hit distinct banks)

constexpr uint8_t f8_mul_c(uint8_t a, uint8_t b) {
    uint32_t r = 0;
    for (int i = 0; i < 8; i++)
        if ((b >> i) & 1) r ^= (uint32_t)a << i;
    for (int bit = 14; bit >= 8; bit--)
        if ((r >> bit) & 1) r ^= 0x11Bu << (bit - 8);
    return (uint8_t)r;
}
constexpr uint8_t f8_inv_c(uint8_t a) {  // a^254
    uint8_t r = 1, sq = a;
    for (int i = 1; i < 8; i++) {
        sq = f8_mul_c(sq, sq);
        r = f8_mul_c(r, sq);
    }
    return r;
}
// f(l) = prod_t (l_t ? s_t : 1 + s_t) = C * prod_t w_t^(l_t) with w_t = s_t / (1 + s_t) and
// C = prod_t (1 + s_t); s = 0xF7, 0x53, 0xB5. Maps 0..2 multiply by w_0..w_2, map 3 by C.
// m[b] = the set of input planes a such that bit b of v * x^a is set.
struct F8Maps {
    uint8_t m[4][8];
};
constexpr F8Maps make_skip3_maps() {
    F8Maps s{};
    const uint8_t small[3] = {0xF7, 0x53, 0xB5};
    uint8_t v[4] = {0, 0, 0, 1};
    for (int t = 0; t < 3; t++) {
        const uint8_t one_s = (uint8_t)(small[t] ^ 1);
        v[t] = f8_mul_c(small[t], f8_inv_c(one_s));
        v[3] = f8_mul_c(v[3], one_s);
    }
    for (int k = 0; k < 4; k++)
        for (int a = 0; a < 8; a++) {
            const uint8_t p = f8_mul_c(v[k], (uint8_t)(1u << a));
            for (int b = 0; b < 8; b++)
                if ((p >> b) & 1) s.m[k][b] |= (uint8_t)(1u << a);
        }
    return s;
}
constexpr F8Maps kSkip3 = make_skip3_maps();

// out = v_K * in + add, bit-sliced (16 words = 8 planes x 2 halves); out may alias in or add.
template <int K>
__device__ __forceinline__ void f8_scale_add_bs(const uint32_t in[16], const uint32_t add[16],
                                                uint32_t out[16]) {
    uint32_t o[16];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int b = 0; b < 8; b++) {
            uint32_t v = add[2 * b + h];
#pragma unroll
            for (int a = 0; a < 8; a++)
                if ((kSkip3.m[K][b] >> a) & 1) v ^= in[2 * a + h];
            o[2 * b + h] = v;
        }
#pragma unroll
    for (int w = 0; w < 16; w++) out[w] = o[w];
}

// y = bit-sliced extension of the 64-bit row (w0 = bits 0..31, w1 = bits 32..63).
__device__ __forceinline__ void ext_bitsliced(const uint4* ext, uint32_t w0, uint32_t w1,
                                              uint32_t y[16]) {
#pragma unroll
    for (int w = 0; w < 16; w++) y[w] = 0u;
#pragma unroll 2
    for (int g = 0; g < 16; g++) {
        const uint32_t n = ((g < 8 ? w0 : w1) >> (4 * (g & 7))) & 15u;
        const uint4* e = ext + g * 64 + n;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint4 v = e[16 * k];
            y[4 * k] ^= v.x;
            y[4 * k + 1] ^= v.y;
            y[4 * k + 2] ^= v.z;
            y[4 * k + 3] ^= v.w;
        }
    }
}

// p = A(lambda) * B(lambda) for the 64 points, bit-sliced: schoolbook product of the planes,
// then x^8 = x^4 + x^3 + x + 1.
__device__ __forceinline__ void row_product_bs(const uint4* ext, uint2 ra, uint2 rb,
                                               uint32_t p[16]) {
    uint32_t ya[16], yb[16];
    ext_bitsliced(ext, ra.x, ra.y, ya);
    ext_bitsliced(ext, rb.x, rb.y, yb);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        uint32_t t[15];
#pragma unroll
        for (int k = 0; k < 15; k++) t[k] = 0u;
#pragma unroll
        for (int i = 0; i < 8; i++)
#pragma unroll
            for (int j = 0; j < 8; j++) t[i + j] ^= ya[2 * i + h] & yb[2 * j + h];
#pragma unroll
        for (int k = 14; k >= 8; k--) {
            t[k - 4] ^= t[k];
            t[k - 5] ^= t[k];
            t[k - 7] ^= t[k];
            t[k - 8] ^= t[k];
        }
#pragma unroll
        for (int b = 0; b < 8; b++) p[2 * b + h] = t[b];
    }
}

// Q(G) = C * sum_l (w_0^l0 w_1^l1 w_2^l2) P_l by Horner over l = 7..0: acc0 closes each row
// pair with w_0, acc1 each quad with w_1, acc2 the group with w_2.
__device__ __forceinline__ void group_vector(const uint4* ext, const uint2* a, const uint2* b,
                                             uint32_t q[16]) {
    uint32_t acc0[16], acc1[16], acc2[16];
#pragma unroll
    for (int w = 0; w < 16; w++) acc1[w] = acc2[w] = 0u;
#pragma unroll 1
    for (int l = 7; l >= 0; l--) {
        uint32_t p[16];
        row_product_bs(ext, a[l], b[l], p);
        if (l & 1) {
#pragma unroll
            for (int w = 0; w < 16; w++) acc0[w] = p[w];
            continue;
        }
        f8_scale_add_bs<0>(acc0, p, acc0);
        if ((l >> 1) & 1) {
#pragma unroll
            for (int w = 0; w < 16; w++) acc1[w] = acc0[w];
            continue;
        }
        f8_scale_add_bs<1>(acc1, acc0, acc1);
        if ((l >> 2) & 1) {
#pragma unroll
            for (int w = 0; w < 16; w++) acc2[w] = acc1[w];
            continue;
        }
        f8_scale_add_bs<2>(acc2, acc1, acc2);
    }
    uint32_t zero[16];
#pragma unroll
    for (int w = 0; w < 16; w++) zero[w] = 0u;
    f8_scale_add_bs<3>(acc2, zero, q);
}

// Group G covers rows 8G..8G+7 = 64 contiguous bytes of a and of b.
__global__ __launch_bounds__(256) void k_zc_round1_groups(
    const uint2* __restrict__ a, const uint2* __restrict__ b, const f128* __restrict__ eq9,
    uint64_t n_groups, uint32_t groups_per_block, const uint4* __restrict__ ext_g,
    f128* __restrict__ partials) {
    __shared__ uint4 ext[16 * 16 * 4];                       // 16 KiB, k-major as in k_zc_round1
    __shared__ __align__(16) uint32_t ys[256 * kQStride];   // 20 KiB
    __shared__ f128 eqs[256];
    __shared__ f128 tabs[32 * 32];                          // 16 KiB
    for (int i = threadIdx.x; i < 16 * 16 * 4; i += blockDim.x) {
        const int g = i >> 6, n = (i >> 2) & 15, k = i & 3;
        ext[(g * 4 + k) * 16 + n] = ext_g[i];
    }
    __syncthreads();
    f128 acc[4];
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k] = f128_zero();
    const int t = threadIdx.x;
    const uint64_t g0 = (uint64_t)blockIdx.x * groups_per_block;
    uint64_t g_end = g0 + groups_per_block;
    if (g_end > n_groups) g_end = n_groups;
    for (uint64_t base = g0; base < g_end; base += 256) {
        const uint64_t G = base + t;
        uint32_t q[16];
        if (G < g_end) {
            group_vector(ext, a + 8 * G, b + 8 * G, q);
            eqs[t] = eq9[G];
        } else {
#pragma unroll
            for (int w = 0; w < 16; w++) q[w] = 0u;
            eqs[t] = f128_zero();
        }
        uint4* yrow = reinterpret_cast<uint4*>(ys + t * kQStride);
#pragma unroll
        for (int k = 0; k < 4; k++) yrow[k] = make_uint4(q[4 * k], q[4 * k + 1], q[4 * k + 2], q[4 * k + 3]);
        __syncthreads();
        m4rm_tabs(eqs, tabs);
        __syncthreads();
        const uint64_t live = g_end - base;
        const int ngroups = (int)(((live < 256u ? live : 256u) + 7) >> 3);
        m4rm_accumulate<kQW, kQStride>(ys, tabs, ngroups, acc);
        __syncthreads();
    }
    m4rm_store<kQW>(acc, partials + (size_t)blockIdx.x * 1024);
}

// red[j] for the 1024 columns of the block partials (column reduce, see k_col_xor_reduce).
// p_ab[i] = sum_b phi8(2^b) (red[64 b + i] + red[512 + 64 b + i]): thread (b, i) forms one
// product, then 64 threads sum the eight.
__global__ __launch_bounds__(512) void k_zc_round1_groups_finish(const f128* __restrict__ red,
                                                                 const f128* __restrict__ phi_pow,
                                                                 f128* __restrict__ out) {
    __shared__ f128 part[8][64];
    const int bb = threadIdx.x >> 6, i = threadIdx.x & 63;
    part[bb][i] = f128_mul(phi_pow[bb], f128_add(red[64 * bb + i], red[512 + 64 * bb + i]));
    __syncthreads();
    if (bb == 0) {
        f128 r = part[0][i];
#pragma unroll
        for (int k = 1; k < 8; k++) r = f128_add(r, part[k][i]);
        out[i] = r;
    }
}

// p_c[i] = sum_s phi8(M[s][i]) S_c[s] with S_c[s] = sum_x eq(r[6..m), x) c_s(x)
//        = (1 + r6) shat[s] + r6 shat[64 + s], shat = fold_1b_rows(c, eq(r[7..m))):
// row y of c as 128 bits holds rows x = 2y (bits 0..63) and 2y + 1 (bits 64..127), and
// eq(r[6..m), 2y + b) = (b ? r6 : 1 + r6) eq(r[7..m), y). One block per i; 128 threads form the
// two halves of S_c, then 64 threads one product each.
__global__ __launch_bounds__(128) void k_zc_round1_c(const f128* __restrict__ shat,
                                                      const f128* __restrict__ phim, f128 r6,
                                                      f128* __restrict__ out) {
    __shared__ f128 half[128];
    __shared__ f128 part[64];
    const int i = blockIdx.x, t = threadIdx.x;
    f128 w = r6;
    if (t < 64) w.x ^= 1u;  // 1 + r6 for shat[s], r6 for shat[64 + s]
    half[t] = f128_mul(w, shat[t]);
    __syncthreads();
    if (t < 64) part[t] = f128_mul(phim[t * 64 + i], f128_add(half[t], half[64 + t]));
    __syncthreads();
    if (t == 0) {
        f128 r = f128_zero();
        for (int k = 0; k < 64; k++) r = f128_add(r, part[k]);
        out[i] = r;
    }
}

// out[c] = XOR over rows g < n of in[g * ncols + c]; block b takes columns 64b .. 64b+63 with
// 16 row lanes.
__global__ __launch_bounds__(1024) void k_col_xor_reduce(const f128* __restrict__ in, int n,
                                                          int ncols, f128* __restrict__ out) {
    __shared__ f128 part[16][64];
    const int col = threadIdx.x & 63, r0 = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + col;
    f128 acc = f128_zero();
    if (c < ncols)
        for (int g = r0; g < n; g += 16) acc = f128_add(acc, in[(size_t)g * ncols + c]);
    part[r0][col] = acc;
    __syncthreads();
    if (r0 == 0 && c < ncols) {
        f128 r = part[0][col];
#pragma unroll
        for (int k = 1; k < 16; k++) r = f128_add(r, part[k][col]);
        out[c] = r;
    }
}

// fold_1b_rows (crates/flock-core/src/pcs/ring_switch.rs:182-230, scalar form :410-470):
// out[j] = sum over rows y of t[y] * bit j of the 128-bit row y, j < 128, by the method of
// four Russians over 8-row groups (8 slots of 32 threads per 256-row batch).
__global__ __launch_bounds__(256) void k_fold_1b_rows(const uint4* __restrict__ rows_g,
                                                       const f128* __restrict__ t_g,
                                                       uint64_t rows, uint32_t rows_per_group,
                                                       f128* __restrict__ partials) {
    __shared__ __align__(16) uint32_t ys[256 * 4];
    __shared__ f128 eqs[256];
    __shared__ f128 tabs[32 * 32];
    const int t = threadIdx.x;
    f128 acc[4];
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k] = f128_zero();
    const uint64_t row0 = (uint64_t)blockIdx.x * rows_per_group;
    uint64_t row_end = row0 + rows_per_group;
    if (row_end > rows) row_end = rows;
    for (uint64_t base = row0; base < row_end; base += 256) {
        const uint64_t y = base + t;
        const bool live = y < row_end;
        reinterpret_cast<uint4*>(ys)[t] = live ? rows_g[y] : make_uint4(0u, 0u, 0u, 0u);
        eqs[t] = live ? t_g[y] : f128_zero();
        __syncthreads();
        m4rm_tabs(eqs, tabs);
        __syncthreads();
        const uint64_t n = row_end - base;
        m4rm_accumulate<4, 4>(ys, tabs, (int)(((n < 256u ? n : 256u) + 7) >> 3), acc);
        __syncthreads();
    }
    m4rm_store<4>(acc, partials + (size_t)blockIdx.x * 1024);
}

// out[j] = sum over the 8 slots of red[128 slot + j].
__global__ __launch_bounds__(128) void k_fold_1b_rows_finish(const f128* __restrict__ red,
                                                              f128* __restrict__ out) {
    const int j = threadIdx.x;
    f128 acc = f128_zero();
#pragma unroll
    for (int s = 0; s < 8; s++) acc = f128_add(acc, red[128 * s + j]);
    out[j] = acc;
}

uint32_t zc_round1_groups(uint64_t rows, uint32_t* rows_per_group, uint32_t max_groups) {
    // at most max_groups groups; each group walks whole 256-row batches
    uint64_t per = (rows + max_groups - 1) / max_groups;
    per = ((per + kR1Threads - 1) / kR1Threads) * kR1Threads;
    if (per < kR1Threads) per = kR1Threads;
    *rows_per_group = (uint32_t)per;
    return (uint32_t)((rows + per - 1) / per);
}

static hipError_t launch_col_reduce(const uint4* partials, uint32_t n, uint4* red, hipStream_t s) {
    k_col_xor_reduce<<<1024 / 64, 1024, 0, s>>>(partials, (int)n, 1024, red);
    return hipGetLastError();
}

hipError_t launch_zc_round1(const uint8_t* a, const uint8_t* b, const uint8_t* c, const uint4* eq,
                            uint64_t rows, const uint4* ext, const uint8_t* logexp,
                            const uint4* phi_pow, uint4* partials, uint4* red, uint4* out,
                            hipStream_t s) {
    uint32_t per = 0;
    const uint32_t groups = zc_round1_groups(rows, &per, 512);
    k_zc_round1<<<groups, kR1Threads, 0, s>>>(a, b, c, eq, rows, per, ext, logexp, partials);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = launch_col_reduce(partials, groups, red, s);
    if (e != hipSuccess) return e;
    k_zc_round1_finish<<<1, kR1Acc, 0, s>>>(red, phi_pow, out);
    return hipGetLastError();
}

hipError_t launch_zc_round1_groups(const uint8_t* a, const uint8_t* b, const uint4* eq9,
                                   uint64_t n_groups, const uint4* ext_bs, const uint4* phi_pow,
                                   uint4* partials, uint4* red, uint4* out, hipStream_t s) {
    uint32_t per = 0;
    const uint32_t blocks = zc_round1_groups(n_groups, &per, 512);  // 2 workgroups per CU (LDS)
    k_zc_round1_groups<<<blocks, 256, 0, s>>>(reinterpret_cast<const uint2*>(a),
                                              reinterpret_cast<const uint2*>(b), eq9, n_groups,
                                              per, ext_bs, partials);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = launch_col_reduce(partials, blocks, red, s);
    if (e != hipSuccess) return e;
    k_zc_round1_groups_finish<<<1, 512, 0, s>>>(red, phi_pow, out);
    return hipGetLastError();
}

hipError_t launch_zc_round1_c(const uint4* shat, const uint4* phim, uint4 r6, uint4* out,
                              hipStream_t s) {
    k_zc_round1_c<<<64, 128, 0, s>>>(shat, phim, r6, out);
    return hipGetLastError();
}

hipError_t launch_fold_1b_rows(const uint4* rows_g, const uint4* t, uint64_t rows,
                              uint4* partials, uint4* red, uint4* out, hipStream_t s) {
    uint32_t per = 0;
    // up to 2048 workgroups of 256-row batches (six fit a CU)
    const uint32_t groups = zc_round1_groups(rows, &per, kZcMaxBlocks);
    k_fold_1b_rows<<<groups, 256, 0, s>>>(rows_g, t, rows, per, partials);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = launch_col_reduce(partials, groups, red, s);
    if (e != hipSuccess) return e;
    k_fold_1b_rows_finish<<<1, 128, 0, s>>>(red, out);
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
