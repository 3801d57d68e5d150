// F128 batch products (the test surface for the device multiply) and the
// block-partials XOR reduction shared by the sumcheck kernels.
#include "f128_device.h"
#include "flock_internal.h"

namespace flock {

__global__ __launch_bounds__(256) void k_f128_mul_batch(const f128* __restrict__ a,
                                                         const f128* __restrict__ b,
                                                         f128* __restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = f128_mul(a[i], b[i]);
}

// a[i] * t through an LDS Shoup table (the fixed-multiplier path used by the NTT and folds).
__global__ __launch_bounds__(256) void k_f128_mul_const_batch(const f128* __restrict__ a, f128 t,
                                                               f128* __restrict__ out, size_t n) {
    __shared__ f128 table[SHOUP_ENTRIES];
    if (threadIdx.x < 32) shoup_build_row(table, t, threadIdx.x);
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = shoup_mul(table, a[i]);
}

// The same fixed-multiplier product on the matrix cores, for measuring the two against each
// other: out[i] = a[i] * t^reps. A product by a fixed t is the GF(2)-linear map M
// (M[r][k] = bit r of t x^k) applied to the 128 bits of a, which v_mfma_i32_32x32x32_i8 runs
// over 0/1 bytes as integer dot products whose bit 0 is the GF(2) sum. A wave takes 32
// elements at a time: B[k][j] = bit k of element j (lane j and j + 32 hold its bits 16-bit
// chunk by chunk, expanded to bytes), A[R][Kc] = the 32 x 32 block of M at rows 32R, columns
// 32Kc (one 16-byte fragment per lane, in LDS), D[R] = 4 MFMAs over Kc; lane l holds
// D[8q + 4(l/32) + u][l%32] in c[4q + u] (the lane layout of this instruction), so four counts
// are packed into one word by bytes and their four parities gathered by a multiply.
typedef int v4i_t __attribute__((ext_vector_type(4)));
typedef int v16i_t __attribute__((ext_vector_type(16)));

__global__ __launch_bounds__(256) void k_f128_mul_const_pow_mfma(const f128* __restrict__ a,
                                                                  f128 t, uint32_t reps,
                                                                  f128* __restrict__ out,
                                                                  size_t n) {
    __shared__ f128 tx[128];           // t * x^k
    __shared__ uint4 As[4][4][64];     // A fragments [R][Kc][lane], 16 KiB
    if (threadIdx.x == 0) {
        f128 v = t;
        for (int k = 0; k < 128; k++) {
            tx[k] = v;
            v = f128_mul_x(v);
        }
    }
    __syncthreads();
    for (uint32_t f = threadIdx.x; f < 4 * 4 * 64; f += blockDim.x) {
        const uint32_t R = f >> 8, kc = (f >> 6) & 3, fl = f & 63, fh = fl >> 5, fc = fl & 31;
        uint32_t w4[4];
        for (int d = 0; d < 4; d++) {
            uint32_t w = 0;
            for (int q = 0; q < 4; q++) {
                const f128 v = tx[32 * kc + 16 * fh + 4 * d + q];
                const uint32_t word = R == 0 ? v.x : R == 1 ? v.y : R == 2 ? v.z : v.w;
                w |= ((word >> fc) & 1u) << (8 * q);
            }
            w4[d] = w;
        }
        As[R][kc][fl] = make_uint4(w4[0], w4[1], w4[2], w4[3]);
    }
    __syncthreads();
    const uint32_t l = threadIdx.x & 63, half = l >> 5, col = l & 31;
    const size_t wave = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const size_t n_waves = ((size_t)gridDim.x * blockDim.x) >> 6;
    for (size_t tile = wave; tile * 32 < n; tile += n_waves) {
        const size_t j = tile * 32 + col;
        f128 x = j < n ? a[j] : f128_zero();
        for (uint32_t rep = 0; rep < reps; rep++) {
            const uint32_t xw[4] = {x.x, x.y, x.z, x.w};
            v4i_t B[4];
#pragma unroll
            for (int kc = 0; kc < 4; kc++) {
                const uint32_t h = (xw[kc] >> (16 * half)) & 0xffffu;
#pragma unroll
                for (int d = 0; d < 4; d++)
                    B[kc][d] = (int)(__umul24((h >> (4 * d)) & 15u, 0x00204081u) & 0x01010101u);
            }
            uint32_t y[4];
#pragma unroll
            for (int R = 0; R < 4; R++) {
                v16i_t c = {};
#pragma unroll
                for (int kc = 0; kc < 4; kc++) {
                    const uint4 af = As[R][kc][l];
                    const v4i_t av = {(int)af.x, (int)af.y, (int)af.z, (int)af.w};
                    c = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, B[kc], c, 0, 0, 0);
                }
                uint32_t w = 0;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    uint32_t s = (uint32_t)c[4 * q] | ((uint32_t)c[4 * q + 1] << 8) |
                                 ((uint32_t)c[4 * q + 2] << 16) | ((uint32_t)c[4 * q + 3] << 24);
                    s &= 0x01010101u;
                    s |= s >> 7;   // bit 1 <- byte 1
                    s |= s >> 14;  // bits 2, 3 <- bytes 2, 3
                    w |= (s & 15u) << (8 * q + 4 * half);
                }
                y[R] = w | __shfl_xor(w, 32, 64);
            }
            x = make_uint4(y[0], y[1], y[2], y[3]);
        }
        if (half == 0 && j < n) out[j] = x;
    }
}

// out[i] = a[i] * t^reps through the LDS Shoup table (the NTT's and folds' product).
__global__ __launch_bounds__(256) void k_f128_mul_const_pow_shoup(const f128* __restrict__ a,
                                                                   f128 t, uint32_t reps,
                                                                   f128* __restrict__ out,
                                                                   size_t n) {
    __shared__ f128 table[SHOUP_ENTRIES];
    if (threadIdx.x < 32) shoup_build_row(table, t, threadIdx.x);
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        f128 x = a[i];
        for (uint32_t rep = 0; rep < reps; rep++) x = shoup_mul(table, x);
        out[i] = x;
    }
}

static inline unsigned grid_for(size_t n, unsigned cap) {
    size_t g = (n + 255) / 256;
    if (g < 1) g = 1;
    return (unsigned)(g > cap ? cap : g);
}

hipError_t launch_f128_mul_batch(const uint4* a, const uint4* b, uint4* out, size_t n,
                                 hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_f128_mul_batch<<<grid_for(n, 65535), 256, 0, s>>>(a, b, out, n);
    return hipGetLastError();
}

hipError_t launch_f128_mul_const_batch(const uint4* a, uint4 t, uint4* out, size_t n,
                                       hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_f128_mul_const_batch<<<grid_for(n, 4096), 256, 0, s>>>(a, t, out, n);
    return hipGetLastError();
}

hipError_t launch_f128_mul_const_pow(const uint4* a, uint4 t, uint32_t reps, int method,
                                     uint4* out, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (method == 1)
        k_f128_mul_const_pow_mfma<<<grid_for((n + 7) / 8, 2048), 256, 0, s>>>(a, t, reps, out, n);
    else
        k_f128_mul_const_pow_shoup<<<grid_for(n, 4096), 256, 0, s>>>(a, t, reps, out, n);
    return hipGetLastError();
}

}  // namespace flock
