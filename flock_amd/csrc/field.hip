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

}  // namespace flock
