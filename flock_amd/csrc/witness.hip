// Witness data movement: the union dense-stack gather (compact_witness,
// crates/flock-core/src/union.rs:418-490).
//
// The padded union buffer is batch-major: chunk-column c of slot t holds its 2^nu rows at
// words [start_t + (c << nu), +2^nu). The committed stack q takes the first n_t rows of each
// used column, columns of a slot back to back, slots in order, then zeros up to the
// committed length. Every word moves once: an HBM copy, 32 bytes per kept word.
#include "f128_device.h"
#include "flock_internal.h"

namespace flock {

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
