// Merkle trees over codeword rows (crates/flock-core/src/merkle.rs), SHA-256 or BLAKE3
// (HashKind, hash.rs:23-27).
//
// Flat layout (merkle.rs:15-22): tree[0..N) = leaf hashes, then each level above, the
// root at tree[2N-2]. A leaf is one codeword position: t lanes x 16 B, little-endian
// {lo, hi} words (pcs/commit.rs:601-614).
//   SHA-256: hash_leaf = SHA-256(leaf bytes) (:205-216), hash_pair = SHA-256(left || right)
//            (:220-232), no domain separation.
//   BLAKE3:  hash_leaf = the leaf's non-root chaining value, hash_pair = a PARENT
//            compression of the two chaining values (:195-205).
//
// Kernels: k_merkle_leaves hashes one leaf per thread (both hashes are sequential within a
// message) and then builds the workgroup's subtree in LDS, so a 256-thread group emits
// its 256 leaves and the 8 levels above them in one launch. k_merkle_reduce does the same
// for an internal level (2 children per thread). A node's digest is kept as the eight
// message words its parent reads: big-endian for SHA-256, little-endian for BLAKE3.
#include "blake3_common.h"
#include "f128_device.h"
#include "flock_internal.h"
#include "sha256_device.h"

namespace flock {

// The second block of a 64-byte message is the fixed padding (0x80000000, zeros, bit length
// 512). Its message schedule is written with plain shifts and fully unrolled, so the compiler
// folds all 64 words to constants and only the rounds remain.
__device__ __forceinline__ uint32_t rotr_plain(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
__device__ __forceinline__ void sha256_compress_pad64(uint32_t h[8]) {
    uint32_t w[64];
#pragma unroll
    for (int i = 0; i < 16; i++) w[i] = i == 0 ? 0x80000000u : (i == 15 ? 512u : 0u);
#pragma unroll
    for (int i = 16; i < 64; i++) {
        const uint32_t w15 = w[i - 15], w2 = w[i - 2];
        const uint32_t s0 = rotr_plain(w15, 7) ^ rotr_plain(w15, 18) ^ (w15 >> 3);
        const uint32_t s1 = rotr_plain(w2, 17) ^ rotr_plain(w2, 19) ^ (w2 >> 10);
        w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
#pragma unroll
    for (int i = 0; i < 64; i++) sha256_round(a, b, c, d, e, f, g, hh, kSha256K[i], w[i]);
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
}

// SHA-256 of a 64-byte message given as 16 big-endian words (a parent node).
__device__ __forceinline__ void sha256_64(const uint32_t m[16], uint32_t out[8]) {
    sha256_iv(out);
    uint32_t w[16];
#pragma unroll
    for (int i = 0; i < 16; i++) w[i] = m[i];
    sha256_compress(out, w);
    sha256_compress_pad64(out);
}

// SHA-256 nodes: digests are the big-endian state words.
struct Sha256Node {
    static __device__ __forceinline__ uint32_t word(uint32_t x) { return bswap32(x); }
    static __device__ __forceinline__ void pair(const uint32_t m[16], uint32_t out[8]) {
        sha256_64(m, out);
    }
    static __device__ __forceinline__ void leaf(const uint4* __restrict__ src, uint64_t leaf_bytes,
                                                uint32_t h[8]) {
        sha256_iv(h);
        const uint64_t full = leaf_bytes / 64;
        for (uint64_t blk = 0; blk < full; blk++) {
            uint32_t w[16];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint4 v = src[blk * 4 + q];
                w[4 * q + 0] = bswap32(v.x);
                w[4 * q + 1] = bswap32(v.y);
                w[4 * q + 2] = bswap32(v.z);
                w[4 * q + 3] = bswap32(v.w);
            }
            sha256_compress(h, w);
        }
        // final block: the remaining 0/16/32/48 bytes, 0x80, zeros, 64-bit bit length
        const int rem_q = (int)((leaf_bytes % 64) / 16);
        uint32_t w[16];
#pragma unroll
        for (int i = 0; i < 16; i++) w[i] = 0u;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            if (q < rem_q) {
                const uint4 v = src[full * 4 + q];
                w[4 * q + 0] = bswap32(v.x);
                w[4 * q + 1] = bswap32(v.y);
                w[4 * q + 2] = bswap32(v.z);
                w[4 * q + 3] = bswap32(v.w);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (q == rem_q) w[4 * q] = 0x80000000u;
        const uint64_t bits = leaf_bytes * 8;
        w[14] = (uint32_t)(bits >> 32);
        w[15] = (uint32_t)bits;
        sha256_compress(h, w);
    }
};

// One BLAKE3 chunk (at most 1024 bytes, a multiple of 16) from the chaining value `cv`:
// blocks of 64 bytes with `counter` = the chunk index, CHUNK_START first, CHUNK_END last.
__device__ __forceinline__ void b3_chunk(const uint4* __restrict__ src, uint64_t bytes,
                                         uint64_t counter, uint32_t cv[8]) {
    const uint64_t nb = bytes == 0 ? 1 : (bytes + 63) / 64;
#pragma unroll 1
    for (uint64_t blk = 0; blk < nb; blk++) {
        const uint64_t len = blk + 1 < nb ? 64 : bytes - 64 * blk;
        const int quads = (int)(len / 16);
        uint32_t w[16];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (q < quads) v = src[blk * 4 + q];
            w[4 * q + 0] = v.x;
            w[4 * q + 1] = v.y;
            w[4 * q + 2] = v.z;
            w[4 * q + 3] = v.w;
        }
        const uint32_t flags = (blk == 0 ? kB3ChunkStart : 0u) | (blk + 1 == nb ? kB3ChunkEnd : 0u);
        b3_compress(cv, w, counter, (uint32_t)len, flags, cv);
    }
}

// BLAKE3 nodes: digests are the little-endian chaining-value words. MultiChunk leaves
// (over 1024 bytes) join their chunk chaining values in BLAKE3's left-balanced tree with a
// per-thread stack of completed subtrees; the single-chunk instance carries no stack.
template <bool MultiChunk>
struct Blake3Node {
    static __device__ __forceinline__ uint32_t word(uint32_t x) { return x; }
    static __device__ __forceinline__ void pair(const uint32_t m[16], uint32_t out[8]) {
        b3_compress(kB3Iv, m, 0, 64, kB3Parent, out);
    }
    static __device__ __forceinline__ void leaf(const uint4* __restrict__ src, uint64_t leaf_bytes,
                                                uint32_t h[8]) {
#pragma unroll
        for (int i = 0; i < 8; i++) h[i] = kB3Iv[i];
        if constexpr (!MultiChunk) {
            b3_chunk(src, leaf_bytes, 0, h);
        } else {
            const uint64_t n_chunks = (leaf_bytes + kB3ChunkLen - 1) / kB3ChunkLen;
            uint32_t stack[kMerkleB3MaxDepth][8];
            int depth = 0;
            for (uint64_t c = 0; c + 1 < n_chunks; c++) {
                uint32_t cv[8];
#pragma unroll
                for (int i = 0; i < 8; i++) cv[i] = kB3Iv[i];
                b3_chunk(src + c * (kB3ChunkLen / 16), kB3ChunkLen, c, cv);
                for (uint64_t total = c + 1; (total & 1) == 0; total >>= 1)
                    b3_parent(stack[--depth], cv, 0, cv);
#pragma unroll
                for (int i = 0; i < 8; i++) stack[depth][i] = cv[i];
                depth++;
            }
            const uint64_t last = n_chunks - 1;
            b3_chunk(src + last * (kB3ChunkLen / 16), leaf_bytes - last * kB3ChunkLen, last, h);
            while (depth > 0) b3_parent(stack[--depth], h, 0, h);
        }
    }
};

template <class H>
__device__ __forceinline__ void store_digest(uint8_t* tree, uint64_t node, const uint32_t h[8]) {
    uint4* dst = reinterpret_cast<uint4*>(tree + node * 32);
    dst[0] = make_uint4(H::word(h[0]), H::word(h[1]), H::word(h[2]), H::word(h[3]));
    dst[1] = make_uint4(H::word(h[4]), H::word(h[5]), H::word(h[6]), H::word(h[7]));
}

// Level offset in the flat tree: the level holding L nodes starts at 2N - 2L.
__device__ __forceinline__ uint64_t level_start(uint64_t n_leaves, uint64_t level_nodes) {
    return 2 * n_leaves - 2 * level_nodes;
}

// Build the subtree above `width` digests held in lds[0..width) (this group's slice of a
// level with `level_nodes` nodes), writing every produced node into the tree.
template <class H>
__device__ __forceinline__ void subtree_up(uint32_t (*lds)[8], int width, uint64_t group,
                                           uint64_t level_nodes, uint64_t n_leaves,
                                           uint8_t* tree) {
    int w = width;
    uint64_t ln = level_nodes;
    while (w > 1) {
        w >>= 1;
        ln >>= 1;
        uint32_t out[8];
        const int i = threadIdx.x;
        if (i < w) {
            uint32_t m[16];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                m[j] = lds[2 * i][j];
                m[8 + j] = lds[2 * i + 1][j];
            }
            H::pair(m, out);
        }
        __syncthreads();
        if (i < w) {
#pragma unroll
            for (int j = 0; j < 8; j++) lds[i][j] = out[j];
            store_digest<H>(tree, level_start(n_leaves, ln) + group * w + i, out);
        }
        __syncthreads();
    }
}

template <class H>
__global__ __launch_bounds__(256) void k_merkle_leaves(const uint8_t* __restrict__ data,
                                                        uint64_t n_leaves, uint64_t leaf_bytes,
                                                        uint8_t* __restrict__ tree) {
    __shared__ uint32_t lds[256][8];
    const int width = blockDim.x;
    const uint64_t leaf = (uint64_t)blockIdx.x * width + threadIdx.x;
    const uint4* src = reinterpret_cast<const uint4*>(data + leaf * leaf_bytes);
    uint32_t h[8];
    H::leaf(src, leaf_bytes, h);
#pragma unroll
    for (int j = 0; j < 8; j++) lds[threadIdx.x][j] = h[j];
    store_digest<H>(tree, leaf, h);
    __syncthreads();
    subtree_up<H>(lds, width, blockIdx.x, n_leaves, n_leaves, tree);
}

// Internal level with `level_nodes` nodes starting at tree node `in_start`; each group
// takes 2*blockDim.x children.
template <class H>
__global__ __launch_bounds__(256) void k_merkle_reduce(uint8_t* __restrict__ tree,
                                                        uint64_t n_leaves, uint64_t level_nodes) {
    __shared__ uint32_t lds[256][8];
    const int width = blockDim.x;  // parents produced by this group at the first step
    const uint64_t in_start = level_start(n_leaves, level_nodes);
    const uint64_t parent = (uint64_t)blockIdx.x * width + threadIdx.x;
    const uint4* src = reinterpret_cast<const uint4*>(tree + (in_start + 2 * parent) * 32);
    uint32_t m[16];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint4 v = src[q];
        m[4 * q + 0] = H::word(v.x);
        m[4 * q + 1] = H::word(v.y);
        m[4 * q + 2] = H::word(v.z);
        m[4 * q + 3] = H::word(v.w);
    }
    uint32_t h[8];
    H::pair(m, h);
    const uint64_t parents = level_nodes >> 1;
#pragma unroll
    for (int j = 0; j < 8; j++) lds[threadIdx.x][j] = h[j];
    store_digest<H>(tree, level_start(n_leaves, parents) + parent, h);
    __syncthreads();
    subtree_up<H>(lds, width, blockIdx.x, parents, n_leaves, tree);
}

template <class H>
static hipError_t merkle_tree(const uint8_t* data, uint64_t n_leaves, uint64_t leaf_bytes,
                              uint8_t* tree, uint64_t top, hipStream_t s) {
    const uint64_t width = n_leaves < 256 ? n_leaves : 256;
    k_merkle_leaves<H><<<(unsigned)(n_leaves / width), (unsigned)width, 0, s>>>(data, n_leaves,
                                                                               leaf_bytes, tree);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    uint64_t level = n_leaves / width;  // nodes at the current top
    while (level > top) {
        const uint64_t parents = level / 2;
        // a group of w threads ends at parents / w nodes: stop at the top layer asked for
        uint64_t w = parents < 256 ? parents : 256;
        if (parents / w < top) w = parents / top;
        k_merkle_reduce<H><<<(unsigned)(parents / w), (unsigned)w, 0, s>>>(tree, n_leaves, level);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        level = parents / w;
    }
    return hipSuccess;
}

hipError_t launch_merkle_tree(const uint8_t* data, uint64_t n_leaves, uint64_t leaf_bytes,
                              uint32_t kind, uint8_t* tree, hipStream_t s, uint64_t top_nodes) {
    if (n_leaves == 0 || (n_leaves & (n_leaves - 1)) || leaf_bytes % 16 || leaf_bytes == 0 ||
        top_nodes == 0 || (top_nodes & (top_nodes - 1)) || top_nodes > n_leaves)
        return hipErrorInvalidValue;
    const uint64_t top = top_nodes;
    if (kind == kMerkleSha256) return merkle_tree<Sha256Node>(data, n_leaves, leaf_bytes, tree, top, s);
    if (kind != kMerkleBlake3) return hipErrorInvalidValue;
    if (leaf_bytes <= kB3ChunkLen)
        return merkle_tree<Blake3Node<false>>(data, n_leaves, leaf_bytes, tree, top, s);
    // the per-thread stack holds one chaining value per level of the chunk tree
    if ((leaf_bytes + kB3ChunkLen - 1) / kB3ChunkLen > (1ull << kMerkleB3MaxDepth))
        return hipErrorInvalidValue;
    return merkle_tree<Blake3Node<true>>(data, n_leaves, leaf_bytes, tree, top, s);
}

}  // namespace flock
