// BLAKE3 compression and hashing modes, from the BLAKE3 specification (O'Connor, Aumasson,
// Neves, Wilcox-O'Hearn, 2020, section 2). The compression is __host__ __device__: the
// device Merkle kernels (merkle.hip) and the host registry digest (api.hip) share it.
//
// Uses in the reference (crate blake3 1.8.5):
//   merkle.rs:195-205   leaf = non-root chaining value of the leaf bytes; parent =
//                       PARENT compression of (left cv || right cv), no ROOT flag
//   schedule.rs:542-580 Registry::digest = root hash of the absorbed stream
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

namespace flock {

constexpr uint32_t kB3Iv[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                               0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
constexpr uint32_t kB3ChunkStart = 1, kB3ChunkEnd = 2, kB3Parent = 4, kB3Root = 8;
constexpr uint64_t kB3ChunkLen = 1024;

// Message word order of each of the 7 rounds: round r reads m[sched[r][i]], the fixed word
// permutation applied r times.
struct B3Sched {
    uint8_t s[7][16];
};
constexpr B3Sched make_b3_sched() {
    constexpr uint8_t perm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
    B3Sched t{};
    for (int i = 0; i < 16; i++) t.s[0][i] = (uint8_t)i;
    for (int r = 1; r < 7; r++)
        for (int i = 0; i < 16; i++) t.s[r][i] = t.s[r - 1][perm[i]];
    return t;
}
constexpr B3Sched kB3Sched = make_b3_sched();

__host__ __device__ __forceinline__ uint32_t b3_rotr(uint32_t x, int n) {
    return (x >> n) | (x << (32 - n));
}

__host__ __device__ __forceinline__ void b3_g(uint32_t& a, uint32_t& b, uint32_t& c,
                                              uint32_t& d, uint32_t x, uint32_t y) {
    a = a + b + x;
    d = b3_rotr(d ^ a, 16);
    c = c + d;
    b = b3_rotr(b ^ c, 12);
    a = a + b + y;
    d = b3_rotr(d ^ a, 8);
    c = c + d;
    b = b3_rotr(b ^ c, 7);
}

// The compression's first 8 output words (a chaining value, or a 32-byte root hash).
__host__ __device__ __forceinline__ void b3_compress(const uint32_t cv[8], const uint32_t m[16],
                                                     uint64_t counter, uint32_t block_len,
                                                     uint32_t flags, uint32_t out[8]) {
    uint32_t v0 = cv[0], v1 = cv[1], v2 = cv[2], v3 = cv[3], v4 = cv[4], v5 = cv[5],
             v6 = cv[6], v7 = cv[7];
    uint32_t v8 = kB3Iv[0], v9 = kB3Iv[1], v10 = kB3Iv[2], v11 = kB3Iv[3];
    uint32_t v12 = (uint32_t)counter, v13 = (uint32_t)(counter >> 32), v14 = block_len,
             v15 = flags;
#pragma unroll
    for (int r = 0; r < 7; r++) {
        const uint8_t* s = kB3Sched.s[r];
        b3_g(v0, v4, v8, v12, m[s[0]], m[s[1]]);
        b3_g(v1, v5, v9, v13, m[s[2]], m[s[3]]);
        b3_g(v2, v6, v10, v14, m[s[4]], m[s[5]]);
        b3_g(v3, v7, v11, v15, m[s[6]], m[s[7]]);
        b3_g(v0, v5, v10, v15, m[s[8]], m[s[9]]);
        b3_g(v1, v6, v11, v12, m[s[10]], m[s[11]]);
        b3_g(v2, v7, v8, v13, m[s[12]], m[s[13]]);
        b3_g(v3, v4, v9, v14, m[s[14]], m[s[15]]);
    }
    out[0] = v0 ^ v8; out[1] = v1 ^ v9; out[2] = v2 ^ v10; out[3] = v3 ^ v11;
    out[4] = v4 ^ v12; out[5] = v5 ^ v13; out[6] = v6 ^ v14; out[7] = v7 ^ v15;
}

__host__ __device__ __forceinline__ void b3_parent(const uint32_t l[8], const uint32_t r[8],
                                                   uint32_t flags, uint32_t out[8]) {
    uint32_t m[16];
    for (int i = 0; i < 8; i++) {
        m[i] = l[i];
        m[8 + i] = r[i];
    }
    b3_compress(kB3Iv, m, 0, 64, kB3Parent | flags, out);
}

// Host streaming hasher: root hash (finish) or non-root chaining value (finish_non_root).
// The newest full block stays buffered until more input arrives, since only the input's
// last block of its last chunk may carry ROOT.
class Blake3Host {
  public:
    Blake3Host() { reset(); }
    void reset() {
        memcpy(cv_, kB3Iv, sizeof(cv_));
        chunk_ = 0;
        blocks_ = 0;
        buf_len_ = 0;
        depth_ = 0;
    }
    void update(const void* data, size_t len) {
        const uint8_t* p = static_cast<const uint8_t*>(data);
        while (len > 0) {
            if (buf_len_ == 64) flush_block();
            size_t take = 64 - buf_len_;
            if (take > len) take = len;
            memcpy(buf_ + buf_len_, p, take);
            buf_len_ += (uint32_t)take;
            p += take;
            len -= take;
        }
    }
    void u32(uint32_t v) {
        uint8_t b[4];
        for (int i = 0; i < 4; i++) b[i] = (uint8_t)(v >> (8 * i));
        update(b, 4);
    }
    void u64(uint64_t v) {
        uint8_t b[8];
        for (int i = 0; i < 8; i++) b[i] = (uint8_t)(v >> (8 * i));
        update(b, 8);
    }
    void finish(uint8_t out[32]) const { finish_flags(true, out); }
    void finish_non_root(uint8_t out[32]) const { finish_flags(false, out); }

  private:
    void block_words(uint32_t w[16]) const {
        uint8_t blk[64] = {0};
        memcpy(blk, buf_, buf_len_);
        for (int i = 0; i < 16; i++)
            w[i] = (uint32_t)blk[4 * i] | (uint32_t)blk[4 * i + 1] << 8 |
                   (uint32_t)blk[4 * i + 2] << 16 | (uint32_t)blk[4 * i + 3] << 24;
    }
    void chunk_output(uint32_t flags, uint32_t out[8]) const {
        uint32_t w[16];
        block_words(w);
        b3_compress(cv_, w, chunk_, buf_len_,
                    (blocks_ == 0 ? kB3ChunkStart : 0u) | kB3ChunkEnd | flags, out);
    }
    // The buffered 64-byte block is known not to end the input.
    void flush_block() {
        if (blocks_ == 15) {
            // it ends its chunk: push the chunk's chaining value, merging one completed
            // subtree per trailing zero bit of the number of chunks so far
            uint32_t c[8];
            chunk_output(0, c);
            for (uint64_t total = chunk_ + 1; (total & 1) == 0; total >>= 1)
                b3_parent(stack_[--depth_], c, 0, c);
            memcpy(stack_[depth_++], c, sizeof(c));
            chunk_++;
            blocks_ = 0;
            memcpy(cv_, kB3Iv, sizeof(cv_));
        } else {
            uint32_t w[16];
            block_words(w);
            b3_compress(cv_, w, chunk_, 64, blocks_ == 0 ? kB3ChunkStart : 0u, cv_);
            blocks_++;
        }
        buf_len_ = 0;
    }
    void finish_flags(bool root, uint8_t out[32]) const {
        uint32_t c[8];
        if (depth_ == 0) {
            chunk_output(root ? kB3Root : 0u, c);
        } else {
            chunk_output(0, c);
            for (uint32_t i = depth_; i-- > 0;)
                b3_parent(stack_[i], c, (root && i == 0) ? kB3Root : 0u, c);
        }
        for (int i = 0; i < 8; i++)
            for (int j = 0; j < 4; j++) out[4 * i + j] = (uint8_t)(c[i] >> (8 * j));
    }

    uint32_t cv_[8];
    uint64_t chunk_;
    uint32_t blocks_;
    uint8_t buf_[64];
    uint32_t buf_len_;
    uint32_t stack_[64][8];
    uint32_t depth_;
};

}  // namespace flock
