/*
 * blake3_hash.c — BLAKE3 hashing modes and the registry digest, restated on the CPU.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the device BLAKE3 Merkle tree and for the host
 * registry digest of libflock_hip.so.
 *
 * BLAKE3 follows its specification (O'Connor, Aumasson, Neves, Wilcox-O'Hearn, "BLAKE3: one
 * function, fast everywhere", 2020, section 2): the input is cut into 1024-byte chunks; a
 * chunk is compressed block by block with the chunk index as the counter, CHUNK_START on its
 * first block and CHUNK_END on its last; the chunk chaining values are joined by a
 * left-balanced binary tree of PARENT compressions; ROOT marks the last compression of the
 * root node only. The compression itself is or_blake3_compress (blake3_r1cs.c).
 *
 * Reference call sites (the `blake3` crate 1.8.5 is not vendored):
 *   merkle.rs:195-205  leaf = Hasher::update(leaf).finalize_non_root(),
 *                      parent = merge_subtrees_non_root(l, r) (PARENT, no ROOT)
 *   schedule.rs:542-580 Registry::digest = root hash of the absorbed byte stream
 *   r1cs.rs:419-428    absorb_matrix
 */
#include <string.h>

#include "flock_oracle.h"

enum { B3_CHUNK_START = 1, B3_CHUNK_END = 2, B3_PARENT = 4, B3_ROOT = 8 };

static const uint32_t kIV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};

static void words_le(const uint8_t* b, uint32_t n_bytes, uint32_t w[16]) {
    uint8_t blk[64] = {0};
    memcpy(blk, b, n_bytes);
    for (int i = 0; i < 16; i++)
        w[i] = (uint32_t)blk[4 * i] | (uint32_t)blk[4 * i + 1] << 8 |
               (uint32_t)blk[4 * i + 2] << 16 | (uint32_t)blk[4 * i + 3] << 24;
}

static void cv_bytes(const uint32_t cv[8], uint8_t out[32]) {
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 4; j++) out[4 * i + j] = (uint8_t)(cv[i] >> (8 * j));
}

static void parent(const uint32_t l[8], const uint32_t r[8], uint32_t flags, uint32_t out[8]) {
    uint32_t m[16], o[16];
    memcpy(m, l, 32);
    memcpy(m + 8, r, 32);
    or_blake3_compress(kIV, m, 0, 64, B3_PARENT | flags, o);
    memcpy(out, o, 32);
}

/* Incremental hasher. The last buffered block is compressed only once more input arrives,
 * because whether it ends the input (and then takes ROOT) is unknown until then. */
typedef struct {
    uint32_t cv[8];
    uint64_t chunk;      /* index of the chunk in progress */
    uint32_t blocks;     /* blocks already compressed in it */
    uint8_t buf[64];
    uint32_t buf_len;
    uint32_t stack[64][8];
    uint32_t depth;
} b3;

static void b3_init(b3* h) {
    memset(h, 0, sizeof(*h));
    memcpy(h->cv, kIV, 32);
}

static void b3_chunk_out(const b3* h, uint32_t flags, uint32_t out[8]) {
    uint32_t w[16], o[16];
    words_le(h->buf, h->buf_len, w);
    or_blake3_compress(h->cv, w, h->chunk, h->buf_len,
                       (h->blocks == 0 ? B3_CHUNK_START : 0) | B3_CHUNK_END | flags, o);
    memcpy(out, o, 32);
}

static void b3_update(b3* h, const uint8_t* p, size_t len) {
    while (len > 0) {
        if (h->buf_len == 64) {
            if (h->blocks == 15) {
                /* the chunk's 16th block: close it and merge completed subtrees; after
                 * `total` chunks, one merge per trailing zero bit of `total` */
                uint32_t cv[8];
                b3_chunk_out(h, 0, cv);
                uint64_t total = h->chunk + 1;
                while ((total & 1) == 0) {
                    parent(h->stack[--h->depth], cv, 0, cv);
                    total >>= 1;
                }
                memcpy(h->stack[h->depth++], cv, 32);
                h->chunk++;
                h->blocks = 0;
                memcpy(h->cv, kIV, 32);
            } else {
                uint32_t w[16], o[16];
                words_le(h->buf, 64, w);
                or_blake3_compress(h->cv, w, h->chunk, 64, h->blocks == 0 ? B3_CHUNK_START : 0,
                                   o);
                memcpy(h->cv, o, 32);
                h->blocks++;
            }
            h->buf_len = 0;
        }
        size_t take = 64 - h->buf_len;
        if (take > len) take = len;
        memcpy(h->buf + h->buf_len, p, take);
        h->buf_len += (uint32_t)take;
        p += take;
        len -= take;
    }
}

static void b3_finalize(const b3* h, int root, uint8_t out[32]) {
    uint32_t cv[8];
    if (h->depth == 0) {
        b3_chunk_out(h, root ? B3_ROOT : 0, cv);
    } else {
        b3_chunk_out(h, 0, cv);
        for (uint32_t i = h->depth; i-- > 0;) parent(h->stack[i], cv, (root && i == 0) ? B3_ROOT : 0, cv);
    }
    cv_bytes(cv, out);
}

void or_blake3_hash(const uint8_t* data, size_t len, uint8_t out[32]) {
    b3 h;
    b3_init(&h);
    b3_update(&h, data, len);
    b3_finalize(&h, 1, out);
}

/* the same stream fed in pieces of the given lengths (the buffering must not matter) */
void or_blake3_hash_pieces(const uint8_t* data, const size_t* piece_lens, size_t n_pieces,
                           uint8_t out[32]) {
    b3 h;
    b3_init(&h);
    for (size_t i = 0; i < n_pieces; i++) {
        b3_update(&h, data, piece_lens[i]);
        data += piece_lens[i];
    }
    b3_finalize(&h, 1, out);
}

/* blake3_leaf_cv, merkle.rs:195-199 */
void or_blake3_leaf_cv(const uint8_t* data, size_t len, uint8_t out[32]) {
    b3 h;
    b3_init(&h);
    b3_update(&h, data, len);
    b3_finalize(&h, 0, out);
}

static void cv_words(const uint8_t* b, uint32_t w[8]) {
    for (int i = 0; i < 8; i++)
        w[i] = (uint32_t)b[4 * i] | (uint32_t)b[4 * i + 1] << 8 | (uint32_t)b[4 * i + 2] << 16 |
               (uint32_t)b[4 * i + 3] << 24;
}

/* blake3_parent_cv, merkle.rs:201-205 */
void or_blake3_parent_cv(const uint8_t l[32], const uint8_t r[32], uint8_t out[32]) {
    uint32_t a[8], b[8], o[8];
    cv_words(l, a);
    cv_words(r, b);
    parent(a, b, 0, o);
    cv_bytes(o, out);
}

/* merkle_tree_impl over hash_leaf / hash_pair (merkle.rs:205-232, 543-617), HashKind::Blake3 */
static int merkle_blake3(const uint8_t* data, size_t data_len, size_t num_leaves,
                         uint8_t* tree) {
    size_t leaf = data_len / num_leaves;
    for (size_t i = 0; i < num_leaves; i++) or_blake3_leaf_cv(data + i * leaf, leaf, tree + 32 * i);
    size_t rs = 0, rl = num_leaves;
    while (rl > 1) {
        size_t nl = rl >> 1;
        for (size_t i = 0; i < nl; i++)
            or_blake3_parent_cv(tree + 32 * (rs + 2 * i), tree + 32 * (rs + 2 * i + 1),
                                tree + 32 * (rs + rl + i));
        rs += rl;
        rl = nl;
    }
    return 0;
}

int or_merkle_tree_kind(const uint8_t* data, size_t data_len, size_t num_leaves, uint32_t kind,
                        uint8_t* tree) {
    if (kind == 0) return or_merkle_tree(data, data_len, num_leaves, tree);
    if (kind != 1) return -1;
    if (num_leaves == 0 || (num_leaves & (num_leaves - 1)) || data_len % num_leaves) return -1;
    return merkle_blake3(data, data_len, num_leaves, tree);
}

/* ---- Registry::digest, schedule.rs:542-580 ---- */

static void put_u32(b3* h, uint32_t v) {
    uint8_t b[4];
    for (int i = 0; i < 4; i++) b[i] = (uint8_t)(v >> (8 * i));
    b3_update(h, b, 4);
}

static void put_u64(b3* h, uint64_t v) {
    uint8_t b[8];
    for (int i = 0; i < 8; i++) b[i] = (uint8_t)(v >> (8 * i));
    b3_update(h, b, 8);
}

/* absorb_matrix, r1cs.rs:419-428 */
static int absorb_matrix(b3* h, const or_sparse* m) {
    put_u64(h, m->num_rows);
    put_u64(h, m->num_cols);
    for (uint64_t r = 0; r < m->num_rows; r++) {
        if (m->identity) {
            put_u64(h, 1);
            put_u64(h, r);
            continue;
        }
        if (!m->row_ptr) return -1;
        const uint64_t a = m->row_ptr[r], b = m->row_ptr[r + 1];
        if (b < a) return -1;
        put_u64(h, b - a);
        for (uint64_t i = a; i < b; i++) put_u64(h, m->cols[i]);
    }
    return 0;
}

int or_registry_digest(const or_type_stmt* types, size_t n_types, uint32_t nu, uint8_t out[32]) {
    if (n_types == 0) return -1;
    b3 h;
    b3_init(&h);
    static const uint8_t label[] = "flock-registry-v1";
    b3_update(&h, label, sizeof(label) - 1);
    const uint8_t version = 1;
    b3_update(&h, &version, 1);
    put_u32(&h, nu);
    put_u32(&h, (uint32_t)n_types);
    /* Registry::new's stable sort by k_log, descending (schedule.rs:386-389) */
    for (int k = 64; k >= 0; k--) {
        for (size_t t = 0; t < n_types; t++) {
            const or_type_stmt* ty = &types[t];
            if ((int)ty->k_log != k) continue;
            put_u32(&h, ty->k_log);
            put_u64(&h, ty->useful_bits);
            const uint8_t present = ty->has_const_pin ? 1 : 0;
            b3_update(&h, &present, 1);
            put_u64(&h, ty->has_const_pin ? ty->const_pin : 0);
            if (absorb_matrix(&h, &ty->a) || absorb_matrix(&h, &ty->b) ||
                absorb_matrix(&h, &ty->c))
                return -1;
        }
    }
    b3_finalize(&h, 1, out);
    return 0;
}
