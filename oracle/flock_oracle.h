/*
 * flock_oracle.h — CPU restatement of the Flock `prove_fast` data-parallel core.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity checker for the HIP
 * kernels in flock_amd/csrc and the timed CPU baseline in bench.py. Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The shipped library (libflock_hip.so) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference checkout). Field elements are F128 = {lo, hi}, lo holding the
 * coefficients of x^0..x^63, reduction polynomial x^128 + x^7 + x^2 + x + 1
 * (crates/flock-core/src/field/gf2_128.rs:10-27).
 */
#ifndef FLOCK_ORACLE_H
#define FLOCK_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint64_t lo, hi; } or_f128;

/* ---- field (crates/flock-core/src/field/gf2_128.rs) ---- */
or_f128 or_f128_mul(or_f128 a, or_f128 b);            /* Mul, :174-201 */
or_f128 or_f128_mul_portable(or_f128 a, or_f128 b);   /* software::ghash_mul, gf2_128/portable.rs:40-43 */
or_f128 or_f128_inv(or_f128 a);                       /* inv, :67-80 */
or_f128 or_f128_mul_by_x(or_f128 a);                  /* mul_by_x, :213-220 */
void or_f128_mul_batch(const or_f128* a, const or_f128* b, or_f128* out, size_t n);
int or_has_pclmul(void);
/* AVX-512 with VPCLMULQDQ (the 4-lane multiply of the batch products and the NTT rows) */
int or_has_vpclmul(void);

/* ---- SplitMix64 (crates/flock-prover/src/bin/dump_commit_vectors.rs:42-68) ---- */
void or_splitmix_f128(uint64_t seed, or_f128* out, size_t n);

/* ---- additive NTT (crates/flock-core/src/ntt/additive_ntt_f128.rs) ---- */
/* evals table for the standard basis {1, x, .., x^(L-1)}: row i has L-i entries,
 * stored row-major in an L x L array (unused tail zero). :67-95, :127-133 */
int or_ntt_evals(int L, or_f128* evals_LxL);
or_f128 or_ntt_twiddle(const or_f128* evals_LxL, int L, int layer, uint64_t block); /* :141-144 */
/* interleaved_scalar_live_from_layer, :287-319. data has 2^log_d * num_ntts words. */
int or_ntt_forward_interleaved(or_f128* data, size_t num_ntts, size_t live, int log_d, int L,
                               int start_layer);
int or_ntt_inverse_scalar(or_f128* data, int log_d, int L); /* :651-673 */

/* ---- SHA-256 (FIPS 180-4) and the Merkle tree (crates/flock-core/src/merkle.rs) ---- */
void or_sha256(const uint8_t* data, size_t len, uint8_t out[32]);
/* merkle_tree_sequential, :593-617 — tree has 2*num_leaves-1 nodes of 32 bytes */
int or_merkle_tree(const uint8_t* data, size_t data_len, size_t num_leaves, uint8_t* tree);
/* the same tree under HashKind (hash.rs:23-27): 0 = SHA-256, 1 = BLAKE3 (merkle.rs:195-232) */
int or_merkle_tree_kind(const uint8_t* data, size_t data_len, size_t num_leaves, uint32_t kind,
                        uint8_t* tree);

/* ---- BLAKE3 modes and Registry::digest (blake3_hash.c) ---- */
void or_blake3_hash(const uint8_t* data, size_t len, uint8_t out[32]);
void or_blake3_hash_pieces(const uint8_t* data, const size_t* piece_lens, size_t n_pieces,
                           uint8_t out[32]);
void or_blake3_leaf_cv(const uint8_t* data, size_t len, uint8_t out[32]);     /* merkle.rs:195-199 */
void or_blake3_parent_cv(const uint8_t l[32], const uint8_t r[32], uint8_t out[32]); /* :201-205 */
/* SparseBinaryMatrix (r1cs.rs:14-18) as CSR; identity = 1 stands for rows [[0], [1], ...]
 * (common.rs:231-237) without materialising them. */
typedef struct {
    uint64_t num_rows, num_cols;
    const uint64_t* row_ptr;
    const uint32_t* cols;
    uint32_t identity;
} or_sparse;
/* a boolean TableType's digested fields (schedule.rs:150-186) */
typedef struct {
    uint32_t k_log;
    uint32_t has_const_pin;
    uint64_t useful_bits;
    uint64_t const_pin;
    or_sparse a, b, c;
} or_type_stmt;
/* Registry::digest (schedule.rs:542-580) for boolean types without IO schemas */
int or_registry_digest(const or_type_stmt* types, size_t n_types, uint32_t nu, uint8_t out[32]);

/* ---- PCS commit (crates/flock-core/src/pcs/commit.rs) ---- */
typedef struct {
    uint32_t m, log_inv_rate, log_batch_size;
    uint32_t num_lanes;   /* 0 = None (all 2^log_batch_size lanes) */
    uint32_t merkle_hash; /* HashKind: 0 = SHA-256, 1 = BLAKE3 */
} or_pcs_params;
int or_pcs_sizes(const or_pcs_params* p, size_t* msg_len, size_t* cw_len, size_t* n_leaves,
                 int* k_code, size_t* num_ntts);
/* commit_lane_major, :474-510 (+ replicate_lane_major_fill :516-536, finalize :572-633) */
int or_commit_lane_major(const or_f128* q, size_t q_len, const or_pcs_params* p, or_f128* codeword,
                         uint8_t* tree);
/* commit / commit_into, :345-397 (+ replicate_message_fill :543-568) */
int or_commit(const or_f128* msg, size_t msg_len, const or_pcs_params* p, or_f128* codeword,
              uint8_t* tree);

/* ---- Fiat-Shamir challenger, SHA-256 discipline (crates/flock-core/src/challenger.rs:609-1044) ---- */
typedef struct or_fs or_fs;
or_fs* or_fs_new(const uint8_t* domain, size_t len);
void or_fs_free(or_fs* fs);
void or_fs_observe_label(or_fs* fs, const uint8_t* label, size_t len);
void or_fs_observe_f128(or_fs* fs, or_f128 v);
void or_fs_observe_f128_slice(or_fs* fs, const or_f128* v, size_t n);
void or_fs_observe_bytes(or_fs* fs, const uint8_t* b, size_t n);
or_f128 or_fs_sample_f128(or_fs* fs);
void or_fs_sample_f128_vec(or_fs* fs, or_f128* out, size_t n);
uint64_t or_fs_grind_pow(or_fs* fs, uint32_t bits);
void or_fs_digest(or_fs* fs, uint8_t out[32]);

/* ---- sumcheck rounds ---- */
/* round_msg_lsb, crates/flock-core/src/pcs/ligerito.rs:3204-3240 */
void or_round_msg_lsb(const or_f128* f, const or_f128* b, size_t n, or_f128 msg[2]);
/* fold_and_msg_lsb, ligerito.rs:3372-3414; nf/nb get n/2 words */
void or_fold_and_msg_lsb(const or_f128* f, const or_f128* b, size_t n, or_f128 r, or_f128* nf,
                         or_f128* nb, or_f128 msg[2]);
/* build_eq, crates/flock-core/src/zerocheck/univariate_skip.rs:31-47 */
void or_build_eq(const or_f128* r, size_t n, or_f128* out);
/* round_pair_naive, crates/flock-core/src/zerocheck/multilinear.rs:331-355 */
void or_round_pair(const or_f128* a, const or_f128* b, size_t n, const or_f128* r, or_f128 out[2]);
/* fold_in_place_pair, multilinear.rs:1148-1163 (out-of-place: a_out/b_out get n/2) */
void or_fold_pair(const or_f128* a, const or_f128* b, size_t n, or_f128 c, or_f128* a_out,
                  or_f128* b_out);

/* Full product sumcheck + eq-weighted zerocheck tail over one transcript. Mirrors the
 * SumcheckProver loop at ligerito.rs:4986-5011 and the tail loop at
 * crates/flock-core/src/zerocheck.rs:790-975. msgs: 2*log_n product messages then
 * 2*log_n zerocheck messages; finals: f, b (product), a, b (zerocheck). */
int or_sumcheck_transcript(const uint8_t* domain, size_t dlen, const or_f128* a, const or_f128* b,
                           int log_n, or_f128* msgs, or_f128 finals[4]);

/* ---- GF(2^8) and the embedding phi_8 (crates/flock-core/src/field/{gf2_8,phi8}.rs) ---- */
uint8_t or_f8_mul(uint8_t a, uint8_t b);   /* AES polynomial 0x11B, gf2_8.rs:60-137 */
uint8_t or_f8_inv(uint8_t a);              /* gf2_8.rs:26-37 */
or_f128 or_phi8(uint8_t v);                /* PHI_8_TABLE[v], phi8.rs:19-1050 */
/* GF(2^8) additive NTT (crates/flock-core/src/ntt.rs:20-165): domain beta + span{1..2^(k-1)} */
void or_f8_ntt_forward(uint8_t* v, int k, uint8_t beta);
void or_f8_ntt_inverse(uint8_t* v, int k, uint8_t beta);

/* ---- zerocheck front (crates/flock-core/src/zerocheck/*.rs) ---- */
/* round1_naive (univariate_skip.rs:68-124) on LSB-first packed bits of length 2^m;
 * r has m entries (only r[6..] used). p_ab, p_c: 64 values on Lambda. */
void or_zc_round1(const uint8_t* a, const uint8_t* b, const uint8_t* c, int m, const or_f128* r,
                  or_f128 p_ab[64], or_f128 p_c[64]);
void or_zc_round1_naive(const uint8_t* a, const uint8_t* b, const uint8_t* c, int m, const or_f128* r,
                  or_f128 p_ab[64], or_f128 p_c[64]);
/* OpenMP threads of the parallel loops (n <= 0: the runtime default) */
void or_set_threads(int n);
/* s_hat_v_c of round1_extract_c_packed_with_s_hat_v (univariate_skip.rs:384-489), 128 values */
void or_zc_s_hat_v_c(const uint8_t* c, int m, const or_f128* r, or_f128 out[128]);
/* fold_1b_rows (pcs/ring_switch.rs:182-230, 410-470) on n 128-bit rows */
void or_fold_1b_rows(const or_f128* rows, const or_f128* t, size_t n, or_f128 out[128]);
/* Lagrange weights on phi_8(offset .. offset+63) at z (multilinear.rs:190-210) */
void or_lagrange_phi8(int offset, or_f128 z, or_f128 w[64]);
/* fold_at_z_naive (multilinear.rs:290-310) on packed bits: out[x] = sum_s bit(x*64+s) L_s(z) */
void or_fold_at_z(const uint8_t* bits, int m, or_f128 z, or_f128* out);
/* The fixed inner eq constants r[6..13) (univariate_skip_optimized.rs:91-139). */
void or_zc_inner_constants(or_f128 out[7]);
/* prove_packed_padded_inner (zerocheck.rs:618-1011), dense padding, naive round 1.
 * Outputs: round1[128] = (round1_ab, round1_c); ml[2*(m-6)]; finals[3] = (a, b, c);
 * nonces[2+m-6] (if grinding); claim[1 + 2*(m-6)] = (z, mlv_challenges, r_rest[1..]). */
int or_zerocheck_prove(or_fs* fs, const uint8_t* a, const uint8_t* b, const uint8_t* c, int m,
                       int grinding, or_f128* round1, or_f128* ml, or_f128 finals[3],
                       uint64_t* nonces, or_f128* claim);

/* ---- lincheck (crates/flock-core/src/lincheck.rs, lincheck/union.rs) ---- */
/* build_quirky_eq_table (lincheck.rs:1355-1378): out has 64 * 2^n_rest words */
void or_quirky_eq(or_f128 z_skip, const or_f128* x_rest, size_t n_rest, or_f128* out);
/* CscCircuit::fold_split (lincheck.rs:463-497) */
void or_csc_fold_split(size_t n_cols, const uint32_t* a_ptr, const uint32_t* a_rows,
                       const uint32_t* b_ptr, const uint32_t* b_rows, const or_f128* eq,
                       or_f128* sa, or_f128* sb);
/* partial_fold_packed_z{,_rows_padded} (lincheck.rs:804-835, 865-990) */
void or_partial_fold(const uint8_t* z, int k_log, int nu, size_t useful_bits, size_t n_rows,
                     const or_f128* eq_outer, or_f128* out);
/* sumcheck_round_eval_par / sumcheck_bind_top_in_place_par (lincheck.rs:1479-1523) */
void or_lc_round_eval(const or_f128* c, const or_f128* z, size_t len, or_f128 out[2]);
void or_lc_bind_top(or_f128* v, size_t len, or_f128 r);
/* prove_union_capture_z_vec_with_grinding, single boolean slot (lincheck/union.rs:203-430)
 * + column_sumcheck_prove (lincheck.rs:1907-2028). claim = (r_inner_skip, w, r_inner_rest). */
int or_lincheck_prove(or_fs* fs, int k_log, const uint32_t* a_ptr, const uint32_t* a_rows,
                      const uint32_t* b_ptr, const uint32_t* b_rows, long const_pin,
                      const uint8_t* z, int nu, size_t useful_bits, size_t n_rows,
                      or_f128 z_skip, const or_f128* x_inner_rest, const or_f128* x_outer,
                      int grinding, or_f128* rounds, or_f128 z_partial[64],
                      or_f128 matrix_evals[2], uint64_t* nonces, or_f128* claim,
                      or_f128* z_vec_out);
/* the same with KeccakLincheckCircuit (keccak.rs:678-835): k_log 16, const pin 4096 */
int or_lincheck_prove_keccak(or_fs* fs, const uint8_t* z, int nu, size_t n_rows,
                             or_f128 z_skip, const or_f128* x_inner_rest, const or_f128* x_outer,
                             int grinding, or_f128* rounds, or_f128 z_partial[64],
                             or_f128 matrix_evals[2], uint64_t* nonces, or_f128* claim,
                             or_f128* z_vec_out);

/* one boolean slot of the union lincheck: a_ptr = NULL selects the Keccak-f[1600] walker */
typedef struct {
    int k_log;
    uint64_t prefix;
    size_t useful_bits, n_rows;
    long const_pin;
    const uint8_t* z; /* the slot's lincheck stripe */
    const uint32_t *a_ptr, *a_rows, *b_ptr, *b_rows;
} or_lc_slot;
int or_lincheck_prove_union(or_fs* fs, const or_lc_slot* slots, size_t n_slots, int col_vars,
                            int nu, or_f128 z_skip, const or_f128* x_inner_rest,
                            const or_f128* x_outer, int grinding, or_f128* rounds,
                            or_f128 z_partial[64], or_f128* matrix_evals, uint64_t* nonces,
                            or_f128* claim, or_f128* z_vec_out);

/* ---- union instance host glue (crates/flock-core/src/{schedule,union}.rs) ---- */
#define OR_MAX_TYPES 16
typedef struct {
    uint32_t k_log, is_element;
    uint64_t useful_bits;
} or_table_type;
typedef struct {
    size_t type_index;
    uint32_t m_slot, prefix_bits;
    uint64_t offset, prefix;
} or_slot;
typedef struct {
    size_t n, num_boolean;
    uint32_t nu, m_total, m_bool, m_elem;
    uint64_t element_base;
    or_slot slots[OR_MAX_TYPES];
} or_registry;
typedef struct {
    uint64_t dense_words, committed_words, packed_len, commit_lanes;
    uint32_t dense_m, compaction_is_identity;
} or_union_size_info;
int or_registry_layout(const or_table_type* types, size_t n, unsigned nu, or_registry* reg);
int or_union_sizes(const or_table_type* types, const or_registry* reg, const uint64_t* counts,
                   unsigned dense_floor_m, unsigned log_batch_size, or_union_size_info* out);
void or_compact_witness(const or_table_type* types, const or_registry* reg,
                        const uint64_t* counts, const or_f128* z_padded, size_t committed,
                        or_f128* q);
void or_x_ab_from_mlv(const or_f128* mlv, size_t n_mlv, unsigned nu, or_f128* x_inner_rest,
                      or_f128* x_outer);
void or_bind_statement(or_fs* fs, const uint8_t digest[32], const uint64_t* counts, size_t n,
                       const uint8_t* cap, size_t cap_len);

/* ---- SHA-256 compression R1CS (crates/flock-prover/src/r1cs_hashes/sha2.rs), sha2_r1cs.c ---- */
/* layout constants: {SCHED_CARRY_BASE, E_NEW_BASE, A_NEW_BASE, OUT_CARRY_BASE, Z_CONST_POS,
 * USEFUL_BITS, ROUND_BASE[1], 0} (sha2.rs:95-194) */
void or_sha2_layout(uint32_t out[8]);
/* build_block_ab_packed_into (sha2.rs:944-1128): 512 u64 each */
void or_sha2_instance(const uint32_t h_in[8], const uint32_t m[16], uint64_t* z, uint64_t* a,
                      uint64_t* b);
/* drive_witness_batch_major_partial (common.rs:805-960); inputs 24 u32 per declared row */
int or_sha2_witness_batch_major(const uint32_t* inputs, size_t n_decl, int n_log, or_f128* z,
                                or_f128* a, or_f128* b, uint8_t* stripe);
/* build_matrices (sha2.rs:548-701) as CSR; a_ptr/b_ptr have 32769 entries */
int or_sha2_matrices(uint64_t* a_ptr, uint32_t* a_cols, uint64_t* b_ptr, uint32_t* b_cols);
void or_csr_to_csc(size_t k, const uint64_t* ptr, const uint32_t* cols, uint32_t* col_ptr,
                   uint32_t* rows);
void or_csr_apply_bits(size_t k, const uint64_t* ptr, const uint32_t* cols, const uint64_t* z,
                       uint64_t* y);

/* ---- Keccak-f[1600] R1CS (crates/flock-prover/src/r1cs_hashes/keccak.rs), keccak_r1cs.c ---- */
void or_keccak_f(uint64_t s[25]); /* lanes s[x + 5y], keccak.rs:184-330 */
/* build_chain_witness_ab_packed_into (keccak.rs:565-643): 1024 u64 each */
void or_keccak_instance(const uint64_t initial[25], uint64_t* z, uint64_t* a, uint64_t* b);
/* generate_witness_batch_major (keccak.rs:1228-1285): padding rows run keccak_f(0) */
int or_keccak_witness_batch_major(const uint64_t* inputs, size_t n_decl, int n_log, or_f128* z,
                                  or_f128* a, or_f128* b, uint8_t* stripe);
/* KeccakLincheckCircuit column marginals (keccak.rs:692-835) as fold_split (lincheck.rs:258):
 * sa = A_0^T eq, sb = B_0^T eq over 2^16 columns; fold_alpha_batched = alpha * sa + sb */
void or_keccak_lc_fold_split(const or_f128* eq, or_f128* sa, or_f128* sb);

/* ---- BLAKE3 compression R1CS (crates/flock-prover/src/r1cs_hashes/blake3.rs), blake3_r1cs.c ---- */
/* {Z_CONST_POS, USEFUL_BITS, G_BASE[4], G_BASE[1]} (blake3.rs:233-247) */
void or_blake3_layout(uint32_t out[4]);
/* blake3_compress (crates/flock-core/src/hash.rs:101-142): the 16-word output state */
void or_blake3_compress(const uint32_t cv[8], const uint32_t m[16], uint64_t counter,
                        uint32_t blen, uint32_t flags, uint32_t out[16]);
/* build_block_witness_ab_packed_into (blake3.rs:1299-1452): 256 u64 each */
void or_blake3_instance(const uint32_t cv[8], const uint32_t m[16], uint64_t counter,
                        uint32_t blen, uint32_t flags, uint64_t* z, uint64_t* a, uint64_t* b);
/* generate_witness_batch_major_partial (blake3.rs:2072-2088); 28 u32 per declared row */
int or_blake3_witness_batch_major(const uint32_t* inputs, size_t n_decl, int n_log, or_f128* z,
                                  or_f128* a, or_f128* b, uint8_t* stripe);
/* build_matrices (blake3.rs:659-783) as CSR; a_ptr/b_ptr have 16385 entries */
int or_blake3_matrices(uint64_t* a_ptr, uint32_t* a_cols, uint64_t* b_ptr, uint32_t* b_cols);

/* ---- merged PCS opening (crates/flock-core/src/pcs.rs:1433-1822), opening.c ---- */
/* initial_k of the embedded `fast` Ligerito config for m (configs/ligerito/m*_fast.toml), or -1 */
int or_lig_initial_k(unsigned m);
/* PcsParams::l0_cap_depth (pcs/commit.rs:266-277): the L0 schedule's cap depth, or -1 */
int or_cap_depth_for(unsigned m);
/* f̂_t(z_row, z_col, x) of the jagged layout with the given column heights (jagged.rs:243-262),
 * through the branching program; n = row vars, m = dense vars */
int or_jag_fhat(const uint64_t* heights, size_t nh, int n, int m, const or_f128* z_row,
                const or_f128* z_col, const or_f128* x, or_f128* out);
/* open_batch_merged for n_claims ring-switched claims with precomputed s_hat_v (128 each).
 * q: the committed dense stack (2^(m-7) words); cw/tree: the L0 commitment (full tree);
 * x_outers: n_claims points of x_len = 1 + n_log + log2(n_heights) words (quirky x_outer_full);
 * heights: jagged column heights. On success *out_bytes (free with or_free_bytes) holds the
 * bincode of MergedOpenProof (pcs.rs:1362-1377). */
int or_open_merged(or_fs* fs, const or_f128* q, const or_pcs_params* p, const or_f128* cw,
                   const uint8_t* tree, size_t n_claims, const or_f128* x_outers, size_t x_len,
                   const or_f128* s_hat_v, const uint64_t* heights, size_t n_heights,
                   unsigned n_log, uint8_t** out_bytes, size_t* out_len, or_f128* q_eval_out);
void or_free_bytes(uint8_t* p);

#ifdef __cplusplus
}
#endif
#endif
