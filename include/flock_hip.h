/*
 * flock_hip.h — C ABI of libflock_hip.so, the MI355X (gfx950) implementation of the
 * data-parallel core of Flock's `prove_fast`: the F_{2^128} multiply, the LCH additive
 * NTT + SHA-256 Merkle commit, and the product-sumcheck / eq-weighted zerocheck rounds.
 *
 * It replaces the reference's CUDA bridge (cuda-ghash/prove_ffi.cu:363-1010, bound from
 * Rust by crates/flock-cuda-ffi) with phase-level entry points that map one-to-one onto
 * the Rust seams the prover calls (citations are reference file:line).
 *
 * Conventions
 *  - Every function returns 0 on success or a FLOCK_HIP_E_* code; it never aborts on bad
 *    input. flock_hip_last_error() describes the last failure on the calling thread.
 *  - Pointers named d_* are device pointers (HBM, e.g. from flock_hip_malloc); h_* are
 *    host pointers. Field elements are flock_f128 {lo, hi}: lo = coefficients of
 *    x^0..x^63, reduction polynomial x^128 + x^7 + x^2 + x + 1
 *    (crates/flock-core/src/field/gf2_128.rs:10-27). Same bytes as the Rust F128.
 *  - A context owns one HIP stream and its scratch memory. Distinct contexts may be used
 *    concurrently from different threads; one context must be externally synchronized.
 *    Calls are stream-ordered; functions that return values to the host synchronize.
 *  - Fixed-size outputs are caller-allocated.
 */
#ifndef FLOCK_HIP_H
#define FLOCK_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLOCK_HIP_OK 0
#define FLOCK_HIP_E_INVALID 1       /* bad argument or shape */
#define FLOCK_HIP_E_NOMEM 2         /* device or host allocation failed */
#define FLOCK_HIP_E_NODEVICE 3      /* no usable gfx950 device */
#define FLOCK_HIP_E_HIP 100         /* HIP runtime error (prove_ffi.cu:44-46 uses 100 for CUDA) */
#define FLOCK_HIP_E_M_TOO_SMALL 101 /* m below the commit minimum (prove_ffi.cu:429) */
#define FLOCK_HIP_E_INTERNAL 102    /* a device result failed its host-side check */
#define FLOCK_HIP_E_REJECT 103      /* a verifier rejected the proof */

typedef struct {
    uint64_t lo, hi;
} flock_f128;

typedef struct flock_hip_ctx flock_hip_ctx;
typedef struct flock_fs flock_fs;

/* ---- device / context -------------------------------------------------------- */

/* Number of visible devices. Replaces flock_cuda_device_count (prove_ffi.cu:410-415). */
int flock_hip_device_count(void);
/* Message for the last error returned on this thread ("" if none). */
const char* flock_hip_last_error(void);
/* One HIP stream + scratch per context; replaces the reference's process-global mutex
 * and arena (prove_ffi.cu:52-131, 421-422). */
int flock_hip_ctx_create(int device, flock_hip_ctx** out);
void flock_hip_ctx_destroy(flock_hip_ctx* ctx);
/* The context's hipStream_t (as void*), for callers that order their own work on it. */
void* flock_hip_ctx_stream(flock_hip_ctx* ctx);
int flock_hip_sync(flock_hip_ctx* ctx);
/* How the round loops of flock_hip_zerocheck_prove, flock_hip_lincheck_prove{,_keccak} and
 * flock_hip_prove_sumcheck meet the transcript between rounds. The proofs and the final
 * transcript are the same in both modes.
 *  FLOCK_ROUNDS_HOST:   (default) per round, one stream synchronisation, the transcript step
 *                       on the host, and one launch.
 *  FLOCK_ROUNDS_DEVICE: a device-resident copy of the transcript observes each message, grinds
 *                       and samples the next challenge inside the round kernel, so the whole
 *                       loop is enqueued at once and runs without the host; the host then
 *                       replays the logged rounds through its transcript and checks every nonce
 *                       and challenge (FLOCK_HIP_E_INTERNAL on a mismatch). The analogue of the
 *                       reference's device challenger (cuda-ghash/zc_challenger_device.cuh). */
#define FLOCK_ROUNDS_HOST 0
#define FLOCK_ROUNDS_DEVICE 2
int flock_hip_ctx_set_round_mode(flock_hip_ctx* ctx, int mode);

/* Process-wide launch tunables: the tile shapes and thresholds `flock_amd/tune.py` (the
 * autotuner) picks. Every value in range gives the same bytes; set them before proving (a
 * launch reads them as it goes). Names, defaults and ranges:
 *   ntt_cols_per_thread  2   [1, 4]   columns per thread of the interior NTT pass
 *   ntt_deep_cols        256 [32, 256] passes with at most this many columns per block run deep
 *   zc_split_min_vars    16  [10, 24] zerocheck rounds whose eq has this many variables split it
 *   zc_la_max_log        17  [10, 22] zerocheck tail: two rounds per launch from 2^this words
 *   sc_la_max_log        18  [10, 24] the opening's product sumchecks: the same, from 2^this
 *   open_split_fold_log  16  [10, 22] F256 ladder: fold and message as two passes from 2^this
 * No counterpart in the reference (its CUDA launch shapes are compile-time constants).
 * Unknown names and values out of range are FLOCK_HIP_E_INVALID. */
int flock_hip_set_tunable(const char* name, int64_t value);
int flock_hip_get_tunable(const char* name, int64_t* value);
int flock_hip_malloc(flock_hip_ctx* ctx, void** d_ptr, size_t bytes);
int flock_hip_dfree(flock_hip_ctx* ctx, void* d_ptr);
int flock_hip_memcpy_h2d(flock_hip_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int flock_hip_memcpy_d2h(flock_hip_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);
/* Free a host buffer the library allocated. Replaces flock_cuda_free (prove_ffi.cu:417). */
void flock_hip_free(void* p);

/* ---- field ------------------------------------------------------------------- */

/* d_out[i] = d_a[i] * d_b[i]. Test surface for `impl Mul for F128` (gf2_128.rs:174-201). */
int flock_hip_f128_mul_batch(flock_hip_ctx* ctx, const flock_f128* d_a, const flock_f128* d_b,
                             flock_f128* d_out, size_t n);
/* d_out[i] = d_a[i] * t through the fixed-multiplier (table) path used by the NTT/folds. */
int flock_hip_f128_mul_const_batch(flock_hip_ctx* ctx, const flock_f128* d_a, flock_f128 t,
                                   flock_f128* d_out, size_t n);

/* d_out[i] = d_a[i] * t^reps by `reps` fixed-multiplier products per element, on the LDS Shoup
 * table (FLOCK_MUL_SHOUP, the NTT's and folds' product) or on the int8 matrix cores
 * (FLOCK_MUL_MFMA_I8: the 128 x 128 GF(2) map of t as v_mfma_i32_32x32x32_i8 over 0/1 bytes,
 * parity from bit 0). No counterpart in the reference: the measurement surface on which the
 * two product forms are timed and profiled against each other. */
enum { FLOCK_MUL_SHOUP = 0, FLOCK_MUL_MFMA_I8 = 1 };
int flock_hip_f128_mul_const_pow_batch(flock_hip_ctx* ctx, const flock_f128* d_a, flock_f128 t,
                                       uint32_t reps, int method, flock_f128* d_out, size_t n);

/* ---- additive NTT ------------------------------------------------------------ */

/* In-place forward interleaved NTT of AdditiveNttF128::standard(log_d) over layers
 * [start_layer, log_d), transforming only lanes < live (the rest must be zero):
 * forward_transform_interleaved_live_from_layer (ntt/additive_ntt_f128.rs:220-266).
 * d_data holds 2^log_d positions x num_ntts lanes, data[pos * num_ntts + lane]. */
int flock_hip_ntt_forward_interleaved(flock_hip_ctx* ctx, flock_f128* d_data, size_t num_ntts,
                                      size_t live, uint32_t log_d, uint32_t start_layer);

/* One fused pass of the same transform: layers [first_layer, first_layer + layers) with
 * layers in {1, 2, 3}, in place. flock_hip_ntt_forward_interleaved is a sequence of these
 * passes (3-layer passes, then 2-layer passes for a remainder); exposed so callers can
 * schedule or time passes individually. */
int flock_hip_ntt_pass(flock_hip_ctx* ctx, flock_f128* d_data, size_t num_ntts, size_t live,
                       uint32_t log_d, uint32_t first_layer, uint32_t layers);

/* ---- PCS commit -------------------------------------------------------------- */

/* HashKind (hash.rs:23-27): the Merkle tree's node hash. */
#define FLOCK_MERKLE_SHA256 0u
#define FLOCK_MERKLE_BLAKE3 1u

/* PcsParams (pcs/commit.rs:31-65). num_lanes = 0 means None (all 2^log_batch_size lanes).
 * merkle_hash: FLOCK_MERKLE_SHA256 (the default) or FLOCK_MERKLE_BLAKE3. */
typedef struct {
    uint32_t m, log_inv_rate, log_batch_size;
    uint32_t num_lanes;
    uint32_t merkle_hash;
} flock_pcs_params;

/* Derived sizes (PcsParams::{msg_len_f128, codeword_len_f128, n_leaves, k_code, num_ntts},
 * pcs/commit.rs:187-228). Any output pointer may be NULL. */
int flock_hip_pcs_sizes(const flock_pcs_params* p, size_t* msg_len, size_t* cw_len,
                        size_t* n_leaves, uint32_t* k_code, size_t* num_ntts);

/* pcs::commit_lane_major (pcs/commit.rs:474-510): d_q is the full 2^(m-7)-word lane-major
 * stack (lane l = q[l*D .. (l+1)*D), lanes >= num_ntts zero). Writes the codeword
 * (cw_len words, codeword[pos * num_ntts + lane]) and the flat Merkle tree
 * (2*n_leaves - 1 nodes x 32 B, merkle.rs:15-22) to caller-allocated device buffers, and,
 * if h_cap != NULL, the 2^cap_depth cap nodes (cap_layer, merkle.rs:649-656) to h_cap. */
int flock_hip_commit_lane_major(flock_hip_ctx* ctx, const flock_f128* d_q, size_t q_len,
                                const flock_pcs_params* p, flock_f128* d_codeword,
                                uint8_t* d_tree, uint32_t cap_depth, uint8_t* h_cap);

/* pcs::commit / commit_into (pcs/commit.rs:345-397) on the interleaved message
 * d_msg[p * num_ntts + lane] of msg_len = num_ntts * 2^log_dim words. */
int flock_hip_commit(flock_hip_ctx* ctx, const flock_f128* d_msg, size_t msg_len,
                     const flock_pcs_params* p, flock_f128* d_codeword, uint8_t* d_tree,
                     uint32_t cap_depth, uint8_t* h_cap);

/* The RS-encoding half of commit / commit_lane_major, without the Merkle tree:
 * replicate_message_fill + forward_transform_interleaved_live_from_layer from layer
 * log_inv_rate (pcs/commit.rs:374-397, 583-588; lane-major fill :516-536). d_codeword
 * receives cw_len words. */
int flock_hip_encode(flock_hip_ctx* ctx, const flock_f128* d_msg, size_t msg_len,
                     const flock_pcs_params* p, flock_f128* d_codeword);
int flock_hip_encode_lane_major(flock_hip_ctx* ctx, const flock_f128* d_q, size_t q_len,
                                const flock_pcs_params* p, flock_f128* d_codeword);

/* merkle::merkle_tree with HashKind::Sha256 (merkle.rs:543-589): n_leaves (a power of two)
 * equal leaves of data_len / n_leaves bytes (a multiple of 16). */
int flock_hip_merkle_tree(flock_hip_ctx* ctx, const uint8_t* d_data, size_t data_len,
                          size_t n_leaves, uint8_t* d_tree);
/* merkle::merkle_tree(data, num_leaves, kind) (merkle.rs:543-589) for either HashKind. Under
 * BLAKE3 a leaf is its bytes' non-root chaining value and a node the PARENT compression of
 * its children (merkle.rs:195-232); leaves may span up to 2^16 chunks of 1024 bytes. */
int flock_hip_merkle_tree_kind(flock_hip_ctx* ctx, const uint8_t* d_data, size_t data_len,
                               size_t n_leaves, uint32_t kind, uint8_t* d_tree);

/* ---- sumcheck rounds ----------------------------------------------------------- */

/* round_msg_lsb (pcs/ligerito.rs:3204-3240): h_msg = {u_0, u_2}. n >= 2, power of two. */
int flock_hip_sumcheck_round_msg(flock_hip_ctx* ctx, const flock_f128* d_f,
                                 const flock_f128* d_b, size_t n, flock_f128 h_msg[2]);
/* fold_and_msg_lsb (pcs/ligerito.rs:3372-3414): d_nf/d_nb (n/2 words) = v0 + r(v0 + v1);
 * h_msg = next round's {u_0, u_2} (zero when n == 2). */
int flock_hip_sumcheck_fold_msg(flock_hip_ctx* ctx, const flock_f128* d_f,
                                const flock_f128* d_b, size_t n, flock_f128 r, flock_f128* d_nf,
                                flock_f128* d_nb, flock_f128 h_msg[2]);
/* The same fold and message without any allocation: no context, no hipMalloc / hipFree, no
 * host synchronisation. No counterpart in the reference; it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops). The grid reduction's scratch
 * is a fixed 16,400-byte device array per GPU, held until the process ends, so the size
 * limit is only the kernel's own (n a power of two in [2, 2^40]). Enqueued on `stream` (a
 * hipStream_t of the current device, NULL for the null stream); d_msg (2 device words) gets
 * {u_0, u_2} when the stream reaches it. Calls that overlap on one GPU race on the scratch. */
int flock_hip_sumcheck_fold_msg_static(void* stream, const flock_f128* d_f,
                                       const flock_f128* d_b, size_t n, flock_f128 r,
                                       flock_f128* d_nf, flock_f128* d_nb, flock_f128* d_msg);
/* build_eq (zerocheck/univariate_skip.rs:31-47): d_out[x] = prod_i (bit_i(x) ? r_i : 1+r_i)
 * for x < 2^nv. */
int flock_hip_build_eq(flock_hip_ctx* ctx, const flock_f128* h_r, size_t nv, flock_f128* d_out);
/* round_pair_naive (zerocheck/multilinear.rs:331-355) with d_eq = build_eq(r[1..]) (n/2
 * words) and r0 = r[0]: h_msg = {r0 * G(1), G(inf)}. If d_eq_next != NULL and n >= 4 it
 * receives the next round's table d_eq_next[z] = d_eq[2z] + d_eq[2z+1] (n/4 words). */
int flock_hip_zc_round_pair(flock_hip_ctx* ctx, const flock_f128* d_a, const flock_f128* d_b,
                            size_t n, const flock_f128* d_eq, flock_f128 r0,
                            flock_f128* d_eq_next, flock_f128 h_msg[2]);
/* fold_and_compute_round_pair_into (zerocheck/multilinear.rs:1197-1420): fold (a, b) of size
 * n >= 4 at rho into d_ao/d_bo (n/2), then round_pair on the folded tables with d_eq
 * (n/4 words) and r0. d_eq_next (n/8 words, may be NULL) as above. */
int flock_hip_zc_fold_round_pair(flock_hip_ctx* ctx, const flock_f128* d_a,
                                 const flock_f128* d_b, size_t n, flock_f128 rho,
                                 flock_f128* d_ao, flock_f128* d_bo, const flock_f128* d_eq,
                                 flock_f128 r0, flock_f128* d_eq_next, flock_f128 h_msg[2]);
/* fold_in_place_pair (zerocheck/multilinear.rs:1148-1163), out of place: n -> n/2. */
int flock_hip_fold_pair(flock_hip_ctx* ctx, const flock_f128* d_a, const flock_f128* d_b,
                        size_t n, flock_f128 c, flock_f128* d_ao, flock_f128* d_bo);

/* ---- zerocheck front --------------------------------------------------------------- */

/* Univariate-skip round 1 (k_skip = 6) with the naive-convention output of
 * round1_naive (zerocheck/univariate_skip.rs:68-124), which is what
 * prove_packed_padded_inner sends after its C_s rescale (zerocheck.rs:700-704).
 * d_a/d_b/d_c: 2^m LSB-first packed bits (2^(m-3) bytes); d_eq = build_eq(r[6..m])
 * (2^(m-6) words). h_out[0..64) = P^AB on Lambda, h_out[64..128) = P^C on Lambda. */
int flock_hip_zc_round1(flock_hip_ctx* ctx, const uint8_t* d_a, const uint8_t* d_b,
                        const uint8_t* d_c, uint32_t m, const flock_f128* d_eq,
                        flock_f128 h_out[128]);
/* Round-2 fold at the skip challenge z (fold_at_z_naive, zerocheck/multilinear.rs:290-310;
 * UniSkipFoldTable, :414-457) of two packed tables into 2^(m-6) words each. */
int flock_hip_zc_fold_at_z(flock_hip_ctx* ctx, const uint8_t* d_a, const uint8_t* d_b,
                           uint32_t m, flock_f128 z, flock_f128* d_az, flock_f128* d_bz);
/* Zerocheck prover for a.b + c = 0 on packed bits, dense padding:
 * prove_packed_padded_inner (zerocheck.rs:618-1011) with the transcript ops in its order:
 * label, [PoW] r_skip, r_outer, round 1, [PoW] z, round 2 fused with the first
 * multilinear message, the tail, final folds. grinding != 0 selects
 * ZerocheckGrinding::per_challenge_128 (zerocheck.rs:67-158), 0 disabled. Outputs
 * (caller-allocated): h_round1[128] = (round1_ab, round1_c); h_ml[2*(m-6)] = multilinear
 * round messages; h_finals[3] = (final_a_eval, final_b_eval, final_c_eval);
 * h_nonces[2 + m-6] when grinding (may be NULL otherwise); h_claim[1 + 2*(m-6)] =
 * (z, mlv_challenges, r_rest = r[6..m)) — ZerocheckClaim (zerocheck.rs:405-425). m >= 13.
 * h_s_hat_v_c[128], when not NULL, receives s_hat_v_c
 * (prove_packed_padded_capture_s_hat_v_c_with_grinding, zerocheck.rs:598-615;
 * round1_extract_c_packed_with_s_hat_v, zerocheck/univariate_skip.rs:354-489):
 * s_hat_v_c[lane | b7 << 6] = sum_y eq(r[7..m), y) * bit(c, y*128 + b7*64 + lane). */
int flock_hip_zerocheck_prove(flock_hip_ctx* ctx, flock_fs* fs, const uint8_t* d_a,
                              const uint8_t* d_b, const uint8_t* d_c, uint32_t m, int grinding,
                              flock_f128* h_round1, flock_f128* h_ml, flock_f128 h_finals[3],
                              uint64_t* h_nonces, flock_f128* h_claim, flock_f128* h_s_hat_v_c);
/* zerocheck::verify_with_grinding (zerocheck.rs:1032-1225): replays a zerocheck proof (the
 * outputs of flock_hip_zerocheck_prove, same layouts) on the verifier's transcript `fs`. It
 * re-derives r and z, checks round 1's c evaluation on Lambda, rebuilds the AB claim from the
 * zero-on-S assumption, runs the multilinear chain and checks the final a*b claim. Every PoW
 * nonce is checked when grinding != 0. Returns 0 and, when h_claim is not NULL, the
 * ZerocheckClaim (the prover's h_claim layout), or FLOCK_HIP_E_REJECT naming the failed check.
 * Host only. */
int flock_zerocheck_verify(flock_fs* fs, uint32_t m, int grinding, const flock_f128* h_round1,
                           const flock_f128* h_ml, const flock_f128 h_finals[3],
                           const uint64_t* h_nonces, flock_f128* h_claim);
/* fold_1b_rows (crates/flock-core/src/pcs/ring_switch.rs:182-230, :410-470): out[j] =
 * sum_y d_t[y] * bit j of the 128-bit row d_rows[y], j < 128 (d_rows and d_t n_rows words). */
int flock_hip_fold_1b_rows(flock_hip_ctx* ctx, const flock_f128* d_rows, size_t n_rows,
                           const flock_f128* d_t, flock_f128 h_out[128]);

/* ---- lincheck (crates/flock-core/src/lincheck.rs, lincheck/union.rs) ---------------- */

/* A boolean R1CS base pair (A_0, B_0), 2^k_log x 2^k_log, in CSC form on the device: the rows
 * holding a 1 in column c are rows[col_ptr[c] .. col_ptr[c+1]) (CscCircuit and csc_from_rows,
 * lincheck.rs:353-431). const_pin = the constant-wire pin column, -1 for none
 * (LincheckCircuit::const_pin_col). */
typedef struct {
    uint32_t k_log;
    uint32_t reserved;
    const uint32_t* a_col_ptr; /* 2^k_log + 1 entries */
    const uint32_t* a_rows;    /* nnz_a entries, each < 2^k_log */
    const uint32_t* b_col_ptr;
    const uint32_t* b_rows;
    uint64_t nnz_a, nnz_b;
    int64_t const_pin;
} flock_csc;

/* CscCircuit::fold_split / fold_alpha_batched (lincheck.rs:432-497) on a device eq table of
 * 2^k_log words: d_sa = A^T eq, d_sb = B^T eq, d_comb = alpha * d_sa + d_sb. Any output may be
 * NULL (not all three). Out-of-range CSC entries give FLOCK_HIP_E_INVALID. */
int flock_hip_lc_csc_fold(flock_hip_ctx* ctx, const flock_csc* circuit, const flock_f128* d_eq,
                          flock_f128 alpha, flock_f128* d_sa, flock_f128* d_sb,
                          flock_f128* d_comb);
/* Partial fold of the lincheck stripe (byte d_z[s*2^k_log + i] = outer rows 8s..8s+7 of column
 * i, 2^(nu+k_log-3) bytes, 4-byte aligned) at the outer eq table d_eq_outer (2^nu words):
 * partial_fold_packed_z (lincheck.rs:804-835) with the support-proportional restriction of
 * partial_fold_packed_z_rows_best (:865-1100): columns < useful_bits, stripes <
 * ceil(n_rows / 8). d_out gets 2^k_log words. */
int flock_hip_lc_partial_fold(flock_hip_ctx* ctx, const uint8_t* d_z, uint32_t k_log, uint32_t nu,
                              size_t useful_bits, size_t n_rows, const flock_f128* d_eq_outer,
                              flock_f128* d_out);
/* Boolean lincheck prover for a registry with one boolean type filling the boolean region:
 * prove_union_capture_z_vec_with_grinding (lincheck/union.rs:203-430) with column_sumcheck_prove
 * (lincheck.rs:1907-2028). x_ab = (z_skip, h_x_inner_rest[k_log-6], h_x_outer[nu]); d_z the
 * stripe as in flock_hip_lc_partial_fold. grinding != 0 selects
 * LincheckGrinding::per_challenge_128 (lincheck.rs:590-603). Outputs (caller-allocated, host):
 * h_rounds[2 (k_log-6)] = (q(1), q(inf)) per round; h_z_partial[64]; h_matrix_evals[2] =
 * (<W, A_0>, <W, B_0>); h_nonces[k_log-6 + 3] when grinding; h_claim[2 + k_log-6] =
 * (r_inner_skip, w, r_inner_rest) (LincheckClaim, lincheck.rs:646-657); d_z_vec (device,
 * 2^k_log words, may be NULL) = the captured pre-sumcheck fold. */
int flock_hip_lincheck_prove(flock_hip_ctx* ctx, flock_fs* fs, const flock_csc* circuit,
                             const uint8_t* d_z, uint32_t nu, size_t useful_bits, size_t n_rows,
                             flock_f128 z_skip, const flock_f128* h_x_inner_rest,
                             const flock_f128* h_x_outer, int grinding, flock_f128* h_rounds,
                             flock_f128 h_z_partial[64], flock_f128 h_matrix_evals[2],
                             uint64_t* h_nonces, flock_f128* h_claim, flock_f128* d_z_vec);
/* Keccak-f[1600] circuit marginals without forming its matrices: KeccakLincheckCircuit::
 * fold_alpha_batched (r1cs_hashes/keccak.rs:692-835) as fold_split (lincheck.rs:258-266).
 * Over the 2^16 columns, d_sa = A_0^T eq and d_sb = B_0^T eq for the row weights d_eq
 * (2^16 words); d_comb = alpha d_sa + d_sb when non-NULL. */
int flock_hip_lc_keccak_fold(flock_hip_ctx* ctx, const flock_f128* d_eq, flock_f128 alpha,
                             flock_f128* d_sa, flock_f128* d_sb, flock_f128* d_comb);
/* flock_hip_lincheck_prove with KeccakLincheckCircuit (keccak.rs:678-835) in place of a CSC
 * circuit: k_log 16, useful_bits 42,560, constant pin at column 4096 (keccak.rs:436-451,
 * 688-690); the lincheck call site for Keccak batches (prover.rs:1154-1160). */
int flock_hip_lincheck_prove_keccak(flock_hip_ctx* ctx, flock_fs* fs, const uint8_t* d_z,
                                    uint32_t nu, size_t n_rows, flock_f128 z_skip,
                                    const flock_f128* h_x_inner_rest, const flock_f128* h_x_outer,
                                    int grinding, flock_f128* h_rounds,
                                    flock_f128 h_z_partial[64], flock_f128 h_matrix_evals[2],
                                    uint64_t* h_nonces, flock_f128* h_claim, flock_f128* d_z_vec);
/* One boolean slot of the union lincheck. circuit: the slot type's device CSC circuit (its
 * k_log and const_pin are used), or NULL for the Keccak-f[1600] walker (k_log 16, pin 4096,
 * useful_bits 42,560). prefix: Slot::prefix (schedule.rs:247-268); the slot's columns are
 * [prefix << k_log, (prefix + 1) << k_log) of the 2^col_vars union columns. d_z: the slot's
 * lincheck stripe (as in flock_hip_lc_partial_fold) with n_rows rows; useful_bits: the
 * type's USEFUL_BITS (ignored for the walker). */
typedef struct {
    const flock_csc* circuit;
    const uint8_t* d_z;
    uint64_t prefix;
    uint64_t useful_bits;
    uint64_t n_rows;
} flock_lc_slot;
/* The union lincheck over several boolean slots: prove_union_capture_z_vec_with_grinding
 * (lincheck/union.rs:203-430) with column_sumcheck_prove (lincheck.rs:1907-2028), the call
 * site of prove_union_with_binding_zc for a registry of several boolean types (prover.rs:
 * 1140-1160). col_vars = M_bool - nu (at least every slot's k_log); per slot the alpha-batched
 * comb through its quirky table over x_inner_rest[..k_t-6], scaled by eq_prefix_weight(
 * x_inner_rest[k_t-6..], prefix_t) (union.rs:90-101); the pins' betas in slot order; one
 * eq(x_outer) table for every slot's partial fold; the column sumcheck over 2^col_vars; then
 * per slot the unscaled (<W_t, A_0>, <W_t, B_0>) with W_t = z_partial (x)
 * eq(r_inner_rest[..k_t-6]). Outputs as flock_hip_lincheck_prove with L = col_vars - 6:
 * h_rounds[2L], h_z_partial[64], h_matrix_evals[2 n_slots] (slot order), h_nonces[L + 2 +
 * pins] when grinding, h_claim[2 + L], d_z_vec (2^col_vars words, may be NULL). One slot
 * with k_log = col_vars is exactly flock_hip_lincheck_prove{,_keccak}. */
int flock_hip_lincheck_prove_union(flock_hip_ctx* ctx, flock_fs* fs, const flock_lc_slot* slots,
                                   size_t n_slots, uint32_t col_vars, uint32_t nu,
                                   flock_f128 z_skip, const flock_f128* h_x_inner_rest,
                                   const flock_f128* h_x_outer, int grinding, flock_f128* h_rounds,
                                   flock_f128 h_z_partial[64], flock_f128* h_matrix_evals,
                                   uint64_t* h_nonces, flock_f128* h_claim, flock_f128* d_z_vec);
/* verify_union_with_grinding (lincheck/union.rs:662-960) for a registry with one boolean type
 * filling the boolean region, the prover of flock_hip_lincheck_prove: replays the proof (same
 * layouts) on the verifier's transcript `fs` — alpha, beta when const_pin >= 0 (the target
 * gains beta * sum_{row < n_rows} eq(x_outer, row)), the rounds' chain from
 * alpha v_a + v_b, z_partial, the fresh skip challenge — checking every PoW nonce when
 * grinding != 0 (h_nonces[k_log-6 + 2 + (const_pin >= 0)]). Then the matrix assertion:
 *  h_circuit != NULL: MatrixAssertion::check (:579-628), the O(nnz) comb collapse at the bound
 *    point against z_partial; h_circuit is the CSC on the HOST (host pointers) and must carry
 *    k_log and const_pin; h_matrix_evals is not read (as in the reference);
 *  h_circuit == NULL: the deferred form with MatrixAssertion::check_reported (:538-576),
 *    alpha <W,A_0> + <W,B_0> + beta W_col[pin] against the claim from h_matrix_evals.
 * (v_a, v_b) is the lincheck's input claim at x_ab = (z_skip, x_inner_rest, x_outer). Returns
 * 0 and, when h_claim is not NULL, the LincheckClaim in the prover's layout, or
 * FLOCK_HIP_E_REJECT naming the failed check. Host only. */
int flock_lincheck_verify(flock_fs* fs, const flock_csc* h_circuit, uint32_t k_log,
                          int64_t const_pin, uint32_t nu, size_t n_rows, flock_f128 z_skip,
                          const flock_f128* h_x_inner_rest, const flock_f128* h_x_outer,
                          flock_f128 v_a, flock_f128 v_b, int grinding,
                          const flock_f128* h_rounds, const flock_f128 h_z_partial[64],
                          const flock_f128 h_matrix_evals[2], const uint64_t* h_nonces,
                          flock_f128* h_claim);
/* One boolean slot of the union lincheck verifier (slot order): h_circuit the type's CSC on
 * the HOST; else walker != 0 for the Keccak-f[1600] host walk (k_log 16, pin 4096); else the
 * slot is checked from the reported matrix evals with the given k_log and const_pin (-1: none).
 * prefix: Slot::prefix; n_rows: the slot's statement count. */
typedef struct {
    const flock_csc* h_circuit;
    uint32_t k_log, walker;
    int64_t const_pin;
    uint64_t prefix, n_rows;
} flock_lc_vslot;
/* verify_union_with_grinding (lincheck/union.rs:662-960) over several boolean slots, the
 * verifier of flock_hip_lincheck_prove_union: alpha, one beta per pinned slot in slot order
 * (the target gains beta_t * sum_{row < n_t} eq(x_outer, row)), the rounds' chain, z_partial,
 * the fresh skip challenge; then MatrixAssertion::check (:579-628) — every slot's comb at its
 * row-side prefix weight, plus beta_t at its pin, collapsed at the column-side prefix weight
 * against z_partial — when every slot has a circuit or the walker, else check_reported
 * (:538-576) from h_matrix_evals (2 per slot). col_vars = M_bool - nu; layouts as the
 * prover's. Host only. */
int flock_lincheck_verify_union(flock_fs* fs, const flock_lc_vslot* slots, size_t n_slots,
                                uint32_t col_vars, uint32_t nu, flock_f128 z_skip,
                                const flock_f128* h_x_inner_rest, const flock_f128* h_x_outer,
                                flock_f128 v_a, flock_f128 v_b, int grinding,
                                const flock_f128* h_rounds, const flock_f128 h_z_partial[64],
                                const flock_f128* h_matrix_evals, const uint64_t* h_nonces,
                                flock_f128* h_claim);
/* flock_lincheck_verify for a Keccak-f[1600] slot (KeccakLincheckCircuit, keccak.rs:678-835;
 * k_log 16, constant pin at column 4096), the proofs of flock_hip_lincheck_prove_keccak, with
 * MatrixAssertion::check: the circuit's column marginals come from a host walk of the
 * permutation's rows (no matrices are formed). Host only. */
int flock_lincheck_verify_keccak(flock_fs* fs, uint32_t nu, size_t n_rows, flock_f128 z_skip,
                                 const flock_f128* h_x_inner_rest, const flock_f128* h_x_outer,
                                 flock_f128 v_a, flock_f128 v_b, int grinding,
                                 const flock_f128* h_rounds, const flock_f128 h_z_partial[64],
                                 const uint64_t* h_nonces, flock_f128* h_claim);

/* ---- union instance host glue (crates/flock-core/src/{schedule,union}.rs) ------------ */

#define FLOCK_MAX_TYPES 16
/* TableType (schedule.rs:150-215): k_log in [7, 40], useful_bits <= 2^k_log; is_element
 * selects TableClass::LargeField (useful_bits a multiple of 128). */
typedef struct {
    uint32_t k_log;
    uint32_t is_element;
    uint64_t useful_bits;
} flock_table_type;
/* Slot (schedule.rs:247-268); type_index = the caller's index of the slot's type. */
typedef struct {
    uint64_t offset, prefix;
    uint32_t m_slot, prefix_bits;
    uint32_t type_index, reserved;
} flock_slot;
/* Registry (schedule.rs:275-300) with the types copied in slot order. */
typedef struct {
    uint32_t n_types, num_boolean, nu, m_total, m_bool, m_elem;
    uint64_t element_base;
    flock_slot slots[FLOCK_MAX_TYPES];
    flock_table_type types[FLOCK_MAX_TYPES];
} flock_registry;
/* UnionInstance sizes (union.rs:325-416). commit_lanes 0 = None. */
typedef struct {
    uint64_t dense_words, committed_words, packed_len, commit_lanes;
    uint32_t dense_m, compaction_is_identity;
} flock_union_sizes;

/* Registry::new (schedule.rs:337-495): class-major, k_log-descending stable slot order, aligned
 * offsets, M, M_bool, M_elem, element_base. Host only. */
int flock_registry_new(const flock_table_type* types, size_t n_types, uint32_t nu,
                       flock_registry* out);
/* dense_words / committed_words / dense_m / commit_lanes(log_batch_size) /
 * compaction_is_identity (union.rs:325-416, MIN_DENSE_M = 22) for the declared counts (slot
 * order, each <= 2^nu). dense_floor_m = 0: none (set_dense_floor, :91-105). Host only. */
int flock_union_sizes_compute(const flock_registry* reg, const uint64_t* counts,
                              uint32_t dense_floor_m, uint32_t log_batch_size,
                              flock_union_sizes* out);
/* x_ab_from_mlv (union.rs:541-552): h_mlv has m_bool - 6 entries; h_x_inner_rest gets
 * m_bool - 6 - nu, h_x_outer nu. Host only. */
int flock_union_x_ab_from_mlv(const flock_registry* reg, const flock_f128* h_mlv, size_t n_mlv,
                              flock_f128* h_x_inner_rest, flock_f128* h_x_outer);
/* SparseBinaryMatrix (r1cs.rs:14-18) as CSR: row r's columns are cols[row_ptr[r] ..
 * row_ptr[r+1]) in the reference's row order (ascending for the hash circuits). identity = 1
 * stands for the num_rows x num_rows identity (common.rs:231-237), row_ptr and cols unused. */
typedef struct {
    uint64_t num_rows, num_cols;
    const uint64_t* row_ptr;
    const uint32_t* cols;
    uint32_t identity, reserved;
} flock_sparse_matrix;
/* The digested fields of a boolean TableType (schedule.rs:150-186, from_block_r1cs :175-186). */
typedef struct {
    uint32_t k_log, has_const_pin;
    uint64_t useful_bits, const_pin;
    flock_sparse_matrix a0, b0, c0;
} flock_type_statement;
/* Registry::digest (schedule.rs:542-580): BLAKE3 of "flock-registry-v1", 1u8, nu and the type
 * count (u32 LE), then per type in Registry::new's order (stable, k_log descending): k_log
 * u32, useful_bits u64, const pin (u8 present, u64 value), absorb_matrix of A_0, B_0, C_0
 * (r1cs.rs:419-428). Boolean types without an IO schema: the element and schema payloads
 * append nothing then. Host only. */
int flock_registry_digest(const flock_type_statement* types, size_t n_types, uint32_t nu,
                          uint8_t out[32]);
/* blake3::hash (the 32-byte root hash, any length), the hash behind the registry digest.
 * Host only. */
void flock_blake3_hash(const void* data, size_t len, uint8_t out[32]);
/* ---- built-in hash circuits (crates/flock-prover/src/r1cs_hashes) ------------------ */

/* (A_0, B_0) of one compression's boolean R1CS, C_0 = I, built on the host on first use and
 * kept for the process (read-only, shared, never freed):
 *   FLOCK_HASH_SHA256  sha2.rs:688-701 build_matrices (k_log 15, pin 25,469, 47.4 M nnz)
 *   FLOCK_HASH_BLAKE3  blake3.rs:659-783 build_matrices (k_log 14, pin 11,706)
 *   FLOCK_HASH_KECCAK  the registry's empty-stub pair (keccak.rs:880-888; k_log 16, pin 4096),
 *                      its lincheck runs through the walker.
 * The Rust seams: Sha2LincheckCircuit / Blake3LincheckCircuit (sha2.rs:837, blake3.rs:937) and
 * the TableType of build_block_r1cs (schedule.rs:175-186). Host only. */
typedef struct flock_circuit flock_circuit;
int flock_circuit_get(uint32_t kind, const flock_circuit** out);
int flock_circuit_info(const flock_circuit* c, uint32_t* k_log, uint64_t* useful_bits,
                       uint64_t* const_pin, uint64_t* nnz_a, uint64_t* nnz_b);
/* CSR rows (which = 0: A_0, 1: B_0), 2^k_log + 1 row pointers, ascending columns per row. */
int flock_circuit_csr(const flock_circuit* c, uint32_t which, const uint64_t** row_ptr,
                      const uint32_t** cols);
/* The lincheck's CSC form (csc_from_rows, lincheck.rs:367-388): ascending rows per column. */
int flock_circuit_csc(const flock_circuit* c, uint32_t which, const uint32_t** col_ptr,
                      const uint32_t** rows);
/* The registry TableType of the circuit (views into c). */
int flock_circuit_type_statement(const flock_circuit* c, flock_type_statement* out);
/* Registry::digest of the one-type registry of the built-in circuit at nu (what
 * bind_statement absorbs for a single-hash prove_fast). */
int flock_circuit_registry_digest(uint32_t kind, uint32_t nu, uint8_t out[32]);
/* Registry::digest of the registry of several built-in types at nu (kinds in any order;
 * the digest takes them in Registry::new's order), e.g. BLAKE3 + SHA-256 at nu 10 for the
 * mixed Blake3Sha2Nu10 registry (flock-prover/src/mixed.rs:59, 91-98). Cached per
 * (kinds, nu). */
int flock_circuit_union_digest(const uint32_t* kinds, size_t n_kinds, uint32_t nu,
                               uint8_t out[32]);
/* Upload the circuit's CSC to the device (flock_hip_malloc on ctx); release with
 * flock_hip_circuit_release. */
int flock_hip_circuit_upload(flock_hip_ctx* ctx, uint32_t kind, flock_csc* out);
void flock_hip_circuit_release(flock_hip_ctx* ctx, flock_csc* c);

/* bind_statement (union.rs:619-628): label "flock-mixed-v1", the 32-byte Registry::digest
 * (flock_registry_digest), the counts as u64 LE, the cap bytes. */
int flock_union_bind_statement(flock_fs* fs, const uint8_t registry_digest[32],
                               const uint64_t* counts, size_t n_counts, const uint8_t* cap,
                               size_t cap_len);
/* compact_witness_unchecked (union.rs:449-490): gather the committed dense stack
 * d_q[committed_words] (committed_words for the given dense_floor_m, 0 = none) from the padded
 * batch-major union buffer d_z_padded[packed_len]. */
int flock_hip_union_compact(flock_hip_ctx* ctx, const flock_registry* reg, const uint64_t* counts,
                            uint32_t dense_floor_m, const flock_f128* d_z_padded,
                            flock_f128* d_q);

/* ---- witness generation (crates/flock-prover/src/r1cs_hashes) -------------------------- */

/* SHA-256 compression witness, batch-major: generate_witness_batch_major_partial
 * (r1cs_hashes/sha2.rs:1876-1898 over drive_witness_batch_major_partial, common.rs:805-960).
 * d_inputs holds n_decl compressions of 24 u32 (h_in[8], m[16]); rows n_decl..2^n_log are
 * dummy rows, zero in every output. d_z, d_a, d_b: 2^(n_log+8) F128 each, word (chunk << n_log)
 * + row = bits [128 chunk, 128 chunk + 128) of that row's 2^15-bit (z, A_0 z, B_0 z);
 * d_stripe (may be NULL): the lincheck stripe, 2^(n_log+12) bytes. 3 <= n_log <= 26. */
int flock_hip_witness_sha256(flock_hip_ctx* ctx, const uint32_t* d_inputs, size_t n_decl,
                             uint32_t n_log, flock_f128* d_z, flock_f128* d_a, flock_f128* d_b,
                             uint8_t* d_stripe);

/* BLAKE3 compression witness, batch-major: generate_witness_batch_major_partial
 * (r1cs_hashes/blake3.rs:2072-2088, rows from build_block_witness_ab_packed_into :1299-1452).
 * d_inputs holds n_decl compressions of 28 u32 (cv[8], m[16], counter_lo, counter_hi,
 * block_len, flags); rows n_decl..2^n_log are zero. d_z, d_a, d_b: 2^(n_log+7) F128 each at
 * (chunk << n_log) + row of the 2^14-bit rows; d_stripe (may be NULL): 2^(n_log+11) bytes.
 * 3 <= n_log <= 26. */
int flock_hip_witness_blake3(flock_hip_ctx* ctx, const uint32_t* d_inputs, size_t n_decl,
                             uint32_t n_log, flock_f128* d_z, flock_f128* d_a, flock_f128* d_b,
                             uint8_t* d_stripe);
/* Keccak-f[1600] permutation witness, batch-major: generate_witness_batch_major
 * (r1cs_hashes/keccak.rs:1228-1285, rows from build_chain_witness_ab_packed_into :565-643).
 * d_inputs holds n_decl initial states of 25 u64 lanes (lane x + 5y, bit z); rows n_decl..2^n_log
 * run keccak_f on the zero state (the reference's padding invocations). d_z, d_a, d_b:
 * 2^(n_log+9) F128 each at (chunk << n_log) + row of the 2^16-bit rows; d_stripe (may be NULL):
 * 2^(n_log+13) bytes. 3 <= n_log <= 24. */
int flock_hip_witness_keccak(flock_hip_ctx* ctx, const uint64_t* d_inputs, size_t n_decl,
                             uint32_t n_log, flock_f128* d_z, flock_f128* d_a, flock_f128* d_b,
                             uint8_t* d_stripe);

/* ---- Fiat-Shamir transcript (host; FsChallenger, SHA-256 discipline) ---------------- */

/* FsChallenger::new (challenger.rs:648-659) and the Challenger methods (:806-958). */
flock_fs* flock_fs_new(const uint8_t* domain, size_t len);
void flock_fs_free(flock_fs* fs);
void flock_fs_observe_label(flock_fs* fs, const uint8_t* label, size_t len);
void flock_fs_observe_f128(flock_fs* fs, flock_f128 v);
void flock_fs_observe_f128_slice(flock_fs* fs, const flock_f128* v, size_t n);
void flock_fs_observe_bytes(flock_fs* fs, const uint8_t* b, size_t n);
void flock_fs_sample_f128(flock_fs* fs, flock_f128* out);
void flock_fs_sample_f128_vec(flock_fs* fs, flock_f128* out, size_t n);
/* Challenger::grind_pow (challenger.rs:884-958): the smallest nonce whose SHA-256 with the
 * state digest starts with `bits` zero bits (bits = 0: nonce 0), absorbed as 8 LE bytes. bits
 * above 64 are refused: UINT64_MAX, FLOCK_HIP_E_INVALID in flock_hip_last_error, nothing
 * absorbed. FLOCK_HIP_PORTABLE_SHA256=1 in the environment forces the portable SHA-256 block
 * function over the SHA extensions for every host transcript. */
uint64_t flock_fs_grind_pow(flock_fs* fs, uint32_t bits);
/* Challenger::verify_pow (challenger.rs:960-981): 1 if nonce passes (bits = 0: only nonce 0),
 * else 0; the nonce bytes are absorbed either way. */
int flock_fs_verify_pow(flock_fs* fs, uint64_t nonce, uint32_t bits);

/* The transcript's running state digest: SHA-256 of everything absorbed so far (the value
 * grind_pow hashes with its nonce, challenger.rs:884-958). Two transcripts that absorbed the
 * same bytes give the same digest. */
void flock_fs_digest(flock_fs* fs, uint8_t out[32]);

/* ---- protocol drivers ------------------------------------------------------------ */

/* Product sumcheck followed by an eq-weighted zerocheck tail on the same two tables of
 * 2^log_n words, over transcript fs. Product phase: observe label
 * "flock-product-sumcheck", then SumcheckProver::new + the sample/fold/observe loop of
 * pcs/ligerito.rs:4986-5011 for log_n rounds, the last fold without message, then
 * observe the two folded values. Zerocheck phase: observe label "flock-zerocheck-v0",
 * sample log_n eq challenges r, then the message/sample/fold loop of
 * zerocheck.rs:790-975 with Convention A (r_next[0] = ONE), a final fold, and observe
 * the two folded values. h_msgs receives 4*log_n words ({u0,u2} per product round, then
 * {G(1), G(inf)} per zerocheck round); h_finals = {f, b, a, b} final values.
 * d_a and d_b are not modified. */
int flock_hip_prove_sumcheck(flock_hip_ctx* ctx, flock_fs* fs, const flock_f128* d_a,
                             const flock_f128* d_b, uint32_t log_n, flock_f128* h_msgs,
                             flock_f128 h_finals[4]);

/* ---- merged PCS opening (pcs.rs:1433-1822) ---------------------------------------- */

/* Device workspace of the opening: grow-only buffers reused across calls on one context.
 * No counterpart in the reference (its opening allocates through scratch.rs pools). */
typedef struct flock_open_ws flock_open_ws;
int flock_open_ws_create(flock_open_ws** out);
void flock_open_ws_destroy(flock_open_ws* ws);
/* device bytes the workspace holds */
uint64_t flock_open_ws_bytes(const flock_open_ws* ws);
/* host wall time (ms) of the last flock_hip_open_merged's stages, n <= 6 of them: [0] ring
 * switch and the jagged weight W, [1] the merged sumcheck, [2] the multipoint values and the
 * twisted sumcheck, [3] the Frobenius assist, [4] the inner Ligerito opening, [5] the wire
 * assembly. No counterpart in the reference (a profiling aid). */
int flock_open_ws_stage_ms(const flock_open_ws* ws, double* out, size_t n);

/* initial_k of the embedded `fast` Ligerito config for m (ligerito::embedded_initial_k,
 * ligerito.rs:564-574; configs/ligerito/m{22..35}_fast.toml), or -1 without one. The commit's
 * log_batch_size must equal it. */
int flock_lig_initial_k(uint32_t m);
/* PcsParams::l0_cap_depth (pcs/commit.rs:266-277): the cap depth of the L0 commitment for
 * dense m under the `fast` profile (the top set bit of the L0 query count), or -1. */
int flock_l0_cap_depth(uint32_t m);

/* s_hat_v_from_z_vec (pcs/ring_switch.rs:1278-1330): h_out[b] = sum_k eq(tail, k) z_vec[128 k + b]
 * over the lincheck's captured z_vec (z_vec_len = 128 << tail_len words, device). */
int flock_hip_s_hat_v_from_z_vec(flock_hip_ctx* ctx, flock_open_ws* ws, const flock_f128* d_z_vec,
                                 size_t z_vec_len, const flock_f128* h_tail, size_t tail_len,
                                 flock_f128 h_out[128]);

/* pcs::open_batch_merged (pcs.rs:1433-1822) for n_claims (1..4) ring-switched claims with
 * precomputed s_hat_v, on transcript fs: ring switch, jagged merged weight and dense merged
 * sumcheck, the multipoint twisted assist with its Frobenius anchor on a forked transcript,
 * and the eq-basis Ligerito opening (F256 ladder, stratified queries, capped paths) under the
 * embedded `fast` config for p->m (OpeningGrinding::per_challenge_128, pcs.rs:155-205).
 * d_q: the committed dense stack (2^(m-7) words); d_codeword / d_tree: its commitment (full
 * tree); h_x_outers: n_claims points of x_len = 1 + n_log + log2(n_heights) words (the quirky
 * point's x_inner_rest ++ x_outer, prover.rs:43-48); h_s_hat_v: 128 words per claim;
 * h_heights: the jagged column heights (union.rs:293-310). On success *out_bytes (free with
 * flock_hip_free) holds the bincode of MergedOpenProof (pcs.rs:1362-1377) and h_q_eval, when
 * not NULL, q^(rho). */
int flock_hip_open_merged(flock_hip_ctx* ctx, flock_open_ws* ws, flock_fs* fs, const flock_f128* d_q,
                          const flock_pcs_params* p, const flock_f128* d_codeword, const uint8_t* d_tree,
                          size_t n_claims, const flock_f128* h_x_outers, size_t x_len,
                          const flock_f128* h_s_hat_v, const uint64_t* h_heights, size_t n_heights,
                          uint32_t n_log, uint8_t** out_bytes, size_t* out_len, flock_f128* h_q_eval);

/* The merged opening's verifier: verify_batch_merged (pcs.rs:1971-2300) for ring-switched
 * claims — per claim the ring switch (ring_switch.rs:2966-3005: the claim against the
 * skip (x) eq(x[0]) weights of s_hat_v, 3 PoW bits), the batching gammas, the dense merged
 * sumcheck replay (2 PoW bits a round); on the forked transcript the multipoint twisted
 * evaluation and its Frobenius anchor (pcs/jagged.rs:2034-2230, 4335-4585: the anchor's
 * boundary-bit sumcheck checked against the branching program at its endpoint, the 128 dual
 * values recombined through the linearized coefficients of each claim's fold map); on the
 * parent transcript the succinct Ligerito verifier of q^(rho) (pcs.rs:1129-1320,
 * pcs/ligerito/extension.rs:2409-2880: OOD samples, stratified queries, capped SHA-256
 * Merkle paths, induced consistency sums, the residual basis against yr); then the merged
 * endpoint q^(rho) V. Inputs as flock_hip_open_merged's (values = the claims' values,
 * h_z_skips = their skip points), proof = the MergedOpenProof bytes; *consumed gets the bytes
 * read. 0 or FLOCK_HIP_E_REJECT naming the failed check. SHA-256 Merkle commitments only.
 * Host only. */
int flock_open_merged_verify(flock_fs* fs, const flock_pcs_params* p, const uint8_t* cap, size_t cap_len,
                             size_t n_claims, const flock_f128* h_values, const flock_f128* h_z_skips,
                             const flock_f128* h_x_outers, size_t x_len, const uint64_t* h_heights,
                             size_t n_heights, uint32_t n_log, const uint8_t* proof, size_t proof_len,
                             size_t* consumed);

/* ---- prove core and serving (host orchestration over the entry points above) ----------- */

/* The three hash circuits of the batched and served workloads (r1cs_hashes/{sha2,blake3,
 * keccak}.rs). */
#define FLOCK_HASH_SHA256 0u
#define FLOCK_HASH_BLAKE3 1u
#define FLOCK_HASH_KECCAK 2u

/* One proving request. h_inputs: n_decl instances of 24 u32 (SHA-256: h_in[8], m[16]), 28 u32
 * (BLAKE3: cv[8], m[16], counter_lo, counter_hi, block_len, flags) or 25 u64 (Keccak state
 * lanes). circuit: the device CSC circuit (SHA-256, BLAKE3), or NULL for the built-in one
 * (flock_hip_circuit_upload, once per prover); unused for Keccak (walker). digest: the
 * 32-byte Registry::digest bound into the transcript, or NULL for the built-in one-type
 * registry's (flock_circuit_registry_digest(kind, n_log)). arrival_s: arrival offset
 * for flock_hip_serve (non-decreasing across a request array). inputs_on_device: the inputs
 * are already in HBM (no host-to-device copy). 3 <= n_log <= 24. */
typedef struct {
    uint32_t kind;
    uint32_t n_log;
    uint64_t n_decl;
    const void* h_inputs;
    const flock_csc* circuit;
    const uint8_t* digest;
    const uint8_t* domain;
    size_t domain_len;
    double arrival_s;
    uint32_t inputs_on_device; /* 1: h_inputs is a device pointer on the prover's device */
    uint32_t reserved;
} flock_prove_request;

/* root: the commitment root (cap depth 0); transcript: flock_fs_digest after the lincheck,
 * which has absorbed the statement, every zerocheck and lincheck message and nonce;
 * zc_z / lc_w: the zerocheck skip point and the lincheck claim value; hashes = n_decl.
 * queued_ms, run_ms, done_s: set by flock_hip_serve (wait after arrival, run time, completion
 * time from the start of the call). stage_ms: host wall time of the stages, each ending at a
 * stream synchronisation — [0] upload, witness, dense stack and commit, [1] zerocheck,
 * [2] lincheck, [3] the merged opening and the wire assembly (0 for the prove core). */
typedef struct {
    uint8_t root[32];
    uint8_t transcript[32];
    flock_f128 zc_z, lc_w;
    uint64_t hashes;
    double queued_ms, run_ms, done_s;
    double stage_ms[4];
    int status;
    int reserved;
} flock_prove_result;

/* A prover: one context (HIP stream) plus grow-only device buffers reused across requests. */
typedef struct flock_hip_prover flock_hip_prover;
int flock_hip_prover_create(int device, flock_hip_prover** out);
void flock_hip_prover_destroy(flock_hip_prover* p);
flock_hip_ctx* flock_hip_prover_ctx(flock_hip_prover* p);
/* The prover's opening workspace (null before its first prove_fast). */
flock_open_ws* flock_hip_prover_open_ws(flock_hip_prover* p);
/* Device bytes the prover holds between requests (grow-only request buffers, the opening's
 * workspace, the context's driver scratch, uploaded built-in circuits), and their release
 * (all but the circuits; the next request allocates again). */
uint64_t flock_hip_prover_held_bytes(const flock_hip_prover* p);
void flock_hip_prover_release(flock_hip_prover* p);
/* Device bytes one request needs (the admission size of flock_hip_serve). */
int flock_hip_prove_footprint(const flock_prove_request* req, uint64_t* bytes);
/* The GPU prove core of prove_fast for a registry with one boolean hash type
 * (prover.rs:926-1160): witness, dense stack, commit (rate 1/2, log_batch_size = the
 * Ligerito config's initial_k; lane-major for an integer lane count), bind_statement of the
 * cap at flock_l0_cap_depth, zerocheck with s_hat_v_c, x_ab split and the union lincheck,
 * with grinding. Stops before the PCS opening (pcs.rs:1433), so it is not a complete proof;
 * below dense m 22 (no Ligerito config) it commits at log_batch_size 6 and binds the root. */
int flock_hip_prove_core(flock_hip_prover* p, const flock_prove_request* req,
                         flock_prove_result* res);
/* The whole prove_fast for a registry with one boolean hash type: prove_union_with_binding_zc
 * (prover.rs:831-1468) through prove_fast_ligerito_union (prover.rs:587-624) — the prove core
 * above, binding the commitment CAP at flock_l0_cap_depth, then s_hat_v_ab and the merged
 * opening. On success out->proof (free with flock_hip_free) is bincode(R1csProofMergedLigerito)
 * (proof.rs:42-47): ZerocheckProof, LincheckProof (one matrix-eval pair), MergedOpenProof;
 * out->cap (free with flock_hip_free) the commitment cap; claim_ab / claim_c the R1csClaim
 * values. The SHA-256 of proof || cap || claim values is the digest the reference's
 * byte-identity anchors pin (union_m6_fixtures.rs:156-172). */
typedef struct {
    uint8_t* proof;
    size_t proof_len;
    uint8_t* cap;
    size_t cap_len;
    flock_f128 claim_ab, claim_c;
} flock_proof_bytes;
int flock_hip_prove_fast(flock_hip_prover* p, const flock_prove_request* req, flock_prove_result* res,
                         flock_proof_bytes* out);
/* One boolean hash type of a mixed request: kind, n_decl instances (0 allowed; Keccak
 * counts 2^n_log whatever n_decl), inputs as in flock_prove_request (host, or device with
 * inputs_on_device), circuit NULL for the built-in one. */
typedef struct {
    uint32_t kind;
    uint32_t reserved;
    uint64_t n_decl;
    const void* inputs;
    const flock_csc* circuit;
} flock_union_part;
/* A prove_fast request over a registry of several boolean hash types at one batch capacity
 * 2^n_log (3 <= n_log <= 24): Registry::new sorts the parts into slots (k_log descending,
 * stable, schedule.rs:337-495). digest NULL: flock_circuit_union_digest of the parts' kinds. */
typedef struct {
    uint32_t n_log;
    uint32_t n_parts; /* 1 .. FLOCK_MAX_TYPES */
    const flock_union_part* parts;
    const uint8_t* digest;
    const uint8_t* domain;
    size_t domain_len;
    uint32_t inputs_on_device;
    uint32_t reserved;
} flock_union_request;
/* What a mixed bundle records beside the proof (MixedProofBundleLigerito, proof_io.rs:
 * 230-282): slot_part[s] = the part in slot s, counts[s] its statement count (slot order),
 * the PCS parameters (dense_m, log_batch_size, num_lanes 0 = None), M_bool, whether a
 * Ligerito config exists (dense m >= 22), and the request's device footprint. */
typedef struct {
    uint32_t n_slots, m_bool, dense_m, log_batch_size;
    uint64_t num_lanes;
    uint32_t has_opening, reserved;
    uint64_t device_bytes;
    uint32_t slot_part[FLOCK_MAX_TYPES];
    uint64_t counts[FLOCK_MAX_TYPES];
} flock_union_info;
int flock_union_request_info(const flock_union_request* req, flock_union_info* out);
/* The whole prove_fast for a registry of several boolean hash types: prove_fast_ligerito_union
 * (prover.rs:587-624) over prove_union_with_binding_zc (prover.rs:831-1468) with one
 * UnionSlotProverInput per slot, as tests/union_m6_fixtures.rs:236-270 drives it: each slot's
 * witness at its offset in the packed boolean region (the rest zero), the dense stack over
 * the slots' counts, commit, bind_statement with the counts in slot order, zerocheck over
 * M_bool, the union lincheck (flock_hip_lincheck_prove_union), s_hat_v_ab, the jagged
 * heights of every slot, the merged opening; out->proof = bincode(R1csProofMergedLigerito)
 * with one matrix-eval pair per slot. A request of one part equals flock_hip_prove_fast. */
int flock_hip_prove_fast_union(flock_hip_prover* p, const flock_union_request* req,
                               flock_prove_result* res, flock_proof_bytes* out);
/* The whole-proof verifier: verify_ligerito_union (verifier.rs:156-190, as the mixed
 * bundle's MixedSetup::verify drives it, mixed.rs:281-301) for a request's registry of
 * built-in types (circuit NULL; inputs are not read): the PCS parameters re-derived from the
 * counts, bind_statement of the cap, the zerocheck verifier over M_bool, x_ab, the union
 * lincheck verifier with the O(nnz) matrix check on the host circuits, and the merged
 * opening's verifier at the (ab, c) claim points; every byte of the proof must be read.
 * Returns 0 with the R1csClaim values, or FLOCK_HIP_E_REJECT naming the failed check.
 * Host only (no GPU). */
int flock_verify_fast(const flock_prove_request* req, const uint8_t* cap, size_t cap_len,
                      const uint8_t* proof, size_t proof_len, flock_f128* claim_ab, flock_f128* claim_c);
int flock_verify_fast_union(const flock_union_request* req, const uint8_t* cap, size_t cap_len,
                            const uint8_t* proof, size_t proof_len, flock_f128* claim_ab,
                            flock_f128* claim_c);
/* Serving loop: `workers` host threads, one prover each, take the requests in arrival order;
 * a request starts once its arrival time has passed and its footprint fits hbm_budget beside
 * what the other workers hold (running ones: max(held, footprint)); a worker gives back its
 * own buffers when that makes the request fit, and with nothing running the idle workers'
 * buffers are freed and the request runs alone. Results and timings go to res[i]; with
 * proofs != NULL each request is a whole prove_fast and proofs[i] gets its bytes (free each
 * with flock_hip_free), else the prove core runs. wall_ms gets the time from the call's start
 * to the last completion. No counterpart in the reference Rust (its CUDA FFI serialises calls
 * behind one mutex, prove_ffi.cu:421-422). */
int flock_hip_serve(flock_hip_prover* const* provers, uint32_t workers,
                    const flock_prove_request* reqs, size_t n, uint64_t hbm_budget,
                    flock_prove_result* res, flock_proof_bytes* proofs, double* wall_ms);

#ifdef __cplusplus
}
#endif
#endif
