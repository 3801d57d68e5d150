// Internal interface between the host runtime (api.hip) and the kernels
// (field.hip, ntt.hip, merkle.hip, sumcheck.hip). Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/flock_hip.h"
#include "transcript_device.h"

struct flock_hip_ctx;

namespace flock {

// Two F128 per block partial (e.g. u0, u2) — 32 bytes.
constexpr int kPartialWords = 2;
constexpr int kMaxReduceBlocks = 512;  // two 256-thread groups per CU: the whole grid is resident

// ---- field.hip ----
hipError_t launch_f128_mul_batch(const uint4* a, const uint4* b, uint4* out, size_t n,
                                 hipStream_t s);
// out[i] = a[i] * t^reps: method 0 the LDS Shoup table, 1 the int8 matrix cores.
hipError_t launch_f128_mul_const_pow(const uint4* a, uint4 t, uint32_t reps, int method,
                                     uint4* out, size_t n, hipStream_t s);
hipError_t launch_f128_mul_const_batch(const uint4* a, uint4 t, uint4* out, size_t n,
                                       hipStream_t s);

// ---- ntt.hip ----
// Lane-major replicate fill: cw[(rep*D + p)*t + l] = q[l*D + p] for l < t.
hipError_t launch_fill_lane_major(const uint4* q, uint4* cw, uint64_t d, uint32_t t,
                                  uint32_t reps, hipStream_t s);
// Interleaved replicate fill: cw[i] = msg[i % msg_len].
hipError_t launch_fill_replicate(const uint4* msg, uint4* cw, uint64_t msg_len, uint32_t reps,
                                 hipStream_t s);
// Forward interleaved LCH NTT over layers [start_layer, k) of 2^k positions x t lanes,
// transforming only lanes < live. evals: k x k basis table (device). If msg != nullptr the
// first pass reads the 2^(k-start_layer) x t message instead of data (the state after
// start_layer copy layers is the message replicated 2^start_layer times), so data needs
// no prior fill. lm_d != 0: the message is lane-major, lane l at msg[l * lm_d ..] (lm_d =
// 2^(k-start_layer) positions), as commit_lane_major's stack.
hipError_t launch_ntt_forward(uint4* data, uint32_t t, uint32_t live, int k, int start_layer,
                              const uint4* evals, const uint4* msg, hipStream_t s,
                              uint64_t lm_d = 0);

// One fused pass of `layers` in {1,2,3} layers starting at l0, in place.
hipError_t launch_ntt_pass(uint4* data, uint32_t t, uint32_t live, int k, int l0, int layers,
                           const uint4* evals, hipStream_t s);

// ---- merkle.hip ----
// SHA-256 Merkle tree, flat layout (leaves first, root last). leaf_bytes % 16 == 0.
// HashKind (hash.rs:23-27) as numbered in flock_pcs_params.merkle_hash
constexpr uint32_t kMerkleSha256 = 0, kMerkleBlake3 = 1;
// BLAKE3 leaves of up to 2^16 chunks (64 MiB): the per-thread subtree stack's depth
constexpr int kMerkleB3MaxDepth = 16;
// top_nodes (a power of two): the build may stop once a layer has this many nodes (the cap
// layer); the layers above it are then left unwritten. 1 builds the whole tree.
hipError_t launch_merkle_tree(const uint8_t* data, uint64_t n_leaves, uint64_t leaf_bytes,
                              uint32_t kind, uint8_t* tree, hipStream_t s, uint64_t top_nodes = 1);
// A context whose commits build their trees up to the cap layer only (the opening's level
// commits, whose paths end at the cap). Off by default.
void ctx_set_tree_to_cap(flock_hip_ctx* c, bool on);

// ---- sumcheck.hip ----
// round_msg_lsb: out[0] = sum f0 b0, out[1] = sum (f0+f1)(b0+b1). partials >= kMaxReduceBlocks*2 + 1 words (the last holds a zeroed ticket counter).
// Round kernels that emit a message pair send it where a RoundOut says (transcript_device.h:
// a plain buffer, or the device transcript's round step); fold kernels take the challenge from
// the device pointer rin instead of the value argument when rin is set.
hipError_t launch_sc_round_msg(const uint4* f, const uint4* b, uint64_t n, uint4* partials,
                               RoundOut ro, hipStream_t s);
// fold_and_msg_lsb: nf/nb (n/2) = v0 + r(v0+v1); out = next message (zero if n/2 == 1).
hipError_t launch_sc_fold_msg(const uint4* f, const uint4* b, uint64_t n, uint4 r, uint4* nf,
                              uint4* nb, uint4* partials, RoundOut ro, hipStream_t s,
                              const uint4* rin = nullptr);
// Zerocheck first message on (a, b) of size n with eq table e (n/2): out = (G(1), G(inf)).
// If e_next != nullptr, e_next[z] = e[2z] ^ e[2z+1] (n/4 entries).
hipError_t launch_zc_round(const uint4* a, const uint4* b, uint64_t n, const uint4* e,
                           uint4* e_next, uint4* partials, RoundOut ro, hipStream_t s);
// Fused: fold (a, b) of size n at rho into (ao, bo) of size n/2, then the message on the
// folded tables with eq table e (n/4), and e_next (n/8) as above.
hipError_t launch_zc_fold_round(const uint4* a, const uint4* b, uint64_t n, uint4 rho,
                                uint4* ao, uint4* bo, const uint4* e, uint4* e_next,
                                uint4* partials, RoundOut ro, hipStream_t s,
                                const uint4* rin = nullptr);
// Process-wide launch tunables (set through flock_hip_set_tunable, read at each launch).
enum Tunable : int {
    kTuneNttColsPerThread = 0,  // columns per thread of the interior NTT pass (1..4)
    kTuneNttDeepCols,           // a pass with at most this many columns per block runs deep (32..256)
    kTuneZcSplitMinVars,        // zerocheck rounds whose eq has >= this many variables split it (10..24)
    kTuneZcLaMaxLog,            // zerocheck tail: two rounds per launch from 2^this words (10..22)
    kTuneScLaMaxLog,            // opening product sumchecks: two rounds per launch from 2^this (10..24)
    kTuneOpenSplitFoldLog,      // F256 ladder: fold and message in two passes from 2^this (10..22)
    kTuneCount
};
int64_t tunable(Tunable t);

// Marks a proof in flight for its lifetime (the host grinds use their worker pool only while a
// single proof runs in the process).
struct ProofScope {
    ProofScope();
    ~ProofScope();
    ProofScope(const ProofScope&) = delete;
    ProofScope& operator=(const ProofScope&) = delete;
};

// Where a grid reduction of several words goes: the workgroups fold into acc (the words, then
// a ticket counter; zero between launches), and the last one moves the sums to out (host-
// visible), clears acc and stores seq into flag (system scope).
struct LaOut {
    uint32_t* acc;
    uint32_t* out;
    uint32_t* flag;
    uint32_t seq;
};
// Two zerocheck tail rounds per launch: fold (a, b) of size n by nfold (1 or 2) challenges
// into (ao, bo), then round j's pair with eq table e ((n >> nfold) / 2 entries) and round
// j+1's pair as quadratics in its challenge (values at 0, 1 and leading coefficients); e2
// gets round j+2's eq table ((n >> nfold) / 8 entries, written when that is >= 1). Eight F128
// go to lo.out: [G(1), G(inf), next G(1) at 0, 1, inf, next G(inf) at 0, 1, inf].
hipError_t launch_zc_fold_la(const uint4* a, const uint4* b, uint64_t n, int nfold, uint4 r1,
                             uint4 r2, uint4* ao, uint4* bo, const uint4* e, uint4* e2, LaOut lo,
                             hipStream_t s);
// Split-eq message for rounds whose eq table has >= 16 variables: eq(x) = e_lo[x & 255] *
// e_hi[x >> 8]; n (the table length) is a multiple of 512.
hipError_t launch_zc_round_split(const uint4* a, const uint4* b, uint64_t n, const uint4* e_lo,
                                 const uint4* e_hi, uint4* partials, RoundOut ro, hipStream_t s);
// Fused fold at rho (n -> n/2) and the split-eq message on the folded tables; n is a multiple
// of 1024.
hipError_t launch_zc_fold_round_split(const uint4* a, const uint4* b, uint64_t n, uint4 rho,
                                      uint4* ao, uint4* bo, const uint4* e_lo, const uint4* e_hi,
                                      uint4* partials, RoundOut ro, hipStream_t s,
                                      const uint4* rin = nullptr);
// out[j * 256 + t] = eq(r[j .. j+8), t) for j < count (r: count + 7 device words).
hipError_t launch_eq_windows(const uint4* r, uint32_t count, uint4* out, hipStream_t s);
// Levels 1..levels-1 of the eq table e0 over nv variables (level j drops the j lowest
// variables), packed one after another into out.
hipError_t launch_eq_levels(const uint4* e0, uint32_t nv, uint32_t levels, uint4* out,
                            hipStream_t s);
// Plain fold (fold_in_place_pair): out-of-place, n -> n/2.
hipError_t launch_fold_pair(const uint4* a, const uint4* b, uint64_t n, uint4 c, uint4* ao,
                            uint4* bo, hipStream_t s, const uint4* cin = nullptr);
// eq table build_eq(r[0..nv)) into out (2^nv entries). r is a device array; scratch holds
// at least 2^(nv/2) + 2^(nv - nv/2) entries.
hipError_t launch_build_eq(const uint4* r, int nv, uint4* out, uint4* scratch, hipStream_t s);
// copies n <= kEqArgMax host coordinates to dst through the launch's arguments
constexpr int kEqArgMax = 40;
hipError_t launch_eq_coords(const uint4* h_r, uint32_t n, uint4* dst, hipStream_t s);
// build_eq of host coordinates queued on the context's stream without a wait: the coordinates
// travel in a launch's arguments, so h_r may be released on return; d_out is ready when the
// stream reaches it
int build_eq_async(flock_hip_ctx* c, const flock_f128* h_r, size_t nv, flock_f128* d_out);
// The two halves of eq(r, .) over nv <= 40 host coordinates: lo = eq over r[0 .. nv/2) (2^(nv/2)
// entries), hi = eq over the rest, so that eq(r, x) = lo[x mod 2^(nv/2)] * hi[x >> nv/2].
int build_eq_halves_async(flock_hip_ctx* c, const flock_f128* h_r, size_t nv, flock_f128* d_lo,
                          flock_f128* d_hi);
// eq halves of nv <= 40 device coordinates r (k_eq_direct alone)
hipError_t launch_eq_halves(const uint4* r, int nv, uint4* lo, uint4* hi, hipStream_t s);
// the device a context was made on (-1 for none)
int ctx_device(const flock_hip_ctx* c);
// out[x_lo + 2^nlo x_hi] = lo[x_lo] * hi[x_hi] (2^(nlo+nhi) words).
hipError_t launch_outer_product(const uint4* lo, int nlo, const uint4* hi, int nhi, uint4* out,
                               hipStream_t s);

// ---- zerocheck.hip ----
// Round 1 on packed bit rows (2^(m-6) rows of 64 bits): out[0..64) = p_ab, out[64..128) = p_c.
// eq: 2^(m-6) words; ext: 16x16x64-byte extension table; logexp: log[256] + exp[512];
// phi_pow: phi8(2^b), b < 8; partials >= kZcMaxBlocks * 1024 words; red: 1024 words.
hipError_t launch_zc_round1(const uint8_t* a, const uint8_t* b, const uint8_t* c, const uint4* eq,
                            uint64_t rows, const uint4* ext, const uint8_t* logexp,
                            const uint4* phi_pow, uint4* partials, uint4* red, uint4* out,
                            hipStream_t s);
// Round 1's p_ab when r[6..9) are the protocol constants phi8(0xF7, 0x53, 0xB5): groups of 8
// rows (n_groups = 2^(m-9)) weighted by eq9 = eq(r[9..m), .); ext_bs: the bit-sliced
// extension table; out[0..64) = p_ab.
hipError_t launch_zc_round1_groups(const uint8_t* a, const uint8_t* b, const uint4* eq9,
                                   uint64_t n_groups, const uint4* ext_bs, const uint4* phi_pow,
                                   uint4* partials, uint4* red, uint4* out, hipStream_t s);
// Round 1's p_c from shat = fold_1b_rows(c, eq(r[7..m))) (128 words) and phim[s*64+i] =
// phi8(L_s(64+i)): out[i] for i < 64.
hipError_t launch_zc_round1_c(const uint4* shat, const uint4* phim, uint4 r6, uint4* out,
                              hipStream_t s);
uint32_t zc_round1_groups(uint64_t rows, uint32_t* rows_per_group, uint32_t max_groups);
constexpr uint32_t kZcMaxBlocks = 2048;  // zerocheck partials hold kZcMaxBlocks x 1024 words
// Fold at z of two packed bit tables with the 8 x 256 fold table.
hipError_t launch_zc_fold_at_z(const uint8_t* a, const uint8_t* b, uint64_t rows,
                               const uint4* table, uint4* az, uint4* bz, hipStream_t s);
// out[z] = in[2z] ^ in[2z+1].
hipError_t launch_eq_fold(const uint4* in, uint64_t n_out, uint4* out, hipStream_t s);
// out[j] = sum_y t[y] * bit j of 128-bit row y (128 words; partials: kZcMaxBlocks x 1024; red:
// 1024 words).
hipError_t launch_fold_1b_rows(const uint4* rows_g, const uint4* t, uint64_t rows,
                              uint4* partials, uint4* red, uint4* out, hipStream_t s);

// ---- lincheck ----
// Keccak-f[1600] circuit marginals over 2^16 columns: sa = A_0^T eq, sb = B_0^T eq and, when
// comb is non-null, comb = alpha * sa + sb. inv: 1600 x 11 phi^T gather indices; trow: 2 x
// 1600 x 11 chi row offsets (A then B); rcm: 25 x 50 words of round-constant vectors.
hipError_t launch_kc_fold(const uint4* eq, const uint16_t* inv, const uint16_t* trow,
                          const uint32_t* rcm, uint4 alpha, uint4* sa, uint4* sb, uint4* comb,
                          hipStream_t s);
hipError_t launch_lc_csc_fold(uint64_t n_cols, const uint32_t* a_ptr, const uint32_t* a_rows,
                              uint64_t nnz_a, const uint32_t* b_ptr, const uint32_t* b_rows,
                              uint64_t nnz_b, const uint4* eq, uint4 alpha, uint4* sa, uint4* sb,
                              uint4* comb, int* err, hipStream_t s);
hipError_t launch_lc_add_at(uint4* v, uint64_t idx, uint4 beta, hipStream_t s);
// v[0..n) *= w
hipError_t launch_lc_scale(uint4* v, uint64_t n, uint4 w, hipStream_t s);
// upper bound of the partial rows (gy) for any stripe count
uint32_t lc_partial_fold_rows(uint64_t K);
void lc_partial_fold_plan(uint64_t K, uint64_t n_stripes, uint32_t* spg, uint64_t* n_groups,
                          uint32_t* gy);
// partials: gy * K words (lc_partial_fold_plan)
hipError_t launch_lc_partial_fold(const uint8_t* z, uint64_t K, uint64_t useful_bits,
                                  uint64_t n_stripes, const uint4* eq_outer, uint4* partials,
                                  uint4* out, hipStream_t s);
hipError_t launch_lc_eval(const uint4* c, const uint4* z, uint64_t len, uint4* partials,
                          RoundOut ro, hipStream_t s);
hipError_t launch_lc_bind_eval(uint4* c, uint4* z, uint64_t len, uint4 r, uint4* partials,
                               RoundOut ro, hipStream_t s, const uint4* rin = nullptr);
hipError_t launch_lc_bind(uint4* c, uint4* z, uint64_t len, uint4 r, hipStream_t s,
                          const uint4* rin = nullptr);
hipError_t launch_dot2(const uint4* w, const uint4* a, const uint4* b, uint64_t n,
                       uint4* partials, uint4* out, hipStream_t s);

// ---- witness data movement ----
hipError_t launch_compact_slot(const uint4* z, uint64_t start, uint32_t nu, uint64_t cols,
                               uint64_t n_t, uint4* q, hipStream_t s);
hipError_t launch_zero_words(uint4* q, uint64_t n, hipStream_t s);
// SHA-256 compression witness, batch-major: z/a/b have 256 << n_log words; stripe (may be null)
// 2^n_log * 4096 bytes. n_log >= 3, rows >= n_decl are zero.
// Keccak-f[1600] witness, batch-major: z/a/b have 512 << n_log words; stripe (may be null)
// 2^n_log * 8192 bytes. Rows >= n_decl run keccak_f on the zero state.
hipError_t launch_keccak_witness(const uint64_t* inputs, uint64_t n_decl, uint32_t n_log,
                                 uint4* z, uint4* a, uint4* b, uint8_t* stripe, hipStream_t s);
// BLAKE3 compression witness, batch-major: z/a/b have 128 << n_log words; stripe (may be null)
// 2^n_log * 2048 bytes. Rows >= n_decl are zero.
hipError_t launch_blake3_witness(const uint32_t* inputs, uint64_t n_decl, uint32_t n_log,
                                 uint4* z, uint4* a, uint4* b, uint8_t* stripe, hipStream_t s);
hipError_t launch_sha2_witness(const uint32_t* inputs, uint64_t n_decl, uint32_t n_log,
                               uint4* z, uint4* a, uint4* b, uint8_t* stripe, hipStream_t s);

// Record an error for flock_hip_last_error (implemented in api.hip); returns code.
int set_error(int code, const char* msg);
// Bytes of the context's grow-only driver scratch, and its release (after a stream sync).
uint64_t ctx_scratch_bytes(const flock_hip_ctx* c);
void ctx_release_scratch(flock_hip_ctx* c);
// Host SHA-256 of n bytes (the x86 SHA extensions where present).
void sha256_host(const void* data, size_t n, uint8_t out[32]);
// Host Lagrange basis on the 64 nodes phi8(offset + i) at z (offset 0: S, 64: Lambda).
void lagrange_phi8_host(int offset, flock_f128 z, flock_f128 w[64]);

}  // namespace flock
