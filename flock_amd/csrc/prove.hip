// Host orchestration above the phase-level ABI: one whole prove_fast (or its prove core) per
// call, and a serving loop that runs many such requests concurrently on one GPU.
//
// prove_fast for a registry with one boolean hash type (crates/flock-prover/src/prover.rs:
// 831-1468): batch-major witness, dense stack, commit, bind_statement of the cap, zerocheck
// (capturing s_hat_v_c), x_ab split, the union lincheck, then s_hat_v_ab and the merged PCS
// opening, assembled into the bincode proof. The prove core stops before the opening.
//
// Serving: each worker owns a prover (context = HIP stream + grow-only device buffers) on its
// own host thread. Requests are taken in arrival order; one is admitted when its arrival time
// has passed and its device footprint fits the budget beside what the other workers hold (a
// request is always admitted when nothing is in flight). Workers never share a stream, so
// one request's latency-bound rounds (host transcript steps between kernels) overlap the
// others' kernels. Every request's transcript is independent, so a result does not depend
// on what ran beside it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/flock_hip.h"
#include "flock_internal.h"

namespace {

struct HashShape {
    uint32_t k_log;
    uint64_t useful_bits;
    size_t input_bytes;  // per instance
};

// K_LOG / USEFUL_BITS of the three circuits: r1cs_hashes/sha2.rs:189-194 (2^15, 25,470),
// blake3.rs:246-247 (2^14, 11,707), keccak.rs:436-451 (2^16, 42,560).
bool shape_of(uint32_t kind, HashShape* s) {
    switch (kind) {
        case FLOCK_HASH_SHA256: *s = {15, 25470, 24 * 4}; return true;
        case FLOCK_HASH_BLAKE3: *s = {14, 11707, 28 * 4}; return true;
        case FLOCK_HASH_KECCAK: *s = {16, 42560, 25 * 8}; return true;
        default: return false;
    }
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(flock_hip_ctx* c, size_t n) {
        if (bytes >= n) return 0;
        if (p) flock_hip_dfree(c, p);
        p = nullptr;
        bytes = 0;
        if (int rc = flock_hip_malloc(c, &p, n)) return rc;
        bytes = n;
        return 0;
    }
    void release(flock_hip_ctx* c) {
        if (p) flock_hip_dfree(c, p);
        p = nullptr;
        bytes = 0;
    }
};

enum { B_IN, B_Z, B_A, B_B, B_STRIPE, B_Q, B_CW, B_TREE, B_ZVEC, B_COUNT };

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// One boolean hash type of a request and the request as a registry of such types (a single
// prove_fast request is a registry of one).
struct Part {
    uint32_t kind;
    uint64_t n_decl;
    const void* inputs;
    const flock_csc* circuit;
};
struct Job {
    uint32_t nu;
    std::vector<Part> parts;
    const uint8_t* digest;
    const uint8_t* domain;
    size_t domain_len;
    bool on_device;
    uint64_t min_decl;  // 1 for a single-type request, 0 for a part of a union
    bool need_inputs = true;  // false for a verifier, which reads no inputs
};

Job job_of(const flock_prove_request* r) {
    return Job{r->n_log, {{r->kind, r->n_decl, r->h_inputs, r->circuit}}, r->digest, r->domain,
               r->domain_len, r->inputs_on_device != 0, 1};
}

Job job_of(const flock_union_request* r) {
    Job j{r->n_log, {}, r->digest, r->domain, r->domain_len, r->inputs_on_device != 0, 0};
    for (uint32_t i = 0; i < r->n_parts; i++)
        j.parts.push_back({r->parts[i].kind, r->parts[i].n_decl, r->parts[i].inputs,
                           r->parts[i].circuit});
    return j;
}

// Sizes derived for one request; an error for an unsupported shape. Slot s holds part
// reg.slots[s].type_index.
struct Plan {
    std::vector<HashShape> hs;  // per part
    flock_registry reg;
    std::vector<uint64_t> counts;  // statement counts, slot order
    flock_union_sizes sz;
    flock_pcs_params pcs;
    int cap_depth;
    bool has_config;
    size_t cw_len, n_leaves;
    uint32_t m_zc;                       // M_bool
    uint64_t in_bytes, stripe_bytes;     // staged inputs, every slot's stripe
    std::vector<uint64_t> in_off, stripe_off;  // per slot
};

int plan_job(const Job& j, Plan* p) {
    const size_t n = j.parts.size();
    if (n == 0 || n > FLOCK_MAX_TYPES || j.nu < 3 || j.nu > 24 || (j.domain_len && !j.domain))
        return FLOCK_HIP_E_INVALID;
    std::vector<flock_table_type> types(n);
    p->hs.resize(n);
    for (size_t i = 0; i < n; i++) {
        const Part& t = j.parts[i];
        if (!shape_of(t.kind, &p->hs[i]) || t.n_decl < j.min_decl || t.n_decl > (1ull << j.nu) ||
            (t.n_decl && !t.inputs && j.need_inputs))
            return FLOCK_HIP_E_INVALID;
        types[i] = {p->hs[i].k_log, 0, p->hs[i].useful_bits};
    }
    if (int rc = flock_registry_new(types.data(), n, j.nu, &p->reg)) return rc;
    // Keccak batches always count the full capacity: padding rows run keccak_f(0)
    // (keccak.rs KeccakSetup); the compression circuits count the declared instances.
    p->counts.resize(n);
    p->in_off.resize(n);
    p->stripe_off.resize(n);
    p->in_bytes = p->stripe_bytes = 0;
    for (size_t s = 0; s < n; s++) {
        const uint32_t ti = p->reg.slots[s].type_index;
        const Part& t = j.parts[ti];
        p->counts[s] = t.kind == FLOCK_HASH_KECCAK ? (1ull << j.nu) : t.n_decl;
        p->in_off[s] = p->in_bytes;
        if (!j.on_device) p->in_bytes += (t.n_decl * p->hs[ti].input_bytes + 255) & ~255ull;
        p->stripe_off[s] = p->stripe_bytes;
        p->stripe_bytes += 1ull << (j.nu + p->hs[ti].k_log - 3);
    }
    if (int rc = flock_union_sizes_compute(&p->reg, p->counts.data(), 0, 6, &p->sz)) return rc;
    // log_batch_size is the Ligerito config's initial_k (ligerito.rs:564-574: 6, but 4 at
    // m = 28 and 5 at m = 29), and commit_lanes depends on it (union.rs:386-390)
    int lbs = flock_lig_initial_k(p->sz.dense_m);
    p->cap_depth = flock_l0_cap_depth(p->sz.dense_m);
    // below the m22 configs there is no opening (the reference's prover_config_for fails,
    // ligerito.rs:590-618); the prove core still runs there and binds the root
    p->has_config = lbs >= 0 && p->cap_depth >= 0;
    if (!p->has_config) {
        lbs = 6;
        p->cap_depth = 0;
    }
    if (lbs != 6)
        if (int rc = flock_union_sizes_compute(&p->reg, p->counts.data(), 0, (uint32_t)lbs,
                                               &p->sz))
            return rc;
    p->pcs = {p->sz.dense_m, 1, (uint32_t)lbs, (uint32_t)p->sz.commit_lanes, 0};
    uint32_t k_code = 0;
    if (int rc = flock_hip_pcs_sizes(&p->pcs, nullptr, &p->cw_len, &p->n_leaves, &k_code, nullptr))
        return rc;
    p->m_zc = p->reg.m_bool;
    return 0;
}

int plan_request(const flock_prove_request* r, Plan* p) {
    if (!r) return FLOCK_HIP_E_INVALID;
    return plan_job(job_of(r), p);
}

// Device bytes of one request: witness (z, a, b over the 2^M_bool-bit boolean region), the
// slots' stripes, dense stack, codeword, tree, the zerocheck driver's scratch (2N + 2h + 2q +
// h + q + N/8 = 4.375 N words for N = 2^(m-6) rows, api.hip flock_hip_zerocheck_prove), the
// lincheck's z_vec, and the opening's workspace (about 9.5 words per committed word: W, q
// copy, the ping-pong halves, the eq bases, and the F256 ladder's four half-size tables of
// 32 B words).
uint64_t plan_footprint(const Plan& p) {
    const uint64_t words = 1ull << (p.m_zc - 7);  // F128 words of z (a, b alike)
    const uint64_t zc_rows = 1ull << (p.m_zc - 6);
    return 3 * 16 * words + p.in_bytes + p.stripe_bytes + 16 * p.sz.committed_words +
           16 * (uint64_t)p.cw_len + 32 * (2 * (uint64_t)p.n_leaves) + 16 * zc_rows * 35 / 8 +
           16 * (1ull << (p.m_zc - p.reg.nu)) + 16 * p.sz.committed_words * 19 / 2;
}

}  // namespace

struct flock_hip_prover {
    flock_hip_ctx* ctx = nullptr;
    DevBuf buf[B_COUNT];
    flock_open_ws* ws = nullptr;  // the opening's workspace, created on first use
    flock_csc circ[2] = {};       // built-in SHA-256 / BLAKE3 circuits on the device, on first use
};

namespace {

// Device bytes a prover holds between requests: its request buffers, the opening's workspace,
// the context's driver scratch and the built-in circuits it uploaded.
uint64_t held_bytes(const flock_hip_prover* p) {
    uint64_t s = 0;
    for (const auto& b : p->buf) s += b.bytes;
    s += flock_open_ws_bytes(p->ws) + flock::ctx_scratch_bytes(p->ctx);
    for (const auto& c : p->circ)
        if (c.a_col_ptr) s += 4 * (2 * ((1ull << c.k_log) + 1) + c.nnz_a + c.nnz_b);
    return s;
}

// Give back everything but the circuits (the next request grows what it needs again).
void release_buffers(flock_hip_prover* p) {
    flock_hip_sync(p->ctx);
    for (auto& b : p->buf) b.release(p->ctx);
    flock_open_ws_destroy(p->ws);
    p->ws = nullptr;
    flock::ctx_release_scratch(p->ctx);
}

}  // namespace

extern "C" {

int flock_hip_prover_create(int device, flock_hip_prover** out) {
    if (!out) return FLOCK_HIP_E_INVALID;
    *out = nullptr;
    flock_hip_prover* p = new (std::nothrow) flock_hip_prover();
    if (!p) return FLOCK_HIP_E_NOMEM;
    if (int rc = flock_hip_ctx_create(device, &p->ctx)) {
        delete p;
        return rc;
    }
    *out = p;
    return 0;
}

void flock_hip_prover_destroy(flock_hip_prover* p) {
    if (!p) return;
    if (p->ctx) {
        flock_hip_sync(p->ctx);
        for (auto& b : p->buf) b.release(p->ctx);
        for (auto& c : p->circ)
            if (c.a_col_ptr) flock_hip_circuit_release(p->ctx, &c);
        flock_open_ws_destroy(p->ws);
        flock_hip_ctx_destroy(p->ctx);
    }
    delete p;
}

flock_hip_ctx* flock_hip_prover_ctx(flock_hip_prover* p) { return p ? p->ctx : nullptr; }

int flock_hip_prove_footprint(const flock_prove_request* req, uint64_t* bytes) {
    Plan p;
    if (!bytes) return FLOCK_HIP_E_INVALID;
    if (int rc = plan_request(req, &p)) return rc;
    *bytes = plan_footprint(p);
    return 0;
}

}  // extern "C"

namespace {

struct BinW {
    std::vector<uint8_t> b;
    void u64(uint64_t v) {
        for (int i = 0; i < 8; i++) b.push_back((uint8_t)(v >> (8 * i)));
    }
    void f(flock_f128 v) {
        u64(v.lo);
        u64(v.hi);
    }
    void vec_f(const flock_f128* v, size_t n) {
        u64(n);
        for (size_t i = 0; i < n; i++) f(v[i]);
    }
    void vec_pairs(const flock_f128* v, size_t n_pairs) {
        u64(n_pairs);
        for (size_t i = 0; i < 2 * n_pairs; i++) f(v[i]);
    }
    void vec_u64(const uint64_t* v, size_t n) {
        u64(n);
        for (size_t i = 0; i < n; i++) u64(v[i]);
    }
};

uint8_t* dup_bytes(const uint8_t* p, size_t n) {
    uint8_t* o = static_cast<uint8_t*>(std::malloc(n ? n : 1));
    if (o && n) std::memcpy(o, p, n);
    return o;
}

// The shared body: the prove core, and when `out` is set the opening and the wire bytes.
int prove_impl(flock_hip_prover* pr, const Job& job, flock_prove_result* res,
               flock_proof_bytes* out) {
    if (!pr || !pr->ctx || !res) return FLOCK_HIP_E_INVALID;
    std::memset(res, 0, sizeof(*res));
    if (out) std::memset(out, 0, sizeof(*out));
    double t_mark = now_s();
    auto stage = [&](int i) {
        const double t = now_s();
        res->stage_ms[i] = (t - t_mark) * 1e3;
        t_mark = t;
    };
    Plan p;
    if (int rc = plan_job(job, &p)) return res->status = rc;
    if (out && !p.has_config) return res->status = FLOCK_HIP_E_INVALID;
    flock_hip_ctx* c = pr->ctx;
    const uint32_t nl = job.nu, n_slots = (uint32_t)job.parts.size();
    const uint32_t col_vars = p.m_zc - nl;
    const uint64_t zwords = 1ull << (p.m_zc - 7);  // F128 words of z (and of a, b)
    int rc = 0;
    const size_t need[B_COUNT] = {(size_t)(p.in_bytes ? p.in_bytes : 16),
                                  (size_t)(16 * zwords),
                                  (size_t)(16 * zwords),
                                  (size_t)(16 * zwords),
                                  (size_t)p.stripe_bytes,
                                  (size_t)(16 * p.sz.committed_words),
                                  (size_t)(16 * p.cw_len),
                                  (size_t)(32 * (2 * p.n_leaves - 1)),
                                  (size_t)(out ? 16ull << col_vars : 16)};
    for (int i = 0; i < B_COUNT && !rc; i++) rc = pr->buf[i].ensure(c, need[i]);
    if (!rc && out && !pr->ws) rc = flock_open_ws_create(&pr->ws);
    // each slot's circuit: the request's, else the built-in one
    std::vector<const flock_csc*> circuit(n_slots, nullptr);
    for (uint32_t s = 0; s < n_slots && !rc; s++) {
        const Part& t = job.parts[p.reg.slots[s].type_index];
        circuit[s] = t.circuit;
        if (!circuit[s] && t.kind != FLOCK_HASH_KECCAK) {
            flock_csc& bc = pr->circ[t.kind];
            if (!bc.a_col_ptr) rc = flock_hip_circuit_upload(c, t.kind, &bc);
            circuit[s] = &bc;
        }
    }
    // the request's registry digest, else the built-in types' Registry::digest at nu
    uint8_t digest[32];
    if (!rc) {
        if (job.digest) {
            std::memcpy(digest, job.digest, 32);
        } else {
            std::vector<uint32_t> kinds(n_slots);
            for (uint32_t i = 0; i < n_slots; i++) kinds[i] = job.parts[i].kind;
            rc = flock_circuit_union_digest(kinds.data(), n_slots, nl, digest);
        }
    }
    if (rc) return res->status = rc;
    auto* d_in = static_cast<uint8_t*>(pr->buf[B_IN].p);
    auto* z = static_cast<flock_f128*>(pr->buf[B_Z].p);
    auto* a = static_cast<flock_f128*>(pr->buf[B_A].p);
    auto* b = static_cast<flock_f128*>(pr->buf[B_B].p);
    auto* stripe = static_cast<uint8_t*>(pr->buf[B_STRIPE].p);
    auto* q = static_cast<flock_f128*>(pr->buf[B_Q].p);
    auto* cw = static_cast<flock_f128*>(pr->buf[B_CW].p);
    auto* tree = static_cast<uint8_t*>(pr->buf[B_TREE].p);
    auto* zvec = out ? static_cast<flock_f128*>(pr->buf[B_ZVEC].p) : nullptr;
    hipStream_t st = static_cast<hipStream_t>(flock_hip_ctx_stream(c));
    // 1. witness (generate_witness_batch_major{,_partial}) of each slot at its offset in the
    // packed boolean region; the columns no slot covers are zero
    {
        std::vector<std::pair<uint64_t, uint64_t>> spans;  // in F128 words
        for (uint32_t s = 0; s < n_slots; s++) {
            const uint64_t w0 = p.reg.slots[s].offset >> 7;
            spans.emplace_back(w0, w0 + (1ull << (p.reg.slots[s].m_slot - 7)));
        }
        std::sort(spans.begin(), spans.end());
        uint64_t at = 0;
        for (size_t i = 0; i <= spans.size() && !rc; i++) {
            const uint64_t end = i < spans.size() ? spans[i].first : zwords;
            if (end > at)
                for (flock_f128* v : {z, a, b})
                    if (flock::launch_zero_words(reinterpret_cast<uint4*>(v + at), end - at, st) !=
                        hipSuccess)
                        rc = FLOCK_HIP_E_HIP;
            if (i < spans.size()) at = std::max(at, spans[i].second);
        }
    }
    for (uint32_t s = 0; s < n_slots && !rc; s++) {
        const Part& t = job.parts[p.reg.slots[s].type_index];
        const HashShape& hs = p.hs[p.reg.slots[s].type_index];
        const void* in = t.inputs;
        if (!job.on_device && t.n_decl) {
            in = d_in + p.in_off[s];
            if (hipMemcpyAsync(d_in + p.in_off[s], t.inputs, t.n_decl * hs.input_bytes,
                               hipMemcpyHostToDevice, st) != hipSuccess) {
                rc = FLOCK_HIP_E_HIP;
                break;
            }
        }
        const uint64_t w0 = p.reg.slots[s].offset >> 7;
        uint8_t* sp = stripe + p.stripe_off[s];
        switch (t.kind) {
            case FLOCK_HASH_SHA256:
                rc = flock_hip_witness_sha256(c, static_cast<const uint32_t*>(in), t.n_decl, nl,
                                              z + w0, a + w0, b + w0, sp);
                break;
            case FLOCK_HASH_BLAKE3:
                rc = flock_hip_witness_blake3(c, static_cast<const uint32_t*>(in), t.n_decl, nl,
                                              z + w0, a + w0, b + w0, sp);
                break;
            default:
                rc = flock_hip_witness_keccak(c, static_cast<const uint64_t*>(in), t.n_decl, nl,
                                              z + w0, a + w0, b + w0, sp);
        }
    }
    // 2. dense stack and 3. commit (prover.rs:979-983): lane-major for an integer lane count,
    // the interleaved commit at full width; the cap at the L0 cap depth
    std::vector<uint8_t> cap(32ull << p.cap_depth);
    if (!rc) rc = flock_hip_union_compact(c, &p.reg, p.counts.data(), 0, z, q);
    if (!rc) {
        if (p.pcs.num_lanes)
            rc = flock_hip_commit_lane_major(c, q, p.sz.committed_words, &p.pcs, cw, tree,
                                             (uint32_t)p.cap_depth, cap.data());
        else
            rc = flock_hip_commit(c, q, p.sz.committed_words, &p.pcs, cw, tree,
                                  (uint32_t)p.cap_depth, cap.data());
    }
    if (!rc && hipMemcpyAsync(res->root, tree + 32 * (2 * p.n_leaves - 2), 32, hipMemcpyDeviceToHost,
                              st) != hipSuccess)
        rc = FLOCK_HIP_E_HIP;
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = FLOCK_HIP_E_HIP;
    if (rc) return res->status = rc;
    stage(0);
    // 4. statement binding (union.rs:619-628), zerocheck with s_hat_v_c (prover.rs:1094-1103)
    flock_fs* fs = flock_fs_new(job.domain, job.domain_len);
    if (!fs) return res->status = FLOCK_HIP_E_NOMEM;
    const uint32_t n_mlv = p.m_zc - 6;
    std::vector<flock_f128> r1(128), ml(2 * n_mlv), claim(1 + 2 * n_mlv), shv(2 * 128);
    std::vector<uint64_t> nonces(2 + n_mlv);
    flock_f128 finals[3];
    rc = flock_union_bind_statement(fs, digest, p.counts.data(), n_slots, cap.data(), cap.size());
    if (!rc)
        rc = flock_hip_zerocheck_prove(c, fs, reinterpret_cast<const uint8_t*>(a),
                                       reinterpret_cast<const uint8_t*>(b),
                                       reinterpret_cast<const uint8_t*>(z), p.m_zc, 1, r1.data(),
                                       ml.data(), finals, nonces.data(), claim.data(),
                                       shv.data() + 128);
    stage(1);
    // 5. x_ab and the union lincheck over the slots (prover.rs:1140-1160)
    const uint32_t L = col_vars - 6;
    std::vector<flock_lc_slot> lc(n_slots);
    uint32_t n_pins = 0;
    for (uint32_t s = 0; s < n_slots; s++) {
        const HashShape& hs = p.hs[p.reg.slots[s].type_index];
        lc[s] = {circuit[s], stripe + p.stripe_off[s], p.reg.slots[s].prefix, hs.useful_bits,
                 p.counts[s]};
        n_pins += circuit[s] ? (circuit[s]->const_pin >= 0) : 1;
    }
    std::vector<flock_f128> x_rest(L ? L : 1), x_outer(nl), rounds(2 * (L ? L : 1)),
        lclaim(2 + L), mev(2 * n_slots);
    std::vector<uint64_t> lnonces(L + 2 + n_pins);
    flock_f128 zp[64];
    if (!rc) rc = flock_union_x_ab_from_mlv(&p.reg, claim.data() + 1, n_mlv, x_rest.data(),
                                            x_outer.data());
    if (!rc)
        rc = flock_hip_lincheck_prove_union(c, fs, lc.data(), n_slots, col_vars, nl, claim[0],
                                            x_rest.data(), x_outer.data(), 1, rounds.data(), zp,
                                            mev.data(), lnonces.data(), lclaim.data(), zvec);
    stage(2);
    if (!rc && out) {
        // 6. s_hat_v_ab from the lincheck's z_vec (prover.rs:1186-1200), the claim points
        // (union.rs:564-600 as quirky x_outer_full, prover.rs:43-48) and the jagged heights
        // (union.rs:293-310: each slot's used columns at its count); then open_batch_merged
        // (prover.rs:1412-1426)
        const flock_f128* r_inner_rest = lclaim.data() + 2;
        rc = flock_hip_s_hat_v_from_z_vec(c, pr->ws, zvec, 1ull << col_vars, r_inner_rest + 1,
                                          L - 1, shv.data());
        const size_t x_len = p.m_zc - 6;
        std::vector<flock_f128> xs(2 * x_len);
        xs[0] = r_inner_rest[0];
        for (uint32_t i = 0; i < nl; i++) xs[1 + i] = x_outer[i];
        for (uint32_t i = 1; i < L; i++) xs[nl + i] = r_inner_rest[i];
        for (size_t i = 0; i < x_len; i++) xs[x_len + i] = claim[1 + n_mlv + i];
        std::vector<uint64_t> heights(1ull << (col_vars - 7), 0);
        for (uint32_t s = 0; s < n_slots; s++) {
            const HashShape& hs = p.hs[p.reg.slots[s].type_index];
            const uint64_t used =
                std::min<uint64_t>((hs.useful_bits + 127) / 128, 1ull << (hs.k_log - 7));
            const uint64_t c_off = p.reg.slots[s].offset >> (7 + nl);
            for (uint64_t y = 0; y < used; y++) heights[c_off + y] = p.counts[s];
        }
        uint8_t* ob = nullptr;
        size_t on = 0;
        if (!rc)
            rc = flock_hip_open_merged(c, pr->ws, fs, q, &p.pcs, cw, tree, 2, xs.data(), x_len,
                                       shv.data(), heights.data(), heights.size(), nl, &ob, &on,
                                       nullptr);
        if (!rc) {
            // R1csProofMergedLigerito (proof.rs:42-47): ZerocheckProof (zerocheck.rs:427-447),
            // LincheckProof (lincheck.rs:520-549, one matrix-eval pair per boolean type),
            // MergedOpenProof
            BinW w;
            w.vec_f(r1.data(), 64);
            w.vec_f(r1.data() + 64, 64);
            w.vec_pairs(ml.data(), n_mlv);
            for (int i = 0; i < 3; i++) w.f(finals[i]);
            w.vec_u64(nonces.data(), nonces.size());
            w.vec_pairs(rounds.data(), L);
            w.vec_f(zp, 64);
            w.vec_pairs(mev.data(), n_slots);
            w.vec_u64(lnonces.data(), lnonces.size());
            w.b.insert(w.b.end(), ob, ob + on);
            out->proof = dup_bytes(w.b.data(), w.b.size());
            out->proof_len = w.b.size();
            out->cap = dup_bytes(cap.data(), cap.size());
            out->cap_len = cap.size();
            out->claim_ab = lclaim[1];
            out->claim_c = finals[2];
            if (!out->proof || !out->cap) rc = FLOCK_HIP_E_NOMEM;
        }
        flock_hip_free(ob);
        stage(3);
    }
    if (!rc) {
        flock_fs_digest(fs, res->transcript);
        res->zc_z = claim[0];
        res->lc_w = lclaim[1];
        for (const Part& t : job.parts) res->hashes += t.n_decl;
    }
    flock_fs_free(fs);
    return res->status = rc;
}

}  // namespace

extern "C" {

flock_open_ws* flock_hip_prover_open_ws(flock_hip_prover* pr) { return pr ? pr->ws : nullptr; }

uint64_t flock_hip_prover_held_bytes(const flock_hip_prover* pr) { return pr ? held_bytes(pr) : 0; }

void flock_hip_prover_release(flock_hip_prover* pr) {
    if (pr && pr->ctx) release_buffers(pr);
}

int flock_hip_prove_core(flock_hip_prover* pr, const flock_prove_request* req,
                         flock_prove_result* res) {
    if (!req) return FLOCK_HIP_E_INVALID;
    return prove_impl(pr, job_of(req), res, nullptr);
}

int flock_hip_prove_fast(flock_hip_prover* pr, const flock_prove_request* req,
                         flock_prove_result* res, flock_proof_bytes* out) {
    if (!out || !req) return FLOCK_HIP_E_INVALID;
    return prove_impl(pr, job_of(req), res, out);
}

int flock_union_request_info(const flock_union_request* req, flock_union_info* out) {
    if (!req || !out || !req->parts || req->n_parts == 0 || req->n_parts > FLOCK_MAX_TYPES)
        return FLOCK_HIP_E_INVALID;
    const Job job = job_of(req);
    Plan p;
    if (int rc = plan_job(job, &p)) return rc;
    std::memset(out, 0, sizeof(*out));
    out->n_slots = req->n_parts;
    out->m_bool = p.m_zc;
    out->dense_m = p.sz.dense_m;
    out->log_batch_size = p.pcs.log_batch_size;
    out->num_lanes = p.pcs.num_lanes;
    out->has_opening = p.has_config ? 1 : 0;
    out->device_bytes = plan_footprint(p);
    for (uint32_t s = 0; s < req->n_parts; s++) {
        out->slot_part[s] = p.reg.slots[s].type_index;
        out->counts[s] = p.counts[s];
    }
    return 0;
}

int flock_hip_prove_fast_union(flock_hip_prover* pr, const flock_union_request* req,
                               flock_prove_result* res, flock_proof_bytes* out) {
    if (!out || !req || !req->parts || req->n_parts == 0 || req->n_parts > FLOCK_MAX_TYPES)
        return FLOCK_HIP_E_INVALID;
    return prove_impl(pr, job_of(req), res, out);
}

int flock_hip_serve(flock_hip_prover* const* provers, uint32_t workers,
                    const flock_prove_request* reqs, size_t n, uint64_t hbm_budget,
                    flock_prove_result* res, flock_proof_bytes* proofs, double* wall_ms) {
    if (!provers || workers == 0 || (n && (!reqs || !res))) return FLOCK_HIP_E_INVALID;
    for (uint32_t w = 0; w < workers; w++)
        if (!provers[w] || !provers[w]->ctx) return FLOCK_HIP_E_INVALID;
    std::vector<uint64_t> fp(n);
    for (size_t i = 0; i < n; i++) {
        Plan p;
        if (int rc = plan_request(&reqs[i], &p)) return rc;
        if (i && reqs[i].arrival_s < reqs[i - 1].arrival_s) return FLOCK_HIP_E_INVALID;
        fp[i] = plan_footprint(p);
    }
    // Admission counts what every worker holds: a running worker reserves max(held, its
    // request's footprint), an idle one what its grow-only buffers keep. A request that does
    // not fit makes its worker give back its own buffers when that suffices; with nothing
    // running it frees the idle workers' buffers and runs alone.
    std::vector<uint64_t> reserved(workers);
    for (uint32_t w = 0; w < workers; w++) reserved[w] = held_bytes(provers[w]);
    std::vector<char> busy(workers, 0);
    std::mutex mu;
    std::condition_variable cv;
    size_t next = 0;
    uint32_t running = 0;
    int first_err = 0;
    const double t0 = now_s();
    auto worker = [&](uint32_t w) {
        flock_hip_prover* pr = provers[w];
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            if (next >= n || first_err) return;
            const size_t i = next;
            const double due = t0 + reqs[i].arrival_s;
            const double t = now_s();
            if (t < due) {
                cv.wait_for(lk, std::chrono::duration<double>(due - t));
                continue;
            }
            uint64_t mine = held_bytes(pr);
            // a due request goes to an idle worker whose buffers already hold it when there is
            // one: growing buffers (hipMalloc, and hipFree's device-wide wait) can stall a large
            // request for hundreds of milliseconds
            if (mine < fp[i]) {
                bool other_fits = false;
                for (uint32_t v = 0; v < workers && !other_fits; v++)
                    other_fits = v != w && !busy[v] && held_bytes(provers[v]) >= fp[i];
                if (other_fits) {
                    cv.notify_all();
                    cv.wait(lk);
                    continue;
                }
            }
            uint64_t others = 0;
            for (uint32_t v = 0; v < workers; v++)
                if (v != w) others += reserved[v];
            if (others + std::max(mine, fp[i]) > hbm_budget) {
                if (running == 0) {
                    for (uint32_t v = 0; v < workers; v++)
                        if (v != w && reserved[v]) {
                            release_buffers(provers[v]);
                            reserved[v] = held_bytes(provers[v]);
                        }
                    if (mine > fp[i]) release_buffers(pr);
                } else if (others + fp[i] <= hbm_budget && mine > fp[i]) {
                    release_buffers(pr);
                } else {
                    cv.wait(lk);
                    continue;
                }
                mine = held_bytes(pr);
            }
            next++;
            reserved[w] = std::max(mine, fp[i]);
            running++;
            busy[w] = 1;
            cv.notify_all();  // the idle workers look at the next request
            lk.unlock();
            const double start = now_s();
            const int rc = proofs ? flock_hip_prove_fast(pr, &reqs[i], &res[i], &proofs[i])
                                  : flock_hip_prove_core(pr, &reqs[i], &res[i]);
            const double end = now_s();
            res[i].queued_ms = (start - due) * 1e3;
            res[i].run_ms = (end - start) * 1e3;
            res[i].done_s = end - t0;
            lk.lock();
            reserved[w] = held_bytes(pr);
            running--;
            busy[w] = 0;
            if (rc && !first_err) first_err = rc;
            cv.notify_all();
        }
    };
    std::vector<std::thread> th;
    th.reserve(workers);
    for (uint32_t w = 0; w < workers; w++) th.emplace_back(worker, w);
    for (auto& x : th) x.join();
    if (wall_ms) *wall_ms = (now_s() - t0) * 1e3;
    return first_err;
}

}  // extern "C"

namespace {

// bincode 1.3 reader over the proof bytes; an overrun or an impossible length marks it bad
struct BinR {
    const uint8_t* p;
    size_t n, off = 0;
    bool ok = true;
    uint64_t u64() {
        if (!ok || n - off < 8) {
            ok = false;
            return 0;
        }
        uint64_t v = 0;
        for (int i = 0; i < 8; i++) v |= (uint64_t)p[off + i] << (8 * i);
        off += 8;
        return v;
    }
    flock_f128 f() {
        flock_f128 v;
        v.lo = u64();
        v.hi = u64();
        return v;
    }
    // a vector of `count` F128 words, its length prefix checked against `expect`
    bool vec_f(std::vector<flock_f128>& v, size_t words_per, size_t expect) {
        const uint64_t k = u64();
        if (!ok || k != expect) return ok = false;
        v.resize(k * words_per);
        for (auto& x : v) x = f();
        return ok;
    }
    bool vec_u64(std::vector<uint64_t>& v, size_t expect) {
        const uint64_t k = u64();
        if (!ok || k != expect) return ok = false;
        v.resize(k);
        for (auto& x : v) x = u64();
        return ok;
    }
};

int verify_impl(const Job& job, const uint8_t* cap, size_t cap_len, const uint8_t* proof,
                size_t proof_len, flock_f128* claim_ab, flock_f128* claim_c) {
    auto reject = [](const char* w) {
        return flock::set_error(FLOCK_HIP_E_REJECT, (std::string("proof rejected: ") + w).c_str());
    };
    Plan p;
    if (!cap || !proof) return flock::set_error(FLOCK_HIP_E_INVALID, "verify: null cap or proof");
    for (const Part& t : job.parts)
        if (t.circuit)
            return flock::set_error(FLOCK_HIP_E_INVALID, "verify: the built-in circuits only (circuit "
                                                         "must be NULL)");
    if (int rc = plan_job(job, &p)) return flock::set_error(rc, "verify: bad request shape");
    if (!p.has_config) return flock::set_error(FLOCK_HIP_E_INVALID, "verify: no Ligerito config");
    if (cap_len != (32ull << p.cap_depth)) return reject("commitment cap size");
    const uint32_t nl = job.nu, n_slots = (uint32_t)job.parts.size();
    const uint32_t m = p.m_zc, n_mlv = m - 6, col_vars = m - nl, L = col_vars - 6;
    // host circuits of the slots
    std::vector<flock_csc> host(n_slots);
    std::vector<flock_lc_vslot> vs(n_slots);
    uint32_t n_pins = 0;
    for (uint32_t s = 0; s < n_slots; s++) {
        const Part& t = job.parts[p.reg.slots[s].type_index];
        vs[s] = {nullptr, p.hs[p.reg.slots[s].type_index].k_log, 0, -1, p.reg.slots[s].prefix,
                 p.counts[s]};
        if (t.kind == FLOCK_HASH_KECCAK) {
            vs[s].walker = 1;
        } else {
            const flock_circuit* c = nullptr;
            uint32_t k_log = 0;
            uint64_t ub = 0, pin = 0;
            flock_csc& h = host[s];
            std::memset(&h, 0, sizeof(h));
            if (int rc = flock_circuit_get(t.kind, &c)) return rc;
            if (int rc = flock_circuit_info(c, &k_log, &ub, &pin, &h.nnz_a, &h.nnz_b)) return rc;
            if (int rc = flock_circuit_csc(c, 0, &h.a_col_ptr, &h.a_rows)) return rc;
            if (int rc = flock_circuit_csc(c, 1, &h.b_col_ptr, &h.b_rows)) return rc;
            h.k_log = k_log;
            h.const_pin = (int64_t)pin;
            vs[s].h_circuit = &h;
        }
        n_pins++;  // every built-in type carries a constant-wire pin
    }
    // R1csProofMergedLigerito (proof.rs:42-47): ZerocheckProof, LincheckProof, MergedOpenProof
    BinR r{proof, proof_len};
    std::vector<flock_f128> r1a, r1c, ml, rounds, zp, mev;
    std::vector<uint64_t> zn, ln;
    flock_f128 fin[3];
    r.vec_f(r1a, 1, 64);
    r.vec_f(r1c, 1, 64);
    r.vec_f(ml, 2, n_mlv);
    for (auto& x : fin) x = r.f();
    r.vec_u64(zn, 2 + n_mlv);
    r.vec_f(rounds, 2, L);
    r.vec_f(zp, 1, 64);
    r.vec_f(mev, 2, n_slots);
    r.vec_u64(ln, L + 2 + n_pins);
    if (!r.ok) return reject("malformed zerocheck / lincheck bytes");
    uint8_t digest[32];
    if (job.digest) {
        std::memcpy(digest, job.digest, 32);
    } else {
        std::vector<uint32_t> kinds(n_slots);
        for (uint32_t i = 0; i < n_slots; i++) kinds[i] = job.parts[i].kind;
        if (int rc = flock_circuit_union_digest(kinds.data(), n_slots, nl, digest)) return rc;
    }
    flock_fs* fs = flock_fs_new(job.domain, job.domain_len);
    if (!fs) return flock::set_error(FLOCK_HIP_E_NOMEM, "transcript");
    struct FsGuard {
        flock_fs* f;
        ~FsGuard() { flock_fs_free(f); }
    } guard{fs};
    // statement binding, zerocheck over M_bool (verify_union_piops, verifier.rs:806-870)
    if (int rc = flock_union_bind_statement(fs, digest, p.counts.data(), n_slots, cap, cap_len)) return rc;
    std::vector<flock_f128> r1(128), claim(1 + 2 * n_mlv);
    std::copy(r1a.begin(), r1a.end(), r1.begin());
    std::copy(r1c.begin(), r1c.end(), r1.begin() + 64);
    if (int rc = flock_zerocheck_verify(fs, m, 1, r1.data(), ml.data(), fin, zn.data(), claim.data()))
        return rc;
    // x_ab and the union lincheck with the O(nnz) matrix check (MatrixAssertion::check)
    std::vector<flock_f128> x_rest(L ? L : 1), x_outer(nl), lclaim(2 + L);
    if (int rc = flock_union_x_ab_from_mlv(&p.reg, claim.data() + 1, n_mlv, x_rest.data(), x_outer.data()))
        return rc;
    if (int rc = flock_lincheck_verify_union(fs, vs.data(), n_slots, col_vars, nl, claim[0], x_rest.data(),
                                             x_outer.data(), fin[0], fin[1], 1, rounds.data(), zp.data(),
                                             mev.data(), ln.data(), lclaim.data()))
        return rc;
    // the merged opening of (ab, c) at their claim points (verifier.rs:604-650)
    const flock_f128* rir = lclaim.data() + 2;
    const size_t x_len = m - 6;
    std::vector<flock_f128> xs(2 * x_len), values = {lclaim[1], fin[2]}, skips = {lclaim[0], claim[0]};
    xs[0] = rir[0];
    for (uint32_t i = 0; i < nl; i++) xs[1 + i] = x_outer[i];
    for (uint32_t i = 1; i < L; i++) xs[nl + i] = rir[i];
    for (size_t i = 0; i < x_len; i++) xs[x_len + i] = claim[1 + n_mlv + i];
    std::vector<uint64_t> heights(1ull << (col_vars - 7), 0);
    for (uint32_t s = 0; s < n_slots; s++) {
        const HashShape& hs = p.hs[p.reg.slots[s].type_index];
        const uint64_t used = std::min<uint64_t>((hs.useful_bits + 127) / 128, 1ull << (hs.k_log - 7));
        const uint64_t c_off = p.reg.slots[s].offset >> (7 + nl);
        for (uint64_t y = 0; y < used; y++) heights[c_off + y] = p.counts[s];
    }
    size_t used = 0;
    if (int rc = flock_open_merged_verify(fs, &p.pcs, cap, cap_len, 2, values.data(), skips.data(), xs.data(),
                                          x_len, heights.data(), heights.size(), nl, proof + r.off,
                                          proof_len - r.off, &used))
        return rc;
    if (r.off + used != proof_len) return reject("trailing bytes after the opening");
    if (claim_ab) *claim_ab = lclaim[1];
    if (claim_c) *claim_c = fin[2];
    return 0;
}

}  // namespace

extern "C" {

int flock_verify_fast(const flock_prove_request* req, const uint8_t* cap, size_t cap_len,
                      const uint8_t* proof, size_t proof_len, flock_f128* claim_ab, flock_f128* claim_c) {
    if (!req || req->n_decl == 0) return flock::set_error(FLOCK_HIP_E_INVALID, "verify: bad request");
    Job j = job_of(req);
    j.need_inputs = false;
    if (req->circuit)
        return flock::set_error(FLOCK_HIP_E_INVALID, "verify: the built-in circuits only (circuit "
                                                     "must be NULL)");
    return verify_impl(j, cap, cap_len, proof, proof_len, claim_ab, claim_c);
}

int flock_verify_fast_union(const flock_union_request* req, const uint8_t* cap, size_t cap_len,
                            const uint8_t* proof, size_t proof_len, flock_f128* claim_ab,
                            flock_f128* claim_c) {
    if (!req || !req->parts || req->n_parts == 0 || req->n_parts > FLOCK_MAX_TYPES)
        return flock::set_error(FLOCK_HIP_E_INVALID, "verify: bad request");
    Job j = job_of(req);
    j.need_inputs = false;
    return verify_impl(j, cap, cap_len, proof, proof_len, claim_ab, claim_c);
}

}  // extern "C"
