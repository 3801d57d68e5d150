// Host orchestration above the phase-level ABI: one prove core per call, and a serving loop
// that runs many such requests concurrently on one GPU.
//
// The prove core is the GPU part of prove_fast for a registry with one boolean hash type
// (crates/flock-prover/src/prover.rs:926-1160): batch-major witness, dense stack, lane-major
// commit, bind_statement, zerocheck (capturing s_hat_v_c), x_ab split and the union lincheck.
// It stops before the PCS opening, so the output is the commitment root and the transcript
// state (which has absorbed every message of the two sumcheck phases), not a proof.
//
// Serving: each worker owns a prover (context = HIP stream + grow-only device buffers) on its
// own host thread. Requests are taken in arrival order; one is admitted when its arrival time
// has passed and its device footprint fits the budget beside the requests in flight (a
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
#include <thread>
#include <vector>

#include "../../include/flock_hip.h"

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

enum { B_IN, B_Z, B_A, B_B, B_STRIPE, B_Q, B_CW, B_TREE, B_COUNT };

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Sizes derived for one request; false for an unsupported shape.
struct Plan {
    HashShape hs;
    flock_registry reg;
    flock_union_sizes sz;
    flock_pcs_params pcs;
    size_t cw_len, n_leaves;
    uint64_t count, n_rows;
    uint32_t m_zc;
};

int plan_request(const flock_prove_request* r, Plan* p) {
    if (!r || !shape_of(r->kind, &p->hs) || r->n_log < 3 || r->n_log > 24 || r->n_decl == 0 ||
        r->n_decl > (1ull << r->n_log) || !r->h_inputs || !r->digest ||
        (r->kind != FLOCK_HASH_KECCAK && !r->circuit) || (r->domain_len && !r->domain))
        return FLOCK_HIP_E_INVALID;
    const flock_table_type t = {p->hs.k_log, 0, p->hs.useful_bits};
    if (int rc = flock_registry_new(&t, 1, r->n_log, &p->reg)) return rc;
    // Keccak batches always count the full capacity: padding rows run keccak_f(0)
    // (keccak.rs KeccakSetup); the compression circuits count the declared instances.
    p->count = r->kind == FLOCK_HASH_KECCAK ? (1ull << r->n_log) : r->n_decl;
    p->n_rows = p->count;
    if (int rc = flock_union_sizes_compute(&p->reg, &p->count, 0, 6, &p->sz)) return rc;
    p->pcs = {p->sz.dense_m, 1, 6, (uint32_t)p->sz.commit_lanes, 0};
    uint32_t k_code = 0;
    if (int rc = flock_hip_pcs_sizes(&p->pcs, nullptr, &p->cw_len, &p->n_leaves, &k_code, nullptr))
        return rc;
    p->m_zc = r->n_log + p->hs.k_log;
    return 0;
}

// Device bytes of one request: witness, stripe, dense stack, codeword, tree, and the
// zerocheck scratch (about 3.9 x 2^(m-6) words), the largest driver scratch.
uint64_t plan_footprint(const Plan& p, uint32_t n_log) {
    const uint64_t words = 1ull << (n_log + p.hs.k_log - 7);  // F128 words of z (a, b alike)
    return 3 * 16 * words + 16 * words /* stripe */ + 16 * p.sz.committed_words +
           16 * (uint64_t)p.cw_len + 32 * (2 * (uint64_t)p.n_leaves) +
           (uint64_t)(3.9 * 16.0 * (double)(1ull << (p.m_zc - 6)));
}

}  // namespace

struct flock_hip_prover {
    flock_hip_ctx* ctx = nullptr;
    DevBuf buf[B_COUNT];
};

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
        flock_hip_ctx_destroy(p->ctx);
    }
    delete p;
}

flock_hip_ctx* flock_hip_prover_ctx(flock_hip_prover* p) { return p ? p->ctx : nullptr; }

int flock_hip_prove_footprint(const flock_prove_request* req, uint64_t* bytes) {
    Plan p;
    if (!bytes) return FLOCK_HIP_E_INVALID;
    if (int rc = plan_request(req, &p)) return rc;
    *bytes = plan_footprint(p, req->n_log);
    return 0;
}

int flock_hip_prove_core(flock_hip_prover* pr, const flock_prove_request* req,
                         flock_prove_result* res) {
    if (!pr || !pr->ctx || !res) return FLOCK_HIP_E_INVALID;
    std::memset(res, 0, sizeof(*res));
    Plan p;
    if (int rc = plan_request(req, &p)) return res->status = rc;
    flock_hip_ctx* c = pr->ctx;
    const uint32_t nl = req->n_log, kl = p.hs.k_log;
    const uint64_t n = 1ull << nl;
    const uint64_t zwords = n << (kl - 7);      // F128 words of z (and of a, b)
    const uint64_t stripe_bytes = n << (kl - 3);
    int rc = 0;
    const size_t need[B_COUNT] = {(size_t)(req->n_decl * p.hs.input_bytes),
                                  (size_t)(16 * zwords),
                                  (size_t)(16 * zwords),
                                  (size_t)(16 * zwords),
                                  (size_t)stripe_bytes,
                                  (size_t)(16 * p.sz.committed_words),
                                  (size_t)(16 * p.cw_len),
                                  (size_t)(32 * (2 * p.n_leaves - 1))};
    for (int i = 0; i < B_COUNT && !rc; i++) rc = pr->buf[i].ensure(c, need[i]);
    if (rc) return res->status = rc;
    void* d_in = pr->buf[B_IN].p;
    auto* z = static_cast<flock_f128*>(pr->buf[B_Z].p);
    auto* a = static_cast<flock_f128*>(pr->buf[B_A].p);
    auto* b = static_cast<flock_f128*>(pr->buf[B_B].p);
    auto* stripe = static_cast<uint8_t*>(pr->buf[B_STRIPE].p);
    auto* q = static_cast<flock_f128*>(pr->buf[B_Q].p);
    auto* cw = static_cast<flock_f128*>(pr->buf[B_CW].p);
    auto* tree = static_cast<uint8_t*>(pr->buf[B_TREE].p);
    hipStream_t st = static_cast<hipStream_t>(flock_hip_ctx_stream(c));
    if (hipMemcpyAsync(d_in, req->h_inputs, need[B_IN], hipMemcpyHostToDevice, st) !=
        hipSuccess)
        return res->status = FLOCK_HIP_E_HIP;
    // 1. witness (generate_witness_batch_major{,_partial})
    switch (req->kind) {
        case FLOCK_HASH_SHA256:
            rc = flock_hip_witness_sha256(c, static_cast<const uint32_t*>(d_in), req->n_decl, nl,
                                          z, a, b, stripe);
            break;
        case FLOCK_HASH_BLAKE3:
            rc = flock_hip_witness_blake3(c, static_cast<const uint32_t*>(d_in), req->n_decl, nl,
                                          z, a, b, stripe);
            break;
        default:
            rc = flock_hip_witness_keccak(c, static_cast<const uint64_t*>(d_in), req->n_decl, nl,
                                          z, a, b, stripe);
    }
    // 2. dense stack and 3. commit (prover.rs:979-983)
    if (!rc) rc = flock_hip_union_compact(c, &p.reg, &p.count, 0, z, q);
    if (!rc) rc = flock_hip_commit_lane_major(c, q, p.sz.committed_words, &p.pcs, cw, tree, 0,
                                              res->root);
    if (rc) return res->status = rc;
    // 4. statement binding, zerocheck with s_hat_v_c capture (prover.rs:1094-1103)
    flock_fs* fs = flock_fs_new(req->domain, req->domain_len);
    if (!fs) return res->status = FLOCK_HIP_E_NOMEM;
    const uint32_t n_mlv = p.m_zc - 6;
    std::vector<flock_f128> r1(128), ml(2 * n_mlv), claim(1 + 2 * n_mlv), shv(128);
    std::vector<uint64_t> nonces(2 + n_mlv);
    flock_f128 finals[3];
    rc = flock_union_bind_statement(fs, req->digest, &p.count, 1, res->root, 32);
    if (!rc)
        rc = flock_hip_zerocheck_prove(c, fs, reinterpret_cast<const uint8_t*>(a),
                                       reinterpret_cast<const uint8_t*>(b),
                                       reinterpret_cast<const uint8_t*>(z), p.m_zc, 1, r1.data(),
                                       ml.data(), finals, nonces.data(), claim.data(), shv.data());
    // 5. x_ab and the union lincheck (prover.rs:1140-1160)
    const uint32_t L = kl - 6;
    std::vector<flock_f128> x_rest(L ? L : 1), x_outer(nl), rounds(2 * (L ? L : 1)),
        lclaim(2 + L);
    std::vector<uint64_t> lnonces(L + 3);
    flock_f128 zp[64], mev[2];
    if (!rc) rc = flock_union_x_ab_from_mlv(&p.reg, claim.data() + 1, n_mlv, x_rest.data(),
                                            x_outer.data());
    if (!rc) {
        if (req->kind == FLOCK_HASH_KECCAK)
            rc = flock_hip_lincheck_prove_keccak(c, fs, stripe, nl, p.n_rows, claim[0],
                                                 x_rest.data(), x_outer.data(), 1, rounds.data(),
                                                 zp, mev, lnonces.data(), lclaim.data(), nullptr);
        else
            rc = flock_hip_lincheck_prove(c, fs, req->circuit, stripe, nl, p.hs.useful_bits,
                                          p.n_rows, claim[0], x_rest.data(), x_outer.data(), 1,
                                          rounds.data(), zp, mev, lnonces.data(), lclaim.data(),
                                          nullptr);
    }
    if (!rc) {
        flock_fs_digest(fs, res->transcript);
        res->zc_z = claim[0];
        res->lc_w = lclaim[1];
        res->hashes = req->n_decl;
    }
    flock_fs_free(fs);
    return res->status = rc;
}

int flock_hip_serve(flock_hip_prover* const* provers, uint32_t workers,
                    const flock_prove_request* reqs, size_t n, uint64_t hbm_budget,
                    flock_prove_result* res, double* wall_ms) {
    if (!provers || workers == 0 || (n && (!reqs || !res))) return FLOCK_HIP_E_INVALID;
    for (uint32_t w = 0; w < workers; w++)
        if (!provers[w] || !provers[w]->ctx) return FLOCK_HIP_E_INVALID;
    std::vector<uint64_t> fp(n);
    for (size_t i = 0; i < n; i++) {
        Plan p;
        if (int rc = plan_request(&reqs[i], &p)) return rc;
        if (i && reqs[i].arrival_s < reqs[i - 1].arrival_s) return FLOCK_HIP_E_INVALID;
        fp[i] = plan_footprint(p, reqs[i].n_log);
    }
    std::mutex mu;
    std::condition_variable cv;
    size_t next = 0;
    uint64_t inflight = 0;
    uint32_t running = 0;
    int first_err = 0;
    const double t0 = now_s();
    auto worker = [&](flock_hip_prover* pr) {
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
            if (running && inflight + fp[i] > hbm_budget) {
                cv.wait(lk);
                continue;
            }
            next++;
            inflight += fp[i];
            running++;
            lk.unlock();
            const double start = now_s();
            const int rc = flock_hip_prove_core(pr, &reqs[i], &res[i]);
            const double end = now_s();
            res[i].queued_ms = (start - due) * 1e3;
            res[i].run_ms = (end - start) * 1e3;
            res[i].done_s = end - t0;
            lk.lock();
            inflight -= fp[i];
            running--;
            if (rc && !first_err) first_err = rc;
            cv.notify_all();
        }
    };
    std::vector<std::thread> th;
    th.reserve(workers);
    for (uint32_t w = 0; w < workers; w++) th.emplace_back(worker, provers[w]);
    for (auto& x : th) x.join();
    if (wall_ms) *wall_ms = (now_s() - t0) * 1e3;
    return first_err;
}

}  // extern "C"
