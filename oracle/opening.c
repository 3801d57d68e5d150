/*
 * opening.c — CPU restatement of the merged PCS opening of Flock's `prove_fast`
 * (crates/flock-core/src/pcs.rs:1433-1822) and its wire encoding.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP opening in
 * flock_amd/csrc. The shipped library never links or loads this file.
 *
 * Restated here, each at the cited reference lines:
 *   - the ring switch of the two boolean claims (pcs/ring_switch.rs:2521-2880),
 *   - the jagged merged weight and the dense merged sumcheck (pcs.rs:1590-1700,
 *     pcs/jagged.rs:1521-1650, 4710-4760),
 *   - the multipoint twisted assist and its Frobenius anchor
 *     (pcs/jagged.rs:1712-2025, 2266-2420, 2547-2590, 2884-3290),
 *   - the eq-basis Ligerito opening over the F256 ladder with stratified queries and
 *     capped paths (pcs.rs:214-360, 423-560; pcs/ligerito/extension.rs:1636-2290;
 *     pcs/ligerito.rs:3086-3140, 4502-4616; pcs/stratified.rs:73-130; merkle.rs:649-682),
 *   - bincode 1.3 (fixint, little-endian, u64 lengths) of MergedOpenProof
 *     (pcs.rs:1362-1377, pcs/jagged.rs:1679-1686, 2790-2803, pcs/ligerito.rs:2156-2218).
 *
 * Every formula is the definitional one; the reference's block trees, lookahead
 * skips, virtual bases and split eq tables are value-identical reassociations and are
 * not reproduced. What pins the result: the reference's own SHA-256 digests of whole
 * prove_fast proofs (crates/flock-prover/tests/union_m6_fixtures.rs:289-323), which
 * tests/test_oracle_prove_fast.py reproduces from the seeds.
 */
#include "flock_oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* field helpers                                                              */
/* ------------------------------------------------------------------------- */
static const or_f128 Z0 = {0, 0};
static const or_f128 O1 = {1, 0};

static inline or_f128 fa(or_f128 a, or_f128 b) {
    or_f128 r = {a.lo ^ b.lo, a.hi ^ b.hi};
    return r;
}
static inline or_f128 fm(or_f128 a, or_f128 b) { return or_f128_mul(a, b); }
static inline int fz(or_f128 a) { return !(a.lo | a.hi); }
static inline int feq(or_f128 a, or_f128 b) { return a.lo == b.lo && a.hi == b.hi; }
static inline or_f128 fbit(int t) {
    or_f128 r = {t < 64 ? (uint64_t)1 << t : 0, t >= 64 ? (uint64_t)1 << (t - 64) : 0};
    return r;
}
static inline int getbit(or_f128 x, int t) {
    return (int)(((t < 64 ? x.lo : x.hi) >> (t & 63)) & 1);
}

/* F256 = F128[u]/(u^2 + u + x^-1), crates/flock-core/src/field/gf2_256.rs:1-160 */
typedef struct {
    or_f128 c0, c1;
} f256;
static const f256 E0 = {{0, 0}, {0, 0}};

/* mul_by_x_inv, gf2_256.rs:24-31 */
static inline or_f128 mul_x_inv(or_f128 z) {
    uint64_t mask = 0 - (z.lo & 1);
    or_f128 r = {((z.lo >> 1) | (z.hi << 63)) ^ (0x43 & mask),
                 (z.hi >> 1) ^ (((uint64_t)1 << 63) & mask)};
    return r;
}
static inline f256 ea(f256 a, f256 b) {
    f256 r = {fa(a.c0, b.c0), fa(a.c1, b.c1)};
    return r;
}
/* Mul for F256, gf2_256.rs:106-122 */
static inline f256 em(f256 a, f256 b) {
    or_f128 p0 = fm(a.c0, b.c0), p1 = fm(a.c1, b.c1), p2 = fm(fa(a.c0, a.c1), fa(b.c0, b.c1));
    f256 r = {fa(p0, mul_x_inv(p1)), fa(p2, p0)};
    return r;
}
static inline f256 emb(f256 a, or_f128 b) { /* Mul<F128>, :132-142 */
    f256 r = {fm(a.c0, b), fm(a.c1, b)};
    return r;
}
static inline f256 ef(or_f128 a) {
    f256 r = {a, {0, 0}};
    return r;
}
/* XOR sums split over OpenMP threads (exact: the field sum is associative and commutative) */
#pragma omp declare reduction(fx : or_f128 : omp_out = fa(omp_out, omp_in)) initializer(omp_priv = (or_f128){0, 0})
#pragma omp declare reduction(ex : f256 : omp_out = ea(omp_out, omp_in)) initializer(omp_priv = (f256){{0, 0}, {0, 0}})
static const f256 E1 = {{1, 0}, {0, 0}};

/* ------------------------------------------------------------------------- */
/* growable buffers and the bincode writer                                    */
/* ------------------------------------------------------------------------- */
typedef struct {
    uint8_t* p;
    size_t n, cap;
} buf_t;

static void buf_put(buf_t* b, const void* d, size_t n) {
    if (b->n + n > b->cap) {
        size_t c = b->cap ? b->cap : 1024;
        while (c < b->n + n) c *= 2;
        b->p = (uint8_t*)realloc(b->p, c);
        b->cap = c;
    }
    memcpy(b->p + b->n, d, n);
    b->n += n;
}
static void bw_u64(buf_t* b, uint64_t v) {
    uint8_t x[8];
    for (int i = 0; i < 8; i++) x[i] = (uint8_t)(v >> (8 * i));
    buf_put(b, x, 8);
}
static void bw_f128(buf_t* b, or_f128 v) { /* F128 {lo, hi}, gf2_128.rs:22-27 */
    bw_u64(b, v.lo);
    bw_u64(b, v.hi);
}
static void bw_f256(buf_t* b, f256 v) { /* F256 {c0, c1}, gf2_256.rs:43-46 */
    bw_f128(b, v.c0);
    bw_f128(b, v.c1);
}
static void bw_vec_f128(buf_t* b, const or_f128* v, size_t n) {
    bw_u64(b, n);
    for (size_t i = 0; i < n; i++) bw_f128(b, v[i]);
}
static void bw_vec_u64(buf_t* b, const uint64_t* v, size_t n) {
    bw_u64(b, n);
    for (size_t i = 0; i < n; i++) bw_u64(b, v[i]);
}
static void bw_vec_hash(buf_t* b, const uint8_t* h, size_t n) { /* Vec<[u8; 32]> */
    bw_u64(b, n);
    buf_put(b, h, 32 * n);
}

typedef struct {
    uint64_t* p;
    size_t n, cap;
} u64v;
static void u64v_push(u64v* v, uint64_t x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 16;
        v->p = (uint64_t*)realloc(v->p, v->cap * 8);
    }
    v->p[v->n++] = x;
}
typedef struct {
    or_f128* p;
    size_t n, cap;
} f128v;
static void f128v_push(f128v* v, or_f128 x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 16;
        v->p = (or_f128*)realloc(v->p, v->cap * sizeof(or_f128));
    }
    v->p[v->n++] = x;
}

void or_free_bytes(uint8_t* p) { free(p); }

/* ------------------------------------------------------------------------- */
/* challenger conveniences (challenger.rs:57-216 default methods)              */
/* ------------------------------------------------------------------------- */
static f256 sample_f256(or_fs* fs) { /* sample_f256 = two-word vector squeeze, :98-101 */
    or_f128 w[2];
    or_fs_sample_f128_vec(fs, w, 2);
    f256 r = {w[0], w[1]};
    return r;
}
static void observe_f256(or_fs* fs, f256 v) { /* observe_f256 = slice of coordinates, :84-86 */
    or_f128 w[2] = {v.c0, v.c1};
    or_fs_observe_f128_slice(fs, w, 2);
}
static void observe_msg256(or_fs* fs, f256 u0, f256 u2) { /* extension.rs:100-104 */
    observe_f256(fs, u0);
    observe_f256(fs, u2);
}
/* grind_pow_and_sample_f128 (default composition), challenger.rs:133-139 */
static or_f128 grind_sample(or_fs* fs, uint32_t bits, uint64_t* nonce) {
    *nonce = or_fs_grind_pow(fs, bits);
    return or_fs_sample_f128(fs);
}
/* grinding_bits_for_degree, challenger.rs:45-51 */
static uint32_t bits_for_degree(size_t d) {
    uint32_t b = 0;
    while (d) {
        b++;
        d >>= 1;
    }
    return b;
}

/* ------------------------------------------------------------------------- */
/* eq tables                                                                  */
/* ------------------------------------------------------------------------- */
/* build_eq_table (lincheck.rs:713-733); LSB-first, same values as build_eq */
static or_f128* eq_table(const or_f128* r, size_t n) {
    or_f128* t = (or_f128*)malloc(sizeof(or_f128) * ((size_t)1 << n));
    or_build_eq(r, n, t);
    return t;
}
/* build_eq_table256 (extension.rs:56-68) */
static f256* eq_table256(const f256* r, size_t n) {
    f256* t = (f256*)malloc(sizeof(f256) << n);
    t[0] = E1;
    for (size_t j = 0; j < n; j++) {
        size_t len = (size_t)1 << j;
        f256 rp = ea(E1, r[j]);
        for (size_t i = 0; i < len; i++) {
            f256 v = t[i];
            t[i + len] = em(v, r[j]);
            t[i] = em(v, rp);
        }
    }
    return t;
}
/* eq(a, b) = prod (a_t b_t + (1+a_t)(1+b_t)), jagged.rs:2332-2337 */
static or_f128 eq_at(const or_f128* a, const or_f128* b, size_t n) {
    or_f128 acc = O1;
    for (size_t i = 0; i < n; i++)
        acc = fm(acc, fa(fm(a[i], b[i]), fm(fa(O1, a[i]), fa(O1, b[i]))));
    return acc;
}

/* ------------------------------------------------------------------------- */
/* Ligerito configurations: the embedded `fast` security configs               */
/* crates/flock-core/configs/ligerito/m{22..35}_fast.toml (ligerito.rs:542-620)  */
/* ------------------------------------------------------------------------- */
typedef struct {
    uint8_t rate, cols, lanes, krec;
    uint16_t queries;
    uint8_t grind, claim, cons, ood;
} lvl_t;
typedef struct {
    uint8_t m, initial_k, n_levels;
    lvl_t lv[7];
} lig_t;

#define L1_ {1, 0, 0, 0, 279, 0, 6, 9, 1}
static const lig_t LIG_FAST[] = {
    {22, 6, 3, {{1, 9, 6, 6, 279, 0, 6, 9, 1}, {2, 6, 4, 4, 136, 0, 6, 9, 2}, {4, 3, 4, 4, 68, 0, 7, 10, 2}}},
    {23, 6, 3, {{1, 10, 6, 6, 279, 0, 6, 9, 1}, {2, 7, 4, 4, 136, 0, 6, 9, 2}, {3, 4, 4, 4, 91, 0, 7, 9, 2}}},
    {24, 6, 3, {{1, 11, 6, 6, 279, 0, 6, 9, 1}, {2, 8, 4, 4, 136, 0, 6, 9, 2}, {3, 5, 4, 4, 91, 0, 7, 9, 2}}},
    {25, 6, 4, {{1, 12, 6, 6, 279, 0, 6, 9, 1}, {2, 9, 4, 4, 136, 0, 6, 9, 2}, {3, 6, 4, 4, 91, 0, 7, 9, 2}, {4, 3, 4, 4, 68, 0, 7, 10, 2}}},
    {26, 6, 4, {{1, 13, 6, 6, 279, 0, 6, 9, 1}, {2, 10, 4, 4, 136, 0, 6, 9, 2}, {3, 7, 4, 4, 91, 0, 7, 9, 2}, {4, 4, 4, 4, 68, 0, 7, 10, 2}}},
    {27, 6, 4, {{1, 14, 6, 6, 279, 0, 6, 9, 1}, {2, 11, 4, 4, 136, 0, 6, 9, 2}, {3, 8, 4, 4, 91, 0, 7, 9, 2}, {4, 5, 4, 4, 68, 0, 7, 10, 2}}},
    {28, 4, 5, {{1, 17, 4, 4, 279, 0, 6, 9, 1}, {2, 14, 4, 4, 136, 0, 6, 9, 2}, {3, 11, 4, 4, 91, 0, 7, 9, 2}, {4, 8, 4, 4, 68, 0, 7, 10, 2}, {5, 5, 4, 4, 55, 0, 8, 10, 2}}},
    {29, 5, 5, {{1, 17, 5, 5, 279, 0, 6, 9, 1}, {2, 14, 4, 4, 136, 0, 6, 9, 2}, {3, 11, 4, 4, 91, 0, 7, 9, 2}, {4, 8, 4, 4, 68, 0, 7, 10, 2}, {5, 5, 4, 4, 55, 0, 8, 10, 2}}},
    {30, 6, 5, {{1, 17, 6, 6, 279, 0, 6, 9, 1}, {2, 14, 4, 4, 136, 0, 6, 9, 2}, {3, 11, 4, 4, 91, 0, 7, 9, 2}, {4, 8, 4, 4, 68, 0, 7, 10, 2}, {5, 5, 4, 4, 55, 0, 8, 10, 2}}},
    {31, 6, 6, {{1, 18, 6, 6, 279, 0, 6, 9, 1}, {2, 15, 4, 4, 136, 0, 6, 9, 2}, {3, 12, 4, 4, 91, 0, 7, 9, 2}, {4, 9, 4, 4, 68, 0, 7, 10, 2}, {5, 6, 4, 4, 55, 0, 8, 10, 2}, {6, 3, 4, 4, 46, 0, 8, 11, 2}}},
    {32, 6, 6, {{1, 19, 6, 6, 279, 0, 6, 9, 1}, {2, 16, 4, 4, 136, 0, 6, 9, 2}, {3, 13, 4, 4, 91, 0, 7, 9, 2}, {4, 10, 4, 4, 68, 0, 7, 10, 2}, {5, 7, 4, 4, 55, 0, 8, 10, 2}, {6, 4, 4, 4, 46, 0, 8, 11, 2}}},
    {33, 6, 6, {{1, 20, 6, 6, 279, 0, 6, 9, 1}, {2, 17, 4, 4, 136, 0, 6, 9, 2}, {3, 14, 4, 4, 91, 0, 7, 9, 2}, {4, 11, 4, 4, 68, 0, 7, 10, 2}, {5, 8, 4, 4, 55, 0, 8, 10, 2}, {6, 5, 4, 4, 46, 0, 8, 11, 2}}},
    {34, 6, 7, {{1, 21, 6, 6, 279, 0, 6, 9, 1}, {2, 18, 4, 4, 136, 0, 6, 9, 2}, {3, 15, 4, 4, 91, 0, 7, 9, 2}, {4, 12, 4, 4, 68, 0, 7, 10, 2}, {5, 9, 4, 4, 55, 0, 8, 10, 2}, {6, 6, 4, 4, 46, 0, 8, 11, 2}, {7, 3, 4, 4, 40, 0, 9, 11, 2}}},
    {35, 6, 7, {{1, 22, 6, 6, 279, 0, 6, 9, 1}, {2, 19, 4, 4, 136, 0, 6, 9, 2}, {3, 16, 4, 4, 91, 0, 7, 9, 2}, {4, 13, 4, 4, 68, 0, 7, 10, 2}, {5, 10, 4, 4, 55, 0, 8, 10, 2}, {6, 7, 4, 4, 46, 0, 8, 11, 2}, {7, 4, 4, 4, 40, 0, 9, 11, 2}}},
};
#undef L1_

static const lig_t* lig_config(unsigned m) {
    for (size_t i = 0; i < sizeof(LIG_FAST) / sizeof(LIG_FAST[0]); i++)
        if (LIG_FAST[i].m == m) return &LIG_FAST[i];
    return NULL;
}

int or_lig_initial_k(unsigned m) {
    const lig_t* c = lig_config(m);
    return c ? c->initial_k : -1;
}

/* per-level committed block log, stratified.rs:172-190 */
static int level_block_log(const lig_t* c, int l) { return c->lv[l].cols + c->lv[l].rate; }

/* ------------------------------------------------------------------------- */
/* stratified schedules (pcs/stratified.rs:73-130)                             */
/* ------------------------------------------------------------------------- */
typedef struct {
    int d, n;
    int depth[48];
} sched_t;

static void sched_make(sched_t* s, size_t q, int d) { /* decompose, :73-85 */
    s->d = d;
    s->n = 0;
    while (q) {
        int c = 63 - __builtin_clzll((unsigned long long)q);
        if (c > d) c = d;
        s->depth[s->n++] = c;
        q -= (size_t)1 << c;
    }
}
static int sched_cap(const sched_t* s) { return s->n ? s->depth[0] : 0; } /* :97-99 */
static size_t sched_queries(const sched_t* s) {
    size_t q = 0;
    for (int i = 0; i < s->n; i++) q += (size_t)1 << s->depth[i];
    return q;
}
/* queries_from_words (ligerito.rs:4530-4547): the low d - c bits come from the word, the
 * top c bits are the stratum index, strata in query_strata order (stratified.rs:114-118) */
static void queries_from_words(const sched_t* s, const or_f128* w, size_t* out) {
    size_t k = 0;
    for (int i = 0; i < s->n; i++) {
        int c = s->depth[i], lo = s->d - c;
        for (size_t st = 0; st < ((size_t)1 << c); st++, k++)
            out[k] = (st << lo) | ((size_t)w[k].lo & (((size_t)1 << lo) - 1));
    }
}

/* ------------------------------------------------------------------------- */
/* Merkle caps and capped paths (merkle.rs:649-682), flat tree, leaves first   */
/* ------------------------------------------------------------------------- */
static const uint8_t* cap_layer(const uint8_t* tree, size_t n_leaves, int c) {
    size_t l = (size_t)1 << c;
    return tree + 32 * (2 * n_leaves - 2 * l);
}
static void path_capped(const uint8_t* tree, size_t n_leaves, size_t idx, int c, buf_t* out) {
    size_t start = 0, len = n_leaves;
    while (len > ((size_t)1 << c)) {
        buf_put(out, tree + 32 * (start + (idx ^ 1)), 32);
        start += len;
        len >>= 1;
        idx >>= 1;
    }
}

int or_cap_depth_for(unsigned m) { /* PcsParams::l0_cap_depth, commit.rs:266-277 */
    const lig_t* c = lig_config(m);
    if (!c) return -1;
    sched_t s;
    sched_make(&s, c->lv[0].queries, level_block_log(c, 0));
    return sched_cap(&s);
}

/* ------------------------------------------------------------------------- */
/* ligero_commit (ligerito.rs:3086-3140) through the PCS commit: same fill,   */
/* same interleaved NTT from layer log_rate, same row-leaf Merkle tree          */
/* ------------------------------------------------------------------------- */
typedef struct {
    or_f128* mat;
    uint8_t* tree;
    size_t block_len, lanes;
} lw_t;

static int ligero_commit(const or_f128* poly, int log_cols, int log_lanes, int log_rate,
                         uint32_t kind, lw_t* w) {
    or_pcs_params p = {(uint32_t)(log_cols + log_lanes + 7), (uint32_t)log_rate,
                       (uint32_t)log_lanes, 0, kind};
    w->lanes = (size_t)1 << log_lanes;
    w->block_len = (size_t)1 << (log_cols + log_rate);
    w->mat = (or_f128*)malloc(sizeof(or_f128) * w->block_len * w->lanes);
    w->tree = (uint8_t*)malloc(32 * (2 * w->block_len - 1));
    return or_commit(poly, w->lanes << log_cols, &p, w->mat, w->tree);
}
static void lw_free(lw_t* w) {
    free(w->mat);
    free(w->tree);
}

/* ------------------------------------------------------------------------- */
/* jagged layout (pcs/jagged.rs:87-140, 569-660)                               */
/* ------------------------------------------------------------------------- */
typedef struct {
    int n, k, m;
    uint64_t* ps; /* 2^k + 1 cumulative heights */
    /* deduped boundary runs (assist_boundaries, :569-584) */
    size_t nb;
    uint64_t *bc, *bn;
    uint32_t* brun;
} jag_t;

static void jag_make(jag_t* J, const uint64_t* heights, size_t nh, int n, int m) {
    J->n = n;
    J->m = m;
    J->k = __builtin_ctzll(nh);
    J->ps = (uint64_t*)malloc(8 * (nh + 1));
    J->ps[0] = 0;
    for (size_t y = 0; y < nh; y++) J->ps[y + 1] = J->ps[y] + heights[y];
    J->bc = (uint64_t*)malloc(8 * nh);
    J->bn = (uint64_t*)malloc(8 * nh);
    J->brun = (uint32_t*)malloc(4 * nh);
    J->nb = 0;
    for (size_t y = 0; y < nh; y++) {
        uint64_t c = J->ps[y], d = J->ps[y + 1];
        if (J->nb && J->bc[J->nb - 1] == c && J->bn[J->nb - 1] == d)
            J->brun[J->nb - 1]++;
        else {
            J->bc[J->nb] = c;
            J->bn[J->nb] = d;
            J->brun[J->nb] = 1;
            J->nb++;
        }
    }
}
static void jag_free(jag_t* J) {
    free(J->ps);
    free(J->bc);
    free(J->bn);
    free(J->brun);
}
static uint64_t jag_area(const jag_t* J) { return J->ps[(size_t)1 << J->k]; }

/* B[d] = sc * eq(z_row, row(d)) * eq(z_col, col(d)) accumulated into out, d < area */
static void jag_weight_acc(const jag_t* J, const or_f128* er, const or_f128* ec, or_f128 sc,
                           or_f128* out) {
    size_t nc = (size_t)1 << J->k;
    for (size_t y = 0; y < nc; y++) {
        or_f128 cy = fm(sc, ec[y]);
        for (uint64_t d = J->ps[y]; d < J->ps[y + 1]; d++)
            out[d] = fa(out[d], fm(er[d - J->ps[y]], cy));
    }
}

/* ------------------------------------------------------------------------- */
/* the jagged branching program (pcs/jagged.rs:163-240, 690-720)               */
/* ------------------------------------------------------------------------- */
enum { ST_INIT = 0, ST_SUCC = 2 };
static int sparse_tr[4][4][2][2]; /* [cd][state][entry] = (eq4 index, next state) */
static int sparse_ready = 0;

static int bp_transition(int row, int index, int curr, int next, int state) {
    int carry = state & 1, cmp = (state >> 1) & 1;
    int sum = row + carry + curr;
    if (index != (sum & 1)) return -1;
    int ncmp = (index == next) ? cmp : next;
    return (sum >> 1) + (ncmp << 1);
}
static void sparse_init(void) { /* assist_sparse_transitions, :700-716 */
    if (sparse_ready) return;
    for (int cd = 0; cd < 4; cd++) {
        int c = cd & 1, d = (cd >> 1) & 1;
        for (int s = 0; s < 4; s++)
            for (int a = 0; a < 2; a++) {
                int b = (a + (s & 1) + c) & 1;
                sparse_tr[cd][s][a][0] = a + 2 * b;
                sparse_tr[cd][s][a][1] = bp_transition(a, b, c, d, s);
            }
    }
    sparse_ready = 1;
}
static inline or_f128 pbit(const or_f128* z, size_t zl, int layer) { /* point_bit, :146-153 */
    return (size_t)layer < zl ? z[layer] : Z0;
}
/* eq table of [a, b] (index bit 0 = a) */
static void eq4(or_f128 a, or_f128 b, or_f128 out[4]) {
    or_f128 a1 = fa(O1, a), b1 = fa(O1, b);
    out[0] = fm(a1, b1);
    out[1] = fm(a, b1);
    out[2] = fm(a1, b);
    out[3] = fm(a, b);
}
/* S' = M_cd S: dst[s] = eq4[i0] src[o0] + eq4[i1] src[o1] */
static void bp_step(const or_f128 e[4], int cd, const or_f128 src[4], or_f128 dst[4]) {
    for (int s = 0; s < 4; s++) {
        or_f128 acc = Z0;
        for (int a = 0; a < 2; a++)
            acc = fa(acc, fm(e[sparse_tr[cd][s][a][0]], src[sparse_tr[cd][s][a][1]]));
        dst[s] = acc;
    }
}
static inline int col_cd(uint64_t tc, uint64_t tn, int layer) {
    return (int)(((tc >> layer) & 1) + 2 * ((tn >> layer) & 1));
}
/* per-run column weights Σ_{y in run} eq(z_col, y) (assist_columns_at, :592-637) */
static or_f128* run_weights(const jag_t* J, const or_f128* z_col) {
    or_f128* ec = eq_table(z_col, (size_t)J->k);
    or_f128* w = (or_f128*)malloc(sizeof(or_f128) * J->nb);
    size_t y = 0;
    for (size_t b = 0; b < J->nb; b++) {
        or_f128 acc = Z0;
        for (uint32_t r = 0; r < J->brun[b]; r++) acc = fa(acc, ec[y++]);
        w[b] = acc;
    }
    free(ec);
    return w;
}
/* f̂_t(z_row, z_col, x) = Σ_runs w · ĝ(z_row, x, t_c, t_n) (multipoint_values, :2547-2585) */
static or_f128 jag_fhat(const jag_t* J, const or_f128* z_row, const or_f128* w,
                        const or_f128* x) {
    int m = J->m;
    or_f128(*e)[4] = malloc(sizeof(or_f128[4]) * (size_t)(m + 1));
    for (int l = 0; l <= m; l++)
        eq4(pbit(z_row, (size_t)J->n, l), pbit(x, (size_t)m, l), e[l]);
    or_f128 acc = Z0;
    for (size_t b = 0; b < J->nb; b++) {
        or_f128 S[4] = {Z0, Z0, Z0, Z0}, T[4];
        S[ST_SUCC] = O1;
        for (int l = m; l >= 0; l--) {
            bp_step(e[l], col_cd(J->bc[b], J->bn[b], l), S, T);
            memcpy(S, T, sizeof(S));
        }
        acc = fa(acc, fm(w[b], S[ST_INIT]));
    }
    free(e);
    return acc;
}

int or_jag_fhat(const uint64_t* heights, size_t nh, int n, int m, const or_f128* z_row,
                const or_f128* z_col, const or_f128* x, or_f128* out) {
    sparse_init();
    jag_t J;
    jag_make(&J, heights, nh, n, m);
    or_f128* w = run_weights(&J, z_col);
    *out = jag_fhat(&J, z_row, w, x);
    free(w);
    jag_free(&J);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Frobenius assist (pcs/jagged.rs:1712-2025): sumcheck over the 2(m+1)        */
/* boundary bits, interleaved c_0, d_0, c_1, d_1, ...                          */
/* ------------------------------------------------------------------------- */
typedef struct {
    or_f128 v;
    or_f128* rounds; /* 2 * 2(m+1) */
    uint64_t* nonces;
    size_t n_rounds, n_nonces;
} frob_t;

/* statements: z_rows[s] (n coords), w[s] (per-run weights, coefficient baked) */
static void frobenius_assist(or_fs* fs, const jag_t* J, size_t n_st, or_f128** z_rows,
                             or_f128** w, const or_f128* rho, uint32_t bits, frob_t* out) {
    int m = J->m;
    size_t nb = J->nb;
    /* eq4 per statement per layer */
    or_f128(*E)[4] = malloc(sizeof(or_f128[4]) * n_st * (size_t)(m + 1));
    for (size_t s = 0; s < n_st; s++)
        for (int l = 0; l <= m; l++)
            eq4(pbit(z_rows[s], (size_t)J->n, l), pbit(rho, (size_t)m, l), E[s * (m + 1) + l]);
    /* suffix vectors S[s][b][l] for l in 0..m+1 */
    size_t sl = (size_t)(m + 2);
    or_f128(*S)[4] = malloc(sizeof(or_f128[4]) * n_st * nb * sl);
    for (size_t s = 0; s < n_st; s++)
        for (size_t b = 0; b < nb; b++) {
            or_f128(*Sb)[4] = S + (s * nb + b) * sl;
            for (int i = 0; i < 4; i++) Sb[m + 1][i] = Z0;
            Sb[m + 1][ST_SUCC] = O1;
            for (int l = m; l >= 0; l--)
                bp_step(E[s * (m + 1) + l], col_cd(J->bc[b], J->bn[b], l), Sb[l + 1], Sb[l]);
        }
    or_f128 v = Z0;
    for (size_t s = 0; s < n_st; s++)
        for (size_t b = 0; b < nb; b++) v = fa(v, fm(w[s][b], S[(s * nb + b) * sl][ST_INIT]));
    static const uint8_t LBL[] = "flock-frobenius-assist-v0";
    or_fs_observe_label(fs, LBL, sizeof(LBL) - 1);
    or_fs_observe_f128(fs, v);
    out->v = v;
    out->n_rounds = 2 * (size_t)(m + 1);
    out->rounds = (or_f128*)malloc(sizeof(or_f128) * 2 * out->n_rounds);
    out->nonces = (uint64_t*)malloc(8 * out->n_rounds);
    out->n_nonces = bits ? out->n_rounds : 0;
    or_f128* p = (or_f128*)malloc(sizeof(or_f128) * n_st * nb);
    for (size_t s = 0; s < n_st; s++) memcpy(p + s * nb, w[s], sizeof(or_f128) * nb);
    or_f128(*pre)[4] = malloc(sizeof(or_f128[4]) * n_st);
    for (size_t s = 0; s < n_st; s++) {
        for (int i = 0; i < 4; i++) pre[s][i] = Z0;
        pre[s][ST_INIT] = O1;
    }
    or_f128 ch4[4];
    size_t ri = 0;
    or_f128(*U)[4][4] = malloc(sizeof(or_f128[4][4]) * n_st);
    or_f128(*B)[4][4] = malloc(sizeof(or_f128[4][4]) * n_st);
    for (int l = 0; l <= m; l++) {
        for (size_t s = 0; s < n_st; s++) {
            or_f128* ps = p + s * nb;
            if (l > 0) /* fold_partials with the previous layer's quadrant (:1084-1121) */
                for (size_t b = 0; b < nb; b++)
                    ps[b] = fm(ps[b], ch4[col_cd(J->bc[b], J->bn[b], l - 1)]);
            /* buckets B[cd] = Σ p_b S_b[l+1] (:1132-1183) */
            for (int cd = 0; cd < 4; cd++)
                for (int i = 0; i < 4; i++) B[s][cd][i] = Z0;
            for (size_t b = 0; b < nb; b++) {
                int cd = col_cd(J->bc[b], J->bn[b], l);
                or_f128* sv = S[(s * nb + b) * sl + l + 1];
                for (int i = 0; i < 4; i++) B[s][cd][i] = fa(B[s][cd][i], fm(ps[b], sv[i]));
            }
            /* u rows: u[cd] = prefix^T M_cd (:1185-1200) */
            const or_f128* e = E[s * (m + 1) + l];
            for (int cd = 0; cd < 4; cd++) {
                for (int i = 0; i < 4; i++) U[s][cd][i] = Z0;
                for (int st = 0; st < 4; st++)
                    for (int a = 0; a < 2; a++) {
                        int ix = sparse_tr[cd][st][a][0], o = sparse_tr[cd][st][a][1];
                        U[s][cd][o] = fa(U[s][cd][o], fm(pre[s][st], e[ix]));
                    }
            }
        }
        /* c-round (:1940-1958) */
        or_f128 g1 = Z0, gi = Z0;
        for (size_t s = 0; s < n_st; s++)
            for (int i = 0; i < 4; i++) {
                g1 = fa(g1, fa(fm(U[s][1][i], B[s][1][i]), fm(U[s][3][i], B[s][3][i])));
                gi = fa(gi, fa(fm(fa(U[s][0][i], U[s][1][i]), fa(B[s][0][i], B[s][1][i])),
                               fm(fa(U[s][2][i], U[s][3][i]), fa(B[s][2][i], B[s][3][i]))));
            }
        or_fs_observe_f128(fs, g1);
        or_fs_observe_f128(fs, gi);
        uint64_t nonce = 0;
        or_f128 rc = bits ? grind_sample(fs, bits, &nonce) : or_fs_sample_f128(fs);
        if (bits) out->nonces[ri] = nonce;
        out->rounds[2 * ri] = g1;
        out->rounds[2 * ri + 1] = gi;
        ri++;
        /* d-round (:1960-1990) */
        or_f128 rc1 = fa(O1, rc);
        g1 = Z0;
        gi = Z0;
        or_f128(*UD)[2][4] = malloc(sizeof(or_f128[2][4]) * n_st);
        for (size_t s = 0; s < n_st; s++)
            for (int i = 0; i < 4; i++) {
                or_f128 ud0 = fa(fm(rc1, U[s][0][i]), fm(rc, U[s][1][i]));
                or_f128 ud1 = fa(fm(rc1, U[s][2][i]), fm(rc, U[s][3][i]));
                or_f128 d0 = fa(fm(rc1, B[s][0][i]), fm(rc, B[s][1][i]));
                or_f128 d1 = fa(fm(rc1, B[s][2][i]), fm(rc, B[s][3][i]));
                g1 = fa(g1, fm(ud1, d1));
                gi = fa(gi, fm(fa(ud0, ud1), fa(d0, d1)));
                UD[s][0][i] = ud0;
                UD[s][1][i] = ud1;
            }
        or_fs_observe_f128(fs, g1);
        or_fs_observe_f128(fs, gi);
        or_f128 rd = bits ? grind_sample(fs, bits, &nonce) : or_fs_sample_f128(fs);
        if (bits) out->nonces[ri] = nonce;
        out->rounds[2 * ri] = g1;
        out->rounds[2 * ri + 1] = gi;
        ri++;
        or_f128 rd1 = fa(O1, rd);
        for (size_t s = 0; s < n_st; s++)
            for (int i = 0; i < 4; i++)
                pre[s][i] = fa(fm(rd1, UD[s][0][i]), fm(rd, UD[s][1][i]));
        free(UD);
        ch4[0] = fm(rc1, rd1);
        ch4[1] = fm(rc, rd1);
        ch4[2] = fm(rc1, rd);
        ch4[3] = fm(rc, rd);
    }
    free(U);
    free(B);
    free(pre);
    free(p);
    free(S);
    free(E);
}

/* ------------------------------------------------------------------------- */
/* Moore inverse and linearized coefficients (pcs/ring_switch.rs:1686-1775)    */
/* ------------------------------------------------------------------------- */
static or_f128* g_minv = NULL;

static const or_f128* moore_inverse(void) {
    if (g_minv) return g_minv;
    const int N = 128;
    or_f128* m = (or_f128*)malloc(sizeof(or_f128) * N * N);
    or_f128* inv = (or_f128*)calloc((size_t)N * N, sizeof(or_f128));
    for (int t = 0; t < N; t++) {
        or_f128 x = fbit(t);
        for (int j = 0; j < N; j++) {
            m[t * N + j] = x;
            x = fm(x, x);
        }
    }
    for (int i = 0; i < N; i++) inv[i * N + i] = O1;
    for (int col = 0; col < N; col++) {
        int piv = col;
        while (fz(m[piv * N + col])) piv++;
        if (piv != col)
            for (int k = 0; k < N; k++) {
                or_f128 t = m[col * N + k];
                m[col * N + k] = m[piv * N + k];
                m[piv * N + k] = t;
                t = inv[col * N + k];
                inv[col * N + k] = inv[piv * N + k];
                inv[piv * N + k] = t;
            }
        or_f128 p = or_f128_inv(m[col * N + col]);
        for (int k = 0; k < N; k++) {
            m[col * N + k] = fm(m[col * N + k], p);
            inv[col * N + k] = fm(inv[col * N + k], p);
        }
        for (int r = 0; r < N; r++)
            if (r != col && !fz(m[r * N + col])) {
                or_f128 f = m[r * N + col];
                for (int k = 0; k < N; k++) {
                    m[r * N + k] = fa(m[r * N + k], fm(f, m[col * N + k]));
                    inv[r * N + k] = fa(inv[r * N + k], fm(f, inv[col * N + k]));
                }
            }
    }
    free(m);
    g_minv = inv;
    return inv;
}

/* x^(2^-1) = x^(2^127): square root, F2-linear (jagged.rs:2266-2300) */
static or_f128 g_sqrt_basis[128];
static int g_sqrt_ready = 0;
static or_f128 frob_inv(or_f128 x) {
    if (!g_sqrt_ready) {
        for (int b = 0; b < 128; b++) {
            or_f128 y = fbit(b);
            for (int i = 0; i < 127; i++) y = fm(y, y);
            g_sqrt_basis[b] = y;
        }
        g_sqrt_ready = 1;
    }
    or_f128 acc = Z0;
    for (int b = 0; b < 128; b++)
        if (getbit(x, b)) acc = fa(acc, g_sqrt_basis[b]);
    return acc;
}

/* an F2-linear map given by 128 basis images, applied through 16 byte tables */
typedef struct {
    or_f128 t[16][256];
} lin_t;
static void lin_make(lin_t* L, const or_f128* images) {
    for (int k = 0; k < 16; k++) {
        L->t[k][0] = Z0;
        for (int v = 1; v < 256; v++) {
            int low = __builtin_ctz((unsigned)v);
            L->t[k][v] = fa(L->t[k][v & (v - 1)], images[8 * k + low]);
        }
    }
}
static inline or_f128 lin_apply(const lin_t* L, or_f128 x) {
    or_f128 acc = Z0;
    for (int k = 0; k < 8; k++) {
        acc = fa(acc, L->t[k][(x.lo >> (8 * k)) & 0xFF]);
        acc = fa(acc, L->t[8 + k][(x.hi >> (8 * k)) & 0xFF]);
    }
    return acc;
}

/* ------------------------------------------------------------------------- */
/* sumcheck helpers                                                           */
/* ------------------------------------------------------------------------- */
/* (G(1), G(∞)) over LSB pairs (jagged.rs:4605-4620) */
static void msg_g1(const or_f128* a, const or_f128* b, size_t n, or_f128* g1, or_f128* gi) {
    or_f128 x = Z0, y = Z0;
#pragma omp parallel for schedule(static) reduction(fx : x, y) if (n >= 16384)
    for (size_t i = 0; i < n / 2; i++) {
        x = fa(x, fm(a[2 * i + 1], b[2 * i + 1]));
        y = fa(y, fm(fa(a[2 * i], a[2 * i + 1]), fa(b[2 * i], b[2 * i + 1])));
    }
    *g1 = x;
    *gi = y;
}
/* v'[i] = v[2i] + r (v[2i+1] + v[2i]) (jagged.rs:4693-4707) */
static void fold_lsb(or_f128* v, size_t n, or_f128 r) {
    if (n < 16384) {
        for (size_t i = 0; i < n / 2; i++) v[i] = fa(v[2 * i], fm(r, fa(v[2 * i + 1], v[2 * i])));
        return;
    }
    /* in place, entry i overwrites a slot a lower entry still reads: fold into a copy */
    or_f128* t = (or_f128*)malloc(sizeof(or_f128) * (n / 2));
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n / 2; i++) t[i] = fa(v[2 * i], fm(r, fa(v[2 * i + 1], v[2 * i])));
    memcpy(v, t, sizeof(or_f128) * (n / 2));
    free(t);
}

/* F256 blocked pairing: output o = blk*d + w reads lo = 2 blk d + w, hi = lo + d
 * (extension.rs:145-224) */
static void e_msg_blocked(const f256* f, const f256* b, size_t n, size_t d, f256* u0, f256* u2) {
    f256 x = E0, y = E0;
    if (d == 1 || n == d) d = 1;
#pragma omp parallel for schedule(static) reduction(ex : x, y) if (n >= 8192)
    for (size_t o = 0; o < n / 2; o++) {
        size_t lo = 2 * (o / d) * d + o % d, hi = lo + d;
        x = ea(x, em(f[lo], b[lo]));
        y = ea(y, em(ea(f[lo], f[hi]), ea(b[lo], b[hi])));
    }
    *u0 = x;
    *u2 = y;
}
static void e_fold_blocked(f256* v, size_t n, size_t d, f256 r) {
    f256* t = (f256*)malloc(sizeof(f256) * (n / 2));
#pragma omp parallel for schedule(static) if (n >= 8192)
    for (size_t o = 0; o < n / 2; o++) {
        size_t lo = 2 * (o / d) * d + o % d, hi = lo + d;
        t[o] = ea(v[lo], em(r, ea(v[hi], v[lo])));
    }
    memcpy(v, t, sizeof(f256) * (n / 2));
    free(t);
}

/* ------------------------------------------------------------------------- */
/* the eq-basis Ligerito opening over the F256 ladder                          */
/* (pcs.rs:214-360 + 423-560 inner-open path; extension.rs:1636-2290)          */
/* ------------------------------------------------------------------------- */
typedef struct {
    f128v yr;
    buf_t proof; /* bincode of LigeritoProof */
} lig_out_t;

/* induced_basis (extension.rs:1606-1623 -> ligerito.rs:2676-2770 with empty rows):
 * basis[j] = Σ_i eq(alpha, i) Π_{k in bits(j)} Ŵ_k(q_i), Ŵ_k = s_k / s_k(v_k) */
static or_f128* induced_basis(int log_cols, const size_t* queries, size_t nq,
                              const or_f128* alpha, size_t n_alpha) {
    size_t n = (size_t)1 << log_cols;
    /* eval_sk_at_vks (ligerito.rs:2402-2420) */
    or_f128 skv[64];
    skv[0] = O1;
    {
        or_f128 layer[64];
        for (int i = 1; i <= log_cols; i++) layer[i - 1] = fbit(i);
        int cur = log_cols;
        for (int i = 0; i < log_cols; i++) {
            for (int j = 0; j < cur; j++) {
                or_f128 s = layer[j];
                or_f128 nx = fa(fm(s, s), fm(skv[i], s));
                if (j == 0)
                    skv[i + 1] = nx;
                else
                    layer[j - 1] = nx;
            }
            cur--;
        }
    }
    or_f128 inv[64];
    for (int i = 0; i <= log_cols; i++) inv[i] = fz(skv[i]) ? Z0 : or_f128_inv(skv[i]);
    or_f128* aw = eq_table(alpha, n_alpha);
    or_f128* basis = (or_f128*)calloc(n, sizeof(or_f128));
    /* the queries' tensors are summed into per-thread bases, then into basis */
#pragma omp parallel if (n * nq >= 65536)
    {
        or_f128* loc = (or_f128*)malloc(sizeof(or_f128) * n);
        or_f128* part = (or_f128*)calloc(n, sizeof(or_f128));
        or_f128 sk[64];
#pragma omp for schedule(static)
        for (size_t i = 0; i < nq; i++) {
            or_f128 x = {(uint64_t)queries[i], 0};
            if (log_cols > 0) {
                sk[0] = x;
                for (int k = 1; k < log_cols; k++) sk[k] = fa(fm(sk[k - 1], sk[k - 1]), fm(skv[k - 1], sk[k - 1]));
                for (int k = 0; k < log_cols; k++) sk[k] = fm(sk[k], inv[k]);
            }
            loc[0] = aw[i];
            for (int k = 0; k < log_cols; k++) {
                size_t len = (size_t)1 << k;
                for (size_t j = 0; j < len; j++) loc[j + len] = fm(sk[k], loc[j]);
            }
            for (size_t j = 0; j < n; j++) part[j] = fa(part[j], loc[j]);
        }
#pragma omp critical
        for (size_t j = 0; j < n; j++) basis[j] = fa(basis[j], part[j]);
        free(part);
        free(loc);
    }
    free(aw);
    return basis;
}

/* induce_enforced_sum (extension.rs:2893-2916) */
static f256 enforced_sum(const or_f128* rows, size_t nq, size_t row_len, const f256* lane_ch,
                         size_t n_lane_ch, const or_f128* alpha, size_t n_alpha) {
    f256* lw = eq_table256(lane_ch, n_lane_ch);
    or_f128* rw = eq_table(alpha, n_alpha);
    f256 acc = E0;
    for (size_t i = 0; i < nq; i++) {
        f256 d = E0;
        for (size_t l = 0; l < row_len; l++) d = ea(d, emb(lw[l], rows[i * row_len + l]));
        acc = ea(acc, emb(d, rw[i]));
    }
    free(lw);
    free(rw);
    return acc;
}

static size_t ceil_log2(size_t n) { /* ligerito.rs:2461-2468 */
    if (n <= 1) return 0;
    size_t b = 0;
    n -= 1;
    while (n) {
        b++;
        n >>= 1;
    }
    return b;
}

typedef struct {
    f256* f; /* current table */
    f256* b; /* combined basis */
    size_t n;
    f256* pend; /* introduced basis awaiting glue */
    buf_t msgs; /* bincode of the SumcheckMessage256 entries */
    size_t n_msgs;
} sc256_t;

static void sc_push(sc256_t* s, f256 u0, f256 u2) {
    bw_f256(&s->msgs, u0);
    bw_f256(&s->msgs, u2);
    s->n_msgs++;
}
/* code_switch_message (extension.rs:1525-1533): split f into (c0, c1) pairs and the basis
 * into (B, u B), then the LSB message of the split pair */
static void sc_switch(sc256_t* s, f256* u0, f256* u2) {
    size_t n = s->n;
    f256* nf = (f256*)malloc(sizeof(f256) * 2 * n);
    f256* nb = (f256*)malloc(sizeof(f256) * 2 * n);
    const f256 U = {{0, 0}, {1, 0}};
    for (size_t i = 0; i < n; i++) {
        nf[2 * i] = ef(s->f[i].c0);
        nf[2 * i + 1] = ef(s->f[i].c1);
        nb[2 * i] = s->b[i];
        nb[2 * i + 1] = em(U, s->b[i]);
    }
    free(s->f);
    free(s->b);
    s->f = nf;
    s->b = nb;
    s->n = 2 * n;
    e_msg_blocked(s->f, s->b, s->n, 1, u0, u2);
}
/* introduce_extension (extension.rs:1535-1549) */
static void sc_introduce(sc256_t* s, f256* basis, f256* u0, f256* u2) {
    e_msg_blocked(s->f, basis, s->n, 1, u0, u2);
    sc_push(s, *u0, *u2);
    s->pend = basis;
}
static void sc_glue(sc256_t* s, or_f128 beta) { /* glue, :1583-1593 */
#pragma omp parallel for schedule(static) if (s->n >= 8192)
    for (size_t i = 0; i < s->n; i++) s->b[i] = ea(s->b[i], emb(s->pend[i], beta));
    free(s->pend);
    s->pend = NULL;
}

/* rows of a level witness at the queries, plus capped paths */
static void open_rows(const or_f128* mat, size_t row_len, const uint8_t* tree, size_t block_len,
                      const size_t* qs, size_t nq, int cap, or_f128* rows, buf_t* paths) {
    for (size_t i = 0; i < nq; i++) {
        memcpy(rows + i * row_len, mat + qs[i] * row_len, sizeof(or_f128) * row_len);
        path_capped(tree, block_len, qs[i], cap, paths);
    }
}
static void bw_rows(buf_t* b, const or_f128* rows, size_t nq, size_t row_len) {
    bw_u64(b, nq);
    for (size_t i = 0; i < nq; i++) bw_vec_f128(b, rows + i * row_len, row_len);
}

/* The inner open: one packed-direct EqPoint claim q̂(rho) = q_eval on the lane-major L0
 * commitment (pcs.rs:214-360 with compute_combined_basis_and_target :423-560), then
 * recursive_prover_with_basis_impl (extension.rs:1636-2290). Writes bincode of
 * BatchOpeningProofLigerito into out. */
static int inner_open(or_fs* fs, const or_f128* q, const or_pcs_params* P, const or_f128* cw,
                      const uint8_t* tree, const or_f128* rho, or_f128 q_eval, buf_t* out) {
    const lig_t* C = lig_config(P->m);
    if (!C || C->initial_k != P->log_batch_size) return -10;
    const int log_n = (int)P->m - 7, ik = C->initial_k, R = C->n_levels - 1;
    const size_t N = (size_t)1 << log_n, t = P->num_lanes ? P->num_lanes : ((size_t)1 << ik);
    const int n1 = log_n - ik;
    /* lane-major (integer-lane) commitments fold lane BLOCKS; a full-width commit keeps the
     * lanes on the low index bits (pcs.rs:272-276, extension.rs:1672-1676) */
    const size_t D = t < ((size_t)1 << ik) ? (size_t)1 << n1 : 1;
    const int rate0 = C->lv[0].rate;
    const size_t bl0 = (size_t)1 << (n1 + rate0);
    sched_t sch[8];
    for (int l = 0; l <= R; l++) sched_make(&sch[l], C->lv[l].queries, level_block_log(C, l));
    for (int l = 0; l <= R; l++)
        if (sched_queries(&sch[l]) != C->lv[l].queries) return -11;
    const uint32_t kind = P->merkle_hash;

    /* compute_combined_basis_and_target, single EqPoint claim (pcs.rs:423-560) */
    static const uint8_t L_OB[] = "flock-pcs-open-batch-v0", L_PD[] = "flock-pcs-packed-direct-v0";
    or_fs_observe_label(fs, L_OB, sizeof(L_OB) - 1);
    or_fs_observe_label(fs, L_PD, sizeof(L_PD) - 1);
    or_fs_observe_f128(fs, q_eval);
    or_f128 gamma;
    or_fs_sample_f128_vec(fs, &gamma, 1); /* batch bits 0 for one claim */
    or_f128 target = fm(gamma, q_eval);
    or_f128* b0 = eq_table(rho, (size_t)log_n);
#pragma omp parallel for schedule(static) if (N >= 8192)
    for (size_t i = 0; i < N; i++) b0[i] = fm(gamma, b0[i]);

    /* recursive_prover_with_basis_impl */
    static const uint8_t L_LG[] = "flock-ligerito-basis-f256-split-v0";
    or_fs_observe_label(fs, L_LG, sizeof(L_LG) - 1);
    or_fs_observe_f128(fs, target);
    const int c0 = sched_cap(&sch[0]);
    const uint8_t* icap = cap_layer(tree, bl0, c0);
    or_fs_observe_bytes(fs, icap, 32 * ((size_t)1 << c0));

    f128v ood = {0}, yr = {0};
    u64v claim_n = {0}, cons_n = {0}, grind_n = {0};
    /* L0 OOD samples: y = q̂(z), basis += beta eq(z) */
    for (int o = 0; o < C->lv[0].ood; o++) {
        or_f128* z = (or_f128*)malloc(sizeof(or_f128) * (size_t)log_n);
        or_fs_sample_f128_vec(fs, z, (size_t)log_n);
        or_f128* ez = eq_table(z, (size_t)log_n);
        or_f128 y = Z0;
#pragma omp parallel for schedule(static) reduction(fx : y) if (N >= 8192)
        for (size_t i = 0; i < N; i++) y = fa(y, fm(q[i], ez[i]));
        or_fs_observe_f128(fs, y);
        f128v_push(&ood, y);
        uint64_t nonce;
        or_f128 beta = grind_sample(fs, C->lv[0].claim, &nonce);
        u64v_push(&claim_n, nonce);
        target = fa(target, fm(beta, y));
#pragma omp parallel for schedule(static) if (N >= 8192)
        for (size_t i = 0; i < N; i++) b0[i] = fa(b0[i], fm(beta, ez[i]));
        free(z);
        free(ez);
    }
    sc256_t s;
    memset(&s, 0, sizeof(s));
    s.n = N;
    s.f = (f256*)malloc(sizeof(f256) * N);
    s.b = (f256*)malloc(sizeof(f256) * N);
    for (size_t i = 0; i < N; i++) {
        s.f[i] = ef(q[i]);
        s.b[i] = ef(b0[i]);
    }
    free(b0);
    f256 u0, u2;
    e_msg_blocked(s.f, s.b, s.n, D, &u0, &u2); /* first message, blocked pairs */
    sc_push(&s, u0, u2);
    observe_msg256(fs, u0, u2);
    /* the initial_k lane folds, the last one into the code switch */
    f256 lane_ch[16];
    for (int j = 0; j < ik; j++) {
        f256 r = sample_f256(fs);
        lane_ch[j] = r;
        e_fold_blocked(s.f, s.n, D, r);
        e_fold_blocked(s.b, s.n, D, r);
        s.n /= 2;
        if (j + 1 < ik) {
            e_msg_blocked(s.f, s.b, s.n, D, &u0, &u2);
        } else {
            sc_switch(&s, &u0, &u2);
        }
        sc_push(&s, u0, u2);
        observe_msg256(fs, u0, u2);
    }
    int split_dim = n1 + 1;
    /* commit level 1 from the split table */
    lw_t prev;
    {
        or_f128* base = (or_f128*)malloc(sizeof(or_f128) * s.n);
        for (size_t i = 0; i < s.n; i++) base[i] = s.f[i].c0;
        int ll = C->lv[1].krec;
        if (ligero_commit(base, split_dim - ll, ll, C->lv[1].rate, kind, &prev)) return -12;
        free(base);
    }
    buf_t caps = {0}; /* Vec<Vec<Hash>> body */
    size_t n_caps = 0;
    {
        int c = sched_cap(&sch[1]);
        const uint8_t* cp = cap_layer(prev.tree, prev.block_len, c);
        or_fs_observe_bytes(fs, cp, 32 * ((size_t)1 << c));
        bw_vec_hash(&caps, cp, (size_t)1 << c);
        n_caps++;
    }
    /* level-1 OOD: introduce eq(z) on the split table, then glue (extension.rs:1982-1996) */
#define LEVEL_OOD(LEVEL)                                                                      \
    for (int o = 0; o < C->lv[LEVEL].ood; o++) {                                              \
        or_f128* z = (or_f128*)malloc(sizeof(or_f128) * (size_t)split_dim);                   \
        or_fs_sample_f128_vec(fs, z, (size_t)split_dim);                                      \
        or_f128* ez = eq_table(z, (size_t)split_dim);                                         \
        or_f128 y = Z0;                                                                       \
        f256* eb = (f256*)malloc(sizeof(f256) * s.n);                                          \
        _Pragma("omp parallel for schedule(static) reduction(fx : y) if (s.n >= 8192)")       \
        for (size_t i = 0; i < s.n; i++) {                                                     \
            y = fa(y, fm(s.f[i].c0, ez[i]));                                                  \
            eb[i] = ef(ez[i]);                                                                \
        }                                                                                     \
        sc_introduce(&s, eb, &u0, &u2);                                                       \
        or_fs_observe_f128(fs, y);                                                            \
        f128v_push(&ood, y);                                                                  \
        observe_msg256(fs, u0, u2);                                                           \
        uint64_t nonce;                                                                       \
        or_f128 beta = grind_sample(fs, C->lv[LEVEL].claim, &nonce);                           \
        u64v_push(&claim_n, nonce);                                                           \
        sc_glue(&s, beta);                                                                    \
        free(z);                                                                              \
        free(ez);                                                                             \
    }
    LEVEL_OOD(1)
    /* L0 queries, consistency batch, rows, induced basis (extension.rs:1999-2038) */
    buf_t l0_proof = {0};
    {
        size_t nq = C->lv[0].queries;
        uint64_t nonce = or_fs_grind_pow(fs, C->lv[0].grind);
        u64v_push(&grind_n, nonce);
        or_f128* w = (or_f128*)malloc(sizeof(or_f128) * nq);
        or_fs_sample_f128_vec(fs, w, nq);
        size_t* qs = (size_t*)malloc(sizeof(size_t) * nq);
        queries_from_words(&sch[0], w, qs);
        size_t na = ceil_log2(nq);
        or_f128 alpha[64];
        nonce = or_fs_grind_pow(fs, C->lv[0].cons);
        u64v_push(&cons_n, nonce);
        or_fs_sample_f128_vec(fs, alpha, na);
        or_f128* rows = (or_f128*)malloc(sizeof(or_f128) * nq * t);
        buf_t paths = {0};
        open_rows(cw, t, tree, bl0, qs, nq, c0, rows, &paths);
        bw_rows(&l0_proof, rows, nq, t);
        bw_u64(&l0_proof, paths.n / 32);
        buf_put(&l0_proof, paths.p, paths.n);
        free(paths.p);
        or_f128* basis = induced_basis(n1, qs, nq, alpha, na);
        f256 enf = enforced_sum(rows, nq, t, lane_ch, (size_t)ik, alpha, na);
        (void)enf;
        size_t bn = (size_t)1 << n1;
        f256* sb = (f256*)malloc(sizeof(f256) * 2 * bn);
        const f256 U = {{0, 0}, {1, 0}};
        for (size_t i = 0; i < bn; i++) {
            sb[2 * i] = ef(basis[i]);
            sb[2 * i + 1] = em(U, ef(basis[i]));
        }
        sc_introduce(&s, sb, &u0, &u2);
        observe_msg256(fs, u0, u2);
        or_f128 beta = grind_sample(fs, C->lv[1].claim, &nonce);
        u64v_push(&claim_n, nonce);
        sc_glue(&s, beta);
        free(basis);
        free(rows);
        free(qs);
        free(w);
    }
    buf_t rec_proofs = {0};
    size_t n_rec = 0;
    buf_t final_proof = {0};
    for (int i = 0; i < R; i++) {
        int k = C->lv[i + 1].krec;
        f256 lch[16];
        for (int j = 0; j < k; j++) {
            f256 r = sample_f256(fs);
            lch[j] = r;
            int switching = (j + 1 == k) && (i + 1 != R);
            e_fold_blocked(s.f, s.n, 1, r);
            e_fold_blocked(s.b, s.n, 1, r);
            s.n /= 2;
            if (switching)
                sc_switch(&s, &u0, &u2);
            else
                e_msg_blocked(s.f, s.b, s.n, 1, &u0, &u2);
            sc_push(&s, u0, u2);
            observe_msg256(fs, u0, u2);
        }
        int ext_dim = split_dim - k;
        int level = i + 1;
        if (i + 1 == R) {
            for (size_t x = 0; x < s.n; x++) {
                f128v_push(&yr, s.f[x].c0);
                f128v_push(&yr, s.f[x].c1);
            }
            for (size_t x = 0; x < yr.n; x++) or_fs_observe_f128(fs, yr.p[x]);
            size_t nq = C->lv[level].queries;
            uint64_t nonce = or_fs_grind_pow(fs, C->lv[level].grind);
            u64v_push(&grind_n, nonce);
            or_f128* w = (or_f128*)malloc(sizeof(or_f128) * nq);
            or_fs_sample_f128_vec(fs, w, nq);
            size_t* qs = (size_t*)malloc(sizeof(size_t) * nq);
            queries_from_words(&sch[level], w, qs);
            size_t na = ceil_log2(nq);
            or_f128 alpha[64];
            nonce = or_fs_grind_pow(fs, C->lv[level].cons);
            u64v_push(&cons_n, nonce);
            or_fs_sample_f128_vec(fs, alpha, na);
            (void)grind_sample(fs, C->lv[level].claim, &nonce);
            u64v_push(&claim_n, nonce);
            or_f128* rows = (or_f128*)malloc(sizeof(or_f128) * nq * prev.lanes);
            buf_t paths = {0};
            open_rows(prev.mat, prev.lanes, prev.tree, prev.block_len, qs, nq,
                      sched_cap(&sch[level]), rows, &paths);
            bw_vec_f128(&final_proof, yr.p, yr.n);
            bw_rows(&final_proof, rows, nq, prev.lanes);
            bw_u64(&final_proof, paths.n / 32);
            buf_put(&final_proof, paths.p, paths.n);
            free(paths.p);
            free(rows);
            free(qs);
            free(w);
            break;
        }
        split_dim = ext_dim + 1;
        int next_level = i + 2;
        lw_t next;
        {
            or_f128* base = (or_f128*)malloc(sizeof(or_f128) * s.n);
            for (size_t x = 0; x < s.n; x++) base[x] = s.f[x].c0;
            int ll = C->lv[next_level].krec;
            if (ligero_commit(base, split_dim - ll, ll, C->lv[next_level].rate, kind, &next))
                return -12;
            free(base);
        }
        {
            int c = sched_cap(&sch[next_level]);
            const uint8_t* cp = cap_layer(next.tree, next.block_len, c);
            or_fs_observe_bytes(fs, cp, 32 * ((size_t)1 << c));
            bw_vec_hash(&caps, cp, (size_t)1 << c);
            n_caps++;
        }
        LEVEL_OOD(next_level)
        {
            size_t nq = C->lv[level].queries;
            uint64_t nonce = or_fs_grind_pow(fs, C->lv[level].grind);
            u64v_push(&grind_n, nonce);
            or_f128* w = (or_f128*)malloc(sizeof(or_f128) * nq);
            or_fs_sample_f128_vec(fs, w, nq);
            size_t* qs = (size_t*)malloc(sizeof(size_t) * nq);
            queries_from_words(&sch[level], w, qs);
            size_t na = ceil_log2(nq);
            or_f128 alpha[64];
            nonce = or_fs_grind_pow(fs, C->lv[level].cons);
            u64v_push(&cons_n, nonce);
            or_fs_sample_f128_vec(fs, alpha, na);
            or_f128* rows = (or_f128*)malloc(sizeof(or_f128) * nq * prev.lanes);
            buf_t paths = {0};
            open_rows(prev.mat, prev.lanes, prev.tree, prev.block_len, qs, nq,
                      sched_cap(&sch[level]), rows, &paths);
            bw_rows(&rec_proofs, rows, nq, prev.lanes);
            bw_u64(&rec_proofs, paths.n / 32);
            buf_put(&rec_proofs, paths.p, paths.n);
            n_rec++;
            free(paths.p);
            or_f128* basis = induced_basis(ext_dim, qs, nq, alpha, na);
            size_t bn = (size_t)1 << ext_dim;
            f256* sb = (f256*)malloc(sizeof(f256) * 2 * bn);
            const f256 U = {{0, 0}, {1, 0}};
            for (size_t x = 0; x < bn; x++) {
                sb[2 * x] = ef(basis[x]);
                sb[2 * x + 1] = em(U, ef(basis[x]));
            }
            sc_introduce(&s, sb, &u0, &u2);
            observe_msg256(fs, u0, u2);
            or_f128 beta = grind_sample(fs, C->lv[next_level].claim, &nonce);
            u64v_push(&claim_n, nonce);
            sc_glue(&s, beta);
            free(basis);
            free(rows);
            free(qs);
            free(w);
        }
        lw_free(&prev);
        prev = next;
    }
#undef LEVEL_OOD
    lw_free(&prev);
    /* BatchOpeningProofLigerito {ring_switches: [], batching_nonces: [], ligerito} */
    bw_u64(out, 0);
    bw_u64(out, 0);
    /* LigeritoProof (ligerito.rs:2177-2218) */
    bw_vec_hash(out, icap, (size_t)1 << c0);
    buf_put(out, l0_proof.p, l0_proof.n);
    bw_u64(out, n_caps);
    buf_put(out, caps.p, caps.n);
    bw_u64(out, n_rec);
    buf_put(out, rec_proofs.p, rec_proofs.n);
    buf_put(out, final_proof.p, final_proof.n);
    bw_u64(out, 0); /* sumcheck_transcript (F128): empty */
    bw_u64(out, s.n_msgs);
    buf_put(out, s.msgs.p, s.msgs.n);
    bw_vec_u64(out, grind_n.p, grind_n.n);
    bw_vec_f128(out, ood.p, ood.n);
    bw_u64(out, 0); /* fold_grinding_nonces */
    bw_vec_u64(out, claim_n.p, claim_n.n);
    bw_vec_u64(out, cons_n.p, cons_n.n);
    free(l0_proof.p);
    free(caps.p);
    free(rec_proofs.p);
    free(final_proof.p);
    free(s.msgs.p);
    free(s.f);
    free(s.b);
    free(ood.p);
    free(yr.p);
    free(claim_n.p);
    free(cons_n.p);
    free(grind_n.p);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* open_batch_merged (pcs.rs:1433-1822) for ring-switched boolean claims       */
/* ------------------------------------------------------------------------- */
int or_open_merged(or_fs* fs, const or_f128* q, const or_pcs_params* P, const or_f128* cw,
                   const uint8_t* tree, size_t n_claims, const or_f128* x_outers, size_t x_len,
                   const or_f128* s_hat_v, const uint64_t* heights, size_t n_heights,
                   unsigned n_log, uint8_t** out_bytes, size_t* out_len, or_f128* q_eval_out) {
    sparse_init();
    if (n_claims < 1 || n_claims > 8) return -1;
    const int dense_log = (int)P->m - 7;
    const size_t N = (size_t)1 << dense_log;
    jag_t J;
    jag_make(&J, heights, n_heights, (int)n_log, dense_log);
    if (x_len != 1 + n_log + (size_t)J.k) {
        jag_free(&J);
        return -2;
    }
    const uint64_t area = jag_area(&J);
    if (area > N) {
        jag_free(&J);
        return -3;
    }
    buf_t out = {0};
    static const uint8_t L_MO[] = "flock-merged-open-v1", L_RS[] = "flock-ring-switch-v0";
    or_fs_observe_label(fs, L_MO, sizeof(L_MO) - 1);
    /* ring switch per claim (ring_switch.rs:2723-2760), ring_switch_bits = 3 */
    or_f128 eqr[8][128], rs_claim[8];
    uint64_t rs_nonce[8];
    for (size_t c = 0; c < n_claims; c++) {
        const or_f128* sv = s_hat_v + 128 * c;
        or_fs_observe_label(fs, L_RS, sizeof(L_RS) - 1);
        or_fs_observe_f128_slice(fs, sv, 128);
        rs_nonce[c] = or_fs_grind_pow(fs, 3);
        or_f128 rdp[7];
        or_fs_sample_f128_vec(fs, rdp, 7);
        or_build_eq(rdp, 7, eqr[c]);
        /* sumcheck_claim = <transpose(s_hat_v), eq(r'')> */
        or_f128 acc = Z0;
        for (int b = 0; b < 128; b++) {
            or_f128 u = Z0;
            for (int i = 0; i < 128; i++)
                if (getbit(sv[i], b)) {
                    if (i < 64)
                        u.lo |= (uint64_t)1 << i;
                    else
                        u.hi |= (uint64_t)1 << (i - 64);
                }
            acc = fa(acc, fm(u, eqr[c][b]));
        }
        rs_claim[c] = acc;
    }
    /* batching coefficients: claim_batch_bits 6 when more than one claim */
    or_f128 gam[8];
    uint64_t batch_nonce = 0;
    const uint32_t bb = n_claims > 1 ? 6 : 0;
    batch_nonce = bb ? or_fs_grind_pow(fs, bb) : 0; /* pcs.rs:193-199, 1514-1521 */
    or_fs_sample_f128_vec(fs, gam, n_claims);
    or_f128 target = Z0;
    lin_t* phi = (lin_t*)malloc(sizeof(lin_t) * n_claims);
    or_f128 ge[8][128];
    for (size_t c = 0; c < n_claims; c++) {
        target = fa(target, fm(gam[c], rs_claim[c]));
        for (int b = 0; b < 128; b++) ge[c][b] = fm(gam[c], eqr[c][b]);
        lin_make(&phi[c], ge[c]); /* the γ-baked fold table (build_fold_byte_table) */
    }
    /* merged weight W[d] = Σ_c Φ_c(eq_row_c[row(d)] eq_col_c[col(d)]), d < area */
    or_f128* W = (or_f128*)calloc(N, sizeof(or_f128));
    or_f128* A = (or_f128*)malloc(sizeof(or_f128) * N);
    memcpy(A, q, sizeof(or_f128) * N);
    const size_t nc = (size_t)1 << J.k;
    for (size_t c = 0; c < n_claims; c++) {
        const or_f128* x = x_outers + c * x_len;
        or_f128* er = eq_table(x + 1, n_log);
        or_f128* ec = eq_table(x + 1 + n_log, (size_t)J.k);
        for (size_t y = 0; y < nc; y++)
            for (uint64_t d = J.ps[y]; d < J.ps[y + 1]; d++)
                W[d] = fa(W[d], lin_apply(&phi[c], fm(er[d - J.ps[y]], ec[y])));
        free(er);
        free(ec);
    }
    /* round-0 prime and the merged sumcheck (pcs.rs:1590-1700) */
    or_f128 u0 = Z0, u2 = Z0;
    for (size_t i = 0; i < N / 2; i++) {
        u0 = fa(u0, fm(A[2 * i], W[2 * i]));
        u2 = fa(u2, fm(fa(A[2 * i], A[2 * i + 1]), fa(W[2 * i], W[2 * i + 1])));
    }
    or_f128 g1 = fa(target, u0), gi = u2;
    or_f128* mrounds = (or_f128*)malloc(sizeof(or_f128) * 2 * (size_t)(dense_log + 1));
    uint64_t* mnon = (uint64_t*)malloc(8 * (size_t)(dense_log + 1));
    or_f128* rho = (or_f128*)malloc(sizeof(or_f128) * (size_t)(dense_log + 1));
    size_t cur = N;
    for (int r = 0; r < dense_log; r++) {
        or_fs_observe_f128(fs, g1);
        or_fs_observe_f128(fs, gi);
        or_f128 ch = grind_sample(fs, 2, &mnon[r]);
        mrounds[2 * r] = g1;
        mrounds[2 * r + 1] = gi;
        rho[r] = ch;
        fold_lsb(A, cur, ch);
        fold_lsb(W, cur, ch);
        cur /= 2;
        if (cur >= 2) msg_g1(A, W, cur, &g1, &gi);
    }
    const or_f128 q_eval = A[0];
    if (q_eval_out) *q_eval_out = q_eval;
    free(W);
    free(A);
    /* linearized coefficients of each γ-baked fold map: c_j = Σ_t Minv[j][t] Φ(e_t) */
    const or_f128* minv = moore_inverse();
    or_f128 coef[8][128];
    for (size_t c = 0; c < n_claims; c++)
        for (int j = 0; j < 128; j++) {
            or_f128 acc = Z0;
            for (int tt = 0; tt < 128; tt++) acc = fa(acc, fm(minv[j * 128 + tt], ge[c][tt]));
            coef[c][j] = acc;
        }
    (void)coef;
    /* fork the assist's transcript (challenger.rs:167-191) */
    or_f128 seed0 = or_fs_sample_f128(fs), seed1 = or_fs_sample_f128(fs);
    static const uint8_t L_PA[] = "flock-par-assist-v1";
    or_fs* ch = or_fs_new(L_PA, sizeof(L_PA) - 1);
    or_fs_observe_f128(ch, seed0);
    or_fs_observe_f128(ch, seed1);

    /* ---- multipoint twisted assist on the child (jagged.rs:2884-3290) ---- */
    const int m = dense_log;
    or_f128* rp = (or_f128*)malloc(sizeof(or_f128) * 128 * (size_t)m); /* rho^(2^-j) */
    memcpy(rp, rho, sizeof(or_f128) * (size_t)m);
    for (int j = 1; j < 128; j++)
        for (int i = 0; i < m; i++) rp[j * m + i] = frob_inv(rp[(j - 1) * m + i]);
    or_f128* vals = (or_f128*)malloc(sizeof(or_f128) * 128 * n_claims);
    for (size_t c = 0; c < n_claims; c++) {
        const or_f128* x = x_outers + c * x_len;
        or_f128* w = run_weights(&J, x + 1 + n_log);
        for (int j = 0; j < 128; j++) vals[c * 128 + j] = jag_fhat(&J, x + 1, w, rp + j * m);
        free(w);
    }
    static const uint8_t L_MP[] = "flock-multipoint-twisted-v1";
    or_fs_observe_label(ch, L_MP, sizeof(L_MP) - 1);
    for (size_t i = 0; i < 128 * n_claims; i++) or_fs_observe_f128(ch, vals[i]);
    uint32_t gbits = bits_for_degree(128 * n_claims - 1);
    if (gbits < 1) gbits = 1;
    uint64_t gnonce;
    or_f128 gamma = grind_sample(ch, gbits, &gnonce);
    or_f128* gpow = (or_f128*)malloc(sizeof(or_f128) * 128 * n_claims);
    gpow[0] = O1;
    for (size_t i = 1; i < 128 * n_claims; i++) gpow[i] = fm(gpow[i - 1], gamma);
    /* g = L(eq(rho, .)), L(x) = Σ_{j<128} γ^j x^(2^-j) */
    or_f128 images[128];
    for (int b = 0; b < 128; b++) {
        or_f128 acc = Z0, y = fbit(b);
        for (int j = 0; j < 128; j++) {
            acc = fa(acc, fm(gpow[j], y));
            y = frob_inv(y);
        }
        images[b] = acc;
    }
    lin_t* Lg = (lin_t*)malloc(sizeof(lin_t));
    lin_make(Lg, images);
    or_f128* G = eq_table(rho, (size_t)m);
#pragma omp parallel for schedule(static) if (N >= 8192)
    for (size_t i = 0; i < N; i++) G[i] = lin_apply(Lg, G[i]);
    free(Lg);
    /* ā = Σ_c γ^(128 c) eq_row_c ⊗ eq_col_c on the jagged area */
    or_f128* AB = (or_f128*)calloc(N, sizeof(or_f128));
    for (size_t c = 0; c < n_claims; c++) {
        const or_f128* x = x_outers + c * x_len;
        or_f128* er = eq_table(x + 1, n_log);
        or_f128* ec = eq_table(x + 1 + n_log, (size_t)J.k);
        jag_weight_acc(&J, er, ec, gpow[128 * c], AB);
        free(er);
        free(ec);
    }
    or_f128* trounds = (or_f128*)malloc(sizeof(or_f128) * 2 * (size_t)m);
    uint64_t* tnon = (uint64_t*)malloc(8 * (size_t)m);
    or_f128* point = (or_f128*)malloc(sizeof(or_f128) * (size_t)m);
    msg_g1(AB, G, N, &g1, &gi);
    cur = N;
    for (int i = 0; i < m; i++) {
        or_fs_observe_f128(ch, g1);
        or_fs_observe_f128(ch, gi);
        or_f128 r = grind_sample(ch, 2, &tnon[i]);
        trounds[2 * i] = g1;
        trounds[2 * i + 1] = gi;
        point[i] = r;
        if (i + 1 == m) break;
        fold_lsb(AB, cur, r);
        fold_lsb(G, cur, r);
        cur /= 2;
        msg_g1(AB, G, cur, &g1, &gi);
    }
    free(AB);
    free(G);
    /* anchor: untwisted assist at the endpoint (jagged.rs:3240-3285) */
    or_f128 g_at = Z0;
    for (int j = 0; j < 128; j++) g_at = fa(g_at, fm(gpow[j], eq_at(rp + j * m, point, (size_t)m)));
    or_f128* zr[8];
    or_f128* ww[8];
    size_t n_st = 0;
    for (size_t c = 0; c < n_claims; c++) {
        const or_f128* x = x_outers + c * x_len;
        or_f128 co = fm(gpow[128 * c], g_at);
        if (fz(co)) continue;
        or_f128* w = run_weights(&J, x + 1 + n_log);
        for (size_t b = 0; b < J.nb; b++) w[b] = fm(w[b], co);
        /* statements sharing a row point merge (frobenius_statements, :1760-1790) */
        size_t s;
        for (s = 0; s < n_st; s++)
            if (!memcmp(zr[s], x + 1, sizeof(or_f128) * n_log)) break;
        if (s < n_st) {
            for (size_t b = 0; b < J.nb; b++) ww[s][b] = fa(ww[s][b], w[b]);
            free(w);
        } else {
            zr[n_st] = (or_f128*)malloc(sizeof(or_f128) * (n_log ? n_log : 1));
            memcpy(zr[n_st], x + 1, sizeof(or_f128) * n_log);
            ww[n_st] = w;
            n_st++;
        }
    }
    frob_t anc;
    frobenius_assist(ch, &J, n_st, zr, ww, point, 2, &anc);
    for (size_t s = 0; s < n_st; s++) {
        free(zr[s]);
        free(ww[s]);
    }
    /* ---- inner open on the parent ---- */
    buf_t inner = {0};
    int rc = inner_open(fs, q, P, cw, tree, rho, q_eval, &inner);
    /* merge the child (challenger.rs:193-203) */
    or_f128 d0 = or_fs_sample_f128(ch), d1 = or_fs_sample_f128(ch);
    or_fs_observe_f128(fs, d0);
    or_fs_observe_f128(fs, d1);
    or_fs_free(ch);
    if (rc == 0) {
        /* MergedOpenProof (pcs.rs:1362-1377) */
        bw_u64(&out, n_claims);
        for (size_t c = 0; c < n_claims; c++) {
            bw_vec_f128(&out, s_hat_v + 128 * c, 128);
            bw_u64(&out, rs_nonce[c]);
        }
        bw_vec_u64(&out, &batch_nonce, bb ? 1 : 0);
        bw_u64(&out, (uint64_t)dense_log); /* merged_rounds: Vec<(F128, F128)> */
        for (int r = 0; r < dense_log; r++) {
            bw_f128(&out, mrounds[2 * r]);
            bw_f128(&out, mrounds[2 * r + 1]);
        }
        bw_vec_u64(&out, mnon, (size_t)dense_log);
        bw_f128(&out, q_eval);
        /* MultipointTwistedProof (jagged.rs:2790-2803) */
        bw_u64(&out, n_claims);
        for (size_t c = 0; c < n_claims; c++) bw_vec_f128(&out, vals + 128 * c, 128);
        bw_u64(&out, 0); /* group_values */
        bw_u64(&out, gnonce);
        bw_u64(&out, (uint64_t)m);
        for (int i = 0; i < m; i++) {
            bw_f128(&out, trounds[2 * i]);
            bw_f128(&out, trounds[2 * i + 1]);
        }
        bw_vec_u64(&out, tnon, (size_t)m);
        /* FrobeniusAssistProof (jagged.rs:1679-1686) */
        bw_f128(&out, anc.v);
        bw_u64(&out, anc.n_rounds);
        for (size_t i = 0; i < anc.n_rounds; i++) {
            bw_f128(&out, anc.rounds[2 * i]);
            bw_f128(&out, anc.rounds[2 * i + 1]);
        }
        bw_vec_u64(&out, anc.nonces, anc.n_nonces);
        buf_put(&out, inner.p, inner.n);
        *out_bytes = out.p;
        *out_len = out.n;
    } else {
        free(out.p);
    }
    free(inner.p);
    free(anc.rounds);
    free(anc.nonces);
    free(trounds);
    free(tnon);
    free(point);
    free(gpow);
    free(vals);
    free(rp);
    free(mrounds);
    free(mnon);
    free(rho);
    free(phi);
    jag_free(&J);
    return rc;
}
