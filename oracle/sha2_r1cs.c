/*
 * sha2_r1cs.c — CPU restatement of the SHA-256 compression R1CS over GF(2): the packed
 * (z, a, b) witness, the batch-major driver with the lincheck byte stripe, and the symbolic
 * row walk that defines (A_0, B_0).
 *
 * TEST INFRASTRUCTURE ONLY (part of liboracle.so; see flock_oracle.h).
 *
 * Reference: crates/flock-prover/src/r1cs_hashes/sha2.rs (layout :95-240, symbolic walk
 * :250-705, packed builder :896-1128), common.rs (parts helpers :95-214, batch-major driver
 * :700-960).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "flock_oracle.h"

/* ---- layout (sha2.rs:60-240) ---- */
enum {
    WB = 32,
    NR = 64,
    NSCHED = 48,
    CPA = 31, /* carries per add */
    RIPB = 30,
    RC_MAJ1 = 0,
    RC_MAJ2 = 31,
    RC_RIP = 62,
    RC_AMAJ = 92,
    RC_ARIP = 123,
    RC_ENEW = 153,
    RC_ADDK = 184,
    SC_MAJ1 = 0,
    SC_MAJ2 = 31,
    SC_RIP = 62,
    SCHED_ADD_BITS = 92,
    EA_PERIOD = 2,
    N_EA_SLOTS = 32,
    H_BASE = 0,
    H_OUT_BASE = 256,
    M_BASE = 512,
    CH_AND_BASE = 1024,
    MAJ_AND_BASE = 3072,
    ROUND_CARRY_BASE = 5120,
};

static const uint32_t SK[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

static unsigned k_seed_t(int r) { return (unsigned)__builtin_ctz(SK[r]) + 1; } /* :136-138 */
static unsigned k_add_rows(int r) { return CPA - k_seed_t(r); }                 /* :140-142 */
static unsigned round_base(int r) {                                              /* :157-170 */
    unsigned acc = ROUND_CARRY_BASE;
    for (int i = 0; i < r; i++) acc += RC_ADDK + k_add_rows(i);
    return acc;
}
static unsigned sched_carry_base(void) { return round_base(NR); }
static unsigned e_new_base(void) { return sched_carry_base() + NSCHED * SCHED_ADD_BITS; }
static unsigned a_new_base(void) { return e_new_base() + N_EA_SLOTS * WB; }
static unsigned out_carry_base(void) { return a_new_base() + N_EA_SLOTS * WB; }
static unsigned z_const_pos(void) { return out_carry_base() + 8 * CPA; }

void or_sha2_layout(uint32_t out[8]) {
    out[0] = sched_carry_base();
    out[1] = e_new_base();
    out[2] = a_new_base();
    out[3] = out_carry_base();
    out[4] = z_const_pos();
    out[5] = z_const_pos() + 1; /* USEFUL_BITS */
    out[6] = round_base(1);
    out[7] = 0;
}

static inline uint32_t ror(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
static uint32_t bs0(uint32_t x) { return ror(x, 2) ^ ror(x, 13) ^ ror(x, 22); }
static uint32_t bs1(uint32_t x) { return ror(x, 6) ^ ror(x, 11) ^ ror(x, 25); }
static uint32_t ss0(uint32_t x) { return ror(x, 7) ^ ror(x, 18) ^ (x >> 3); }
static uint32_t ss1(uint32_t x) { return ror(x, 17) ^ ror(x, 19) ^ (x >> 10); }

/* ---- parts helpers (common.rs:95-214) ---- */
#define LO31 0x7FFFFFFFu
#define LO30 0x3FFFFFFFu

static uint32_t add_parts(uint32_t x, uint32_t y, uint32_t* l, uint32_t* r, uint32_t* c) {
    uint32_t s = x + y, cin = s ^ x ^ y;
    *l = (x ^ cin) & LO31;
    *r = (y ^ cin) & LO31;
    *c = *l & *r;
    return s;
}

/* fused_add3_parts: maj[3] = (left, right, prod), rip[3] */
static uint32_t add3_parts(uint32_t x, uint32_t y, uint32_t m, uint32_t maj[3], uint32_t rip[3]) {
    uint32_t xm = x ^ m, ym = y ^ m, w = xm & ym, p = x ^ y ^ m, bw = (w ^ m) << 1;
    uint32_t s = p + bw, cin = s ^ p ^ bw, rl = p ^ cin, rr = bw ^ cin;
    maj[0] = xm & LO31;
    maj[1] = ym & LO31;
    maj[2] = w & LO31;
    rip[0] = (rl >> 1) & LO30;
    rip[1] = (rr >> 1) & LO30;
    rip[2] = ((rl & rr) >> 1) & LO30;
    return s;
}

static uint32_t add4_parts(uint32_t x, uint32_t y, uint32_t z, uint32_t w, uint32_t m1[3],
                           uint32_t m2[3], uint32_t rip[3]) {
    uint32_t xz = x ^ z, yz = y ^ z, w1 = xz & yz, p1 = x ^ y ^ z, b1 = (w1 ^ z) << 1;
    uint32_t pw = p1 ^ w, bw = b1 ^ w, w2 = pw & bw, p2 = p1 ^ b1 ^ w, b2 = (w2 ^ w) << 1;
    uint32_t s = p2 + b2, cin = s ^ p2 ^ b2, rl = p2 ^ cin, rr = b2 ^ cin;
    m1[0] = xz & LO31;
    m1[1] = yz & LO31;
    m1[2] = w1 & LO31;
    m2[0] = pw & LO31;
    m2[1] = bw & LO31;
    m2[2] = w2 & LO31;
    rip[0] = (rl >> 1) & LO30;
    rip[1] = (rr >> 1) & LO30;
    rip[2] = ((rl & rr) >> 1) & LO30;
    return s;
}

static uint32_t const_parts(uint32_t k, uint32_t y, uint32_t* l, uint32_t* r, uint32_t* p) {
    unsigned t = (unsigned)__builtin_ctz(k) + 1;
    uint32_t ll, rr, cc, s = add_parts(k, y, &ll, &rr, &cc);
    uint32_t mask = (1u << (31 - t)) - 1;
    *l = (ll >> t) & mask;
    *r = (rr >> t) & mask;
    *p = (cc >> t) & mask;
    return s;
}

static void orbits(uint64_t* buf, unsigned pos, uint32_t v, unsigned n) {
    /* bits [0, n) of v at bit position pos (word-wise: one or two words) */
    const uint64_t w = n >= 32 ? (uint64_t)v : ((uint64_t)v & ((1ull << n) - 1));
    const unsigned sh = pos & 63;
    buf[pos >> 6] |= w << sh;
    if (sh && sh + n > 64) buf[(pos >> 6) + 1] |= w >> (64 - sh);
}

static void put3(uint64_t* z, uint64_t* a, uint64_t* b, unsigned pos, const uint32_t tr[3],
                 unsigned n) {
    orbits(z, pos, tr[2], n);
    orbits(a, pos, tr[0], n);
    orbits(b, pos, tr[1], n);
}

/* build_block_ab_packed_into (sha2.rs:944-1128): z, a, b are 512 u64 each, zeroed here. */
void or_sha2_instance(const uint32_t h_in[8], const uint32_t m[16], uint64_t* z, uint64_t* a,
                      uint64_t* b) {
    memset(z, 0, 512 * 8);
    memset(a, 0, 512 * 8);
    memset(b, 0, 512 * 8);
    const unsigned zc = z_const_pos();
    orbits(z, zc, 1, 1);
    orbits(a, zc, 1, 1);
    orbits(b, zc, 1, 1);
    for (int w = 0; w < 8; w++) {
        orbits(z, H_BASE + 32 * w, h_in[w], 32);
        orbits(a, H_BASE + 32 * w, h_in[w], 32);
        orbits(b, H_BASE + 32 * w, 0xFFFFFFFFu, 32);
    }
    for (int i = 0; i < 16; i++) {
        orbits(z, M_BASE + 32 * i, m[i], 32);
        orbits(a, M_BASE + 32 * i, m[i], 32);
        orbits(b, M_BASE + 32 * i, 0xFFFFFFFFu, 32);
    }
    uint32_t ws[64];
    memcpy(ws, m, 64);
    const unsigned sb = sched_carry_base();
    for (int t = 16; t < 64; t++) {
        uint32_t m1[3], m2[3], rp[3];
        ws[t] = add4_parts(ss1(ws[t - 2]), ss0(ws[t - 15]), ws[t - 7], ws[t - 16], m1, m2, rp);
        const unsigned base = sb + (t - 16) * SCHED_ADD_BITS;
        put3(z, a, b, base + SC_MAJ1, m1, CPA);
        put3(z, a, b, base + SC_MAJ2, m2, CPA);
        put3(z, a, b, base + SC_RIP, rp, RIPB);
    }
    uint32_t A = h_in[0], B = h_in[1], C = h_in[2], D = h_in[3], E = h_in[4], F = h_in[5],
             G = h_in[6], H = h_in[7];
    const unsigned eb = e_new_base(), ab = a_new_base();
    for (int r = 0; r < NR; r++) {
        const uint32_t fg = F ^ G, ch = E & fg;
        orbits(z, CH_AND_BASE + 32 * r, ch, 32);
        orbits(a, CH_AND_BASE + 32 * r, E, 32);
        orbits(b, CH_AND_BASE + 32 * r, fg, 32);
        const uint32_t ba = B ^ A, ca = C ^ A, mj = ba & ca;
        orbits(z, MAJ_AND_BASE + 32 * r, mj, 32);
        orbits(a, MAJ_AND_BASE + 32 * r, ba, 32);
        orbits(b, MAJ_AND_BASE + 32 * r, ca, 32);
        const uint32_t ch_out = ch ^ G, maj_out = mj ^ A;
        const unsigned rb = round_base(r);
        uint32_t kl, kr, kp;
        const uint32_t hk = const_parts(SK[r], H, &kl, &kr, &kp);
        {
            const uint32_t tr[3] = {kl, kr, kp};
            put3(z, a, b, rb + RC_ADDK, tr, k_add_rows(r));
        }
        uint32_t m1[3], m2[3], rp[3];
        const uint32_t t1 = add4_parts(hk, bs1(E), ch_out, ws[r], m1, m2, rp);
        put3(z, a, b, rb + RC_MAJ1, m1, CPA);
        put3(z, a, b, rb + RC_MAJ2, m2, CPA);
        put3(z, a, b, rb + RC_RIP, rp, RIPB);
        uint32_t am[3], ar[3];
        const uint32_t a_new = add3_parts(t1, bs0(A), maj_out, am, ar);
        put3(z, a, b, rb + RC_AMAJ, am, CPA);
        put3(z, a, b, rb + RC_ARIP, ar, RIPB);
        uint32_t el, er, ep;
        const uint32_t e_new = add_parts(D, t1, &el, &er, &ep);
        {
            const uint32_t tr[3] = {el, er, ep};
            put3(z, a, b, rb + RC_ENEW, tr, CPA);
        }
        if (r % EA_PERIOD == EA_PERIOD - 1) {
            const unsigned oa = ab + WB * (r / EA_PERIOD), oe = eb + WB * (r / EA_PERIOD);
            orbits(z, oa, a_new, 32);
            orbits(a, oa, a_new, 32);
            orbits(b, oa, 0xFFFFFFFFu, 32);
            orbits(z, oe, e_new, 32);
            orbits(a, oe, e_new, 32);
            orbits(b, oe, 0xFFFFFFFFu, 32);
        }
        H = G;
        G = F;
        F = E;
        E = e_new;
        D = C;
        C = B;
        B = A;
        A = a_new;
    }
    const uint32_t fin[8] = {A, B, C, D, E, F, G, H};
    const unsigned ob = out_carry_base();
    for (int w = 0; w < 8; w++) {
        uint32_t l, r, c;
        const uint32_t s = add_parts(fin[w], h_in[w], &l, &r, &c);
        const uint32_t tr[3] = {l, r, c};
        put3(z, a, b, ob + w * CPA, tr, CPA);
        orbits(z, H_OUT_BASE + 32 * w, s, 32);
        orbits(a, H_OUT_BASE + 32 * w, s, 32);
        orbits(b, H_OUT_BASE + 32 * w, 0xFFFFFFFFu, 32);
    }
}

/* drive_witness_batch_major_partial + stripe_from_rows (common.rs:560-600, 805-960): declared
 * rows o < n_decl hold the compressions inputs[o] = (h_in[8], m[16]); rows past n_decl and chunk
 * columns >= 199 are zero. z, a, b: 2^(n_log + 8) F128 words at (chunk << n_log) + row; stripe:
 * 2^n_log * 4096 bytes, stripe[(row/8) * 32768 + bit] bit (row % 8) = z bit of that row. */
int or_sha2_witness_batch_major(const uint32_t* inputs, size_t n_decl, int n_log, or_f128* z,
                                or_f128* a, or_f128* b, uint8_t* stripe) {
    const size_t n = (size_t)1 << n_log;
    if (n_log < 3 || n_decl > n) return 1;
    const size_t K = 32768, words = K / 64;
    memset(z, 0, n * 256 * sizeof(or_f128));
    memset(a, 0, n * 256 * sizeof(or_f128));
    memset(b, 0, n * 256 * sizeof(or_f128));
    memset(stripe, 0, n * (K / 8));
    /* groups of 8 rows are independent (each owns its stripe bytes): one group per iteration */
    const size_t groups = (n_decl + 7) / 8;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(| : bad)
    for (size_t g = 0; g < groups; g++) {
        uint64_t* bz = (uint64_t*)malloc(8 * words * 8);
        uint64_t* ba = (uint64_t*)malloc(words * 8);
        uint64_t* bb = (uint64_t*)malloc(words * 8);
        if (!bz || !ba || !bb) {
            bad = 1;
        } else {
            const size_t o1 = (g + 1) * 8 < n_decl ? (g + 1) * 8 : n_decl;
            memset(bz, 0, 8 * words * 8);
            for (size_t o = g * 8; o < o1; o++) {
                uint64_t* zr = bz + (o - g * 8) * words;
                or_sha2_instance(inputs + 24 * o, inputs + 24 * o + 8, zr, ba, bb);
                for (size_t c = 0; c < 256; c++) {
                    const size_t idx = (c << n_log) + o;
                    z[idx].lo = zr[2 * c];
                    z[idx].hi = zr[2 * c + 1];
                    a[idx].lo = ba[2 * c];
                    a[idx].hi = ba[2 * c + 1];
                    b[idx].lo = bb[2 * c];
                    b[idx].hi = bb[2 * c + 1];
                }
            }
            /* stripe byte i of the group = bit i of its 8 rows (row r at bit r) */
            uint8_t* sg = stripe + g * K;
            for (size_t w = 0; w < words; w++) {
                uint64_t x[8];
                for (int r = 0; r < 8; r++) x[r] = bz[r * words + w];
                for (int i = 0; i < 64; i++) {
                    uint8_t v = 0;
                    for (int r = 0; r < 8; r++) v |= (uint8_t)(((x[r] >> i) & 1) << r);
                    sg[w * 64 + i] = v;
                }
            }
        }
        free(bz);
        free(ba);
        free(bb);
    }
    return bad;
}

/* ======================================================================== */
/* Symbolic rows (sha2.rs:250-705): supports are sorted, deduplicated       */
/* column lists; XOR of two supports is their symmetric difference.         */
/* ======================================================================== */
#include "r1cs_sup.h"

/* rot_xor3 / sigma_xor (:288-320): bit i = w[(i+r1)%32] ^ w[(i+r2)%32] ^ (w[i+s] or rot) */
static word word_sigma(const word* x, int r1, int r2, int r3, int is_shift) {
    word w;
    sup empty = {NULL, 0};
    for (int i = 0; i < 32; i++) {
        sup third;
        if (is_shift) third = (i + r3 < 32) ? x->b[i + r3] : empty;
        else third = x->b[(i + r3) % 32];
        w.b[i] = sup_xor3(x->b[(i + r1) % 32], x->b[(i + r2) % 32], third);
    }
    return w;
}

typedef struct {
    sup* a_rows;
    sup* b_rows;
    unsigned zc;
} sink;

static void emit(sink* s, unsigned slot, sup a, sup b) { /* MatSink::row, :352-357 */
    sup_free(&s->a_rows[slot]);
    sup_free(&s->b_rows[slot]);
    s->a_rows[slot] = sup_copy(a);
    s->b_rows[slot] = sup_copy(b);
}

/* add32_inline (:383-404): carries at slot(i) = base + i */
static word add32_inline(sink* s, const word* x, const word* y, unsigned base) {
    word sum;
    sup cin = sup_new(0);
    for (int i = 0; i < 32; i++) {
        sum.b[i] = sup_xor3(x->b[i], y->b[i], cin);
        if (i < CPA) {
            const unsigned slot = base + i;
            sup l = sup_xor(x->b[i], cin), r = sup_xor(y->b[i], cin);
            emit(s, slot, l, r);
            sup_free(&l);
            sup_free(&r);
            sup one = sup_one(slot), nc = sup_xor(cin, one);
            sup_free(&one);
            sup_free(&cin);
            cin = nc;
        }
    }
    sup_free(&cin);
    return sum;
}

/* add_const_inline (:406-433) */
static word add_const_inline(sink* s, uint32_t k, const word* y, unsigned base) {
    const unsigned t = (unsigned)__builtin_ctz(k) + 1;
    word sum;
    sup carry = sup_new(0);
    for (unsigned i = 0; i < 32; i++) {
        if (i == t) {
            sup_free(&carry);
            carry = sup_copy(y->b[t - 1]);
        }
        sup ks = ((k >> i) & 1) ? sup_one(s->zc) : sup_new(0);
        sum.b[i] = sup_xor3(ks, y->b[i], carry);
        if (i >= t && i < CPA) {
            const unsigned slot = base + (i - t);
            sup l = sup_xor(ks, carry), r = sup_xor(y->b[i], carry);
            emit(s, slot, l, r);
            sup_free(&l);
            sup_free(&r);
            sup one = sup_one(slot), nc = sup_xor(carry, one);
            sup_free(&one);
            sup_free(&carry);
            carry = nc;
        }
        sup_free(&ks);
    }
    sup_free(&carry);
    return sum;
}

/* csa_layer (:435-457): returns p, writes bw */
static word csa_layer(sink* s, const word* x, const word* y, const word* z, unsigned maj_base,
                      word* bw) {
    word p;
    *bw = word_zero();
    for (int i = 0; i < 32; i++) {
        p.b[i] = sup_xor3(x->b[i], y->b[i], z->b[i]);
        if (i < CPA) {
            const unsigned slot = maj_base + i;
            sup l = sup_xor(x->b[i], z->b[i]), r = sup_xor(y->b[i], z->b[i]);
            emit(s, slot, l, r);
            sup_free(&l);
            sup_free(&r);
            if (i + 1 < 32) {
                sup one = sup_one(slot);
                sup_free(&bw->b[i + 1]);
                bw->b[i + 1] = sup_xor(one, z->b[i]);
                sup_free(&one);
            }
        }
    }
    return p;
}

/* csa_ripple (:459-473) */
static word csa_ripple(sink* s, const word* p, const word* bw, unsigned rip_base) {
    word sum;
    sup g = sup_new(0);
    for (int i = 0; i < 32; i++) {
        sum.b[i] = sup_xor3(p->b[i], bw->b[i], g);
        if (i >= 1 && i <= RIPB) {
            const unsigned slot = rip_base + (i - 1);
            sup l = sup_xor(p->b[i], g), r = sup_xor(bw->b[i], g);
            emit(s, slot, l, r);
            sup_free(&l);
            sup_free(&r);
            sup one = sup_one(slot), ng = sup_xor(g, one);
            sup_free(&one);
            sup_free(&g);
            g = ng;
        }
    }
    sup_free(&g);
    return sum;
}

static word fused_add3(sink* s, const word* x, const word* y, const word* z, unsigned maj_base,
                       unsigned rip_base) {
    word bw, p = csa_layer(s, x, y, z, maj_base, &bw);
    word out = csa_ripple(s, &p, &bw, rip_base);
    word_free(&p);
    word_free(&bw);
    return out;
}

static word fused_add4(sink* s, const word* x, const word* y, const word* z, const word* w,
                       unsigned m1, unsigned m2, unsigned rip) {
    word b1, p1 = csa_layer(s, x, y, z, m1, &b1);
    word b2, p2 = csa_layer(s, &p1, &b1, w, m2, &b2);
    word out = csa_ripple(s, &p2, &b2, rip);
    word_free(&p1);
    word_free(&b1);
    word_free(&p2);
    word_free(&b2);
    return out;
}

/* materialize (:507-515): rows (raw[b]) * z[ZC] = z[slot_b]; returns the slot word */
static word materialize(sink* s, const word* raw, unsigned base) {
    word out;
    sup zc = sup_one(s->zc);
    for (int b = 0; b < 32; b++) {
        emit(s, base + b, raw->b[b], zc);
        out.b[b] = sup_one(base + b);
    }
    sup_free(&zc);
    return out;
}

/* walk_rows_cfg (:548-686) with the production configuration (EA materialized). */
static void walk_rows(sink* s) {
    const unsigned zc = s->zc;
    {
        sup one = sup_one(zc);
        emit(s, zc, one, one);
        for (int w = 0; w < 8; w++)
            for (int b = 0; b < 32; b++) {
                sup x = sup_one(H_BASE + 32 * w + b);
                emit(s, H_BASE + 32 * w + b, x, one);
                sup_free(&x);
            }
        for (int i = 0; i < 16; i++)
            for (int b = 0; b < 32; b++) {
                sup x = sup_one(M_BASE + 32 * i + b);
                emit(s, M_BASE + 32 * i + b, x, one);
                sup_free(&x);
            }
        sup_free(&one);
    }
    word h_in[8];
    word* w_arr = (word*)malloc(sizeof(word) * 64);
    for (int w = 0; w < 8; w++) h_in[w] = word_wire(H_BASE + 32 * w);
    for (int i = 0; i < 16; i++) w_arr[i] = word_wire(M_BASE + 32 * i);
    const unsigned sb = sched_carry_base();
    for (int t = 16; t < 64; t++) {
        word s1 = word_sigma(&w_arr[t - 2], 17, 19, 10, 1);
        word s0 = word_sigma(&w_arr[t - 15], 7, 18, 3, 1);
        const unsigned base = sb + (t - 16) * SCHED_ADD_BITS;
        w_arr[t] = fused_add4(s, &s1, &s0, &w_arr[t - 7], &w_arr[t - 16], base + SC_MAJ1,
                              base + SC_MAJ2, base + SC_RIP);
        word_free(&s1);
        word_free(&s0);
    }
    word st[8];
    for (int i = 0; i < 8; i++) st[i] = word_copy(&h_in[i]);
    const unsigned eb = e_new_base(), ab = a_new_base();
    for (int r = 0; r < NR; r++) {
        word ch_and, maj_and;
        for (int bit = 0; bit < 32; bit++) {
            const unsigned slot = CH_AND_BASE + 32 * r + bit;
            sup fg = sup_xor(st[5].b[bit], st[6].b[bit]);
            emit(s, slot, st[4].b[bit], fg);
            sup_free(&fg);
            ch_and.b[bit] = sup_one(slot);
        }
        for (int bit = 0; bit < 32; bit++) {
            const unsigned slot = MAJ_AND_BASE + 32 * r + bit;
            sup x = sup_xor(st[0].b[bit], st[1].b[bit]), y = sup_xor(st[0].b[bit], st[2].b[bit]);
            emit(s, slot, x, y);
            sup_free(&x);
            sup_free(&y);
            maj_and.b[bit] = sup_one(slot);
        }
        word ch_out = word_xor(&ch_and, &st[6]);
        word maj_out = word_xor(&maj_and, &st[0]);
        word_free(&ch_and);
        word_free(&maj_and);
        const unsigned rb = round_base(r);
        word hk = add_const_inline(s, SK[r], &st[7], rb + RC_ADDK);
        word S1 = word_sigma(&st[4], 6, 11, 25, 0);
        word t1 = fused_add4(s, &hk, &S1, &ch_out, &w_arr[r], rb + RC_MAJ1, rb + RC_MAJ2,
                             rb + RC_RIP);
        word S0 = word_sigma(&st[0], 2, 13, 22, 0);
        word a_raw = fused_add3(s, &t1, &S0, &maj_out, rb + RC_AMAJ, rb + RC_ARIP);
        const int mat = (r % EA_PERIOD) == EA_PERIOD - 1;
        word a_new, e_new;
        if (mat) {
            a_new = materialize(s, &a_raw, ab + WB * (r / EA_PERIOD));
            word_free(&a_raw);
        } else {
            a_new = a_raw;
        }
        word e_raw = add32_inline(s, &st[3], &t1, rb + RC_ENEW);
        if (mat) {
            e_new = materialize(s, &e_raw, eb + WB * (r / EA_PERIOD));
            word_free(&e_raw);
        } else {
            e_new = e_raw;
        }
        word_free(&hk);
        word_free(&S1);
        word_free(&S0);
        word_free(&t1);
        word_free(&ch_out);
        word_free(&maj_out);
        /* (a', b', c', d', e', f', g', h') = (A_NEW, a, b, c, E_NEW, e, f, g) */
        word_free(&st[7]);
        word_free(&st[3]);
        st[7] = st[6];
        st[6] = st[5];
        st[5] = st[4];
        st[4] = e_new;
        st[3] = st[2];
        st[2] = st[1];
        st[1] = st[0];
        st[0] = a_new;
    }
    const unsigned ob = out_carry_base();
    for (int w = 0; w < 8; w++) {
        word raw = add32_inline(s, &st[w], &h_in[w], ob + w * CPA);
        word slots = materialize(s, &raw, H_OUT_BASE + 32 * w);
        word_free(&raw);
        word_free(&slots);
    }
    for (int i = 0; i < 8; i++) {
        word_free(&st[i]);
        word_free(&h_in[i]);
    }
    for (int t = 0; t < 64; t++) word_free(&w_arr[t]);
    free(w_arr);
}

/* build_matrices (sha2.rs:688-701) in CSR form: rows of A_0 at a_cols[a_ptr[r]..a_ptr[r+1]).
 * Pass NULL column arrays to get only the nnz (a_ptr/b_ptr of 32769 entries are filled). */
int or_sha2_matrices(uint64_t* a_ptr, uint32_t* a_cols, uint64_t* b_ptr, uint32_t* b_cols) {
    const unsigned K = 32768;
    sink s;
    s.a_rows = (sup*)calloc(K, sizeof(sup));
    s.b_rows = (sup*)calloc(K, sizeof(sup));
    s.zc = z_const_pos();
    walk_rows(&s);
    a_ptr[0] = b_ptr[0] = 0;
    for (unsigned r = 0; r < K; r++) {
        a_ptr[r + 1] = a_ptr[r] + s.a_rows[r].n;
        b_ptr[r + 1] = b_ptr[r] + s.b_rows[r].n;
        if (a_cols) memcpy(a_cols + a_ptr[r], s.a_rows[r].v, sizeof(uint32_t) * s.a_rows[r].n);
        if (b_cols) memcpy(b_cols + b_ptr[r], s.b_rows[r].v, sizeof(uint32_t) * s.b_rows[r].n);
        sup_free(&s.a_rows[r]);
        sup_free(&s.b_rows[r]);
    }
    free(s.a_rows);
    free(s.b_rows);
    return 0;
}

/* CSR (row supports) -> CSC (csc_from_rows, lincheck.rs:367-388): rows listed per column in
 * increasing row order. */
void or_csr_to_csc(size_t k, const uint64_t* ptr, const uint32_t* cols, uint32_t* col_ptr,
                   uint32_t* rows) {
    memset(col_ptr, 0, sizeof(uint32_t) * (k + 1));
    for (size_t r = 0; r < k; r++)
        for (uint64_t e = ptr[r]; e < ptr[r + 1]; e++) col_ptr[cols[e] + 1]++;
    for (size_t c = 0; c < k; c++) col_ptr[c + 1] += col_ptr[c];
    uint32_t* next = (uint32_t*)malloc(sizeof(uint32_t) * (k + 1));
    memcpy(next, col_ptr, sizeof(uint32_t) * (k + 1));
    for (size_t r = 0; r < k; r++)
        for (uint64_t e = ptr[r]; e < ptr[r + 1]; e++) rows[next[cols[e]]++] = (uint32_t)r;
    free(next);
}

/* y = M z over GF(2) for one instance: y bit r = parity of z over row r's support. */
void or_csr_apply_bits(size_t k, const uint64_t* ptr, const uint32_t* cols, const uint64_t* z,
                       uint64_t* y) {
    memset(y, 0, (k / 64) * 8);
    for (size_t r = 0; r < k; r++) {
        unsigned p = 0;
        for (uint64_t e = ptr[r]; e < ptr[r + 1]; e++) p ^= (unsigned)(z[cols[e] >> 6] >> (cols[e] & 63)) & 1u;
        if (p) y[r >> 6] |= 1ull << (r & 63);
    }
}
