/*
 * blake3_r1cs.c — CPU restatement of the BLAKE3 compression R1CS over GF(2): the compression,
 * the packed (z, a, b) witness, the batch-major driver with the lincheck byte stripe, and the
 * symbolic row walk that defines (A_0, B_0).
 *
 * TEST INFRASTRUCTURE ONLY (part of liboracle.so; see flock_oracle.h).
 *
 * Reference: crates/flock-prover/src/r1cs_hashes/blake3.rs (layout :122-340, symbolic walk
 * :396-783, packed builder build_block_witness_ab_packed_into :1283-1452,
 * generate_witness_batch_major_partial :2072-2088), crates/flock-core/src/hash.rs:101-142
 * (blake3_compress), common.rs (parts helpers :95-142).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "flock_oracle.h"
#include "r1cs_sup.h"

enum {
    B3_CPA = 31,
    B3_RIP = 30,
    B3_FADD = 61,
    B3_CV = 0,
    B3_OUT_LO = 256,
    B3_M = 512,
    B3_T_LO = 1024,
    B3_T_HI = 1056,
    B3_BLEN = 1088,
    B3_FLAGS = 1120,
    B3_OUT_HI = 1152,
    B3_GS = 1408,
};

static const uint32_t B3IV[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                                 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
static const int B3PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
static const int B3LANES[8][4] = {{0, 4, 8, 12}, {1, 5, 9, 13}, {2, 6, 10, 14}, {3, 7, 11, 15},
                                  {0, 5, 10, 15}, {1, 6, 11, 12}, {2, 7, 8, 13}, {3, 4, 9, 14}};

static unsigned c1_rows(int g) { return g < 2 ? 30 : g < 4 ? 29 : 31; } /* g_c1_rows, :154-160 */
static unsigned g_bits(int g) { return 2 * B3_FADD + B3_CPA + c1_rows(g); }
static unsigned g_base(int g) {
    unsigned acc = B3_GS;
    for (int i = 0; i < g; i++) acc += g_bits(i);
    return acc;
}
static unsigned off_maj2(int g) { return B3_FADD + c1_rows(g); }
static unsigned off_rip2(int g) { return off_maj2(g) + B3_CPA; }
static unsigned off_c2(int g) { return off_rip2(g) + B3_RIP; }

void or_blake3_layout(uint32_t out[4]) {
    out[0] = g_base(56); /* Z_CONST_POS */
    out[1] = g_base(56) + 1;
    out[2] = g_base(4); /* end of the round-1 column G's */
    out[3] = g_base(1);
}

/* per_round_msg_idx (:364-389): msg[r][g] = (mx, my) indexes */
static void msg_idx(int out[7][8][2]) {
    int perm[16], next[16];
    for (int i = 0; i < 16; i++) perm[i] = i;
    for (int r = 0; r < 7; r++) {
        for (int g = 0; g < 8; g++) {
            out[r][g][0] = perm[2 * g];
            out[r][g][1] = perm[2 * g + 1];
        }
        for (int i = 0; i < 16; i++) next[i] = perm[B3PERM[i]];
        memcpy(perm, next, sizeof(perm));
    }
}

static inline uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

/* blake3_compress (crates/flock-core/src/hash.rs:101-142) */
void or_blake3_compress(const uint32_t cv[8], const uint32_t m[16], uint64_t counter,
                        uint32_t blen, uint32_t flags, uint32_t out[16]) {
    uint32_t s[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                      B3IV[0], B3IV[1], B3IV[2], B3IV[3], (uint32_t)counter,
                      (uint32_t)(counter >> 32), blen, flags};
    int mi[7][8][2];
    msg_idx(mi);
    for (int r = 0; r < 7; r++)
        for (int g = 0; g < 8; g++) {
            const int* l = B3LANES[g];
            uint32_t a = s[l[0]], b = s[l[1]], c = s[l[2]], d = s[l[3]];
            a = a + b + m[mi[r][g][0]];
            d = rotr32(d ^ a, 16);
            c = c + d;
            b = rotr32(b ^ c, 12);
            a = a + b + m[mi[r][g][1]];
            d = rotr32(d ^ a, 8);
            c = c + d;
            b = rotr32(b ^ c, 7);
            s[l[0]] = a;
            s[l[1]] = b;
            s[l[2]] = c;
            s[l[3]] = d;
        }
    for (int i = 0; i < 8; i++) {
        s[i] ^= s[i + 8];
        s[i + 8] ^= cv[i];
    }
    memcpy(out, s, sizeof(s));
}

#define LO31 0x7FFFFFFFu
#define LO30 0x3FFFFFFFu

static uint32_t add_parts(uint32_t x, uint32_t y, uint32_t t[3]) { /* add_carry_parts */
    const uint32_t s = x + y, cin = s ^ x ^ y;
    t[0] = (x ^ cin) & LO31;
    t[1] = (y ^ cin) & LO31;
    t[2] = t[0] & t[1];
    return s;
}

static uint32_t add3_parts(uint32_t x, uint32_t y, uint32_t m, uint32_t maj[3], uint32_t rip[3]) {
    const uint32_t xm = x ^ m, ym = y ^ m, w = xm & ym, p = x ^ y ^ m, bw = (w ^ m) << 1;
    const uint32_t s = p + bw, cin = s ^ p ^ bw, rl = p ^ cin, rr = bw ^ cin;
    maj[0] = xm & LO31;
    maj[1] = ym & LO31;
    maj[2] = w & LO31;
    rip[0] = (rl >> 1) & LO30;
    rip[1] = (rr >> 1) & LO30;
    rip[2] = ((rl & rr) >> 1) & LO30;
    return s;
}

static void orb(uint64_t* buf, unsigned pos, uint32_t v, unsigned n) {
    for (unsigned i = 0; i < n; i++)
        if ((v >> i) & 1) buf[(pos + i) >> 6] |= 1ull << ((pos + i) & 63);
}
/* (z, a, b) = (t[2], t[0], t[1]) over n bits at pos */
static void put3(uint64_t* z, uint64_t* a, uint64_t* b, unsigned pos, const uint32_t t[3],
                 unsigned n) {
    orb(z, pos, t[2], n);
    orb(a, pos, t[0], n);
    orb(b, pos, t[1], n);
}
static void put_lin(uint64_t* z, uint64_t* a, uint64_t* b, unsigned pos, uint32_t v) {
    orb(z, pos, v, 32);
    orb(a, pos, v, 32);
    orb(b, pos, 0xFFFFFFFFu, 32);
}

/* build_block_witness_ab_packed_into (blake3.rs:1299-1452): 256 u64 each, zeroed here. */
void or_blake3_instance(const uint32_t cv[8], const uint32_t m[16], uint64_t counter,
                        uint32_t blen, uint32_t flags, uint64_t* z, uint64_t* a, uint64_t* b) {
    memset(z, 0, 256 * 8);
    memset(a, 0, 256 * 8);
    memset(b, 0, 256 * 8);
    const unsigned zc = g_base(56);
    orb(z, zc, 1, 1);
    orb(a, zc, 1, 1);
    orb(b, zc, 1, 1);
    for (int w = 0; w < 8; w++) put_lin(z, a, b, B3_CV + 32 * w, cv[w]);
    for (int i = 0; i < 16; i++) put_lin(z, a, b, B3_M + 32 * i, m[i]);
    put_lin(z, a, b, B3_T_LO, (uint32_t)counter);
    put_lin(z, a, b, B3_T_HI, (uint32_t)(counter >> 32));
    put_lin(z, a, b, B3_BLEN, blen);
    put_lin(z, a, b, B3_FLAGS, flags);
    uint32_t s[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                      B3IV[0], B3IV[1], B3IV[2], B3IV[3], (uint32_t)counter,
                      (uint32_t)(counter >> 32), blen, flags};
    int mi[7][8][2];
    msg_idx(mi);
    for (int r = 0; r < 7; r++)
        for (int gi = 0; gi < 8; gi++) {
            const int g = 8 * r + gi;
            const int* l = B3LANES[gi];
            const unsigned gb = g_base(g);
            uint32_t maj1[3], rip1[3], maj2[3], rip2[3], t1[3], t2[3];
            const uint32_t a1 = add3_parts(s[l[0]], s[l[1]], m[mi[r][gi][0]], maj1, rip1);
            put3(z, a, b, gb, maj1, B3_CPA);
            put3(z, a, b, gb + B3_CPA, rip1, B3_RIP);
            const uint32_t d1 = rotr32(s[l[3]] ^ a1, 16);
            uint32_t c1;
            if (g < 4) {
                c1 = add_parts(B3IV[g], d1, t1);
                const unsigned n = c1_rows(g), t = B3_CPA - n;
                const uint32_t mask = (1u << n) - 1;
                const uint32_t tt[3] = {(t1[0] >> t) & mask, (t1[1] >> t) & mask,
                                        (t1[2] >> t) & mask};
                put3(z, a, b, gb + B3_FADD, tt, n);
            } else {
                c1 = add_parts(s[l[2]], d1, t1);
                put3(z, a, b, gb + B3_FADD, t1, B3_CPA);
            }
            const uint32_t b1 = rotr32(s[l[1]] ^ c1, 12);
            const uint32_t a2 = add3_parts(a1, b1, m[mi[r][gi][1]], maj2, rip2);
            const uint32_t d2 = rotr32(d1 ^ a2, 8);
            const uint32_t c2 = add_parts(c1, d2, t2);
            put3(z, a, b, gb + off_maj2(g), maj2, B3_CPA);
            put3(z, a, b, gb + off_rip2(g), rip2, B3_RIP);
            put3(z, a, b, gb + off_c2(g), t2, B3_CPA);
            s[l[0]] = a2;
            s[l[1]] = rotr32(b1 ^ c2, 7);
            s[l[2]] = c2;
            s[l[3]] = d2;
        }
    for (int w = 0; w < 8; w++) {
        put_lin(z, a, b, B3_OUT_LO + 32 * w, s[w] ^ s[w + 8]);
        put_lin(z, a, b, B3_OUT_HI + 32 * w, s[w + 8] ^ cv[w]);
    }
}

/* generate_witness_batch_major_partial (blake3.rs:2072-2088 over common.rs:805-960): inputs
 * 28 u32 per declared row = (cv[8], m[16], counter_lo, counter_hi, block_len, flags); rows past
 * n_decl are zero. z, a, b: 2^(n_log + 7) F128 words; stripe 2^n_log * 2048 bytes. */
int or_blake3_witness_batch_major(const uint32_t* inputs, size_t n_decl, int n_log, or_f128* z,
                                  or_f128* a, or_f128* b, uint8_t* stripe) {
    const size_t n = (size_t)1 << n_log, K = 16384;
    if (n_log < 3 || n_decl > n) return 1;
    memset(z, 0, n * 128 * sizeof(or_f128));
    memset(a, 0, n * 128 * sizeof(or_f128));
    memset(b, 0, n * 128 * sizeof(or_f128));
    memset(stripe, 0, n * (K / 8));
    uint64_t bz[256], ba[256], bb[256];
    for (size_t o = 0; o < n_decl; o++) {
        const uint32_t* in = inputs + 28 * o;
        or_blake3_instance(in, in + 8, (uint64_t)in[24] | ((uint64_t)in[25] << 32), in[26],
                           in[27], bz, ba, bb);
        for (size_t c = 0; c < 128; c++) {
            const size_t idx = (c << n_log) + o;
            z[idx].lo = bz[2 * c];
            z[idx].hi = bz[2 * c + 1];
            a[idx].lo = ba[2 * c];
            a[idx].hi = ba[2 * c + 1];
            b[idx].lo = bb[2 * c];
            b[idx].hi = bb[2 * c + 1];
        }
        for (size_t i = 0; i < K; i++)
            if ((bz[i >> 6] >> (i & 63)) & 1) stripe[(o / 8) * K + i] |= (uint8_t)(1u << (o % 8));
    }
    return 0;
}

/* ---- the symbolic walk (blake3.rs:396-783) ---- */
typedef struct {
    sup* ar;
    sup* br;
    unsigned zc;
} b3sink;

static void b3_emit(b3sink* s, unsigned slot, sup a, sup b) {
    sup_free(&s->ar[slot]);
    sup_free(&s->br[slot]);
    s->ar[slot] = sup_copy(a);
    s->br[slot] = sup_copy(b);
}

/* sup ^= {slot} */
static void sup_toggle(sup* x, unsigned slot) {
    sup one = sup_one(slot), t = sup_xor(*x, one);
    sup_free(&one);
    sup_free(x);
    *x = t;
}

static word word_rotr(const word* w, int n) {
    word o;
    for (int i = 0; i < 32; i++) o.b[i] = sup_copy(w->b[(i + n) % 32]);
    return o;
}

static word word_const(uint32_t v, unsigned zc) {
    word o;
    for (int i = 0; i < 32; i++) o.b[i] = ((v >> i) & 1) ? sup_one(zc) : sup_new(0);
    return o;
}

/* write_add_carry_rows + add_sum (:448-461, 523-545): cin[i] = {base + j : j < i} */
static word add_rows(b3sink* s, const word* x, const word* y, unsigned base) {
    word sum;
    sup cin = sup_new(0);
    for (int i = 0; i < 32; i++) {
        sum.b[i] = sup_xor3(x->b[i], y->b[i], cin);
        if (i < B3_CPA) {
            sup l = sup_xor(x->b[i], cin), r = sup_xor(y->b[i], cin);
            b3_emit(s, base + i, l, r);
            sup_free(&l);
            sup_free(&r);
            sup_toggle(&cin, base + i);
        }
    }
    sup_free(&cin);
    return sum;
}

/* write_fused_add_rows + fused_add_sum (:462-481, 557-604) */
static word fadd_rows(b3sink* s, const word* x, const word* y, const word* m, unsigned maj,
                      unsigned rip) {
    for (int i = 0; i < B3_CPA; i++) {
        sup l = sup_xor(x->b[i], m->b[i]), r = sup_xor(y->b[i], m->b[i]);
        b3_emit(s, maj + i, l, r);
        sup_free(&l);
        sup_free(&r);
    }
    sup g = sup_new(0); /* {rip + l - 1 : 1 <= l < j} */
    for (int j = 1; j <= B3_RIP; j++) {
        sup p = sup_xor3(x->b[j], y->b[j], m->b[j]);
        sup l = sup_xor(p, g);
        sup one = sup_one(maj + j - 1), bw = sup_xor(one, m->b[j - 1]);
        sup r = sup_xor(bw, g);
        b3_emit(s, rip + j - 1, l, r);
        sup_free(&p);
        sup_free(&l);
        sup_free(&one);
        sup_free(&bw);
        sup_free(&r);
        sup_toggle(&g, rip + j - 1);
    }
    sup_free(&g);
    word sum;
    sup acc = sup_new(0); /* {rip + l - 1 : 1 <= l < min(i, 31)} */
    for (int i = 0; i < 32; i++) {
        sup v = sup_xor3(x->b[i], y->b[i], m->b[i]);
        if (i >= 1) {
            sup one = sup_one(maj + i - 1), t1 = sup_xor(v, one), t2 = sup_xor(t1, m->b[i - 1]);
            sup t3 = sup_xor(t2, acc);
            sup_free(&v);
            sup_free(&one);
            sup_free(&t1);
            sup_free(&t2);
            v = t3;
            if (i <= B3_RIP) sup_toggle(&acc, rip + i - 1);
        }
        sum.b[i] = v;
    }
    sup_free(&acc);
    return sum;
}

/* write_const_add_rows + const_add_sum (:482-509, 606-636) */
static word cadd_rows(b3sink* s, uint32_t k, const word* y, unsigned base, unsigned n_rows) {
    const unsigned t = B3_CPA - n_rows;
    sup carry = sup_copy(y->b[t - 1]); /* carry into bit i >= t */
    sup zc = sup_one(s->zc), empty = sup_new(0);
    for (unsigned i = t; i < 31; i++) {
        sup l = sup_xor(carry, ((k >> i) & 1) ? zc : empty);
        sup r = sup_xor(y->b[i], carry);
        b3_emit(s, base + i - t, l, r);
        sup_free(&l);
        sup_free(&r);
        sup_toggle(&carry, base + i - t);
    }
    word sum;
    sup c2 = sup_copy(y->b[t - 1]);
    for (unsigned i = 0; i < 32; i++) {
        sup v = sup_xor(y->b[i], ((k >> i) & 1) ? zc : empty);
        if (i >= t) {
            sup w = sup_xor(v, c2);
            sup_free(&v);
            v = w;
            if (i < 31) sup_toggle(&c2, base + i - t);
        }
        sum.b[i] = v;
    }
    sup_free(&c2);
    sup_free(&carry);
    sup_free(&zc);
    sup_free(&empty);
    return sum;
}

/* build_matrices (:659-783) in CSR form; a_ptr/b_ptr have 16385 entries. NULL column arrays
 * return only the counts. */
int or_blake3_matrices(uint64_t* a_ptr, uint32_t* a_cols, uint64_t* b_ptr, uint32_t* b_cols) {
    const unsigned K = 16384;
    b3sink s;
    s.ar = (sup*)calloc(K, sizeof(sup));
    s.br = (sup*)calloc(K, sizeof(sup));
    s.zc = g_base(56);
    {
        sup one = sup_one(s.zc);
        b3_emit(&s, s.zc, one, one);
        const unsigned bases[6] = {B3_CV, B3_M, B3_T_LO, B3_T_HI, B3_BLEN, B3_FLAGS};
        const unsigned lens[6] = {256, 512, 32, 32, 32, 32};
        for (int k = 0; k < 6; k++)
            for (unsigned j = 0; j < lens[k]; j++) {
                sup x = sup_one(bases[k] + j);
                b3_emit(&s, bases[k] + j, x, one);
                sup_free(&x);
            }
        sup_free(&one);
    }
    word st[16];
    for (int w = 0; w < 8; w++) st[w] = word_wire(B3_CV + 32 * w);
    for (int i = 0; i < 4; i++) st[8 + i] = word_const(B3IV[i], s.zc);
    st[12] = word_wire(B3_T_LO);
    st[13] = word_wire(B3_T_HI);
    st[14] = word_wire(B3_BLEN);
    st[15] = word_wire(B3_FLAGS);
    int mi[7][8][2];
    msg_idx(mi);
    for (int r = 0; r < 7; r++)
        for (int gi = 0; gi < 8; gi++) {
            const int g = 8 * r + gi;
            const int* l = B3LANES[gi];
            const unsigned gb = g_base(g);
            word mx = word_wire(B3_M + 32 * mi[r][gi][0]), my = word_wire(B3_M + 32 * mi[r][gi][1]);
            word a1 = fadd_rows(&s, &st[l[0]], &st[l[1]], &mx, gb, gb + B3_CPA);
            word t = word_xor(&st[l[3]], &a1), d1 = word_rotr(&t, 16);
            word_free(&t);
            word c1 = g < 4 ? cadd_rows(&s, B3IV[g], &d1, gb + B3_FADD, c1_rows(g))
                            : add_rows(&s, &st[l[2]], &d1, gb + B3_FADD);
            t = word_xor(&st[l[1]], &c1);
            word b1 = word_rotr(&t, 12);
            word_free(&t);
            word a2 = fadd_rows(&s, &a1, &b1, &my, gb + off_maj2(g), gb + off_rip2(g));
            t = word_xor(&d1, &a2);
            word d2 = word_rotr(&t, 8);
            word_free(&t);
            word c2 = add_rows(&s, &c1, &d2, gb + off_c2(g));
            t = word_xor(&b1, &c2);
            word bn = word_rotr(&t, 7);
            word_free(&t);
            word_free(&mx);
            word_free(&my);
            word_free(&a1);
            word_free(&d1);
            word_free(&c1);
            word_free(&b1);
            for (int q = 0; q < 4; q++) word_free(&st[l[q]]);
            st[l[0]] = a2;
            st[l[1]] = bn;
            st[l[2]] = c2;
            st[l[3]] = d2;
        }
    {
        sup one = sup_one(s.zc);
        for (int w = 0; w < 8; w++) {
            word lo = word_xor(&st[w], &st[w + 8]);
            word cvw = word_wire(B3_CV + 32 * w);
            word hi = word_xor(&st[w + 8], &cvw);
            for (int i = 0; i < 32; i++) {
                b3_emit(&s, B3_OUT_LO + 32 * w + i, lo.b[i], one);
                b3_emit(&s, B3_OUT_HI + 32 * w + i, hi.b[i], one);
            }
            word_free(&lo);
            word_free(&cvw);
            word_free(&hi);
        }
        sup_free(&one);
    }
    for (int i = 0; i < 16; i++) word_free(&st[i]);
    a_ptr[0] = b_ptr[0] = 0;
    for (unsigned r = 0; r < K; r++) {
        a_ptr[r + 1] = a_ptr[r] + s.ar[r].n;
        b_ptr[r + 1] = b_ptr[r] + s.br[r].n;
        if (a_cols) memcpy(a_cols + a_ptr[r], s.ar[r].v, sizeof(uint32_t) * s.ar[r].n);
        if (b_cols) memcpy(b_cols + b_ptr[r], s.br[r].v, sizeof(uint32_t) * s.br[r].n);
        sup_free(&s.ar[r]);
        sup_free(&s.br[r]);
    }
    free(s.ar);
    free(s.br);
    return 0;
}
