/*
 * r1cs_sup.h — sparse XOR supports for the oracle's symbolic R1CS row walks (test
 * infrastructure; included by sha2_r1cs.c and blake3_r1cs.c). A support is a sorted,
 * duplicate-free column list; the XOR of two supports is their symmetric difference, the
 * char-2 cancellation of crates/flock-prover/src/r1cs_hashes/common.rs xor_dedup.
 */
#ifndef FLOCK_ORACLE_R1CS_SUP_H
#define FLOCK_ORACLE_R1CS_SUP_H
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint32_t* v;
    uint32_t n;
} sup;

static inline sup sup_new(uint32_t cap) {
    sup s;
    s.v = (uint32_t*)malloc(sizeof(uint32_t) * (cap ? cap : 1));
    s.n = 0;
    return s;
}
static inline sup sup_one(uint32_t x) {
    sup s = sup_new(1);
    s.v[0] = x;
    s.n = 1;
    return s;
}
static inline sup sup_copy(sup a) {
    sup s = sup_new(a.n);
    memcpy(s.v, a.v, sizeof(uint32_t) * a.n);
    s.n = a.n;
    return s;
}
static inline void sup_free(sup* s) {
    free(s->v);
    s->v = NULL;
    s->n = 0;
}
static inline sup sup_xor(sup a, sup b) { /* xor_sup, :260-278 */
    sup o = sup_new(a.n + b.n);
    uint32_t i = 0, j = 0;
    while (i < a.n && j < b.n) {
        if (a.v[i] < b.v[j]) o.v[o.n++] = a.v[i++];
        else if (a.v[i] > b.v[j]) o.v[o.n++] = b.v[j++];
        else {
            i++;
            j++;
        }
    }
    while (i < a.n) o.v[o.n++] = a.v[i++];
    while (j < b.n) o.v[o.n++] = b.v[j++];
    return o;
}
static inline sup sup_xor3(sup a, sup b, sup c) {
    sup t = sup_xor(a, b), o = sup_xor(t, c);
    sup_free(&t);
    return o;
}

typedef struct {
    sup b[32];
} word;

static inline void word_free(word* w) {
    for (int i = 0; i < 32; i++) sup_free(&w->b[i]);
}
static inline word word_zero(void) {
    word w;
    for (int i = 0; i < 32; i++) w.b[i] = sup_new(0);
    return w;
}
static inline word word_copy(const word* x) {
    word w;
    for (int i = 0; i < 32; i++) w.b[i] = sup_copy(x->b[i]);
    return w;
}
static inline word word_wire(unsigned base) {
    word w;
    for (int i = 0; i < 32; i++) w.b[i] = sup_one(base + i);
    return w;
}
static inline word word_xor(const word* x, const word* y) {
    word w;
    for (int i = 0; i < 32; i++) w.b[i] = sup_xor(x->b[i], y->b[i]);
    return w;
}

#endif
