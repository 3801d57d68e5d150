/*
 * keccak_r1cs.c — CPU restatement of the Keccak-f[1600] R1CS witness: the packed (z, a, b)
 * builder and the batch-major driver with the lincheck byte stripe.
 *
 * TEST INFRASTRUCTURE ONLY (part of liboracle.so; see flock_oracle.h).
 *
 * Reference: crates/flock-prover/src/r1cs_hashes/keccak.rs (layout :436-510, lane
 * primitives :184-330, packed builder build_chain_witness_ab_packed_into :565-643,
 * generate_witness_batch_major :1228-1285).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "flock_oracle.h"

static const unsigned RHO[5][5] = {
    {0, 36, 3, 41, 18}, {1, 44, 10, 45, 2}, {62, 6, 43, 15, 61}, {28, 55, 25, 21, 56},
    {27, 20, 39, 8, 14}};
static const uint64_t RC[24] = {
    0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
    0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
    0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
    0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
    0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
    0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};

static uint64_t rotl(uint64_t v, unsigned r) { return r ? (v << r) | (v >> (64 - r)) : v; }

/* theta_lanes + rho_pi_lanes (keccak.rs:224-281): B = (pi o rho o theta)(s) */
static void theta_rho_pi(const uint64_t s_in[25], uint64_t b[25]) {
    uint64_t s[25], c[5], d[5];
    memcpy(s, s_in, sizeof(s));
    for (int x = 0; x < 5; x++) c[x] = s[x] ^ s[x + 5] ^ s[x + 10] ^ s[x + 15] ^ s[x + 20];
    for (int x = 0; x < 5; x++) d[x] = c[(x + 4) % 5] ^ rotl(c[(x + 1) % 5], 1);
    for (int i = 0; i < 25; i++) s[i] ^= d[i % 5];
    for (int y = 0; y < 5; y++)
        for (int x = 0; x < 5; x++) {
            const int a = (x + 3 * y) % 5, bb = x;
            b[x + 5 * y] = rotl(s[a + 5 * bb], RHO[a][bb] % 64);
        }
}

void or_keccak_f(uint64_t s[25]) {
    for (int r = 0; r < 24; r++) {
        uint64_t b[25];
        theta_rho_pi(s, b);
        for (int y = 0; y < 5; y++)
            for (int x = 0; x < 5; x++)
                s[x + 5 * y] = b[x + 5 * y] ^ (~b[(x + 1) % 5 + 5 * y] & b[(x + 2) % 5 + 5 * y]);
        s[0] ^= RC[r];
    }
}

/* build_chain_witness_ab_packed_into (keccak.rs:565-643): 1024 u64 each, zeroed here.
 * u64 words: state_0 lanes at 0..24, state_24 lanes at 32..56, the constant at word 64 (bit 0),
 * t_r lanes at 65 + 25 r + lane. */
void or_keccak_instance(const uint64_t initial[25], uint64_t* z, uint64_t* a, uint64_t* b) {
    memset(z, 0, 1024 * 8);
    memset(a, 0, 1024 * 8);
    memset(b, 0, 1024 * 8);
    z[64] = a[64] = b[64] = 1;
    uint64_t s[25];
    memcpy(s, initial, sizeof(s));
    for (int l = 0; l < 25; l++) {
        z[l] = a[l] = s[l];
        b[l] = ~0ull;
    }
    for (int r = 0; r < 24; r++) {
        uint64_t bs[25], t[25];
        theta_rho_pi(s, bs);
        for (int y = 0; y < 5; y++)
            for (int x = 0; x < 5; x++) {
                const int i = x + 5 * y, i1 = (x + 1) % 5 + 5 * y, i2 = (x + 2) % 5 + 5 * y;
                t[i] = ~bs[i1] & bs[i2];
                z[65 + 25 * r + i] = t[i];
                a[65 + 25 * r + i] = ~bs[i1];
                b[65 + 25 * r + i] = bs[i2];
            }
        for (int i = 0; i < 25; i++) s[i] = bs[i] ^ t[i];
        s[0] ^= RC[r];
    }
    for (int l = 0; l < 25; l++) {
        z[32 + l] = a[32 + l] = s[l];
        b[32 + l] = ~0ull;
    }
}

/* generate_witness_batch_major (keccak.rs:1228-1285): rows o < n_decl take inputs[25 o ..];
 * the other rows run keccak_f on the zero state (padding invocations). z, a, b: 2^(n_log + 9)
 * F128 words at (chunk << n_log) + row; stripe: 2^n_log * 8192 bytes. */
int or_keccak_witness_batch_major(const uint64_t* inputs, size_t n_decl, int n_log, or_f128* z,
                                  or_f128* a, or_f128* b, uint8_t* stripe) {
    const size_t n = (size_t)1 << n_log, K = 65536;
    if (n_log < 3 || n_decl > n) return 1;
    uint64_t* bz = (uint64_t*)malloc(1024 * 8);
    uint64_t* ba = (uint64_t*)malloc(1024 * 8);
    uint64_t* bb = (uint64_t*)malloc(1024 * 8);
    const uint64_t zero[25] = {0};
    memset(stripe, 0, n * (K / 8));
    for (size_t o = 0; o < n; o++) {
        or_keccak_instance(o < n_decl ? inputs + 25 * o : zero, bz, ba, bb);
        for (size_t c = 0; c < 512; c++) {
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
    free(bz);
    free(ba);
    free(bb);
    return 0;
}
