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

/* ---- lincheck column marginals (KeccakLincheckCircuit::fold_alpha_batched, keccak.rs:692-835)
 *
 * sa = A_0^T eq and sb = B_0^T eq over the K = 2^16 columns (fold_split, lincheck.rs:258-266),
 * so fold_alpha_batched = alpha * sa + sb. A_0 and B_0 are never formed: state_r (r >= 1) is
 * the linear form phi(state_{r-1}) + t_{r-1} + RC_{r-1}, so the coefficient vector on state_r
 * is carried back round by round (K_{r-1} = phi^T(K_r) + chi_{r-1}) and dropped onto the
 * t_{r-1} columns, and the constant parts (NOT of chi's A operand, the propagated round
 * constants) land on the constant column. */
#define KC_K (1u << 16)
#define KC_ZC 4096u
#define KC_SB 1600

static int kc_idx(int x, int y, int z) { return x + 5 * y + 25 * z; }
static unsigned kc_lane_pos(int j) { return 64u * (unsigned)(j % 25) + (unsigned)(j / 25); }
static unsigned kc_pos_s0(int j) { return kc_lane_pos(j); }
static unsigned kc_pos_s24(int j) { return 2048u + kc_lane_pos(j); }
static unsigned kc_pos_t(int r, int j) { return 4160u + 1600u * (unsigned)r + kc_lane_pos(j); }

/* theta_rho_pi_preimage (keccak.rs:147-176): the 11 state bits whose XOR is B[x, y, z] */
static void kc_preim(int x, int y, int z, int out[11]) {
    const int a = (x + 3 * y) % 5, b = x;
    const int c = (z + 64 - (int)(RHO[a][b] % 64)) % 64, cp = (c + 63) % 64;
    out[0] = kc_idx(a, b, c);
    for (int yp = 0; yp < 5; yp++) {
        out[1 + yp] = kc_idx((a + 4) % 5, yp, c);
        out[6 + yp] = kc_idx((a + 1) % 5, yp, cp);
    }
}

static void kc_preim_of(int s, int out[11]) {
    kc_preim(s % 25 % 5, s % 25 / 5, s / 25, out);
}

/* out = phi^T(v): out[s_in] = sum over s_out with s_in in preim(s_out) of v[s_out] */
static void kc_phi_t(const or_f128* v, or_f128* out) {
    memset(out, 0, sizeof(or_f128) * KC_SB);
    for (int so = 0; so < KC_SB; so++) {
        int p[11];
        kc_preim_of(so, p);
        for (int i = 0; i < 11; i++) {
            out[p[i]].lo ^= v[so].lo;
            out[p[i]].hi ^= v[so].hi;
        }
    }
}

static void kc_acc(or_f128* d, or_f128 v) {
    d->lo ^= v.lo;
    d->hi ^= v.hi;
}

void or_keccak_lc_fold_split(const or_f128* eq, or_f128* sa, or_f128* sb) {
    memset(sa, 0, sizeof(or_f128) * KC_K);
    memset(sb, 0, sizeof(or_f128) * KC_K);
    or_f128* chi_a = calloc(24 * KC_SB, sizeof(or_f128));
    or_f128* chi_b = calloc(24 * KC_SB, sizeof(or_f128));
    or_f128 pin[KC_SB], k[KC_SB], nk[KC_SB];
    /* constant row: A = B = [ZC] */
    kc_acc(&sa[KC_ZC], eq[KC_ZC]);
    kc_acc(&sb[KC_ZC], eq[KC_ZC]);
    /* state_0 input rows: A = [row], B = [ZC] */
    for (int j = 0; j < KC_SB; j++) {
        kc_acc(&sa[kc_pos_s0(j)], eq[kc_pos_s0(j)]);
        kc_acc(&sb[KC_ZC], eq[kc_pos_s0(j)]);
    }
    /* state_24 pin rows: A = L_24[j], B = [ZC] */
    for (int j = 0; j < KC_SB; j++) {
        pin[j] = eq[kc_pos_s24(j)];
        kc_acc(&sb[KC_ZC], pin[j]);
    }
    /* t rows of round r at B-coordinates (x, y, z): A = [ZC] + B_r[x+1, y, z],
     * B = B_r[x+2, y, z], with B_r[.] = XOR of the 11 state_r bits of its preimage */
    or_f128 sum_t = {0, 0};
    for (int r = 0; r < 24; r++)
        for (int z = 0; z < 64; z++)
            for (int y = 0; y < 5; y++)
                for (int x = 0; x < 5; x++) {
                    const or_f128 e = eq[kc_pos_t(r, kc_idx(x, y, z))];
                    int p[11];
                    kc_acc(&sum_t, e);
                    kc_preim((x + 1) % 5, y, z, p);
                    for (int i = 0; i < 11; i++) kc_acc(&chi_a[r * KC_SB + p[i]], e);
                    kc_preim((x + 2) % 5, y, z, p);
                    for (int i = 0; i < 11; i++) kc_acc(&chi_b[r * KC_SB + p[i]], e);
                }
    kc_acc(&sa[KC_ZC], sum_t);
    /* round-constant parts: RC_r is the constant vector of state_r (RC_0 = 0) */
    {
        uint8_t rc[KC_SB], nrc[KC_SB];
        memset(rc, 0, sizeof(rc));
        or_f128 rc_a = {0, 0}, rc_b = {0, 0}, rc_pin = {0, 0};
        for (int r = 0; r < 24; r++) {
            for (int s = 0; s < KC_SB; s++)
                if (rc[s]) {
                    kc_acc(&rc_a, chi_a[r * KC_SB + s]);
                    kc_acc(&rc_b, chi_b[r * KC_SB + s]);
                }
            for (int so = 0; so < KC_SB; so++) {
                int p[11];
                kc_preim_of(so, p);
                uint8_t bit = 0;
                for (int i = 0; i < 11; i++) bit ^= rc[p[i]];
                nrc[so] = bit;
            }
            memcpy(rc, nrc, sizeof(rc));
            for (int z = 0; z < 64; z++)
                if ((RC[r] >> z) & 1) rc[kc_idx(0, 0, z)] ^= 1;
        }
        for (int s = 0; s < KC_SB; s++)
            if (rc[s]) kc_acc(&rc_pin, pin[s]);
        kc_acc(&sa[KC_ZC], rc_a);
        kc_acc(&sa[KC_ZC], rc_pin);
        kc_acc(&sb[KC_ZC], rc_b);
    }
    /* A side: K_24 = pin; t_23 takes K_24, then K_{r-1} = phi^T(K_r) + chi_a[r-1] */
    for (int s = 0; s < KC_SB; s++) kc_acc(&sa[kc_pos_t(23, s)], pin[s]);
    kc_phi_t(pin, k);
    for (int s = 0; s < KC_SB; s++) kc_acc(&k[s], chi_a[23 * KC_SB + s]);
    for (int r = 23; r >= 1; r--) {
        for (int s = 0; s < KC_SB; s++) kc_acc(&sa[kc_pos_t(r - 1, s)], k[s]);
        kc_phi_t(k, nk);
        for (int s = 0; s < KC_SB; s++) {
            k[s] = nk[s];
            kc_acc(&k[s], chi_a[(r - 1) * KC_SB + s]);
        }
    }
    for (int s = 0; s < KC_SB; s++) kc_acc(&sa[kc_pos_s0(s)], k[s]);
    /* B side: K_24 = 0, so K_23 = chi_b[23] */
    memcpy(k, chi_b + 23 * KC_SB, sizeof(k));
    for (int r = 23; r >= 1; r--) {
        for (int s = 0; s < KC_SB; s++) kc_acc(&sb[kc_pos_t(r - 1, s)], k[s]);
        kc_phi_t(k, nk);
        for (int s = 0; s < KC_SB; s++) {
            k[s] = nk[s];
            kc_acc(&k[s], chi_b[(r - 1) * KC_SB + s]);
        }
    }
    for (int s = 0; s < KC_SB; s++) kc_acc(&sb[kc_pos_s0(s)], k[s]);
    free(chi_a);
    free(chi_b);
}
