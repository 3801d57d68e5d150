// Built-in hash circuits: the base matrices (A_0, B_0) of one compression's boolean R1CS
// (C_0 = I: every z column is the output of its own row), built on the host.
//
//   SHA-256  crates/flock-prover/src/r1cs_hashes/sha2.rs:243-701 (build_matrices, the
//            production walk with E_NEW / A_NEW materialized every other round)
//   BLAKE3   crates/flock-prover/src/r1cs_hashes/blake3.rs:380-783 (build_matrices, Option F)
//   Keccak   keccak.rs:880-888: empty-stub matrices (its lincheck runs through the walker)
//
// Every intermediate bit is a GF(2)-linear form over the z columns, kept as a sorted column
// set (XOR = symmetric difference), so a row's support is canonical (ascending, no repeats)
// whatever order the walk combines terms in. The adders are the circuits' shared building
// blocks: a ripple add with materialized carries, a constant-operand add, and carry-save
// layers closed by a ripple of the shifted majority word.
//
// Outputs: CSR rows (the registry digest's SparseBinaryMatrix, r1cs.rs:14-18, absorbed by
// Registry::digest) and the lincheck's CSC form (csc_from_rows, lincheck.rs:367-388), cached
// per kind for the life of the process; device copies per context on request.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <utility>
#include <vector>

#include "../../include/flock_hip.h"
#include "flock_internal.h"

namespace {

using Form = std::vector<uint32_t>;  // sorted z columns whose XOR is the bit

Form operator^(const Form& x, const Form& y) {
    Form out;
    out.reserve(x.size() + y.size());
    size_t i = 0, j = 0;
    while (i < x.size() && j < y.size()) {
        if (x[i] == y[j]) {
            i++;
            j++;
        } else if (x[i] < y[j]) {
            out.push_back(x[i++]);
        } else {
            out.push_back(y[j++]);
        }
    }
    out.insert(out.end(), x.begin() + i, x.end());
    out.insert(out.end(), y.begin() + j, y.end());
    return out;
}
Form& operator^=(Form& x, const Form& y) { return x = x ^ y; }
Form col(uint32_t c) { return Form{c}; }

constexpr int kW = 32;
using Word = std::array<Form, kW>;

Word operator^(const Word& x, const Word& y) {
    Word o;
    for (int i = 0; i < kW; i++) o[i] = x[i] ^ y[i];
    return o;
}
// bit i of rotr(x, n) is bit (i + n) mod 32 of x
Word rotr(const Word& x, int n) {
    Word o;
    for (int i = 0; i < kW; i++) o[i] = x[(i + n) % kW];
    return o;
}
Word shr(const Word& x, int n) {
    Word o;
    for (int i = 0; i + n < kW; i++) o[i] = x[i + n];
    return o;
}
// a word whose bit i is the single column base + i
Word wire(uint32_t base) {
    Word o;
    for (int i = 0; i < kW; i++) o[i] = col(base + i);
    return o;
}

// The circuit under construction: row r of A_0 / B_0 (rows default to empty = padding).
struct Builder {
    uint32_t k, pin;
    std::vector<Form> A, B;
    Builder(uint32_t k_log, uint32_t z_const) : k(1u << k_log), pin(z_const), A(k), B(k) {}
    void row(uint32_t r, Form a, Form b) {
        A[r] = std::move(a);
        B[r] = std::move(b);
    }
    Form one() const { return col(pin); }
    Word constant(uint32_t v) const {
        Word o;
        for (int i = 0; i < kW; i++)
            if ((v >> i) & 1) o[i] = one();
        return o;
    }
    // free input bits: z[s] * 1 = z[s]
    void inputs(uint32_t base, uint32_t n) {
        for (uint32_t s = base; s < base + n; s++) row(s, col(s), one());
    }
    // z[slot(i)] = raw[i] * 1; returns the slot word
    template <class Slot>
    Word materialize(const Word& raw, Slot slot) {
        Word o;
        for (int i = 0; i < kW; i++) {
            row(slot(i), raw[i], one());
            o[i] = col(slot(i));
        }
        return o;
    }
    // x + y mod 2^32 with its 31 carry products at carry(i):
    // carry row i = (x_i + c_i)(y_i + c_i), c_{i+1} = c_i + carry_i, sum_i = x_i + y_i + c_i
    template <class Slot>
    Word ripple_add(const Word& x, const Word& y, Slot carry) {
        Word s;
        Form c;
        for (int i = 0; i < kW; i++) {
            s[i] = x[i] ^ y[i] ^ c;
            if (i < kW - 1) {
                row(carry(i), x[i] ^ c, y[i] ^ c);
                c ^= col(carry(i));
            }
        }
        return s;
    }
    // k + y for a constant k: with t = ctz(k) + 1 the carries below bit t vanish and the carry
    // into bit t is y_{t-1}, so products exist only for bits t..30, at base + (i - t)
    Word const_add(uint32_t kc, const Word& y, uint32_t base) {
        const int t = __builtin_ctz(kc) + 1;
        Word s;
        Form c;
        for (int i = 0; i < kW; i++) {
            if (i == t) c = y[t - 1];
            const Form ki = ((kc >> i) & 1) ? one() : Form{};
            s[i] = ki ^ y[i] ^ c;
            if (i >= t && i < kW - 1) {
                const uint32_t slot = base + (uint32_t)(i - t);
                row(slot, ki ^ c, y[i] ^ c);
                c ^= col(slot);
            }
        }
        return s;
    }
    // one carry-save layer over (x, y) with z on both product sides: 31 majority products
    // w_i = (x_i + z_i)(y_i + z_i) at maj + i; returns the partial sum x + y + z and the
    // shifted majority word (bit 0 empty, bit i+1 = w_i + z_i)
    std::pair<Word, Word> csa(const Word& x, const Word& y, const Word& z, uint32_t maj) {
        Word p, sh;
        for (int i = 0; i < kW; i++) {
            p[i] = x[i] ^ y[i] ^ z[i];
            if (i < kW - 1) {
                row(maj + i, x[i] ^ z[i], y[i] ^ z[i]);
                sh[i + 1] = col(maj + i) ^ z[i];
            }
        }
        return {std::move(p), std::move(sh)};
    }
    // the ripple closing a carry-save pair (p, sh) with sh_0 empty: products for bits 1..30 at
    // rip + (i - 1); sum_i = p_i + sh_i + g_i with g the running XOR of the products
    Word close(const Word& p, const Word& sh, uint32_t rip) {
        Word s;
        Form g;
        for (int i = 0; i < kW; i++) {
            s[i] = p[i] ^ sh[i] ^ g;
            if (i >= 1 && i <= kW - 2) {
                row(rip + (i - 1), p[i] ^ g, sh[i] ^ g);
                g ^= col(rip + (i - 1));
            }
        }
        return s;
    }
    Word add3(const Word& x, const Word& y, const Word& z, uint32_t maj, uint32_t rip) {
        auto [p, sh] = csa(x, y, z, maj);
        return close(p, sh, rip);
    }
    Word add4(const Word& x, const Word& y, const Word& z, const Word& w, uint32_t maj1,
              uint32_t maj2, uint32_t rip) {
        auto [p1, s1] = csa(x, y, z, maj1);
        auto [p2, s2] = csa(p1, s1, w, maj2);
        return close(p2, s2, rip);
    }
};

// ---- SHA-256 (sha2.rs:60-240 layout, :548-686 walk) ----
namespace sha {
constexpr uint32_t kK[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
// region bases
constexpr uint32_t kH = 0, kHOut = 256, kM = 512, kChAnd = 1024, kMajAnd = 3072, kRounds = 5120;
// offsets inside round r's block: maj1, maj2, T1 ripple, a_new maj / ripple, e_new carries,
// then the h + K[r] products (31 - t_r of them)
constexpr uint32_t kMaj1 = 0, kMaj2 = 31, kRip = 62, kAMaj = 92, kARip = 123, kENew = 153,
                   kAddK = 184;
constexpr uint32_t kSchedStep = 92;  // a schedule step's fused 4-operand add: 31 + 31 + 30

struct Layout {
    uint32_t round[65];
    uint32_t sched, e_new, a_new, out_carry, pin;
    Layout() {
        uint32_t at = kRounds;
        for (int r = 0; r < 64; r++) {
            round[r] = at;
            at += kAddK + (31u - (uint32_t)(__builtin_ctz(kK[r]) + 1));
        }
        round[64] = at;
        sched = at;
        e_new = sched + 48 * kSchedStep;
        a_new = e_new + 32 * kW;
        out_carry = a_new + 32 * kW;
        pin = out_carry + 8 * 31;
    }
};

Word bsig0(const Word& x) { return rotr(x, 2) ^ rotr(x, 13) ^ rotr(x, 22); }
Word bsig1(const Word& x) { return rotr(x, 6) ^ rotr(x, 11) ^ rotr(x, 25); }
Word ssig0(const Word& x) { return rotr(x, 7) ^ rotr(x, 18) ^ shr(x, 3); }
Word ssig1(const Word& x) { return rotr(x, 17) ^ rotr(x, 19) ^ shr(x, 10); }

void build(Builder& c) {
    const Layout L;
    c.row(L.pin, c.one(), c.one());
    c.inputs(kH, 8 * kW);
    c.inputs(kM, 16 * kW);
    std::array<Word, 8> h_in;
    for (int w = 0; w < 8; w++) h_in[w] = wire(kH + kW * w);
    std::vector<Word> W(64);
    for (int i = 0; i < 16; i++) W[i] = wire(kM + kW * i);
    // schedule: W[t] = s1(W[t-2]) + s0(W[t-15]) + W[t-7] + W[t-16], never materialized
    for (int t = 16; t < 64; t++) {
        const uint32_t b = L.sched + (uint32_t)(t - 16) * kSchedStep;
        W[t] = c.add4(ssig1(W[t - 2]), ssig0(W[t - 15]), W[t - 7], W[t - 16], b, b + 31, b + 62);
    }
    std::array<Word, 8> s = h_in;  // (a, b, c, d, e, f, g, h)
    for (int r = 0; r < 64; r++) {
        const Word& a = s[0];
        const Word& e = s[4];
        Word ch, mj;
        for (int i = 0; i < kW; i++) {
            const uint32_t sc = kChAnd + kW * r + i, sm = kMajAnd + kW * r + i;
            c.row(sc, e[i], s[5][i] ^ s[6][i]);
            c.row(sm, a[i] ^ s[1][i], a[i] ^ s[2][i]);
            ch[i] = col(sc);
            mj[i] = col(sm);
        }
        const Word ch_out = ch ^ s[6], maj_out = mj ^ a;
        const uint32_t rb = L.round[r];
        const Word hk = c.const_add(kK[r], s[7], rb + kAddK);
        const Word t1 = c.add4(hk, bsig1(e), ch_out, W[r], rb + kMaj1, rb + kMaj2, rb + kRip);
        Word a_new = c.add3(t1, bsig0(a), maj_out, rb + kAMaj, rb + kARip);
        Word e_new = c.ripple_add(s[3], t1, [&](int i) { return rb + kENew + (uint32_t)i; });
        if (r & 1) {  // the cascade breakers, every other round
            const uint32_t j = (uint32_t)r / 2;
            a_new = c.materialize(a_new, [&](int i) { return L.a_new + kW * j + (uint32_t)i; });
            e_new = c.materialize(e_new, [&](int i) { return L.e_new + kW * j + (uint32_t)i; });
        }
        s = {std::move(a_new), s[0], s[1], s[2], std::move(e_new), s[4], s[5], s[6]};
    }
    // feed-forward: H_out[w] = state[w] + H_in[w], carries then the materialized sum
    for (int w = 0; w < 8; w++) {
        const Word sum =
            c.ripple_add(s[w], h_in[w], [&](int i) { return L.out_carry + 31u * w + (uint32_t)i; });
        c.materialize(sum, [&](int i) { return kHOut + kW * w + (uint32_t)i; });
    }
}
}  // namespace sha

// ---- BLAKE3 compression (blake3.rs:122-345 layout, :659-783 walk) ----
namespace b3 {
constexpr uint32_t kIV[4] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a};
constexpr int kPerm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
constexpr int kLanes[8][4] = {{0, 4, 8, 12}, {1, 5, 9, 13}, {2, 6, 10, 14}, {3, 7, 11, 15},
                              {0, 5, 10, 15}, {1, 6, 11, 12}, {2, 7, 8, 13}, {3, 4, 9, 14}};
constexpr uint32_t kCv = 0, kOutLo = 256, kM = 512, kTLo = 1024, kTHi = 1056, kBlen = 1088,
                   kFlags = 1120, kOutHi = 1152, kGs = 1408;

// products of G g's c + d_1 add: round 1's column G's add the constant IV[g] (t = 1 for the
// odd IV[0], IV[1]; t = 2 for the even IV[2], IV[3]), every other G a full ripple
uint32_t c1_rows(int g) { return g < 2 ? 30 : g < 4 ? 29 : 31; }

void build(Builder& c, uint32_t* pin_out) {
    uint32_t gb[57];
    gb[0] = kGs;
    for (int g = 0; g < 56; g++) gb[g + 1] = gb[g] + 61 + c1_rows(g) + 61 + 31;
    const uint32_t pin = gb[56];
    *pin_out = pin;
    c.pin = pin;
    c.row(pin, c.one(), c.one());
    c.inputs(kCv, 8 * kW);
    c.inputs(kM, 16 * kW);
    c.inputs(kTLo, 4 * kW);  // counter lo / hi, block_len, flags
    std::array<Word, 16> st;
    for (int w = 0; w < 8; w++) st[w] = wire(kCv + kW * w);
    for (int i = 0; i < 4; i++) st[8 + i] = c.constant(kIV[i]);
    st[12] = wire(kTLo);
    st[13] = wire(kTHi);
    st[14] = wire(kBlen);
    st[15] = wire(kFlags);
    int perm[16];
    for (int i = 0; i < 16; i++) perm[i] = i;
    for (int r = 0; r < 7; r++) {
        for (int gi = 0; gi < 8; gi++) {
            const int g = 8 * r + gi;
            const int* ln = kLanes[gi];
            const Word mx = wire(kM + kW * (uint32_t)perm[2 * gi]);
            const Word my = wire(kM + kW * (uint32_t)perm[2 * gi + 1]);
            const uint32_t base = gb[g], c1 = c1_rows(g);
            const Word a1 = c.add3(st[ln[0]], st[ln[1]], mx, base, base + 31);
            const Word d1 = rotr(st[ln[3]] ^ a1, 16);
            const Word cc1 = g < 4 ? c.const_add(kIV[g], d1, base + 61)
                                   : c.ripple_add(st[ln[2]], d1, [&](int i) { return base + 61 + (uint32_t)i; });
            const Word b1 = rotr(st[ln[1]] ^ cc1, 12);
            const uint32_t maj2 = base + 61 + c1, rip2 = maj2 + 31, cc2b = rip2 + 30;
            Word a2 = c.add3(a1, b1, my, maj2, rip2);
            Word d2 = rotr(d1 ^ a2, 8);
            Word cc2 = c.ripple_add(cc1, d2, [&](int i) { return cc2b + (uint32_t)i; });
            Word bn = rotr(b1 ^ cc2, 7);
            st[ln[0]] = std::move(a2);
            st[ln[1]] = std::move(bn);
            st[ln[2]] = std::move(cc2);
            st[ln[3]] = std::move(d2);
        }
        int next[16];
        for (int i = 0; i < 16; i++) next[i] = perm[kPerm[i]];
        std::memcpy(perm, next, sizeof(perm));
    }
    // finalization: out_lo[w] = state[w] ^ state[w+8], out_hi[w] = state[w+8] ^ cv[w]
    for (int w = 0; w < 8; w++) {
        c.materialize(st[w] ^ st[w + 8], [&](int i) { return kOutLo + kW * w + (uint32_t)i; });
        c.materialize(st[w + 8] ^ wire(kCv + kW * w), [&](int i) { return kOutHi + kW * w + (uint32_t)i; });
    }
}
}  // namespace b3

}  // namespace

struct flock_circuit {
    uint32_t kind = 0, k_log = 0;
    uint64_t useful_bits = 0, pin = 0;
    std::vector<uint64_t> a_ptr, b_ptr;  // CSR
    std::vector<uint32_t> a_cols, b_cols;
    std::vector<uint32_t> a_cptr, a_rows, b_cptr, b_rows;  // CSC
};

namespace {

void to_csr(const std::vector<Form>& rows, std::vector<uint64_t>& ptr, std::vector<uint32_t>& cols) {
    ptr.assign(rows.size() + 1, 0);
    for (size_t r = 0; r < rows.size(); r++) ptr[r + 1] = ptr[r] + rows[r].size();
    cols.resize(ptr.back());
    for (size_t r = 0; r < rows.size(); r++) std::copy(rows[r].begin(), rows[r].end(), cols.begin() + ptr[r]);
}

// csc_from_rows: per column, the rows holding it, ascending
void to_csc(size_t k, const std::vector<uint64_t>& ptr, const std::vector<uint32_t>& cols,
            std::vector<uint32_t>& cptr, std::vector<uint32_t>& rows) {
    cptr.assign(k + 1, 0);
    for (uint32_t c : cols) cptr[c + 1]++;
    for (size_t c = 0; c < k; c++) cptr[c + 1] += cptr[c];
    rows.resize(cols.size());
    std::vector<uint32_t> at(cptr.begin(), cptr.end() - 1);
    for (size_t r = 0; r < k; r++)
        for (uint64_t i = ptr[r]; i < ptr[r + 1]; i++) rows[at[cols[i]]++] = (uint32_t)r;
}

flock_circuit* make(uint32_t kind) {
    auto* c = new (std::nothrow) flock_circuit();
    if (!c) return nullptr;
    c->kind = kind;
    if (kind == FLOCK_HASH_KECCAK) {
        c->k_log = 16;
        c->useful_bits = 42560;
        c->pin = 4096;
        c->a_ptr.assign((1u << 16) + 1, 0);
        c->b_ptr = c->a_ptr;
        c->a_cptr.assign((1u << 16) + 1, 0);
        c->b_cptr = c->a_cptr;
        return c;
    }
    uint32_t pin;
    if (kind == FLOCK_HASH_SHA256) {
        c->k_log = 15;
        pin = sha::Layout().pin;
        Builder b(15, pin);
        sha::build(b);
        to_csr(b.A, c->a_ptr, c->a_cols);
        to_csr(b.B, c->b_ptr, c->b_cols);
    } else {
        c->k_log = 14;
        Builder b(14, 0);
        b3::build(b, &pin);
        to_csr(b.A, c->a_ptr, c->a_cols);
        to_csr(b.B, c->b_ptr, c->b_cols);
    }
    c->pin = pin;
    c->useful_bits = pin + 1;
    const size_t k = (size_t)1 << c->k_log;
    to_csc(k, c->a_ptr, c->a_cols, c->a_cptr, c->a_rows);
    to_csc(k, c->b_ptr, c->b_cols, c->b_cptr, c->b_rows);
    return c;
}

std::mutex g_mu;
flock_circuit* g_circ[3] = {nullptr, nullptr, nullptr};
// (sorted kinds, nu) -> Registry::digest
std::map<std::pair<std::vector<uint32_t>, uint32_t>, std::array<uint8_t, 32>> g_digest;

}  // namespace

namespace flock {
const flock_circuit* builtin_circuit(uint32_t kind) {
    if (kind > FLOCK_HASH_KECCAK) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_circ[kind]) g_circ[kind] = make(kind);
    return g_circ[kind];
}
}  // namespace flock

extern "C" {

int flock_circuit_get(uint32_t kind, const flock_circuit** out) {
    if (!out || kind > FLOCK_HASH_KECCAK) return flock::set_error(FLOCK_HIP_E_INVALID, "circuit kind");
    *out = flock::builtin_circuit(kind);
    return *out ? 0 : flock::set_error(FLOCK_HIP_E_NOMEM, "circuit build");
}

int flock_circuit_info(const flock_circuit* c, uint32_t* k_log, uint64_t* useful_bits,
                       uint64_t* const_pin, uint64_t* nnz_a, uint64_t* nnz_b) {
    if (!c) return FLOCK_HIP_E_INVALID;
    if (k_log) *k_log = c->k_log;
    if (useful_bits) *useful_bits = c->useful_bits;
    if (const_pin) *const_pin = c->pin;
    if (nnz_a) *nnz_a = c->a_cols.size();
    if (nnz_b) *nnz_b = c->b_cols.size();
    return 0;
}

int flock_circuit_csr(const flock_circuit* c, uint32_t which, const uint64_t** row_ptr,
                      const uint32_t** cols) {
    if (!c || which > 1 || !row_ptr || !cols) return FLOCK_HIP_E_INVALID;
    *row_ptr = which ? c->b_ptr.data() : c->a_ptr.data();
    *cols = which ? c->b_cols.data() : c->a_cols.data();
    return 0;
}

int flock_circuit_csc(const flock_circuit* c, uint32_t which, const uint32_t** col_ptr,
                      const uint32_t** rows) {
    if (!c || which > 1 || !col_ptr || !rows) return FLOCK_HIP_E_INVALID;
    *col_ptr = which ? c->b_cptr.data() : c->a_cptr.data();
    *rows = which ? c->b_rows.data() : c->a_rows.data();
    return 0;
}

int flock_circuit_type_statement(const flock_circuit* c, flock_type_statement* out) {
    if (!c || !out) return FLOCK_HIP_E_INVALID;
    std::memset(out, 0, sizeof(*out));
    const uint64_t k = 1ull << c->k_log;
    out->k_log = c->k_log;
    out->has_const_pin = 1;
    out->useful_bits = c->useful_bits;
    out->const_pin = c->pin;
    out->a0 = {k, k, c->a_ptr.data(), c->a_cols.data(), 0, 0};
    out->b0 = {k, k, c->b_ptr.data(), c->b_cols.data(), 0, 0};
    out->c0 = {k, k, nullptr, nullptr, 1, 0};
    return 0;
}

int flock_circuit_union_digest(const uint32_t* kinds, size_t n_kinds, uint32_t nu,
                               uint8_t out[32]) {
    if (!kinds || !out || n_kinds == 0 || n_kinds > FLOCK_MAX_TYPES || nu > 63)
        return FLOCK_HIP_E_INVALID;
    std::vector<flock_type_statement> t(n_kinds);
    for (size_t i = 0; i < n_kinds; i++) {
        const flock_circuit* c = nullptr;
        if (int rc = flock_circuit_get(kinds[i], &c)) return rc;
        flock_circuit_type_statement(c, &t[i]);
    }
    // types of equal k_log are equal built-in types, so the digest depends on the multiset
    std::vector<uint32_t> key(kinds, kinds + n_kinds);
    std::sort(key.begin(), key.end());
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_digest.find({key, nu});
        if (it != g_digest.end()) {
            std::memcpy(out, it->second.data(), 32);
            return 0;
        }
    }
    if (int rc = flock_registry_digest(t.data(), n_kinds, nu, out)) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    std::array<uint8_t, 32>& d = g_digest[{key, nu}];
    std::memcpy(d.data(), out, 32);
    return 0;
}

int flock_circuit_registry_digest(uint32_t kind, uint32_t nu, uint8_t out[32]) {
    return flock_circuit_union_digest(&kind, 1, nu, out);
}

int flock_hip_circuit_upload(flock_hip_ctx* ctx, uint32_t kind, flock_csc* out) {
    const flock_circuit* c = nullptr;
    if (!ctx || !out) return FLOCK_HIP_E_INVALID;
    if (int rc = flock_circuit_get(kind, &c)) return rc;
    std::memset(out, 0, sizeof(*out));
    out->k_log = c->k_log;
    out->const_pin = (int64_t)c->pin;
    out->nnz_a = c->a_rows.size();
    out->nnz_b = c->b_rows.size();
    const std::vector<uint32_t>* src[4] = {&c->a_cptr, &c->a_rows, &c->b_cptr, &c->b_rows};
    void* dst[4] = {nullptr, nullptr, nullptr, nullptr};
    hipStream_t st = static_cast<hipStream_t>(flock_hip_ctx_stream(ctx));
    int rc = 0;
    for (int i = 0; i < 4 && !rc; i++) {
        const size_t bytes = 4 * src[i]->size();
        rc = flock_hip_malloc(ctx, &dst[i], bytes ? bytes : 4);
        if (!rc && bytes &&
            hipMemcpyAsync(dst[i], src[i]->data(), bytes, hipMemcpyHostToDevice, st) != hipSuccess)
            rc = flock::set_error(FLOCK_HIP_E_HIP, "circuit upload");
    }
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = flock::set_error(FLOCK_HIP_E_HIP, "circuit upload");
    if (rc) {
        for (void* p : dst)
            if (p) flock_hip_dfree(ctx, p);
        return rc;
    }
    out->a_col_ptr = static_cast<const uint32_t*>(dst[0]);
    out->a_rows = static_cast<const uint32_t*>(dst[1]);
    out->b_col_ptr = static_cast<const uint32_t*>(dst[2]);
    out->b_rows = static_cast<const uint32_t*>(dst[3]);
    return 0;
}

void flock_hip_circuit_release(flock_hip_ctx* ctx, flock_csc* c) {
    if (!ctx || !c) return;
    const void* p[4] = {c->a_col_ptr, c->a_rows, c->b_col_ptr, c->b_rows};
    for (const void* x : p)
        if (x) flock_hip_dfree(ctx, const_cast<void*>(x));
    std::memset(c, 0, sizeof(*c));
}

}  // extern "C"
