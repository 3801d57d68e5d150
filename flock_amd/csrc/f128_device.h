// GF(2^128) arithmetic for gfx950 (CDNA4), GHASH polynomial x^128 + x^7 + x^2 + x + 1.
//
// An element is a uint4 {x: bits 0..31, y: 32..63, z: 64..95, w: 96..127}, which is
// byte-identical to the reference's little-endian {lo: u64, hi: u64} layout
// (crates/flock-core/src/field/gf2_128.rs:22-27). CDNA4 has no carry-less multiply,
// so products are built from 32-bit VALU shifts/funnel shifts and XORs:
//
//  * f128_mul        — generic a*b: 32 funnel-shifted copies of b, one per bit offset
//                      s in [0,32), each masked by bit s of every word of a and
//                      XOR-accumulated into a 256-bit product, then reduced.
//  * ShoupTable       — a*t for a fixed t: 32 LDS subtables T_k[n] = (n * x^(4k)) * t
//                      (16 entries of 16 B = one 256-byte row = all 64 LDS banks, so a
//                      64-lane lookup never bank-conflicts); the product is the XOR of
//                      32 lookups, one per nibble of a.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "transcript_device.h"

typedef uint4 f128;

__device__ __forceinline__ f128 f128_zero() { return make_uint4(0u, 0u, 0u, 0u); }
__device__ __forceinline__ f128 f128_one() { return make_uint4(1u, 0u, 0u, 0u); }

__device__ __forceinline__ f128 f128_add(f128 a, f128 b) {
    return make_uint4(a.x ^ b.x, a.y ^ b.y, a.z ^ b.z, a.w ^ b.w);
}
__device__ __forceinline__ f128 f128_add3(f128 a, f128 b, f128 c) {
    return make_uint4(a.x ^ b.x ^ c.x, a.y ^ b.y ^ c.y, a.z ^ b.z ^ c.z, a.w ^ b.w ^ c.w);
}

// (hi:lo) >> (32 - s) in 32 bits, i.e. the word of (x << s) that receives lo's top bits.
__device__ __forceinline__ uint32_t fshl(uint32_t hi, uint32_t lo, int s) {
    return __builtin_amdgcn_alignbit(hi, lo, (uint32_t)(32 - s));
}

// 256-bit unreduced product accumulator (r[0] = bits 0..31).
struct f256u {
    uint32_t r[8];
};

__device__ __forceinline__ void f256u_zero(f256u& a) {
#pragma unroll
    for (int i = 0; i < 8; i++) a.r[i] = 0u;
}
__device__ __forceinline__ void f256u_xor(f256u& a, const f256u& b) {
#pragma unroll
    for (int i = 0; i < 8; i++) a.r[i] ^= b.r[i];
}

// acc ^= a * b (carry-less, unreduced). For every bit offset s in [0, 32) the 160-bit
// value b << s is built with funnel shifts, and bit s of each word w of a selects whether
// it is XORed into acc at word offset w. The s loop is kept rolled so the product needs
// only ~30 registers; each step is 5 shifts, 4 bit-extracts and 20 and-xor (bitop3) ops.
__device__ __forceinline__ void clmul128_acc(f256u& acc, f128 a, f128 b) {
    // s = 0: b itself
    {
        const uint32_t m0 = 0u - (a.x & 1u), m1 = 0u - (a.y & 1u), m2 = 0u - (a.z & 1u),
                       m3 = 0u - (a.w & 1u);
        acc.r[0] ^= b.x & m0;
        acc.r[1] ^= (b.y & m0) ^ (b.x & m1);
        acc.r[2] ^= (b.z & m0) ^ (b.y & m1) ^ (b.x & m2);
        acc.r[3] ^= (b.w & m0) ^ (b.z & m1) ^ (b.y & m2) ^ (b.x & m3);
        acc.r[4] ^= (b.w & m1) ^ (b.z & m2) ^ (b.y & m3);
        acc.r[5] ^= (b.w & m2) ^ (b.z & m3);
        acc.r[6] ^= b.w & m3;
    }
    uint32_t r0 = acc.r[0], r1 = acc.r[1], r2 = acc.r[2], r3 = acc.r[3], r4 = acc.r[4],
             r5 = acc.r[5], r6 = acc.r[6], r7 = acc.r[7];
#pragma unroll 1
    for (int s = 1; s < 32; s++) {
        const uint32_t s0 = b.x << s;
        const uint32_t s1 = fshl(b.y, b.x, s);
        const uint32_t s2 = fshl(b.z, b.y, s);
        const uint32_t s3 = fshl(b.w, b.z, s);
        const uint32_t s4 = b.w >> (32 - s);
        const uint32_t m0 = (uint32_t)(((int32_t)(a.x << (31 - s))) >> 31);
        const uint32_t m1 = (uint32_t)(((int32_t)(a.y << (31 - s))) >> 31);
        const uint32_t m2 = (uint32_t)(((int32_t)(a.z << (31 - s))) >> 31);
        const uint32_t m3 = (uint32_t)(((int32_t)(a.w << (31 - s))) >> 31);
        r0 ^= s0 & m0;
        r1 ^= (s1 & m0) ^ (s0 & m1);
        r2 ^= (s2 & m0) ^ (s1 & m1) ^ (s0 & m2);
        r3 ^= (s3 & m0) ^ (s2 & m1) ^ (s1 & m2) ^ (s0 & m3);
        r4 ^= (s4 & m0) ^ (s3 & m1) ^ (s2 & m2) ^ (s1 & m3);
        r5 ^= (s4 & m1) ^ (s3 & m2) ^ (s2 & m3);
        r6 ^= (s4 & m2) ^ (s3 & m3);
        r7 ^= s4 & m3;
    }
    acc.r[0] = r0; acc.r[1] = r1; acc.r[2] = r2; acc.r[3] = r3;
    acc.r[4] = r4; acc.r[5] = r5; acc.r[6] = r6; acc.r[7] = r7;
}

// Fold the top 128 bits into the bottom 128 mod x^128 + x^7 + x^2 + x + 1
// (the same identity as ghash_reduce, crates/flock-core/src/field/gf2_128.rs:278-297):
// U * x^128 = U + U<<1 + U<<2 + U<<7, and the <=7 bits pushed past x^127 fold again.
__device__ __forceinline__ f128 f256u_reduce(const f256u& a) {
    const uint32_t u0 = a.r[4], u1 = a.r[5], u2 = a.r[6], u3 = a.r[7];
    uint32_t t0 = u0 ^ (u0 << 1) ^ (u0 << 2) ^ (u0 << 7);
    uint32_t t1 = u1 ^ fshl(u1, u0, 1) ^ fshl(u1, u0, 2) ^ fshl(u1, u0, 7);
    uint32_t t2 = u2 ^ fshl(u2, u1, 1) ^ fshl(u2, u1, 2) ^ fshl(u2, u1, 7);
    uint32_t t3 = u3 ^ fshl(u3, u2, 1) ^ fshl(u3, u2, 2) ^ fshl(u3, u2, 7);
    const uint32_t ov = (u3 >> 31) ^ (u3 >> 30) ^ (u3 >> 25);
    const uint32_t corr = ov ^ (ov << 1) ^ (ov << 2) ^ (ov << 7);
    return make_uint4(a.r[0] ^ t0 ^ corr, a.r[1] ^ t1, a.r[2] ^ t2, a.r[3] ^ t3);
}

__device__ __forceinline__ f128 f128_mul(f128 a, f128 b) {
    f256u acc;
    f256u_zero(acc);
    clmul128_acc(acc, a, b);
    return f256u_reduce(acc);
}

// ---------------------------------------------------------------------------
// Generic product with a per-lane 4-bit comb table in LDS (Lopez-Dahab comb).
//
// b = b_lo + b_hi * x^125 with b_lo < 2^125 and b_hi < 8. The lane writes
// T[n] = n(x) * b_lo (n < 16, all < 2^128) into its LDS column, then for nibble
// position j = 7..0 XORs T[nibble_j(a_w)] into acc at word offset w (w = 0..3) and shifts
// acc left by 4 between positions; a * b_hi * x^125 is added with three masked shifts.
// Workspace layout: entry n of lane l at ws[n * 64 + l] (entry-major, lane-minor), so the
// 16 lanes served per LDS cycle hit 16 distinct 16-byte bank groups: no bank conflicts.
// T[0] must be zero (written once by comb_ws_init).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void comb_ws_init(f128* ws_lane) { ws_lane[0] = f128_zero(); }

// Build the table of b (b_lo part) in the lane's column; returns b_hi. STRIDE = distance in
// words between entries: 64 for the per-lane workspace, 1 for a table shared by many lanes
// (16 contiguous entries = 256 B, every bank once).
template <int STRIDE = 64>
__device__ __forceinline__ uint32_t comb_build(f128* ws_lane, f128 b) {
    const uint32_t bh = b.w >> 29;
    const f128 t1 = make_uint4(b.x, b.y, b.z, b.w & 0x1fffffffu);
    const f128 t2 = make_uint4(t1.x << 1, fshl(t1.y, t1.x, 1), fshl(t1.z, t1.y, 1),
                               fshl(t1.w, t1.z, 1));
    const f128 t4 = make_uint4(t1.x << 2, fshl(t1.y, t1.x, 2), fshl(t1.z, t1.y, 2),
                               fshl(t1.w, t1.z, 2));
    const f128 t8 = make_uint4(t1.x << 3, fshl(t1.y, t1.x, 3), fshl(t1.z, t1.y, 3),
                               fshl(t1.w, t1.z, 3));
    const f128 t3 = f128_add(t1, t2), t5 = f128_add(t1, t4), t6 = f128_add(t2, t4);
    const f128 t7 = f128_add(t3, t4), t9 = f128_add(t1, t8), t10 = f128_add(t2, t8);
    const f128 t12 = f128_add(t4, t8);
    ws_lane[1 * STRIDE] = t1;
    ws_lane[2 * STRIDE] = t2;
    ws_lane[3 * STRIDE] = t3;
    ws_lane[4 * STRIDE] = t4;
    ws_lane[5 * STRIDE] = t5;
    ws_lane[6 * STRIDE] = t6;
    ws_lane[7 * STRIDE] = t7;
    ws_lane[8 * STRIDE] = t8;
    ws_lane[9 * STRIDE] = t9;
    ws_lane[10 * STRIDE] = t10;
    ws_lane[11 * STRIDE] = f128_add(t3, t8);
    ws_lane[12 * STRIDE] = t12;
    ws_lane[13 * STRIDE] = f128_add(t5, t8);
    ws_lane[14 * STRIDE] = f128_add(t6, t8);
    ws_lane[15 * STRIDE] = f128_add(t7, t8);
    return bh;
}

// acc ^= a * b where the lane's table holds b_lo and bh = b_hi (unreduced 256-bit).
template <int STRIDE = 64>
__device__ __forceinline__ void comb_apply(f256u& out, const f128* ws_lane, uint32_t bh, f128 a) {
    const uint32_t aw[4] = {a.x, a.y, a.z, a.w};
    uint32_t r[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
    // nibble position j = 7: no shift before it; kept rolled so only 4 lookups are live
#pragma unroll 1
    for (int j = 7; j >= 0; j--) {
        if (j != 7) {
#pragma unroll
            for (int i = 7; i > 0; i--) r[i] = fshl(r[i], r[i - 1], 4);
            r[0] <<= 4;
        }
        const f128 t0 = ws_lane[((aw[0] >> (4 * j)) & 15u) * STRIDE];
        const f128 t1 = ws_lane[((aw[1] >> (4 * j)) & 15u) * STRIDE];
        const f128 t2 = ws_lane[((aw[2] >> (4 * j)) & 15u) * STRIDE];
        const f128 t3 = ws_lane[((aw[3] >> (4 * j)) & 15u) * STRIDE];
        r[0] ^= t0.x;
        r[1] = __builtin_amdgcn_bitop3_b32(r[1], t0.y, t1.x, 0x96);
        r[2] = __builtin_amdgcn_bitop3_b32(r[2], t0.z, t1.y, 0x96) ^ t2.x;
        r[3] = __builtin_amdgcn_bitop3_b32(r[3], t0.w, t1.z, 0x96) ^
               __builtin_amdgcn_bitop3_b32(t2.y, t3.x, 0u, 0x96);
        r[4] = __builtin_amdgcn_bitop3_b32(r[4], t1.w, t2.z, 0x96) ^ t3.y;
        r[5] = __builtin_amdgcn_bitop3_b32(r[5], t2.w, t3.z, 0x96);
        r[6] ^= t3.w;
    }
    // a * b_hi (b_hi < 8) placed at x^125 = word 3, bit 29
    const uint32_t m0 = 0u - (bh & 1u), m1 = 0u - ((bh >> 1) & 1u), m2 = 0u - ((bh >> 2) & 1u);
    uint32_t p[5];
    p[0] = (a.x & m0) ^ ((a.x << 1) & m1) ^ ((a.x << 2) & m2);
    p[1] = (a.y & m0) ^ (fshl(a.y, a.x, 1) & m1) ^ (fshl(a.y, a.x, 2) & m2);
    p[2] = (a.z & m0) ^ (fshl(a.z, a.y, 1) & m1) ^ (fshl(a.z, a.y, 2) & m2);
    p[3] = (a.w & m0) ^ (fshl(a.w, a.z, 1) & m1) ^ (fshl(a.w, a.z, 2) & m2);
    p[4] = ((a.w >> 31) & m1) ^ ((a.w >> 30) & m2);
    r[3] ^= p[0] << 29;
    r[4] ^= fshl(p[1], p[0], 29);
    r[5] ^= fshl(p[2], p[1], 29);
    r[6] ^= fshl(p[3], p[2], 29);
    r[7] ^= fshl(p[4], p[3], 29);
#pragma unroll
    for (int i = 0; i < 8; i++) out.r[i] ^= r[i];
}

__device__ __forceinline__ f128 comb_mul(f128* ws_lane, f128 a, f128 b) {
    const uint32_t bh = comb_build(ws_lane, b);
    f256u acc;
    f256u_zero(acc);
    comb_apply(acc, ws_lane, bh, a);
    return f256u_reduce(acc);
}

// a * x mod p (gf2_128.rs:213-220).
__device__ __forceinline__ f128 f128_mul_x(f128 a) {
    const uint32_t m = (uint32_t)((int32_t)a.w >> 31);
    return make_uint4((a.x << 1) ^ (0x87u & m), fshl(a.y, a.x, 1), fshl(a.z, a.y, 1),
                      fshl(a.w, a.z, 1));
}

// a * x^s mod p for s in [0, 128): shift, then reduce the spilled bits.
__device__ __forceinline__ f128 f128_shl_mod(f128 a, int s) {
    f256u t;
    f256u_zero(t);
    const uint32_t w[4] = {a.x, a.y, a.z, a.w};
    const int q = s >> 5, b = s & 31;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        // word i of a lands in words i+q (low part) and i+q+1 (high part)
        const uint32_t lo = b ? (w[i] << b) : w[i];
        const uint32_t hi = b ? (w[i] >> (32 - b)) : 0u;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (j == i + q) t.r[j] ^= lo;
            if (j == i + q + 1) t.r[j] ^= hi;
        }
    }
    return f256u_reduce(t);
}

// ---------------------------------------------------------------------------
// Fixed-multiplier tables (Shoup, 4-bit windows) in LDS.
// Layout: table[k * 16 + n] = (n(x) * x^(4k)) * t, k in [0,32), n in [0,16). 8 KiB. A row of
// 16 entries is 256 B, every LDS bank once, so random nibble lookups never conflict.
// ---------------------------------------------------------------------------
constexpr int SHOUP_ENTRIES = 512;

__device__ __forceinline__ void shoup_put(f128* table, int idx, f128 v) { table[idx] = v; }

// Build subtable k (16 entries) of the table for multiplier t.
__device__ __forceinline__ void shoup_build_row(f128* table, f128 t, int k) {
    const f128 e1 = f128_shl_mod(t, 4 * k);
    const f128 e2 = f128_mul_x(e1);
    const f128 e4 = f128_mul_x(e2);
    const f128 e8 = f128_mul_x(e4);
    // Rows are 256 B apart, so lanes writing the same entry of different rows at once would
    // hit the same banks. Lane k writes entry g(s) ^ r at step s instead (g = Gray code,
    // r = k mod 16): 16 lanes, 16 distinct entries, every bank once. Entries are linear in
    // n, so entry(g(s) ^ r) = entry(g(s - 1) ^ r) + e_(bit flipped at step s).
    const uint32_t r = (uint32_t)k & 15u;
    const uint32_t m1 = 0u - (r & 1u), m2 = 0u - ((r >> 1) & 1u);
    const uint32_t m4 = 0u - ((r >> 2) & 1u), m8 = 0u - (r >> 3);
    f128 v;
    v.x = (e1.x & m1) ^ (e2.x & m2) ^ (e4.x & m4) ^ (e8.x & m8);
    v.y = (e1.y & m1) ^ (e2.y & m2) ^ (e4.y & m4) ^ (e8.y & m8);
    v.z = (e1.z & m1) ^ (e2.z & m2) ^ (e4.z & m4) ^ (e8.z & m8);
    v.w = (e1.w & m1) ^ (e2.w & m2) ^ (e4.w & m4) ^ (e8.w & m8);
#pragma unroll
    for (int s = 0; s < 16; s++) {
        if (s) {
            const int bit = __builtin_ctz(s);
            v = f128_add(v, bit == 0 ? e1 : bit == 1 ? e2 : bit == 2 ? e4 : e8);
        }
        shoup_put(table, 16 * k + (int)(((uint32_t)(s ^ (s >> 1))) ^ r), v);
    }
}

// (x >> 8K) & 0xF0 in one VALU op: an SDWA and reading byte K (the compiler selects the SDWA
// form by itself for byte 3 only).
template <int K>
__device__ __forceinline__ uint32_t byte_hi_nibble(uint32_t x) {
    if constexpr (K == 0) {
        return x & 0xF0u;
    } else if constexpr (K == 1) {
        uint32_t r;
        asm("v_and_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_1 "
            "src1_sel:DWORD"
            : "=v"(r)
            : "v"(x), "v"(0xF0u));
        return r;
    } else if constexpr (K == 2) {
        uint32_t r;
        asm("v_and_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_2 "
            "src1_sel:DWORD"
            : "=v"(r)
            : "v"(x), "v"(0xF0u));
        return r;
    } else {
        return (x >> 24) & 0xF0u;
    }
}

// XOR of 32 lookups, two per three-input XOR (v_bitop3 0x96). The byte offset of entry n is
// n << 4, so each lookup's offset is one masked byte: (w >> 8k) & 0xF0 is nibble 2k + 1 of w,
// and the same of w << 4 is nibble 2k. Tables at a constant LDS address fold their row offset
// into the read's immediate.
__device__ __forceinline__ f128 shoup_mul(const f128* __restrict__ table, f128 a) {
    const char* tb = reinterpret_cast<const char*>(table);
    const uint32_t aw[4] = {a.x, a.y, a.z, a.w};
    f128 acc = f128_zero();
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const uint32_t odd = aw[w], even = aw[w] << 4;
        const uint32_t o[8] = {byte_hi_nibble<0>(even), byte_hi_nibble<0>(odd),
                               byte_hi_nibble<1>(even), byte_hi_nibble<1>(odd),
                               byte_hi_nibble<2>(even), byte_hi_nibble<2>(odd),
                               byte_hi_nibble<3>(even), byte_hi_nibble<3>(odd)};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const f128 t0 = *reinterpret_cast<const f128*>(tb + (w * 8 + 2 * k) * 256 + o[2 * k]);
            const f128 t1 =
                *reinterpret_cast<const f128*>(tb + (w * 8 + 2 * k + 1) * 256 + o[2 * k + 1]);
            acc.x = __builtin_amdgcn_bitop3_b32(acc.x, t0.x, t1.x, 0x96);
            acc.y = __builtin_amdgcn_bitop3_b32(acc.y, t0.y, t1.y, 0x96);
            acc.z = __builtin_amdgcn_bitop3_b32(acc.z, t0.z, t1.z, 0x96);
            acc.w = __builtin_amdgcn_bitop3_b32(acc.w, t0.w, t1.w, 0x96);
        }
    }
    return acc;
}

// a * t with an 8-row table of t (rows 0..7 of the Shoup table above, 2 KiB): per 32-bit word
// w of a the XOR of 8 lookups gives S_w = a_w * t, and a * t = sum of x^(32 w) S_w by Horner
// from the top word. acc * x^32 moves words up by one and folds the top word h back as
// h * (x^7 + x^2 + x + 1), which spans bits 0..38.
__device__ __forceinline__ f128 shoup8_word(const char* tb, uint32_t x) {
    const uint32_t even = x << 4;
    const uint32_t o[8] = {byte_hi_nibble<0>(even), byte_hi_nibble<0>(x), byte_hi_nibble<1>(even),
                           byte_hi_nibble<1>(x),    byte_hi_nibble<2>(even), byte_hi_nibble<2>(x),
                           byte_hi_nibble<3>(even), byte_hi_nibble<3>(x)};
    f128 acc = f128_zero();
#pragma unroll
    for (int k = 0; k < 8; k += 2) {
        const f128 t0 = *reinterpret_cast<const f128*>(tb + k * 256 + o[k]);
        const f128 t1 = *reinterpret_cast<const f128*>(tb + (k + 1) * 256 + o[k + 1]);
        acc.x = __builtin_amdgcn_bitop3_b32(acc.x, t0.x, t1.x, 0x96);
        acc.y = __builtin_amdgcn_bitop3_b32(acc.y, t0.y, t1.y, 0x96);
        acc.z = __builtin_amdgcn_bitop3_b32(acc.z, t0.z, t1.z, 0x96);
        acc.w = __builtin_amdgcn_bitop3_b32(acc.w, t0.w, t1.w, 0x96);
    }
    return acc;
}

__device__ __forceinline__ f128 shoup8_mul(const f128* __restrict__ table, f128 a) {
    const char* tb = reinterpret_cast<const char*>(table);
    f128 acc = shoup8_word(tb, a.w);
    const uint32_t aw[3] = {a.x, a.y, a.z};
#pragma unroll
    for (int w = 2; w >= 0; w--) {
        const uint32_t h = acc.w;
        const uint32_t lo = __builtin_amdgcn_bitop3_b32(h, h << 1, h << 2, 0x96) ^ (h << 7);
        const uint32_t hi = __builtin_amdgcn_bitop3_b32(h >> 31, h >> 30, h >> 25, 0x96);
        const f128 s = shoup8_word(tb, aw[w]);
        acc = make_uint4(lo ^ s.x, acc.x ^ hi ^ s.y, acc.y ^ s.z, acc.z ^ s.w);
    }
    return acc;
}

// ---------------------------------------------------------------------------
// Wave / block XOR reductions.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_xor(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v ^= __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ f128 wave_xor_f128(f128 v) {
    return make_uint4(wave_xor(v.x), wave_xor(v.y), wave_xor(v.z), wave_xor(v.w));
}

// Per-wave comb workspace: 16 entries x 64 lanes x 16 B (see comb_build).
constexpr int kCombWsBytes = 4 * 16 * 64 * 16;  // four waves per 256-thread block
__device__ __forceinline__ f128* comb_lane(f128* ws) {
    f128* w = ws + (threadIdx.x >> 6) * 1024 + (threadIdx.x & 63);
    comb_ws_init(w);
    return w;
}

// XOR-reduce two values over the grid into out[0], out[1] in one launch. With one workgroup,
// thread 0 writes out directly. Otherwise each workgroup writes its pair to
// partials[2*blockIdx.x + {0,1}] and takes a ticket from a device-scope counter kept at
// partials[2*kGridReduceMax] (zero between launches); the workgroup that draws the last
// ticket folds every pair (volatile loads, past any stale L1 line), writes out and resets
// the counter. gridDim.x <= kGridReduceMax.
constexpr int kGridReduceMax = 512;

// LDS a grid_xor_pair works in: per-wave pairs, the final pair, the last-block flag, and the
// device transcript step's state.
struct PairShared {
    f128 red[2][4];  // workgroups of at most 256 threads
    f128 fin[2];
    int last;
    flock::FsShared fs;
};

// The pair goes where `ro` says (flock::RoundOut): into ro.out when set, then ro.seq into
// ro.flag (system scope) when set; with ro.fs the workgroup that holds the final pair also runs
// the device transcript's round step on it. `sh` is LDS of the workgroup (a kernel whose
// dynamic LDS is free by then passes that, so the reduction adds no static LDS).
__device__ __forceinline__ void grid_xor_pair_in(f128 a0, f128 a1, f128* partials,
                                                 flock::RoundOut ro, PairShared* sh) {
    f128* out = ro.out;
    auto& red = sh->red;
    auto& fin = sh->fin;
    int& last = sh->last;
    a0 = wave_xor_f128(a0);
    a1 = wave_xor_f128(a1);
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nw = (blockDim.x + 63) >> 6;
    if (lane == 0) {
        red[0][wid] = a0;
        red[1][wid] = a1;
    }
    __syncthreads();
    unsigned* ctr = reinterpret_cast<unsigned*>(partials + 2 * kGridReduceMax);
    if (threadIdx.x == 0) {
        f128 r0 = f128_zero(), r1 = f128_zero();
        for (int w = 0; w < nw; w++) {
            r0 = f128_add(r0, red[0][w]);
            r1 = f128_add(r1, red[1][w]);
        }
        if (gridDim.x == 1) {
            fin[0] = r0;
            fin[1] = r1;
            if (out) {
                out[0] = r0;
                out[1] = r1;
            }
            if (ro.flag) {
                __threadfence_system();
                __hip_atomic_store(ro.flag, ro.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        } else {
            partials[2 * blockIdx.x] = r0;
            partials[2 * blockIdx.x + 1] = r1;
            __threadfence();
            last = atomicAdd(ctr, 1u) == gridDim.x - 1;
        }
    }
    __syncthreads();
    if (gridDim.x > 1) {
        if (!last) return;
        __threadfence();
        const volatile uint32_t* vp = reinterpret_cast<const volatile uint32_t*>(partials);
        f128 c0 = f128_zero(), c1 = f128_zero();
        for (uint32_t i = threadIdx.x; i < gridDim.x; i += blockDim.x) {
            const volatile uint32_t* q = vp + 8 * i;
            c0 = f128_add(c0, make_uint4(q[0], q[1], q[2], q[3]));
            c1 = f128_add(c1, make_uint4(q[4], q[5], q[6], q[7]));
        }
        c0 = wave_xor_f128(c0);
        c1 = wave_xor_f128(c1);
        __syncthreads();  // red is reused
        if (lane == 0) {
            red[0][wid] = c0;
            red[1][wid] = c1;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            f128 r0 = f128_zero(), r1 = f128_zero();
            for (int w = 0; w < nw; w++) {
                r0 = f128_add(r0, red[0][w]);
                r1 = f128_add(r1, red[1][w]);
            }
            fin[0] = r0;
            fin[1] = r1;
            if (out) {
                out[0] = r0;
                out[1] = r1;
            }
            if (ro.flag) {
                __threadfence_system();
                __hip_atomic_store(ro.flag, ro.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            atomicExch(ctr, 0u);
        }
    }
    if (ro.fs) {
        __syncthreads();
        flock::fs_round_step(ro.fs, fin[0], fin[1], &sh->fs);
    }
}

__device__ __forceinline__ void grid_xor_pair(f128 a0, f128 a1, f128* partials,
                                              flock::RoundOut ro) {
    __shared__ PairShared sh;
    grid_xor_pair_in(a0, a1, partials, ro, &sh);
}
