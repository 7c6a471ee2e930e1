// Binary-field arithmetic for the Ligerito hot path, host and device.
//
// Fields are plain polynomial-basis GF(2^n) (binary_field.rs:176-189):
//   GF(2^32)  modulus x^32 + x^15 + x^9 + x^7 + x^4 + x^3 + 1   (low part 0x8299)
//   GF(2^128) modulus x^128 + x^7 + x^2 + x + 1                 (low part 0x87)
// A product is the carry-less product reduced modulo that polynomial; every
// routine below computes exactly that value (binary_field.rs:52-105,132-140).
//
// Multiplication by a constant lambda is GF(2)-linear in the other operand,
// so lambda*w = XOR over the 4-bit windows of w of table[k][window_k] with
// table[k][v] = lambda * v * x^(4k).  A table has 8 (GF(2^32)) or 32
// (GF(2^128)) groups of 16 entries; a wave that shares one lambda reads each
// group from 16 distinct LDS banks, so table lookups do not conflict.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LIG_HD __host__ __device__ inline
#define LIG_UNROLL _Pragma("unroll")
#else
#define LIG_HD inline
#define LIG_UNROLL
#endif

namespace lig {

constexpr uint32_t GF32_POLY = 0x8299u;
constexpr uint64_t GF128_POLY = 0x87u;

// ------------------------------------------------------------- GF(2^32) ---

// Reduces a carry-less product of degree <= 62 modulo the GF(2^32) modulus.
// 0x8299 has degree 15, so two folds suffice.
LIG_HD uint32_t gf32_reduce64(uint64_t x) {
    uint64_t hi = x >> 32;
    uint64_t t = hi ^ (hi << 3) ^ (hi << 4) ^ (hi << 7) ^ (hi << 9) ^ (hi << 15);
    uint32_t lo = (uint32_t)x ^ (uint32_t)t;
    uint32_t h2 = (uint32_t)(t >> 32);  // <= 14 bits
    uint32_t t2 = h2 ^ (h2 << 3) ^ (h2 << 4) ^ (h2 << 7) ^ (h2 << 9) ^ (h2 << 15);
    return lo ^ t2;
}

// Portable carry-less 32x32 -> 64 (branch-free shift-and-add).
LIG_HD uint64_t clmul32_portable(uint32_t a, uint32_t b) {
    uint64_t acc = 0, aa = a;
LIG_UNROLL
    for (int i = 0; i < 32; i++) acc ^= (aa << i) & (0ull - (uint64_t)((b >> i) & 1u));
    return acc;
}

// Carry-less 32x32 -> 64 from integer multiplies ("multiplication with
// holes", as in BearSSL's bmul32): each operand splits into the bits of
// four residues mod 4; an integer product of two such parts has at most 8
// terms at any position, so its carries reach at most 3 positions up and
// never the next position of the same residue, which the final masks keep.
// 16 integer products instead of 32 shift-and-mask steps.
LIG_HD uint64_t clmul32_holes(uint32_t a, uint32_t b) {
    const uint32_t a0 = a & 0x11111111u, a1 = a & 0x22222222u, a2 = a & 0x44444444u, a3 = a & 0x88888888u;
    const uint32_t b0 = b & 0x11111111u, b1 = b & 0x22222222u, b2 = b & 0x44444444u, b3 = b & 0x88888888u;
    const uint64_t A0 = a0, A1 = a1, A2 = a2, A3 = a3;
    const uint64_t z0 = (A0 * b0) ^ (A1 * b3) ^ (A2 * b2) ^ (A3 * b1);
    const uint64_t z1 = (A0 * b1) ^ (A1 * b0) ^ (A2 * b3) ^ (A3 * b2);
    const uint64_t z2 = (A0 * b2) ^ (A1 * b1) ^ (A2 * b0) ^ (A3 * b3);
    const uint64_t z3 = (A0 * b3) ^ (A1 * b2) ^ (A2 * b1) ^ (A3 * b0);
    return (z0 & 0x1111111111111111ull) | (z1 & 0x2222222222222222ull) | (z2 & 0x4444444444444444ull) |
           (z3 & 0x8888888888888888ull);
}

// lambda * v * x^(4k) for a 4-bit v: one entry of a multiply-by-lambda table.
LIG_HD uint32_t gf32_table_entry(uint32_t lambda, int k, uint32_t v) {
    uint64_t p = 0, l = lambda;
LIG_UNROLL
    for (int i = 0; i < 4; i++) p ^= (l << i) & (0ull - (uint64_t)((v >> i) & 1u));
    return gf32_reduce64(p << (4 * k));
}

// Word e of the 128-word multiply-by-lambda table the kernels use: entry v
// (0..15) of nibble position k (0..7) sits at word k * 16 + v, so the 16
// entries of one position fall in 16 distinct LDS banks.
LIG_HD uint32_t gf32_tab_word(uint32_t lambda, int e) { return gf32_table_entry(lambda, e >> 4, (uint32_t)(e & 15)); }

// ------------------------------------------------------------ GF(2^128) ---

struct u128v {
    uint64_t lo, hi;
};
LIG_HD u128v operator^(u128v a, u128v b) { return {a.lo ^ b.lo, a.hi ^ b.hi}; }
LIG_HD u128v& operator^=(u128v& a, u128v b) {
    a.lo ^= b.lo;
    a.hi ^= b.hi;
    return a;
}
LIG_HD bool operator==(u128v a, u128v b) { return a.lo == b.lo && a.hi == b.hi; }

// Reduces a 256-bit carry-less product (w3:w2:w1:w0) modulo x^128 + 0x87.
LIG_HD u128v gf128_reduce256(uint64_t w0, uint64_t w1, uint64_t w2, uint64_t w3) {
    // x^128 == 0x87: fold the high 128 bits, multiplied by 0x87 (x^7+x^2+x+1).
    // hi * 0x87 spans up to 135 bits; the 7 bits above 128 fold once more.
    uint64_t c2 = w2 ^ (w2 << 1) ^ (w2 << 2) ^ (w2 << 7);
    uint64_t c3 = w3 ^ (w3 << 1) ^ (w3 << 2) ^ (w3 << 7) ^ (w2 >> 63) ^ (w2 >> 62) ^ (w2 >> 57);
    uint64_t over = (w3 >> 63) ^ (w3 >> 62) ^ (w3 >> 57);  // bits 128.. of hi*0x87
    uint64_t f = over ^ (over << 1) ^ (over << 2) ^ (over << 7);
    return {w0 ^ c2 ^ f, w1 ^ c3};
}

// Portable carry-less 64x64 -> 128.
LIG_HD void clmul64_portable(uint64_t a, uint64_t b, uint64_t* lo, uint64_t* hi) {
    uint64_t l = 0, h = 0;
LIG_UNROLL
    for (int i = 0; i < 64; i++) {
        uint64_t m = 0ull - ((b >> i) & 1ull);
        l ^= (a << i) & m;
        h ^= (i ? (a >> (64 - i)) : 0ull) & m;
    }
    *lo = l;
    *hi = h;
}

// Karatsuba over 64-bit limbs, then reduction.
LIG_HD u128v gf128_mul_portable(u128v a, u128v b) {
    uint64_t l0, l1, h0, h1, m0, m1;
    clmul64_portable(a.lo, b.lo, &l0, &l1);
    clmul64_portable(a.hi, b.hi, &h0, &h1);
    clmul64_portable(a.lo ^ a.hi, b.lo ^ b.hi, &m0, &m1);
    m0 ^= l0 ^ h0;
    m1 ^= l1 ^ h1;
    return gf128_reduce256(l0, l1 ^ m0, h0 ^ m1, h1);
}

// lambda * v * x^(4k), k < 32, v < 16.
LIG_HD u128v gf128_table_entry(u128v lambda, int k, uint32_t v) {
    // p = lambda * v (<= 131 bits) as three words.
    uint64_t p0 = 0, p1 = 0, p2 = 0;
LIG_UNROLL
    for (int i = 0; i < 4; i++) {
        uint64_t m = 0ull - (uint64_t)((v >> i) & 1u);
        p0 ^= (lambda.lo << i) & m;
        p1 ^= ((lambda.hi << i) | (i ? lambda.lo >> (64 - i) : 0ull)) & m;
        p2 ^= (i ? lambda.hi >> (64 - i) : 0ull) & m;
    }
    // shift left by 4k bits into four words.
    int sh = 4 * k;
    uint64_t w[4] = {0, 0, 0, 0};
    int q = sh >> 6, r = sh & 63;
    uint64_t src[3] = {p0, p1, p2};
LIG_UNROLL
    for (int i = 0; i < 3; i++) {
        // place src[i] << r at word i+q (and carry into i+q+1)
        for (int o = 0; o < 4; o++) {
            if (o == i + q) w[o] ^= src[i] << r;
            if (r && o == i + q + 1) w[o] ^= src[i] >> (64 - r);
        }
    }
    return gf128_reduce256(w[0], w[1], w[2], w[3]);
}

}  // namespace lig
