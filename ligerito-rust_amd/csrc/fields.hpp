// Device field traits used by the kernels: element type, multiply-by-table,
// table construction, general multiply and the unreduced accumulator used
// for inner products (reduction is linear, so sums of unreduced carry-less
// products are reduced once).
#pragma once
#include <hip/hip_runtime.h>

#include "gf.hpp"
#include "kernels.hpp"

namespace lig {

template <class F> struct Ops;

template <> struct Ops<F32> {
    using T = uint32_t;
    using Acc = uint64_t;

    __device__ static T zero() { return 0u; }
    __device__ static T one() { return 1u; }
    __device__ static T from_u64(uint64_t v) { return (uint32_t)v; }
    __device__ static T table_entry(T lam, int e) { return gf32_tab_word(lam, e); }
    // table layout of gf32_tab_word: nibble k, value v at word k * 16 + v
    __device__ static T mul_tab(const T* __restrict__ t, T w) {
        T r0 = t[w & 15u] ^ t[16 + ((w >> 4) & 15u)];
        T r1 = t[32 + ((w >> 8) & 15u)] ^ t[48 + ((w >> 12) & 15u)];
        T r2 = t[64 + ((w >> 16) & 15u)] ^ t[80 + ((w >> 20) & 15u)];
        T r3 = t[96 + ((w >> 24) & 15u)] ^ t[112 + (w >> 28)];
        return (r0 ^ r1) ^ (r2 ^ r3);
    }
    __device__ static T mul(T a, T b) { return gf32_reduce64(clmul32_holes(a, b)); }
    __device__ static Acc acc_zero() { return 0ull; }
    __device__ static void acc_mul(Acc& acc, T a, T b) { acc ^= clmul32_holes(a, b); }
    __device__ static T acc_reduce(Acc a) { return gf32_reduce64(a); }
    __device__ static bool eq(T a, T b) { return a == b; }
    // little-endian bytes as 32-bit words (BinaryField::to_bytes)
    static constexpr int WORDS = 1;
    __device__ static uint32_t word(T v, int) { return v; }
    // get_field: the low bits of the squeezed u128 (four little-endian words)
    __device__ static T from_squeeze(const uint32_t o[4]) { return o[0]; }
    // multiplication by the constants 2 and 3 used by the sumcheck (x = 3, 1 + x = 2)
    __device__ static T mul2(T a) { return (a << 1) ^ ((0u - (a >> 31)) & GF32_POLY); }
};

template <> struct Ops<F128> {
    using T = u128v;
    using Acc = u128v;

    __device__ static T zero() { return {0ull, 0ull}; }
    __device__ static T one() { return {1ull, 0ull}; }
    __device__ static T from_u64(uint64_t v) { return {v, 0ull}; }
    __device__ static T table_entry(T lam, int e) { return gf128_table_entry(lam, e >> 4, (uint32_t)(e & 15)); }
    __device__ static T mul_tab(const T* __restrict__ t, T w) {
        T r = {0ull, 0ull};
#pragma unroll
        for (int k = 0; k < 16; k++) r ^= t[16 * k + (int)((w.lo >> (4 * k)) & 15u)];
#pragma unroll
        for (int k = 0; k < 16; k++) r ^= t[256 + 16 * k + (int)((w.hi >> (4 * k)) & 15u)];
        return r;
    }
    __device__ static T mul(T a, T b) { return gf128_mul_portable(a, b); }
    __device__ static Acc acc_zero() { return {0ull, 0ull}; }
    __device__ static void acc_mul(Acc& acc, T a, T b) { acc ^= gf128_mul_portable(a, b); }
    __device__ static T acc_reduce(Acc a) { return a; }
    __device__ static bool eq(T a, T b) { return a.lo == b.lo && a.hi == b.hi; }
    static constexpr int WORDS = 4;
    __device__ static uint32_t word(T v, int i) {
        const uint64_t x = i < 2 ? v.lo : v.hi;
        return (uint32_t)(x >> (32 * (i & 1)));
    }
    __device__ static T from_squeeze(const uint32_t o[4]) {
        return {(uint64_t)o[0] | ((uint64_t)o[1] << 32), (uint64_t)o[2] | ((uint64_t)o[3] << 32)};
    }
    __device__ static T mul2(T a) {
        uint64_t c = a.hi >> 63;
        return {(a.lo << 1) ^ ((0ull - c) & GF128_POLY), (a.hi << 1) | (a.lo >> 63)};
    }
};

}  // namespace lig
