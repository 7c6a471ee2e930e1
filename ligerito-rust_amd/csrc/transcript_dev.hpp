// Fiat–Shamir transcript on the device (emulated_fs.rs:6-75): the SHA-256
// running hash of the seed and every absorbed byte string, where a squeeze
// absorbs its u32 little-endian counter and finalizes a copy.
//
// The state mirrors the host Transcript exactly, so the host can hand a
// transcript to the device, let a chain of kernels absorb and squeeze, and
// replay the same steps afterwards to continue from the identical state.
// Every function here is called with the same arguments by all lanes of a
// wave (wave-uniform), so the compiler keeps the state in scalar registers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gf.hpp"
#include "kernels.hpp"

namespace lig {
namespace dfs {

__device__ __forceinline__ uint32_t rotr(uint32_t x, int n) { return __builtin_amdgcn_alignbit(x, x, n); }
// a ^ b ^ c in one instruction (v_bitop3_b32, truth table 0x96)
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) { return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96); }
__device__ __forceinline__ uint32_t big_s0(uint32_t a) { return xor3(rotr(a, 2), rotr(a, 13), rotr(a, 22)); }
__device__ __forceinline__ uint32_t big_s1(uint32_t e) { return xor3(rotr(e, 6), rotr(e, 11), rotr(e, 25)); }
__device__ __forceinline__ uint32_t small_s0(uint32_t w) { return xor3(rotr(w, 7), rotr(w, 18), w >> 3); }
__device__ __forceinline__ uint32_t small_s1(uint32_t w) { return xor3(rotr(w, 17), rotr(w, 19), w >> 10); }
// Ch and Maj as one three-input bit operation each (truth tables over
// (x, y, z) = (0xF0, 0xCC, 0xAA)).
__device__ __forceinline__ uint32_t ch(uint32_t x, uint32_t y, uint32_t z) { return __builtin_amdgcn_bitop3_b32(x, y, z, 0xCA); }
__device__ __forceinline__ uint32_t maj(uint32_t x, uint32_t y, uint32_t z) { return __builtin_amdgcn_bitop3_b32(x, y, z, 0xE8); }
// Three-operand add; written out because the compiler keeps these sums as
// chains of two-operand adds, and one wave's SHA-256 latency is its
// instruction count.
__device__ __forceinline__ uint32_t add3(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_add3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Values every lane loaded from one address, moved to scalar registers so
// the transcript arithmetic runs on the scalar unit.
__device__ __forceinline__ uint32_t uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t uniform_acc(uint64_t x) {
    return (uint64_t)uniform((uint32_t)x) | ((uint64_t)uniform((uint32_t)(x >> 32)) << 32);
}
__device__ __forceinline__ u128v uniform_acc(u128v x) { return {uniform_acc(x.lo), uniform_acc(x.hi)}; }

// FIPS 180-4 round constants.
__constant__ static const uint32_t K[64] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,
    0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u,
    0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
    0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u,
    0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,
    0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
    0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
    0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};

// One compression of the big-endian message words w[16] into h[8].
__device__ __forceinline__ void compress_inline(uint32_t h[8], const uint32_t w_in[16]) {
    uint32_t w[16];
#pragma unroll
    for (int i = 0; i < 16; i++) w[i] = w_in[i];
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
#pragma unroll
    for (int i = 0; i < 64; i++) {
        uint32_t wi;
        if (i < 16) {
            wi = w[i];
        } else {
            wi = add3(w[i & 15], small_s0(w[(i - 15) & 15]), w[(i - 7) & 15]) + small_s1(w[(i - 2) & 15]);
            w[i & 15] = wi;
        }
        const uint32_t t1 = add3(hh + K[i], big_s1(e), ch(e, f, g)) + wi;
        const uint32_t na = add3(t1, big_s0(a), maj(a, b, c));
        hh = g;
        g = f;
        f = e;
        e = d + t1;
        d = c;
        c = b;
        b = a;
        a = na;
    }
    h[0] += a;
    h[1] += b;
    h[2] += c;
    h[3] += d;
    h[4] += e;
    h[5] += f;
    h[6] += g;
    h[7] += hh;
}

// The compression as one shared out-of-line function: a transcript step
// compresses at several sites (an absorbed block, a squeeze's counter, one
// or two finalization blocks), and inlined copies are each fetched cold the
// first time a launch reaches them (a section's leading step takes paths
// the other rounds do not).  Scalars in and out keep every value in
// registers under the calling convention.
struct Sha8 {
    uint32_t v[8];
};
__device__ __noinline__ Sha8 compress_shared(uint32_t h0, uint32_t h1, uint32_t h2, uint32_t h3, uint32_t h4, uint32_t h5,
                                             uint32_t h6, uint32_t h7, uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3,
                                             uint32_t w4, uint32_t w5, uint32_t w6, uint32_t w7, uint32_t w8, uint32_t w9,
                                             uint32_t w10, uint32_t w11, uint32_t w12, uint32_t w13, uint32_t w14,
                                             uint32_t w15) {
    uint32_t h[8] = {h0, h1, h2, h3, h4, h5, h6, h7};
    const uint32_t w[16] = {w0, w1, w2, w3, w4, w5, w6, w7, w8, w9, w10, w11, w12, w13, w14, w15};
    compress_inline(h, w);
    Sha8 r;
#pragma unroll
    for (int i = 0; i < 8; i++) r.v[i] = h[i];
    return r;
}
__device__ __forceinline__ void compress(uint32_t h[8], const uint32_t w[16]) {
    const Sha8 r = compress_shared(h[0], h[1], h[2], h[3], h[4], h[5], h[6], h[7], w[0], w[1], w[2], w[3], w[4], w[5], w[6],
                                   w[7], w[8], w[9], w[10], w[11], w[12], w[13], w[14], w[15]);
#pragma unroll
    for (int i = 0; i < 8; i++) h[i] = r.v[i];
}

// Working copy of a DevTranscript in registers.
struct State {
    uint32_t h[8];
    uint32_t buf[16];  // pending block, big-endian words, zero past blen
    uint32_t blen;
    uint64_t total;
    uint32_t counter;

    // from a state held in kernel arguments (member reads, no address taken)
    __device__ void load_value(const DevTranscript& v) {
#pragma unroll
        for (int i = 0; i < 8; i++) h[i] = v.h[i];
#pragma unroll
        for (int i = 0; i < 16; i++) buf[i] = v.buf[i];
        blen = v.blen;
        total = (uint64_t)v.total_lo | ((uint64_t)v.total_hi << 32);
        counter = v.counter;
    }
    __device__ void load(const DevTranscript* s) {
#pragma unroll
        for (int i = 0; i < 8; i++) h[i] = uniform(s->h[i]);
#pragma unroll
        for (int i = 0; i < 16; i++) buf[i] = uniform(s->buf[i]);
        blen = uniform(s->blen);
        total = (uint64_t)uniform(s->total_lo) | ((uint64_t)uniform(s->total_hi) << 32);
        counter = uniform(s->counter);
    }
    // written by one lane (vector stores)
    __device__ void store(DevTranscript* s) const {
#pragma unroll
        for (int i = 0; i < 8; i++) s->h[i] = h[i];
#pragma unroll
        for (int i = 0; i < 16; i++) s->buf[i] = buf[i];
        s->blen = blen;
        s->total_lo = (uint32_t)total;
        s->total_hi = (uint32_t)(total >> 32);
        s->counter = counter;
    }
    // Absorbs the 4 little-endian bytes of v.  Every string the protocol
    // absorbs (seed, roots, field elements, counters) is a whole number of
    // 32-bit words, so the pending length stays a multiple of 4.
    __device__ void word_le(uint32_t v) {
        const uint32_t be = __builtin_bswap32(v);
        // words past the pending length are zero: OR the word in with a
        // per-word mask (a conditional store here becomes an indexed store,
        // which puts buf in scratch memory)
#pragma unroll
        for (int i = 0; i < 16; i++) buf[i] |= be & (0u - (uint32_t)((uint32_t)i == (blen >> 2)));
        blen += 4;
        total += 4;
        if (blen == 64) {
            compress(h, buf);
#pragma unroll
            for (int i = 0; i < 16; i++) buf[i] = 0;
            blen = 0;
        }
    }
    // squeeze (emulated_fs.rs:22-32): absorb the counter, finalize a copy;
    // out = the first 16 digest bytes as four little-endian words
    __device__ void squeeze(uint32_t out[4]) {
        word_le(counter);
        counter++;
        uint32_t hc[8], bc[16];
#pragma unroll
        for (int i = 0; i < 8; i++) hc[i] = h[i];
#pragma unroll
        for (int i = 0; i < 16; i++) bc[i] = buf[i];
        // padding: 0x80, zeros, 64-bit big-endian bit length
#pragma unroll
        for (int i = 0; i < 16; i++) bc[i] |= 0x80000000u & (0u - (uint32_t)((uint32_t)i == (blen >> 2)));
        if (blen >= 56) {
            compress(hc, bc);
#pragma unroll
            for (int i = 0; i < 16; i++) bc[i] = 0;
        }
        const uint64_t bits = total * 8;
        bc[14] = (uint32_t)(bits >> 32);
        bc[15] = (uint32_t)bits;
        compress(hc, bc);
#pragma unroll
        for (int i = 0; i < 4; i++) out[i] = __builtin_bswap32(hc[i]);
    }
};

}  // namespace dfs
}  // namespace lig
