// BLAKE3 (hash mode, 32-byte output) for host and device, written from the
// published BLAKE3 specification.  The reference hashes Merkle leaves with
// blake3::hash and parents with blake3::hash(left || right)
// (merkle_tree.rs:15,50,114-117,176).
//
// The device helpers hash one message held in registers/LDS/global memory by
// a single lane: chunks are compressed in order and chunk chaining values
// are merged with the standard stack rule, so any length is supported.
#pragma once
#include <stddef.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define B3_HD __host__ __device__ inline
#else
#define B3_HD inline
#endif

namespace lig {
namespace b3 {

enum : uint32_t { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8 };

B3_HD uint32_t iv(int i) {
    switch (i) {
        case 0: return 0x6A09E667u;
        case 1: return 0xBB67AE85u;
        case 2: return 0x3C6EF372u;
        case 3: return 0xA54FF53Au;
        case 4: return 0x510E527Fu;
        case 5: return 0x9B05688Cu;
        case 6: return 0x1F83D9ABu;
        default: return 0x5BE0CD19u;
    }
}

B3_HD uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

#define LIG_B3_G(a, b, c, d, x, y)      \
    a = a + b + (x);                    \
    d = rotr(d ^ a, 16);                \
    c = c + d;                          \
    b = rotr(b ^ c, 12);                \
    a = a + b + (y);                    \
    d = rotr(d ^ a, 8);                 \
    c = c + d;                          \
    b = rotr(b ^ c, 7);

// One round on the 16-word state with message words already permuted for
// that round (schedule applied by the caller via the index table below).
#define LIG_B3_ROUND(m0, m1, m2, m3, m4, m5, m6, m7, m8, m9, m10, m11, m12, m13, m14, m15) \
    LIG_B3_G(s0, s4, s8, s12, m0, m1)                                                       \
    LIG_B3_G(s1, s5, s9, s13, m2, m3)                                                       \
    LIG_B3_G(s2, s6, s10, s14, m4, m5)                                                      \
    LIG_B3_G(s3, s7, s11, s15, m6, m7)                                                      \
    LIG_B3_G(s0, s5, s10, s15, m8, m9)                                                      \
    LIG_B3_G(s1, s6, s11, s12, m10, m11)                                                    \
    LIG_B3_G(s2, s7, s8, s13, m12, m13)                                                     \
    LIG_B3_G(s3, s4, s9, s14, m14, m15)

// Compression function; writes the 8-word chaining value (first half of
// the output).  The message schedule for rounds 1..6 is the permutation
// [2,6,3,10,7,0,4,13,1,11,12,5,9,14,15,8] applied repeatedly, unrolled.
B3_HD void compress_cv(const uint32_t cv[8], const uint32_t m[16], uint64_t counter, uint32_t block_len,
                       uint32_t flags, uint32_t out[8]) {
    uint32_t s0 = cv[0], s1 = cv[1], s2 = cv[2], s3 = cv[3], s4 = cv[4], s5 = cv[5], s6 = cv[6], s7 = cv[7];
    uint32_t s8 = iv(0), s9 = iv(1), s10 = iv(2), s11 = iv(3);
    uint32_t s12 = (uint32_t)counter, s13 = (uint32_t)(counter >> 32), s14 = block_len, s15 = flags;
    LIG_B3_ROUND(m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], m[9], m[10], m[11], m[12], m[13], m[14], m[15])
    LIG_B3_ROUND(m[2], m[6], m[3], m[10], m[7], m[0], m[4], m[13], m[1], m[11], m[12], m[5], m[9], m[14], m[15], m[8])
    LIG_B3_ROUND(m[3], m[4], m[10], m[12], m[13], m[2], m[7], m[14], m[6], m[5], m[9], m[0], m[11], m[15], m[8], m[1])
    LIG_B3_ROUND(m[10], m[7], m[12], m[9], m[14], m[3], m[13], m[15], m[4], m[0], m[11], m[2], m[5], m[8], m[1], m[6])
    LIG_B3_ROUND(m[12], m[13], m[9], m[11], m[15], m[10], m[14], m[8], m[7], m[2], m[5], m[3], m[0], m[1], m[6], m[4])
    LIG_B3_ROUND(m[9], m[14], m[11], m[5], m[8], m[12], m[15], m[1], m[13], m[3], m[0], m[10], m[2], m[6], m[4], m[7])
    LIG_B3_ROUND(m[11], m[15], m[5], m[0], m[1], m[9], m[8], m[6], m[14], m[10], m[2], m[12], m[3], m[4], m[7], m[13])
    out[0] = s0 ^ s8;
    out[1] = s1 ^ s9;
    out[2] = s2 ^ s10;
    out[3] = s3 ^ s11;
    out[4] = s4 ^ s12;
    out[5] = s5 ^ s13;
    out[6] = s6 ^ s14;
    out[7] = s7 ^ s15;
}

// Hash of a 64-byte message (two 32-byte node hashes): one block, one chunk,
// flags CHUNK_START | CHUNK_END | ROOT.  Used for Merkle parents.
B3_HD void hash64(const uint32_t m[16], uint32_t out[8]) {
    uint32_t cv[8];
    for (int i = 0; i < 8; i++) cv[i] = iv(i);
    compress_cv(cv, m, 0, 64, CHUNK_START | CHUNK_END | ROOT, out);
}

// Streaming hasher over 64-byte blocks supplied by the caller (the last
// block zero-padded).  Keeps a stack of chunk chaining values.
struct Hasher {
    uint32_t cv[8];
    uint32_t stack[8][8];  // up to 2^8 chunks = 256 KiB messages
    int depth;
    uint64_t chunk;     // index of the current chunk
    uint32_t blocks;    // blocks already compressed in the current chunk

    B3_HD void init() {
        for (int i = 0; i < 8; i++) cv[i] = iv(i);
        depth = 0;
        chunk = 0;
        blocks = 0;
    }

    // Absorbs a full 64-byte block that is not the message's last block.
    B3_HD void block(const uint32_t m[16]) {
        if (blocks == 15) {
            // last block of a full chunk (not the message's last): finish chunk
            uint32_t out[8];
            compress_cv(cv, m, chunk, 64, CHUNK_END | (blocks == 0 ? CHUNK_START : 0), out);
            push_chunk(out);
        } else {
            uint32_t out[8];
            compress_cv(cv, m, chunk, 64, blocks == 0 ? CHUNK_START : 0, out);
            for (int i = 0; i < 8; i++) cv[i] = out[i];
            blocks++;
        }
    }

    B3_HD void push_chunk(const uint32_t ccv[8]) {
        uint32_t cur[8];
        for (int i = 0; i < 8; i++) cur[i] = ccv[i];
        uint64_t total = chunk + 1;
        while ((total & 1) == 0) {
            depth--;
            uint32_t m[16];
            for (int i = 0; i < 8; i++) {
                m[i] = stack[depth][i];
                m[8 + i] = cur[i];
            }
            uint32_t k[8];
            for (int i = 0; i < 8; i++) k[i] = iv(i);
            compress_cv(k, m, 0, 64, PARENT, cur);
            total >>= 1;
        }
        for (int i = 0; i < 8; i++) stack[depth][i] = cur[i];
        depth++;
        chunk++;
        blocks = 0;
        for (int i = 0; i < 8; i++) cv[i] = iv(i);
    }

    // Absorbs the message's last block (len bytes, 1..64; zero-padded) and
    // writes the 32-byte hash.  Messages are never empty here.
    B3_HD void finish(const uint32_t m[16], uint32_t len, uint32_t out[8]) {
        uint32_t flags = CHUNK_END | (blocks == 0 ? CHUNK_START : 0);
        if (depth == 0) {
            compress_cv(cv, m, chunk, len, flags | ROOT, out);
            return;
        }
        uint32_t node[8];
        compress_cv(cv, m, chunk, len, flags, node);
        for (int d = depth - 1; d >= 0; d--) {
            uint32_t pm[16];
            for (int i = 0; i < 8; i++) {
                pm[i] = stack[d][i];
                pm[8 + i] = node[i];
            }
            uint32_t k[8];
            for (int i = 0; i < 8; i++) k[i] = iv(i);
            compress_cv(k, pm, 0, 64, PARENT | (d == 0 ? ROOT : 0), node);
        }
        for (int i = 0; i < 8; i++) out[i] = node[i];
    }
};

}  // namespace b3
}  // namespace lig
