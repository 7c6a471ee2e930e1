// Host utilities: error strings, SHA-256, Fiat–Shamir transcript, BLAKE3 over
// byte strings, Merkle proof/verify logic.
#include <cpuid.h>
#include <stdlib.h>
#include <immintrin.h>

#include <algorithm>
#include <array>
#include <unordered_set>

#include "host.hpp"

namespace lig {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

// ----------------------------------------------------------------- SHA-256 --
// FIPS 180-4.
namespace {
const uint32_t kSha[64] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,
    0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u,
    0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
    0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u,
    0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,
    0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
    0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
    0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};

inline uint32_t ror(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

void sha_block(uint32_t st[8], const uint8_t* p) {
    uint32_t w[64];
    for (int i = 0; i < 16; i++) w[i] = (uint32_t)p[4 * i] << 24 | (uint32_t)p[4 * i + 1] << 16 | (uint32_t)p[4 * i + 2] << 8 | p[4 * i + 3];
    for (int i = 16; i < 64; i++)
        w[i] = w[i - 16] + (ror(w[i - 15], 7) ^ ror(w[i - 15], 18) ^ (w[i - 15] >> 3)) + w[i - 7] +
               (ror(w[i - 2], 17) ^ ror(w[i - 2], 19) ^ (w[i - 2] >> 10));
    uint32_t v[8];
    memcpy(v, st, 32);
    for (int i = 0; i < 64; i++) {
        uint32_t t1 = v[7] + (ror(v[4], 6) ^ ror(v[4], 11) ^ ror(v[4], 25)) + ((v[4] & v[5]) ^ (~v[4] & v[6])) + kSha[i] + w[i];
        uint32_t t2 = (ror(v[0], 2) ^ ror(v[0], 13) ^ ror(v[0], 22)) + ((v[0] & v[1]) ^ (v[0] & v[2]) ^ (v[1] & v[2]));
        memmove(v + 1, v, 7 * sizeof(uint32_t));
        v[4] += t1;
        v[0] = t1 + t2;
    }
    for (int i = 0; i < 8; i++) st[i] += v[i];
}
// SHA-256 block function with the x86 SHA extensions (same FIPS 180-4
// compression, four rounds per sha256rnds2 pair).
__attribute__((target("sha,sse4.1,ssse3"))) void sha_blocks_ni(uint32_t st[8], const uint8_t* p, size_t nblocks) {
    const __m128i MASK = _mm_set_epi64x(0x0c0d0e0f08090a0bULL, 0x0405060700010203ULL);
    __m128i tmp = _mm_loadu_si128((const __m128i*)&st[0]);
    __m128i s1 = _mm_loadu_si128((const __m128i*)&st[4]);
    tmp = _mm_shuffle_epi32(tmp, 0xB1);
    s1 = _mm_shuffle_epi32(s1, 0x1B);
    __m128i s0 = _mm_alignr_epi8(tmp, s1, 8);
    s1 = _mm_blend_epi16(s1, tmp, 0xF0);
    while (nblocks--) {
        const __m128i save0 = s0, save1 = s1;
        __m128i w[4];
        for (int g = 0; g < 16; g++) {
            __m128i wg;
            if (g < 4) {
                wg = _mm_shuffle_epi8(_mm_loadu_si128((const __m128i*)(p + 16 * g)), MASK);
            } else {
                __m128i t = _mm_sha256msg1_epu32(w[(g - 4) & 3], w[(g - 3) & 3]);
                t = _mm_add_epi32(t, _mm_alignr_epi8(w[(g - 1) & 3], w[(g - 2) & 3], 4));
                wg = _mm_sha256msg2_epu32(t, w[(g - 1) & 3]);
            }
            w[g & 3] = wg;
            __m128i m = _mm_add_epi32(wg, _mm_loadu_si128((const __m128i*)&kSha[4 * g]));
            s1 = _mm_sha256rnds2_epu32(s1, s0, m);
            m = _mm_shuffle_epi32(m, 0x0E);
            s0 = _mm_sha256rnds2_epu32(s0, s1, m);
        }
        s0 = _mm_add_epi32(s0, save0);
        s1 = _mm_add_epi32(s1, save1);
        p += 64;
    }
    tmp = _mm_shuffle_epi32(s0, 0x1B);
    s1 = _mm_shuffle_epi32(s1, 0xB1);
    s0 = _mm_blend_epi16(tmp, s1, 0xF0);
    s1 = _mm_alignr_epi8(s1, tmp, 8);
    _mm_storeu_si128((__m128i*)&st[0], s0);
    _mm_storeu_si128((__m128i*)&st[4], s1);
}

// Two independent one-block compressions interleaved instruction by
// instruction: a single chain of sha256rnds2 waits on each round's latency,
// two chains keep the unit busy.
__attribute__((target("sha,sse4.1,ssse3"))) void sha_ni_x2(uint32_t sa[8], const uint8_t* pa, uint32_t sb[8],
                                                           const uint8_t* pb) {
    const __m128i MASK = _mm_set_epi64x(0x0c0d0e0f08090a0bULL, 0x0405060700010203ULL);
    __m128i a0, a1, b0, b1;
    {
        __m128i t = _mm_shuffle_epi32(_mm_loadu_si128((const __m128i*)&sa[0]), 0xB1);
        a1 = _mm_shuffle_epi32(_mm_loadu_si128((const __m128i*)&sa[4]), 0x1B);
        a0 = _mm_alignr_epi8(t, a1, 8);
        a1 = _mm_blend_epi16(a1, t, 0xF0);
        t = _mm_shuffle_epi32(_mm_loadu_si128((const __m128i*)&sb[0]), 0xB1);
        b1 = _mm_shuffle_epi32(_mm_loadu_si128((const __m128i*)&sb[4]), 0x1B);
        b0 = _mm_alignr_epi8(t, b1, 8);
        b1 = _mm_blend_epi16(b1, t, 0xF0);
    }
    const __m128i sa0 = a0, sa1 = a1, sb0 = b0, sb1 = b1;
    __m128i wa[4], wb[4];
    for (int g = 0; g < 16; g++) {
        __m128i ga, gb;
        if (g < 4) {
            ga = _mm_shuffle_epi8(_mm_loadu_si128((const __m128i*)(pa + 16 * g)), MASK);
            gb = _mm_shuffle_epi8(_mm_loadu_si128((const __m128i*)(pb + 16 * g)), MASK);
        } else {
            __m128i ta = _mm_sha256msg1_epu32(wa[(g - 4) & 3], wa[(g - 3) & 3]);
            __m128i tb = _mm_sha256msg1_epu32(wb[(g - 4) & 3], wb[(g - 3) & 3]);
            ta = _mm_add_epi32(ta, _mm_alignr_epi8(wa[(g - 1) & 3], wa[(g - 2) & 3], 4));
            tb = _mm_add_epi32(tb, _mm_alignr_epi8(wb[(g - 1) & 3], wb[(g - 2) & 3], 4));
            ga = _mm_sha256msg2_epu32(ta, wa[(g - 1) & 3]);
            gb = _mm_sha256msg2_epu32(tb, wb[(g - 1) & 3]);
        }
        wa[g & 3] = ga;
        wb[g & 3] = gb;
        const __m128i k = _mm_loadu_si128((const __m128i*)&kSha[4 * g]);
        __m128i ma = _mm_add_epi32(ga, k), mb = _mm_add_epi32(gb, k);
        a1 = _mm_sha256rnds2_epu32(a1, a0, ma);
        b1 = _mm_sha256rnds2_epu32(b1, b0, mb);
        ma = _mm_shuffle_epi32(ma, 0x0E);
        mb = _mm_shuffle_epi32(mb, 0x0E);
        a0 = _mm_sha256rnds2_epu32(a0, a1, ma);
        b0 = _mm_sha256rnds2_epu32(b0, b1, mb);
    }
    a0 = _mm_add_epi32(a0, sa0);
    a1 = _mm_add_epi32(a1, sa1);
    b0 = _mm_add_epi32(b0, sb0);
    b1 = _mm_add_epi32(b1, sb1);
    __m128i t = _mm_shuffle_epi32(a0, 0x1B);
    a1 = _mm_shuffle_epi32(a1, 0xB1);
    _mm_storeu_si128((__m128i*)&sa[0], _mm_blend_epi16(t, a1, 0xF0));
    _mm_storeu_si128((__m128i*)&sa[4], _mm_alignr_epi8(a1, t, 8));
    t = _mm_shuffle_epi32(b0, 0x1B);
    b1 = _mm_shuffle_epi32(b1, 0xB1);
    _mm_storeu_si128((__m128i*)&sb[0], _mm_blend_epi16(t, b1, 0xF0));
    _mm_storeu_si128((__m128i*)&sb[4], _mm_alignr_epi8(b1, t, 8));
}

bool have_sha_ni() {
    static int cached = -1;
    if (cached < 0) {
        unsigned a = 0, b = 0, c = 0, d = 0;
        cached = (__get_cpuid_count(7, 0, &a, &b, &c, &d) && (b & (1u << 29))) ? 1 : 0;
        if (getenv("LIG_NO_SHA_NI")) cached = 0;
    }
    return cached == 1;
}

void compress(uint32_t st[8], const uint8_t* p, size_t nblocks) {
    if (have_sha_ni()) {
        sha_blocks_ni(st, p, nblocks);
        return;
    }
    for (size_t i = 0; i < nblocks; i++) sha_block(st, p + 64 * i);
}
}  // namespace

Sha256::Sha256() {
    static const uint32_t h0[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                                   0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
    memcpy(h, h0, 32);
}

void Sha256::update(const void* data, size_t len) {
    const uint8_t* p = (const uint8_t*)data;
    total += len;
    if (blen) {
        size_t t = std::min<size_t>(64 - blen, len);
        memcpy(buf + blen, p, t);
        blen += (uint32_t)t;
        p += t;
        len -= t;
        if (blen == 64) {
            compress(h, buf, 1);
            blen = 0;
        }
    }
    if (len >= 64) {
        compress(h, p, len / 64);
        p += 64 * (len / 64);
        len %= 64;
    }
    if (len) {
        memcpy(buf, p, len);
        blen = (uint32_t)len;
    }
}

void Sha256::digest(uint8_t out[32]) const {
    Sha256 c = *this;
    uint64_t bits = c.total * 8;
    uint8_t pad[72] = {0x80};
    size_t padlen = (c.blen < 56) ? 56 - c.blen : 120 - c.blen;
    c.update(pad, padlen);
    uint8_t lb[8];
    for (int i = 0; i < 8; i++) lb[i] = (uint8_t)(bits >> (56 - 8 * i));
    c.update(lb, 8);
    for (int i = 0; i < 8; i++) {
        out[4 * i] = (uint8_t)(c.h[i] >> 24);
        out[4 * i + 1] = (uint8_t)(c.h[i] >> 16);
        out[4 * i + 2] = (uint8_t)(c.h[i] >> 8);
        out[4 * i + 3] = (uint8_t)c.h[i];
    }
}

// ------------------------------------------------------------- transcript --
Transcript::Transcript(uint32_t seed) {
    uint8_t b[4];
    memcpy(b, &seed, 4);  // to_le_bytes
    hasher.update(b, 4);
}

void Transcript::squeeze(uint64_t* lo, uint64_t* hi) {
    uint8_t b[4];
    memcpy(b, &counter, 4);
    hasher.update(b, 4);
    counter++;
    // finalize a copy: pending bytes, 0x80, zeros, 64-bit big-endian length
    uint8_t blk[128];
    const uint32_t n = hasher.blen;
    memcpy(blk, hasher.buf, n);
    blk[n] = 0x80;
    const size_t len = n < 56 ? 64 : 128;
    memset(blk + n + 1, 0, len - 8 - n - 1);
    const uint64_t bits = hasher.total * 8;
    for (int i = 0; i < 8; i++) blk[len - 8 + i] = (uint8_t)(bits >> (56 - 8 * i));
    uint32_t st[8];
    memcpy(st, hasher.h, 32);
    compress(st, blk, len / 64);
    uint8_t d[16];
    for (int i = 0; i < 4; i++) {
        d[4 * i] = (uint8_t)(st[i] >> 24);
        d[4 * i + 1] = (uint8_t)(st[i] >> 16);
        d[4 * i + 2] = (uint8_t)(st[i] >> 8);
        d[4 * i + 3] = (uint8_t)st[i];
    }
    memcpy(lo, d, 8);
    memcpy(hi, d + 8, 8);
}

size_t Transcript::get_query(size_t n) {
    uint64_t lo, hi;
    squeeze(&lo, &hi);
    return (size_t)(lo % (uint64_t)n);  // (u128 as usize) % n on a 64-bit target
}

namespace {
__attribute__((target("avx2"))) inline __m256i rotr8(__m256i x, int n) {
    return _mm256_or_si256(_mm256_srli_epi32(x, n), _mm256_slli_epi32(x, 32 - n));
}

// Eight independent SHA-256 compressions with AVX2 (lane i: state st[i],
// block blk[i]); the FIPS 180-4 rounds on 8 x 32-bit lanes.
__attribute__((target("avx2"))) void sha_x8(const uint32_t st[8][8], const uint8_t blk[8][64], uint32_t out[8][8]) {
#define rotr rotr8
    __m256i w[16], v[8], s0[8];
    for (int t = 0; t < 16; t++) {
        alignas(32) uint32_t lane[8];
        for (int i = 0; i < 8; i++) {
            const uint8_t* q = blk[i] + 4 * t;
            lane[i] = (uint32_t)q[0] << 24 | (uint32_t)q[1] << 16 | (uint32_t)q[2] << 8 | q[3];
        }
        w[t] = _mm256_load_si256((const __m256i*)lane);
    }
    for (int k = 0; k < 8; k++) {
        alignas(32) uint32_t lane[8];
        for (int i = 0; i < 8; i++) lane[i] = st[i][k];
        v[k] = s0[k] = _mm256_load_si256((const __m256i*)lane);
    }
    for (int t = 0; t < 64; t++) {
        __m256i wt;
        if (t < 16) {
            wt = w[t];
        } else {
            const __m256i w15 = w[(t - 15) & 15], w2 = w[(t - 2) & 15];
            const __m256i sg0 = _mm256_xor_si256(_mm256_xor_si256(rotr(w15, 7), rotr(w15, 18)), _mm256_srli_epi32(w15, 3));
            const __m256i sg1 = _mm256_xor_si256(_mm256_xor_si256(rotr(w2, 17), rotr(w2, 19)), _mm256_srli_epi32(w2, 10));
            wt = _mm256_add_epi32(_mm256_add_epi32(w[t & 15], sg0), _mm256_add_epi32(w[(t - 7) & 15], sg1));
            w[t & 15] = wt;
        }
        const __m256i e = v[4], a = v[0];
        const __m256i S1 = _mm256_xor_si256(_mm256_xor_si256(rotr(e, 6), rotr(e, 11)), rotr(e, 25));
        const __m256i ch = _mm256_xor_si256(_mm256_and_si256(e, v[5]), _mm256_andnot_si256(e, v[6]));
        const __m256i t1 = _mm256_add_epi32(_mm256_add_epi32(_mm256_add_epi32(v[7], S1), _mm256_add_epi32(ch, wt)),
                                            _mm256_set1_epi32((int)kSha[t]));
        const __m256i S0 = _mm256_xor_si256(_mm256_xor_si256(rotr(a, 2), rotr(a, 13)), rotr(a, 22));
        const __m256i mj = _mm256_xor_si256(_mm256_xor_si256(_mm256_and_si256(a, v[1]), _mm256_and_si256(a, v[2])),
                                            _mm256_and_si256(v[1], v[2]));
        const __m256i t2 = _mm256_add_epi32(S0, mj);
        v[7] = v[6];
        v[6] = v[5];
        v[5] = v[4];
        v[4] = _mm256_add_epi32(v[3], t1);
        v[3] = v[2];
        v[2] = v[1];
        v[1] = v[0];
        v[0] = _mm256_add_epi32(t1, t2);
    }
    for (int k = 0; k < 8; k++) {
        alignas(32) uint32_t lane[8];
        _mm256_store_si256((__m256i*)lane, _mm256_add_epi32(v[k], s0[k]));
        for (int i = 0; i < 8; i++) out[i][k] = lane[i];
    }
#undef rotr
}

bool have_avx2() {
    static int cached = -1;
    if (cached < 0) {
        unsigned a = 0, b = 0, c = 0, d = 0;
        cached = (__get_cpuid_count(7, 0, &a, &b, &c, &d) && (b & (1u << 5))) ? 1 : 0;
        if (getenv("LIG_NO_AVX2")) cached = 0;
    }
    return cached == 1;
}
}  // namespace

namespace {
// The first finalization block of a hash whose pending bytes are buf[0, nb)
// (nb a multiple of 4): those bytes, 0x80, zeros to the end of the block
// (the caller adds the length).  Fixed-size masks instead of copies and
// fills of variable length.
struct PadMasks {
    alignas(16) uint8_t keep[16][64];
    alignas(16) uint8_t pad[16][64];
    PadMasks() {
        for (int i = 0; i < 16; i++)
            for (int b = 0; b < 64; b++) {
                keep[i][b] = b < 4 * i ? 0xFF : 0x00;
                pad[i][b] = b == 4 * i ? 0x80 : 0x00;
            }
    }
};
const PadMasks pad_masks;

inline void pad_block(uint8_t* out, const uint8_t* buf, uint32_t nb) {
    const int i = (int)(nb >> 2);
    for (int j = 0; j < 4; j++) {
        __m128i v = _mm_loadu_si128((const __m128i*)(buf + 16 * j));
        v = _mm_and_si128(v, _mm_load_si128((const __m128i*)(pad_masks.keep[i] + 16 * j)));
        v = _mm_or_si128(v, _mm_load_si128((const __m128i*)(pad_masks.pad[i] + 16 * j)));
        _mm_storeu_si128((__m128i*)(out + 16 * j), v);
    }
}
}  // namespace

// get_distinct_queries (emulated_fs.rs:61-75).  Squeezes are generated in
// groups: the running state absorbs each counter in turn and the
// finalizations that fit one block run together (two interleaved with the
// SHA extensions, eight at a time with AVX2 without them), the others one by
// one; afterwards the transcript is rewound to exactly the squeezes the
// rejection sampling consumed.
std::vector<size_t> Transcript::get_distinct_queries(size_t n, size_t s) {
    LIG_REQUIRE(s <= n, LIG_E_INVALID, "get_distinct_queries: s > n never terminates");
    std::vector<size_t> q;
    q.reserve(s);
    // Indices seen so far: an open-addressing set of at least 4s slots
    // (value + 1, 0 = empty), on the stack for the usual s, rather than a
    // bitmap over all n indices.
    size_t cap = 64;
    while (cap < 4 * s) cap <<= 1;
    uint64_t local[1024];
    std::vector<uint64_t> heap;
    uint64_t* slots = local;
    if (cap > 1024) {
        heap.assign(cap, 0);
        slots = heap.data();
    } else {
        memset(local, 0, cap * sizeof(uint64_t));
    }
    const int shift = 64 - __builtin_ctzll(cap);
    // the row counts are powers of two: a mask instead of a 64-bit division
    const bool pow2 = (n & (n - 1)) == 0;
    auto take = [&](uint64_t lo) {
        const size_t v = pow2 ? (size_t)(lo & (uint64_t)(n - 1)) : (size_t)(lo % (uint64_t)n);
        size_t h = (size_t)(((uint64_t)v * 0x9E3779B97F4A7C15ull) >> shift);
        for (;;) {
            if (slots[h] == 0) {
                slots[h] = (uint64_t)v + 1;
                q.push_back(v);
                return;
            }
            if (slots[h] == (uint64_t)v + 1) return;
            h = (h + 1) & (cap - 1);
        }
    };
    if (have_sha_ni() && hasher.blen % 4 == 0) {
        // Squeezes in pairs: each counter goes straight into the running
        // block (a full block is compressed into the running state), each
        // squeeze's finalization is built from a copy of that block, and the
        // pair's first finalization blocks run interleaved (a pending length
        // of 56 or 60 bytes needs a second block, run alone).  If the first
        // squeeze of a pair completes the set, the transcript is rewound to
        // just after it.
        while (q.size() < s) {
            uint32_t st[2][8];
            alignas(16) uint8_t blk[2][2][64];
            bool two[2];
            Sha256 h_mid;
            uint32_t c_mid = 0;
            for (int k = 0; k < 2; k++) {
                if (k == 1) {
                    h_mid = hasher;
                    c_mid = counter;
                }
                // the pending length stays a multiple of 4
                memcpy(hasher.buf + hasher.blen, &counter, 4);
                hasher.blen += 4;
                hasher.total += 4;
                if (hasher.blen == 64) {
                    compress(hasher.h, hasher.buf, 1);
                    hasher.blen = 0;
                }
                counter++;
                const uint32_t nb = hasher.blen;
                memcpy(st[k], hasher.h, 32);
                uint8_t* b0 = blk[k][0];
                pad_block(b0, hasher.buf, nb);
                two[k] = nb >= 56;
                uint8_t* last = b0;
                if (two[k]) {
                    last = blk[k][1];
                    memset(last, 0, 56);
                }
                const uint64_t bits_be = __builtin_bswap64(hasher.total * 8);
                memcpy(last + 56, &bits_be, 8);
            }
            sha_ni_x2(st[0], blk[0][0], st[1], blk[1][0]);
            uint64_t lo[2];
            for (int k = 0; k < 2; k++) {
                if (two[k]) sha_blocks_ni(st[k], blk[k][1], 1);
                lo[k] = (uint64_t)__builtin_bswap32(st[k][0]) | ((uint64_t)__builtin_bswap32(st[k][1]) << 32);
            }
            take(lo[0]);
            if (q.size() == s) {
                hasher = h_mid;
                counter = c_mid;
                break;
            }
            take(lo[1]);
        }
        std::sort(q.begin(), q.end());
        return q;
    }
    if (!have_avx2()) {
        while (q.size() < s) {
            uint64_t lo, hi;
            squeeze(&lo, &hi);
            take(lo);
        }
        std::sort(q.begin(), q.end());
        return q;
    }
    while (q.size() < s) {
        const Sha256 h0 = hasher;
        const uint32_t c0 = counter;
        uint32_t st[8][8];
        alignas(32) uint8_t blk[8][64];
        uint64_t lo[8];
        int single[8], n1 = 0;
        uint32_t out[8][8];
        for (int k = 0; k < 8; k++) {
            uint8_t b[4];
            memcpy(b, &counter, 4);
            hasher.update(b, 4);
            counter++;
            const uint32_t nb = hasher.blen;
            const uint64_t bits = hasher.total * 8;
            if (nb < 56) {
                memcpy(st[n1], hasher.h, 32);
                memcpy(blk[n1], hasher.buf, nb);
                blk[n1][nb] = 0x80;
                memset(blk[n1] + nb + 1, 0, 56 - nb - 1);
                for (int i = 0; i < 8; i++) blk[n1][56 + i] = (uint8_t)(bits >> (56 - 8 * i));
                single[n1++] = k;
            } else {
                // two-block finalization on the scalar path
                uint8_t d[32];
                hasher.digest(d);
                memcpy(&lo[k], d, 8);
            }
        }
        for (int i = n1; i < 8; i++) {
            memcpy(st[i], st[0], 32);
            memcpy(blk[i], blk[0], 64);
        }
        if (n1) sha_x8(st, blk, out);
        for (int i = 0; i < n1; i++) {
            uint8_t d[8];
            for (int w = 0; w < 2; w++) {
                d[4 * w] = (uint8_t)(out[i][w] >> 24);
                d[4 * w + 1] = (uint8_t)(out[i][w] >> 16);
                d[4 * w + 2] = (uint8_t)(out[i][w] >> 8);
                d[4 * w + 3] = (uint8_t)out[i][w];
            }
            memcpy(&lo[single[i]], d, 8);
        }
        int used = 0;
        for (int k = 0; k < 8 && q.size() < s; k++) {
            used = k + 1;
            take(lo[k]);
        }
        if (used < 8) {
            // rewind to the state after exactly `used` squeezes
            hasher = h0;
            counter = c0;
            for (int k = 0; k < used; k++) {
                uint8_t b[4];
                memcpy(b, &counter, 4);
                hasher.update(b, 4);
                counter++;
            }
        }
    }
    std::sort(q.begin(), q.end());
    return q;
}

// ------------------------------------------------------------ BLAKE3 bytes --
void blake3_bytes(const uint8_t* data, size_t len, uint8_t out[32]) {
    b3::Hasher h;
    h.init();
    uint32_t m[16];
    size_t nblocks = len ? (len + 63) / 64 : 1;
    for (size_t b = 0; b + 1 < nblocks; b++) {
        memcpy(m, data + 64 * b, 64);
        h.block(m);
    }
    size_t last = len - 64 * (nblocks - 1);
    memset(m, 0, 64);
    if (last) memcpy(m, data + 64 * (nblocks - 1), last);
    uint32_t o[8];
    h.finish(m, (uint32_t)last, o);
    memcpy(out, o, 32);
}

// ------------------------------------------------------------- Merkle host --
std::vector<uint64_t> merkle_proof_nodes(size_t n_leaves, const std::vector<size_t>& sorted_queries) {
    // merkle_tree.rs:62-111; depth = log2(n_leaves) layers below the root.
    std::vector<uint64_t> out;
    std::vector<size_t> q = sorted_queries;
    size_t cnt = q.size();
    {
        size_t depth = 0;
        while (((size_t)1 << depth) < n_leaves) depth++;
        out.reserve(cnt * depth);
    }
    size_t layer_off = 0, layer_n = n_leaves;
    while (layer_n > 1) {
        size_t next = 0, i = 0;
        while (i < cnt) {
            size_t query = q[i], sib = query ^ 1;
            next++;
            q[next - 1] = query >> 1;
            if (i == cnt - 1) {
                out.push_back(layer_off + sib);
                break;
            }
            if (query % 2 != 0) {
                out.push_back(layer_off + sib);
                i += 1;
            } else if (q[i + 1] != sib) {
                out.push_back(layer_off + sib);
                i += 1;
            } else {
                i += 2;
            }
        }
        cnt = next;
        layer_off += layer_n;
        layer_n >>= 1;
    }
    return out;
}

static void hash_pair(const uint8_t* l, const uint8_t* r, uint8_t out[32]) {
    uint32_t m[16], o[8];
    memcpy(m, l, 32);
    memcpy(m + 8, r, 32);
    b3::hash64(m, o);
    memcpy(out, o, 32);
}

bool merkle_verify_hashes(const uint8_t root[32], const uint8_t* proof, size_t n_proof, size_t depth,
                          std::vector<std::array<uint8_t, 32>> layer, std::vector<size_t> q) {
    size_t cnt = q.size(), pc = 0;
    if (cnt == 0) return false;
    for (size_t d = 0; d < depth; d++) {
        size_t next = 0, i = 0;
        while (i < cnt) {
            size_t query = q[i], sib = query ^ 1;
            next++;
            q[next - 1] = query >> 1;
            uint8_t h[32];
            if (i == cnt - 1) {
                if (++pc > n_proof) return false;
                if (query % 2 != 0) hash_pair(proof + 32 * (pc - 1), layer[i].data(), h);
                else hash_pair(layer[i].data(), proof + 32 * (pc - 1), h);
                memcpy(layer[next - 1].data(), h, 32);
                break;
            }
            if (query % 2 != 0) {
                if (++pc > n_proof) return false;
                hash_pair(proof + 32 * (pc - 1), layer[i].data(), h);
                i += 1;
            } else if (q[i + 1] != sib) {
                if (++pc > n_proof) return false;
                hash_pair(layer[i].data(), proof + 32 * (pc - 1), h);
                i += 1;
            } else {
                hash_pair(layer[i].data(), layer[i + 1].data(), h);
                i += 2;
            }
            memcpy(layer[next - 1].data(), h, 32);
        }
        cnt = next;
    }
    return cnt == 1 && memcmp(layer[0].data(), root, 32) == 0;
}

}  // namespace lig
