// Host-side building blocks of the Ligerito library: error reporting,
// host field arithmetic (PCLMULQDQ), SHA-256 Fiat–Shamir transcript, BLAKE3
// over byte strings and the Merkle query logic.  Everything here runs on the
// CPU; the hot path runs in kernels.hip.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>
#include <smmintrin.h>
#include <wmmintrin.h>

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ligerito.h"
#include "blake3.hpp"
#include "gf.hpp"

namespace lig {

// ----------------------------------------------------------------- errors --
void set_error(const std::string& msg);
const char* last_error();

// Exceptions carry a status code; the C ABI catches them (nothing unwinds
// across the ABI).
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define LIG_HIP(expr)                                                                                   \
    do {                                                                                                \
        hipError_t _e = (expr);                                                                         \
        if (_e != hipSuccess)                                                                           \
            throw ::lig::Error(_e == hipErrorOutOfMemory ? LIG_E_OOM : LIG_E_HIP,                        \
                               std::string(#expr " failed: ") + hipGetErrorString(_e));                  \
    } while (0)

#define LIG_REQUIRE(cond, code, msg)                          \
    do {                                                      \
        if (!(cond)) throw ::lig::Error((code), (msg));       \
    } while (0)

// RAII device allocation.
template <class T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count) LIG_HIP(hipMalloc((void**)&p, count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// ------------------------------------------------------ host field traits --
inline uint64_t clmul64_lo(uint64_t a, uint64_t b, uint64_t* hi) {
    __m128i x = _mm_set_epi64x(0, (long long)a), y = _mm_set_epi64x(0, (long long)b);
    __m128i r = _mm_clmulepi64_si128(x, y, 0x00);
    *hi = (uint64_t)_mm_extract_epi64(r, 1);
    return (uint64_t)_mm_cvtsi128_si64(r);
}

struct H32 {
    using T = uint32_t;
    static constexpr int BITS = 32, BYTES = 4;
    static T zero() { return 0; }
    static T one() { return 1; }
    static T from_int(uint64_t v) { return (T)v; }
    static T mul(T a, T b) {
        uint64_t hi;
        uint64_t p = clmul64_lo(a, b, &hi);
        return gf32_reduce64(p);
    }
    // get_field: keeps the low 4 bytes of the squeezed u128 (emulated_fs.rs:34-54)
    static T from_squeeze(uint64_t lo, uint64_t) { return (T)lo; }
    static void to_bytes(T v, uint8_t* o) { memcpy(o, &v, 4); }
    static T from_bytes(const uint8_t* i) { T v; memcpy(&v, i, 4); return v; }
    static bool eq(T a, T b) { return a == b; }
    static bool is_zero(T a) { return a == 0; }
};

struct H128 {
    using T = u128v;
    static constexpr int BITS = 128, BYTES = 16;
    static T zero() { return {0, 0}; }
    static T one() { return {1, 0}; }
    static T from_int(uint64_t v) { return {v, 0}; }
    static T mul(T a, T b) {
        uint64_t l1, h1, m1;
        uint64_t l0 = clmul64_lo(a.lo, b.lo, &l1);
        uint64_t h0 = clmul64_lo(a.hi, b.hi, &h1);
        uint64_t m0 = clmul64_lo(a.lo ^ a.hi, b.lo ^ b.hi, &m1);
        m0 ^= l0 ^ h0;
        m1 ^= l1 ^ h1;
        return gf128_reduce256(l0, l1 ^ m0, h0 ^ m1, h1);
    }
    static T from_squeeze(uint64_t lo, uint64_t hi) { return {lo, hi}; }
    static void to_bytes(T v, uint8_t* o) { memcpy(o, &v.lo, 8); memcpy(o + 8, &v.hi, 8); }
    static T from_bytes(const uint8_t* i) { T v; memcpy(&v.lo, i, 8); memcpy(&v.hi, i + 8, 8); return v; }
    static bool eq(T a, T b) { return a == b; }
    static bool is_zero(T a) { return a.lo == 0 && a.hi == 0; }
};

template <class H> inline typename H::T hxor(typename H::T a, typename H::T b) { return a ^ b; }

// Fermat inverse a^(2^n - 2) (binary_field.rs:40-50): square-and-multiply.
template <class H> typename H::T h_inverse(typename H::T a) {
    LIG_REQUIRE(!H::is_zero(a), LIG_E_ASSERT, "Cannot compute inverse of zero");
    // 2^n - 2 = 0b11..110: result = prod_{i=1}^{n-1} a^(2^i)
    typename H::T r = H::one(), sq = a;
    for (int i = 1; i < H::BITS; i++) {
        sq = H::mul(sq, sq);
        r = H::mul(r, sq);
    }
    return r;
}

// ------------------------------------------------------------------ SHA-256 --
struct Sha256 {
    uint32_t h[8];
    uint8_t buf[64];
    uint64_t total = 0;
    uint32_t blen = 0;
    Sha256();
    void update(const void* data, size_t len);
    void digest(uint8_t out[32]) const;  // finalizes a copy
};

// Fiat–Shamir transcript (emulated_fs.rs:6-75).
struct Transcript {
    Sha256 hasher;
    uint32_t counter = 0;
    explicit Transcript(uint32_t seed);
    void absorb(const void* b, size_t n) { hasher.update(b, n); }
    void squeeze(uint64_t* lo, uint64_t* hi);
    template <class H> typename H::T get_field() {
        uint64_t lo, hi;
        squeeze(&lo, &hi);
        return H::from_squeeze(lo, hi);
    }
    size_t get_query(size_t n);
    std::vector<size_t> get_distinct_queries(size_t n, size_t s);
    template <class H> void absorb_elems(const typename H::T* e, size_t count) {
        std::vector<uint8_t> b(count * H::BYTES);
        for (size_t i = 0; i < count; i++) H::to_bytes(e[i], &b[i * H::BYTES]);
        absorb(b.data(), b.size());
    }
};

// ------------------------------------------------------------ BLAKE3 bytes --
void blake3_bytes(const uint8_t* data, size_t len, uint8_t out[32]);

// ------------------------------------------------------------- Merkle host --
// Node index (in the contiguous all-layers layout) of every sibling the
// batched proof of merkle_tree.rs:62-111 emits, in emission order.
std::vector<uint64_t> merkle_proof_nodes(size_t n_leaves, const std::vector<size_t>& sorted_queries);
// merkle_tree.rs:164-188 on host.  leaf_hashes: one 32-byte hash per query.
bool merkle_verify_hashes(const uint8_t root[32], const uint8_t* proof, size_t n_proof, size_t depth,
                          std::vector<std::array<uint8_t, 32>> layer, std::vector<size_t> queries);

}  // namespace lig
