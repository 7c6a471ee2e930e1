// HIP kernels for the Ligerito prover hot path on CDNA4 (gfx950).
//
// Data layout in HBM: an encoded matrix is row-major [4m][n] (a Merkle leaf
// is one contiguous row), a Merkle tree is every layer back to back
// (layer 0 = leaf hashes, 32 bytes per node), multilinear tables are plain
// arrays.  All arithmetic is exact binary-field arithmetic.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "blake3.hpp"
#include "fields.hpp"
#include "kernels.hpp"
#include "transcript_dev.hpp"

namespace lig {

template <> struct PassShape<F32> {
    static constexpr int K = 7;   // 128 rows per tile
    static constexpr int C = 64;  // 64 columns (one wave covers one row)
};
template <> struct PassShape<F128> {
    static constexpr int K = 4;   // 16 rows per tile (8 KiB tables)
    static constexpr int C = 64;
};

static inline unsigned grid_for(size_t work, unsigned block, unsigned cap = 65535u * 16u) {
    size_t g = (work + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// Workgroup barrier over LDS traffic only: waits for this wave's LDS
// operations (lgkmcnt), not for its global stores, which may stay in flight
// (__syncthreads would wait for them too).  The empty asm statements keep the
// compiler from moving memory accesses across the barrier.
#define LIG_LDS_BARRIER()                    \
    do {                                     \
        asm volatile("" ::: "memory");       \
        __builtin_amdgcn_s_waitcnt(0xC07F);  \
        __builtin_amdgcn_s_barrier();        \
        asm volatile("" ::: "memory");       \
    } while (0)

// Writes the 16 entries of nibble position k (k = 0..7) of a GF(2^32)
// table at t (128 words, entry v at word 16k + v) for a GF(2)-linear map L
// given the images c_b = L(x^(4k+b)), b = 0..3: entry v = XOR of c_b over
// the bits b of v.
__device__ __forceinline__ void gf32_tab_write(uint32_t* t, int k, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3) {
    const uint32_t c01 = c0 ^ c1, c23 = c2 ^ c3;
    uint4* d4 = reinterpret_cast<uint4*>(t + k * 16);
    d4[0] = make_uint4(0u, c0, c1, c01);
    d4[1] = make_uint4(c2, c2 ^ c0, c2 ^ c1, c2 ^ c01);
    d4[2] = make_uint4(c3, c3 ^ c0, c3 ^ c1, c3 ^ c01);
    d4[3] = make_uint4(c23, c23 ^ c0, c23 ^ c1, c23 ^ c01);
}
// The same entries, the position's four 16-byte chunks (entries 4c..4c+3 =
// X_c ^ {0, c0, c1, c0^c1}, X_c the c-combination of c2, c3) written in an
// order rotated by lane.  Where consecutive lanes build consecutive
// positions 64 bytes apart, lanes 4 apart otherwise write the same LDS
// banks at every step; rotated, 16 consecutive lanes write 16 distinct bank
// groups.  A few more registers than gf32_tab_write.
__device__ __forceinline__ void gf32_tab_write_rot(uint32_t* t, int k, uint32_t c0, uint32_t c1, uint32_t c2,
                                                   uint32_t c3) {
    const uint32_t c01 = c0 ^ c1;
    uint4* d4 = reinterpret_cast<uint4*>(t + k * 16);
    const uint32_t rot = (threadIdx.x >> 2) & 3u;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const uint32_t c = ((uint32_t)s + rot) & 3u;
        const uint32_t x = ((0u - (c & 1u)) & c2) ^ ((0u - (c >> 1)) & c3);
        d4[c] = make_uint4(x, x ^ c0, x ^ c1, x ^ c01);
    }
}
__device__ __forceinline__ uint32_t gf32_dbl(uint32_t c) { return (c << 1) ^ ((0u - (c >> 31)) & GF32_POLY); }
// nibble position k of the multiply-by-lam table: c_b = lam * x^(4k+b)
// (ROT: written with gf32_tab_write_rot)
template <bool ROT = false>
__device__ __forceinline__ void gf32_tab_build(uint32_t* t, uint32_t lam, int k) {
    const uint32_t c0 = gf32_reduce64((uint64_t)lam << (4 * k));
    const uint32_t c1 = gf32_dbl(c0), c2 = gf32_dbl(c1), c3 = gf32_dbl(c2);
    if constexpr (ROT) gf32_tab_write_rot(t, k, c0, c1, c2, c3);
    else gf32_tab_write(t, k, c0, c1, c2, c3);
}
// nibble position k of the squaring table: c_b = x^(8k+2b)
__device__ __forceinline__ void gf32_tab_build_sqr(uint32_t* t, int k) {
    const uint32_t c0 = gf32_reduce64(1ull << (8 * k)), c1 = gf32_reduce64(1ull << (8 * k + 2));
    const uint32_t c2 = gf32_reduce64(1ull << (8 * k + 4)), c3 = gf32_reduce64(1ull << (8 * k + 6));
    gf32_tab_write(t, k, c0, c1, c2, c3);
}

// ------------------------------------------------------------------ tables --

template <class F>
__global__ void k_build_tables(const typename F::T* __restrict__ lams, size_t count, typename F::T* __restrict__ tabs) {
    size_t total = count * (size_t)F::TAB;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        size_t i = idx / F::TAB;
        int e = (int)(idx % F::TAB);
        tabs[idx] = Ops<F>::table_entry(lams[i], e);
    }
}

template <class F> void launch_build_tables(const typename F::T* lams, size_t count, typename F::T* tabs, hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL(k_build_tables<F>, dim3(grid_for(count * F::TAB, 256)), dim3(256), 0, s, lams, count, tabs);
}

// --------------------------------------------------------- transpose+scale --
// Tile of 64 message rows x 64 columns: coalesced column-major reads, LDS
// transpose, pis scaling with two table multiplies, coalesced row writes.
template <class F>
__global__ __launch_bounds__(256) void k_transpose_scale(const typename F::T* __restrict__ poly, size_t m, size_t n,
                                                         const typename F::T* __restrict__ lo_tabs,
                                                         const typename F::T* __restrict__ hi_tabs,
                                                         typename F::T* __restrict__ msg) {
    using T = typename F::T;
    __shared__ T tile[64][65];
    const int rt = m < 64 ? (int)m : 64;
    const int ct = n < 64 ? (int)n : 64;
    const size_t row_tiles = (m + 63) / 64;
    const size_t i0 = (blockIdx.x % row_tiles) * 64;
    const size_t c0 = (blockIdx.x / row_tiles) * 64;
    for (int idx = threadIdx.x; idx < rt * ct; idx += blockDim.x) {
        int r = idx % rt, c = idx / rt;  // lanes walk down a column
        tile[c][r] = poly[(c0 + c) * m + i0 + r];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < rt * ct; idx += blockDim.x) {
        int c = idx % ct, r = idx / ct;  // lanes walk along a row
        size_t i = i0 + r;
        T x = tile[c][r];
        x = Ops<F>::mul_tab(lo_tabs + (size_t)(i & 63) * F::TAB, x);
        x = Ops<F>::mul_tab(hi_tabs + (size_t)(i >> 6) * F::TAB, x);
        msg[i * n + c0 + c] = x;
    }
}

// Largest log2 row length the GF(2^32) induce kernels hold tables for.
constexpr int PREP_MAXELL = 26;

// Copies n words (n <= PER * blockDim.x) from global memory into LDS with
// every load of a thread issued before its stores.
template <int PER>
__device__ __forceinline__ void stage_words(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, int n) {
    uint32_t v[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = threadIdx.x + k * blockDim.x;
        v[k] = i < n ? src[i] : 0u;
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = threadIdx.x + k * blockDim.x;
        if (i < n) dst[i] = v[k];
    }
}

__device__ __forceinline__ uint32_t mul_tab_lds(const uint32_t* tabs, uint32_t tb, uint32_t w);

// GF(2^32) variant: tiles of 2^LTR rows x 2^LCT columns (2^LCT = min(n, 64);
// m >= 2^LTR).  A tile's rows i0..i0+2^LTR-1 are scaled by pis[i0 + r]: the
// workgroup builds one multiply table per row in LDS, so each element costs
// one table multiply.  The launcher gives every tile its own workgroup; a
// smaller grid walks the tiles, issuing the next tile's loads before this
// tile is scaled and stored.  The barriers wait for LDS traffic only.
// With polys, blockIdx.y picks one of several inputs, polys[blockIdx.y],
// and the blockIdx.y-th m x n block of msg.
template <int LCT, int LTR>
__global__ __launch_bounds__(256) void k_transpose_scale32(const uint32_t* __restrict__ poly, size_t m, size_t n,
                                                          const uint32_t* __restrict__ pis, uint32_t* __restrict__ msg,
                                                          const uint32_t* const* __restrict__ polys) {
    if (polys) {
        poly = polys[blockIdx.y];
        msg += (size_t)blockIdx.y * m * n;
    }
    // a pointer read from memory is generic to the compiler: name its
    // address space so the tile loads are global loads, not flat ones
    const __attribute__((address_space(1))) uint32_t* gpoly = (const __attribute__((address_space(1))) uint32_t*)poly;
    constexpr int CT = 1 << LCT, TR = 1 << LTR, KPT = (TR * CT) / 256;  // elements per thread
    constexpr int NH = (TR * 8 + 255) / 256;                            // table positions per thread
    static_assert(KPT >= 1, "at least 256 elements per tile");
    __shared__ __attribute__((aligned(256))) uint32_t rtab[TR * 128];
    __shared__ uint32_t tile[CT][TR + 1];
    const size_t row_tiles = m / TR, ntiles = row_tiles * (n / CT);
    const int tid = threadIdx.x;
    uint32_t tv[KPT], lam[NH];
    auto load = [&](size_t ti) {
        const size_t i0 = (ti % row_tiles) * TR, c0 = (ti / row_tiles) * CT;
#pragma unroll
        for (int k = 0; k < KPT; k++) {
            const int idx = tid + 256 * k;
            tv[k] = gpoly[(c0 + (idx >> LTR)) * m + i0 + (idx & (TR - 1))];
        }
#pragma unroll
        for (int h = 0; h < NH; h++) lam[h] = pis[i0 + ((tid + 256 * h) >> 3)];
    };
    if (blockIdx.x < ntiles) load(blockIdx.x);
    for (size_t tile_i = blockIdx.x; tile_i < ntiles; tile_i += gridDim.x) {
        const size_t i0 = (tile_i % row_tiles) * TR, c0 = (tile_i / row_tiles) * CT;
#pragma unroll
        for (int k = 0; k < KPT; k++) {
            const int idx = tid + 256 * k;
            tile[idx >> LTR][idx & (TR - 1)] = tv[k];
        }
#pragma unroll
        for (int h = 0; h < NH; h++) {
            const int item = tid + 256 * h;
            if (item < TR * 8) gf32_tab_build<true>(rtab + (item >> 3) * 128, lam[h], item & 7);
        }
        LIG_LDS_BARRIER();
        if (tile_i + gridDim.x < ntiles) load(tile_i + gridDim.x);
#pragma unroll
        for (int k = 0; k < KPT; k++) {
            const int idx = tid + 256 * k;
            const int c = idx & (CT - 1), r = idx >> LCT;
            msg[(i0 + r) * n + c0 + c] = mul_tab_lds(rtab, (uint32_t)(r * 512), tile[c][r]);
        }
        LIG_LDS_BARRIER();  // tile and tables are read before the next tile overwrites them
    }
}

constexpr int TS128_ROWS = 8;      // message rows per k_transpose_scale128 tile
constexpr int TS128_TAB_ROWS = 4;  // rows whose tables share LDS at a time (32 KB)
// log2 rows per k_transpose_scale32 tile: 32-row tiles (24.8 KB of LDS) let
// five workgroups share a CU against three of 64 rows (49.4 KB), which hides
// more of the column reads' latency (2^24 encode: 429.7 -> 426.6 us)
constexpr int TS32_LTR = 5;
__global__ __launch_bounds__(256) void k_transpose_scale128(const u128v* __restrict__ poly, size_t m, size_t n,
                                                            const u128v* __restrict__ pis, u128v* __restrict__ msg);

template <class F>
void launch_transpose_scale(const typename F::T* poly, size_t m, size_t n, const typename F::T* pis,
                            const typename F::T* lo_tabs, const typename F::T* hi_tabs, typename F::T* msg, hipStream_t s,
                            const typename F::T* const* polys, int inputs) {
    size_t tiles = ((m + 63) / 64) * ((n + 63) / 64);
    const bool fast32 = F::BITS == 32 && m >= 64 && n >= 4 && (m & (m - 1)) == 0 && (n & (n - 1)) == 0;
    if (inputs != 1 && (!fast32 || !polys || inputs < 1 || inputs > 65535))
        throw std::invalid_argument("several inputs per launch: GF(2^32), m >= 64, n >= 4, a pointer table");
    if constexpr (F::BITS == 32) {
        if (fast32) {
            int lct = 0;
            while ((1 << lct) < (n < 64 ? (int)n : 64)) lct++;
            // one workgroup per tile of 2^LTR rows (64 rows for four columns:
            // a 256-thread tile needs 256 elements)
            constexpr int LTR = TS32_LTR;
            const size_t t32 = (m >> (lct == 2 ? 6 : LTR)) * (n >> lct);
            if (t32 > 0x7fffffffu) throw std::invalid_argument("transpose: too many tiles");
            const unsigned g = (unsigned)t32;
            switch (lct) {
                case 2: hipLaunchKernelGGL((k_transpose_scale32<2, 6>), dim3(g, (unsigned)inputs), dim3(256), 0, s, poly, m, n, pis, msg,
                                           reinterpret_cast<const uint32_t* const*>(polys)); return;
                case 3: hipLaunchKernelGGL((k_transpose_scale32<3, LTR>), dim3(g, (unsigned)inputs), dim3(256), 0, s, poly, m, n, pis, msg,
                                           reinterpret_cast<const uint32_t* const*>(polys)); return;
                case 4: hipLaunchKernelGGL((k_transpose_scale32<4, LTR>), dim3(g, (unsigned)inputs), dim3(256), 0, s, poly, m, n, pis, msg,
                                           reinterpret_cast<const uint32_t* const*>(polys)); return;
                case 5: hipLaunchKernelGGL((k_transpose_scale32<5, LTR>), dim3(g, (unsigned)inputs), dim3(256), 0, s, poly, m, n, pis, msg,
                                           reinterpret_cast<const uint32_t* const*>(polys)); return;
                default: hipLaunchKernelGGL((k_transpose_scale32<6, LTR>), dim3(g, (unsigned)inputs), dim3(256), 0, s, poly, m, n, pis, msg,
                                           reinterpret_cast<const uint32_t* const*>(polys)); return;
            }
        }
    }
    if constexpr (F::BITS == 128) {
        if (n >= 64 && m >= TS128_ROWS && (m & (m - 1)) == 0 && (n & (n - 1)) == 0) {
            const unsigned cw = n < 256 ? (unsigned)n : 256u;
            const size_t g = (m / TS128_ROWS) * (n / cw);
            if (g > 0x7fffffffu) throw std::invalid_argument("transpose: too many tiles");
            hipLaunchKernelGGL(k_transpose_scale128, dim3((unsigned)g), dim3(cw), 0, s, poly, m, n, pis, msg);
            return;
        }
    }
    hipLaunchKernelGGL(k_transpose_scale<F>, dim3((unsigned)tiles), dim3(256), 0, s, poly, m, n, lo_tabs, hi_tabs, msg);
}

template <class F>
__global__ void k_scale_rows(typename F::T* __restrict__ buf, size_t rows, size_t n, const typename F::T* __restrict__ pis) {
    size_t total = rows * n;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x)
        buf[idx] = Ops<F>::mul(buf[idx], pis[idx / n]);
}

template <class F> void launch_scale_rows(typename F::T* buf, size_t rows, size_t n, const typename F::T* pis, hipStream_t s) {
    if (!rows) return;
    hipLaunchKernelGGL(k_scale_rows<F>, dim3(grid_for(rows * n, 256)), dim3(256), 0, s, buf, rows, n, pis);
}

// dst[c*rows + r] = src[r*cols + c]: 32 x 32 tiles through LDS (one padding
// column so a tile column is read across banks), 256 threads per tile, the
// grid striding over tiles.
template <class F>
__global__ __launch_bounds__(256) void k_transpose(const typename F::T* __restrict__ src, size_t rows, size_t cols,
                                                   typename F::T* __restrict__ dst) {
    using T = typename F::T;
    __shared__ T t[32][33];
    const size_t ty = (rows + 31) / 32, tx = (cols + 31) / 32;
    const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;
    for (size_t tile = blockIdx.x; tile < tx * ty; tile += gridDim.x) {
        const size_t r0 = (tile / tx) * 32, c0 = (tile % tx) * 32;
        for (int y = ly; y < 32; y += 8)
            if (r0 + y < rows && c0 + lx < cols) t[y][lx] = src[(r0 + y) * cols + c0 + lx];
        __syncthreads();
        for (int y = ly; y < 32; y += 8)
            if (c0 + y < cols && r0 + lx < rows) dst[(c0 + y) * rows + r0 + lx] = t[lx][y];
        __syncthreads();
    }
}

template <class F> void launch_transpose(const typename F::T* src, size_t rows, size_t cols, typename F::T* dst, hipStream_t s) {
    if (!rows || !cols) return;
    const size_t tiles = ((rows + 31) / 32) * ((cols + 31) / 32);
    hipLaunchKernelGGL(k_transpose<F>, dim3(grid_for(tiles, 1, 1u << 20)), dim3(256), 0, s, src, rows, cols, dst);
}

// The scratch of the entry points that allocate nothing: one static array
// per device (each GPU the library runs on holds its own copy until the
// process ends).  Host code never passes `noalloc_scratch` itself to a
// kernel: the name is a host-side stand-in, and the device address comes
// from hipGetSymbolAddress with the device current.
__device__ __align__(256) uint8_t noalloc_scratch[NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex noalloc_mu;
void* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

void* noalloc_scratch_ptr(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) throw std::invalid_argument("device index beyond the scratch table");
    std::lock_guard<std::mutex> g(noalloc_mu);
    if (!noalloc_ptr[device]) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess || cur != device)
            throw std::runtime_error("noalloc scratch: the device must be current");
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("hipGetSymbolAddress(noalloc_scratch) failed: ") + hipGetErrorString(e));
        noalloc_ptr[device] = p;
    }
    return noalloc_ptr[device];
}

// binary_fft.rs:180-206, one stage over all blocks; twiddle index of block b
// at stage `stage` is 2^stage - 1 + b (heap order, binary_fft.rs:127-139).
template <class F>
__global__ void k_fft_stage(typename F::T* __restrict__ buf, size_t N, size_t n, int stage, const typename F::T* __restrict__ tw,
                            int inverse) {
    using T = typename F::T;
    size_t half = N >> (stage + 1);
    size_t pairs = (N / 2) * n;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < pairs; idx += (size_t)gridDim.x * blockDim.x) {
        size_t c = idx % n, p = idx / n;
        size_t blk = p / half, off = p % half;
        size_t ru = blk * 2 * half + off, rw = ru + half;
        T lam = tw[((size_t)1 << stage) - 1 + blk];
        T u = buf[ru * n + c], w = buf[rw * n + c];
        if (!inverse) {
            u ^= Ops<F>::mul(lam, w);
            w ^= u;
        } else {
            w ^= u;
            u ^= Ops<F>::mul(lam, w);
        }
        buf[ru * n + c] = u;
        buf[rw * n + c] = w;
    }
}

template <class F>
void launch_fft_stage(typename F::T* buf, size_t N, size_t n, int stage, const typename F::T* tw, int inverse, hipStream_t s) {
    hipLaunchKernelGGL(k_fft_stage<F>, dim3(grid_for((N / 2) * n, 256)), dim3(256), 0, s, buf, N, n, stage, tw, inverse);
}

// ------------------------------------------------------------- FFT passes --
// One pass = k consecutive stages d0..d0+k-1 of a quarter's size-m FFT
// (global stages d0+2..).  A tile holds the 2^k rows
//   r(t) = B*2^(ell-d0) + t*S + lo,  S = 2^(ell-d0-k),
// that the butterflies of those stages connect, for up to C columns.  At
// tile stage j the block of row t is B*2^j + (t >> (k-j)); its twiddle is
// twiddle index 2^s - 1 + q*2^d + B*2^j + t'.  Tables for the 2^j blocks of
// a stage are staged into LDS; a wave covering one row pair reads a single
// table (conflict-free 16-entry groups).
template <class F, int K, int C>
__global__ __launch_bounds__(256) void k_fft_pass(const typename F::T* __restrict__ src, typename F::T* __restrict__ dst,
                                                  size_t m, size_t n, int ell, int d0, int k, int src_shared,
                                                  const typename F::T* __restrict__ tw_tabs) {
    using T = typename F::T;
    __shared__ T tile[(1 << K) * C];
    __shared__ T tabs[(1 << (K - 1)) * F::TAB_STRIDE];
    const size_t q = blockIdx.y;
    const int cc = n < (size_t)C ? (int)n : C;
    const size_t nchunks = (n + C - 1) / C;
    const size_t tile_id = blockIdx.x / nchunks;
    const size_t c0 = (blockIdx.x % nchunks) * (size_t)C;
    const int sh = ell - d0 - k;
    const size_t S = (size_t)1 << sh;
    const size_t B = tile_id >> sh, lo = tile_id & (S - 1);
    const size_t rbase = (B << (ell - d0)) + lo;
    const int R = 1 << k;

    for (int idx = threadIdx.x; idx < R * cc; idx += blockDim.x) {
        int t = idx / cc, c = idx % cc;
        size_t r = rbase + (size_t)t * S;
        size_t srow = src_shared ? r : q * m + r;
        tile[t * C + c] = src[srow * n + c0 + c];
    }
    for (int j = 0; j < k; j++) {
        const int d = d0 + j;
        const int s = d + 2;
        const int nb = 1 << j;
        const size_t tw_base = (((size_t)1 << s) - 1) + (q << d) + (B << j);
        for (int idx = threadIdx.x; idx < nb * F::TAB; idx += blockDim.x) {
            int b = idx / F::TAB, e = idx % F::TAB;
            tabs[b * F::TAB_STRIDE + e] = tw_tabs[(tw_base + b) * F::TAB + e];
        }
        __syncthreads();
        const int hs = k - 1 - j;
        const int h = 1 << hs;
        for (int idx = threadIdx.x; idx < (R / 2) * cc; idx += blockDim.x) {
            int p = idx / cc, c = idx % cc;
            int blk = p >> hs;
            int tu = (blk << (hs + 1)) + (p & (h - 1));
            int tv = tu + h;
            T u = tile[tu * C + c], w = tile[tv * C + c];
            u ^= Ops<F>::mul_tab(&tabs[blk * F::TAB_STRIDE], w);
            w ^= u;
            tile[tu * C + c] = u;
            tile[tv * C + c] = w;
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < R * cc; idx += blockDim.x) {
        int t = idx / cc, c = idx % cc;
        size_t r = rbase + (size_t)t * S;
        dst[(q * m + r) * n + c0 + c] = tile[t * C + c];
    }
}

static void fft_quarters32(const uint32_t* msg, uint32_t* mat, size_t m, size_t n, const uint32_t* tw, hipStream_t s,
                           PassObserver* obs, PassTune tune, int codewords);
static void fft_quarters128(const u128v* msg, u128v* mat, size_t m, size_t n, const u128v* tw, hipStream_t s,
                            PassObserver* obs, PassTune tune);

template <class F>
void launch_fft_quarters(const typename F::T* msg, typename F::T* mat, size_t m, size_t n, const typename F::T* tw,
                         const typename F::T* tw_tabs, hipStream_t s, PassObserver* obs, PassTune tune, int codewords) {
    constexpr int K = PassShape<F>::K, C = PassShape<F>::C;
    if (n == 0 || (n & (n - 1)) != 0) throw std::invalid_argument("column count must be a power of two");
    int ell = 0;
    while (((size_t)1 << ell) < m) ell++;
    if (codewords != 1 && (F::BITS != 32 || ell == 0 || codewords < 1 || codewords > 16383))
        throw std::invalid_argument("several codewords per launch: GF(2^32), m >= 2, at most 16383");
    if constexpr (F::BITS == 32) {
        if (ell > 0) {
            fft_quarters32(msg, mat, m, n, tw, s, obs, tune, codewords);
            return;
        }
    } else {
        if (ell > 0) {
            fft_quarters128(msg, mat, m, n, tw, s, obs, tune);
            return;
        }
    }
    int npass = ell == 0 ? 1 : (ell + K - 1) / K;
    int d0 = 0;
    const size_t nchunks = (n + C - 1) / C;
    for (int p = 0; p < npass; p++) {
        int k = ell / npass + (p < ell % npass ? 1 : 0);
        size_t tiles = (size_t)1 << (ell - k);
        dim3 grid((unsigned)(tiles * nchunks), 4);
        const typename F::T* src = (p == 0 && msg) ? msg : mat;
        int shared = (p == 0 && msg) ? 1 : 0;
        hipLaunchKernelGGL((k_fft_pass<F, K, C>), grid, dim3(256), 0, s, src, mat, m, n, ell, d0, k, shared, tw_tabs);
        d0 += k;
    }
}

// ------------------------------------------------ GF(2^128) register pass --
constexpr int FFT128_MAXJ = 16;  // jobs (64-thread column chunks) per workgroup
// Multiply by the table at t (32 nibble positions x 16 entries, position k
// of w at entries 16k..16k+15; the lo word holds positions 0..15).
// Each step folds its two entries into the sums with one three-input XOR
// per word (v_bitop3): 64 XOR instructions per product, not 128.  Entry
// byte offsets: per 32-bit word of w, (w << 4) & 0xF0F0F0F0 holds 16 times
// the even nibbles in its bytes and w & 0xF0F0F0F0 the odd ones, so each
// lookup's lane offset is one byte extract; the table and nibble-position
// offsets are constants the compiler folds into the reads' immediates.
__device__ __forceinline__ uint32_t byte_of(uint32_t q, int j) { return (q >> (8 * j)) & 0xFFu; }
__device__ __forceinline__ u128v mul_tab128_lds(const uint4* __restrict__ t, u128v w) {
    const char* tb = reinterpret_cast<const char*>(t);
    const uint32_t wd[4] = {(uint32_t)w.lo, (uint32_t)(w.lo >> 32), (uint32_t)w.hi, (uint32_t)(w.hi >> 32)};
    uint4 acc = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        // nibble positions 8i .. 8i+7 (entries at 4096 * (i >> 1) + 256 * (8 * (i & 1) + nibble index))
        const uint32_t qa = (wd[i] << 4) & 0xF0F0F0F0u, qb = wd[i] & 0xF0F0F0F0u;
        const int base = 4096 * (i >> 1) + 2048 * (i & 1);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint4 a = *reinterpret_cast<const uint4*>(tb + base + 256 * (2 * j) + byte_of(qa, j));
            const uint4 b = *reinterpret_cast<const uint4*>(tb + base + 256 * (2 * j + 1) + byte_of(qb, j));
            acc.x = __builtin_amdgcn_bitop3_b32(acc.x, a.x, b.x, 0x96);
            acc.y = __builtin_amdgcn_bitop3_b32(acc.y, a.y, b.y, 0x96);
            acc.z = __builtin_amdgcn_bitop3_b32(acc.z, a.z, b.z, 0x96);
            acc.w = __builtin_amdgcn_bitop3_b32(acc.w, a.w, b.w, 0x96);
        }
    }
    return {(uint64_t)acc.x | ((uint64_t)acc.y << 32), (uint64_t)acc.z | ((uint64_t)acc.w << 32)};
}

__device__ __forceinline__ uint4 u128_to4(u128v v) {
    return make_uint4((uint32_t)v.lo, (uint32_t)(v.lo >> 32), (uint32_t)v.hi, (uint32_t)(v.hi >> 32));
}

// lam * x^s (0 <= s < 128): the 256-bit shift of lam, reduced.
__device__ __forceinline__ u128v gf128_mul_xpow(u128v lam, int s4) {
    uint64_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
    if (s4 == 0) {
        w0 = lam.lo;
        w1 = lam.hi;
    } else if (s4 < 64) {
        w0 = lam.lo << s4;
        w1 = (lam.hi << s4) | (lam.lo >> (64 - s4));
        w2 = lam.hi >> (64 - s4);
    } else if (s4 == 64) {
        w1 = lam.lo;
        w2 = lam.hi;
    } else {
        const int r = s4 - 64;
        w1 = lam.lo << r;
        w2 = (lam.hi << r) | (lam.lo >> (64 - r));
        w3 = lam.hi >> (64 - r);
    }
    return gf128_reduce256(w0, w1, w2, w3);
}

// Nibble position k (0..31) of the multiply-by-lam table at t (the layout
// mul_tab128_lds reads): entry v = lam * v * x^(4k), from c_b = lam * x^(4k+b).
__device__ __forceinline__ void gf128_tab_build(uint4* t, u128v lam, int k) {
    u128v c[4];
    c[0] = gf128_mul_xpow(lam, 4 * k);
#pragma unroll
    for (int b = 1; b < 4; b++) {
        const uint64_t top = c[b - 1].hi >> 63;
        c[b] = {(c[b - 1].lo << 1) ^ ((0ull - top) & GF128_POLY), (c[b - 1].hi << 1) | (c[b - 1].lo >> 63)};
    }
    uint4* tdst = t + (k < 16 ? 16 * k : 256 + 16 * (k - 16));
#pragma unroll
    for (int v = 0; v < 16; v++) {
        u128v e = {0ull, 0ull};
#pragma unroll
        for (int b = 0; b < 4; b++)
            if ((v >> b) & 1) e = e ^ c[b];
        tdst[v] = u128_to4(e);
    }
}

// A whole multiply-by-lam table built by one wave: lane pair (2k, 2k+1)
// holds nibble position k, lane 2k + h writes entries 8h..8h+7.  Called
// with a wave-uniform lam (a scalar load), so the build waits for no vector
// load in flight.
__device__ __forceinline__ void gf128_tab_build_wave(uint4* t, u128v lam, int lane) {
    const int k = lane >> 1, h = lane & 1;
    // c[0] = lam * x^(4k): one shift and reduction (a chain of k multiplies
    // by x^4 kept the wave for up to 31 steps)
    u128v c[4];
    c[0] = gf128_mul_xpow(lam, 4 * k);
#pragma unroll
    for (int b = 1; b < 4; b++) {
        const uint64_t top = c[b - 1].hi >> 63;
        c[b] = {(c[b - 1].lo << 1) ^ ((0ull - top) & GF128_POLY), (c[b - 1].hi << 1) | (c[b - 1].lo >> 63)};
    }
    // entries 8h + v: c[3] when h, then the XOR of c[0..2] over the bits of
    // v, written in Gray-code order from one running value (few registers
    // live beside the caller's in-flight loads)
    u128v e = h ? c[3] : u128v{0ull, 0ull};
    uint4* tdst = t + (k < 16 ? 16 * k : 256 + 16 * (k - 16)) + 8 * h;
    tdst[0] = u128_to4(e);
    e = e ^ c[0];
    tdst[1] = u128_to4(e);
    e = e ^ c[1];
    tdst[3] = u128_to4(e);
    e = e ^ c[0];
    tdst[2] = u128_to4(e);
    e = e ^ c[2];
    tdst[6] = u128_to4(e);
    e = e ^ c[0];
    tdst[7] = u128_to4(e);
    e = e ^ c[1];
    tdst[5] = u128_to4(e);
    e = e ^ c[0];
    tdst[4] = u128_to4(e);
}

// GF(2^128) transpose-and-scale: tiles of TR message rows x blockDim
// columns.  A thread reads TR consecutive elements of its column (128 bytes,
// one whole cache line of the column-major input; with half a line per
// thread every line was fetched twice), the workgroup builds one multiply
// table per row from pis in LDS, TB rows at a time, and each wave writes rows
// of 64 consecutive columns, every lane multiplying by its row's table
// (wave-uniform, so the reads of a nibble position hit 16 entries).
__global__ __launch_bounds__(256) void k_transpose_scale128(const u128v* __restrict__ poly, size_t m, size_t n,
                                                            const u128v* __restrict__ pis, u128v* __restrict__ msg) {
    constexpr int TR = TS128_ROWS, TB = TS128_TAB_ROWS;
    static_assert(TR % TB == 0, "table rows divide the tile");
    __shared__ __attribute__((aligned(16))) uint4 tabs[TB * 512];
    const size_t cw = blockDim.x;
    const size_t row_tiles = m / TR;
    const size_t i0 = (blockIdx.x % row_tiles) * TR, c = (blockIdx.x / row_tiles) * cw + threadIdx.x;
    u128v v[TR];
#pragma unroll
    for (int r = 0; r < TR; r++) v[r] = poly[c * m + i0 + r];
    // the column reads stay in flight while a wave builds whole row tables
    // from a scalar load of the row's pis entry
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), nw = (int)(blockDim.x >> 6);
    const int lane = (int)threadIdx.x & 63;
#pragma unroll
    for (int h = 0; h < TR / TB; h++) {
        if (h > 0) __syncthreads();  // the previous rows' tables are read
        for (int r = wave; r < TB; r += nw) gf128_tab_build_wave(tabs + r * 512, pis[i0 + h * TB + r], lane);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < TB; r++) msg[(i0 + h * TB + r) * n + c] = mul_tab128_lds(tabs + r * 512, v[h * TB + r]);
    }
}

// One launch = K stages (d0..d0+K-1) of every quarter's size-m FFT.  Each
// thread owns one column of a 2^K-row tile (rows S apart) and runs the K
// stages in registers; a workgroup of 64*J threads covers J consecutive jobs
// (column chunks / tiles) that share B, hence the 2^K - 1 twiddle tables it
// builds in LDS from the plain twiddles (entry v of position k = lambda *
// v * x^(4k), from c_b = lambda * x^(4k+b)).
// Two 1024-thread workgroups per CU (their 56 KB of tables fit twice in
// LDS) need eight waves per SIMD, so at most 64 registers.
template <int K>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void k_fft_pass128(const u128v* __restrict__ src, u128v* __restrict__ dst, size_t m,
                                                      size_t n, int ell, int d0, int q_arg, int src_shared,
                                                      const u128v* __restrict__ tw) {
    constexpr int R = 1 << K, NT = R - 1;
    __shared__ __attribute__((aligned(16))) uint4 tabs[NT * 512];
    const size_t q = q_arg >= 0 ? (size_t)q_arg : (size_t)blockIdx.y;
    const int cc = n < 64 ? (int)n : 64;
    const size_t nchunks = (n + 63) >> 6;  // a power of two
    const int lgnc = 63 - __clzll((unsigned long long)nchunks);
    const int sh = ell - d0 - K;
    const size_t S = (size_t)1 << sh;
    const int J = (int)(blockDim.x >> 6), sub = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const size_t job0 = (size_t)blockIdx.x * J;
    const size_t B = (job0 >> lgnc) >> sh;
    const size_t job = job0 + sub;
    const size_t tile_id = job >> lgnc;
    const size_t c0 = (job & (nchunks - 1)) << 6;
    const size_t lo = tile_id & (S - 1);
    const size_t rbase = (B << (ell - d0)) + lo;
    const bool active = lane < cc;
    // twiddle of table tt (layer d0+j, block B*2^j + t')
    auto lam_of = [&](int tt) -> u128v {
        const int j = 31 - __clz(tt + 1);
        const size_t tp = (size_t)(tt + 1) - ((size_t)1 << j);
        const int d = d0 + j;
        const size_t first = ((size_t)1 << (d + 2)) - 1;
        return tw[first + (q << d) + (B << j) + tp];
    };
    // The tile's loads are issued first and stay in flight while the
    // tables are built: a wave builds whole tables from a scalar load of
    // their twiddle.  The loads are unconditional (a conditional load makes
    // the compiler wait for every load in flight): lanes past the matrix's
    // columns re-read column c0.
    const size_t qoff = src_shared ? 0 : q * m;
    const size_t col = c0 + (active ? (size_t)lane : 0);
    u128v v[R];
#pragma unroll
    for (int a = 0; a < R; a++) v[a] = src[(qoff + rbase + a * S) * n + col];
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    for (int tt = wave; tt < NT; tt += J) gf128_tab_build_wave(tabs + tt * 512, lam_of(tt), lane);
    __syncthreads();
    if (!active) return;
#pragma unroll
    for (int j = 0; j < K; j++) {
        const int h = R >> (j + 1);
#pragma unroll
        for (int a = 0; a < R; a++) {
            if (a & h) continue;
            const int tt = (1 << j) - 1 + (a >> (K - j));
            const u128v u = v[a] ^ mul_tab128_lds(tabs + tt * 512, v[a + h]);
            v[a] = u;
            v[a + h] = v[a + h] ^ u;
        }
    }
#pragma unroll
    for (int a = 0; a < R; a++) dst[(q * m + rbase + a * S) * n + col] = v[a];
}

static void fft_quarters128(const u128v* msg, u128v* mat, size_t m, size_t n, const u128v* tw, hipStream_t s,
                            PassObserver* obs, PassTune tune) {
    int ell = 0;
    while (((size_t)1 << ell) < m) ell++;
    if (tune.kmax < 0 || tune.kmax > 3 || tune.group < 0 || tune.group > FFT128_MAXJ || (tune.group & (tune.group - 1)))
        throw std::invalid_argument("GF(2^128) passes: kmax 1..3, group a power of two up to 16");
    const int KMAX = tune.kmax ? tune.kmax : 3;
    const int JMAX = tune.group ? tune.group : FFT128_MAXJ;
    const int npass = (ell + KMAX - 1) / KMAX;
    const size_t nchunks = (n + 63) / 64;
    const int cc = n < 64 ? (int)n : 64;
    (void)cc;
    // all four quarters per launch (grid.y), as for GF(2^32)
    int d0 = 0;
    for (int p = 0; p < npass; p++) {
        const int K = ell / npass + (p < ell % npass ? 1 : 0);
        const size_t S = (size_t)1 << (ell - d0 - K);
        const size_t jobs = ((size_t)1 << (ell - K)) * nchunks;
        // up to 16 jobs (1024 threads) share one workgroup's tables: the
        // 56 KB of tables then serve 16 waves, and two such workgroups fill
        // a CU with eight waves per SIMD to cover the table reads' latency
        int J = 1;
        while (J < JMAX && (size_t)(J * 2) <= S * nchunks && jobs % (size_t)(J * 2) == 0) J *= 2;
        const u128v* src = (p == 0 && msg) ? msg : mat;
        const int shared = (p == 0 && msg) ? 1 : 0;
        dim3 grid((unsigned)(jobs / J), 4);
        if (obs) obs->before(K, m * n * 4);
        switch (K) {
            case 1: hipLaunchKernelGGL(k_fft_pass128<1>, grid, dim3(64 * J), 0, s, src, mat, m, n, ell, d0, -1, shared, tw); break;
            case 2: hipLaunchKernelGGL(k_fft_pass128<2>, grid, dim3(64 * J), 0, s, src, mat, m, n, ell, d0, -1, shared, tw); break;
            default: hipLaunchKernelGGL(k_fft_pass128<3>, grid, dim3(64 * J), 0, s, src, mat, m, n, ell, d0, -1, shared, tw); break;
        }
        if (obs) obs->after(K, m * n * 4);
        d0 += K;
    }
}

// ------------------------------------------------ GF(2^32) LDS-table helpers --
// Tables live in LDS 512 bytes apart (256-byte aligned base).  For a table
// starting `tb` bytes into the array, the address of the entry for nibble k
// of w is tb | (nibble << 2) plus the immediate 64*k: v_perm_b32 builds it in
// one instruction from tb and a byte of the pre-masked nibble words.
__device__ __forceinline__ uint32_t lds_ld32(const uint32_t* base, uint32_t byte_off) {
    return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(base) + byte_off);
}

// a ^ b ^ c in one instruction (v_bitop3_b32, truth table 0x96)
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// One GF(2^32) table lookup address per nibble: qa = (w << 2) & 0x3C3C3C3C
// holds nibble 2j * 4 in byte j, qb = (w >> 2) & 0x3C3C3C3C nibble 2j+1 * 4;
// a byte permute puts byte j under the table base tb (a multiple of 256),
// and the nibble position's 16-word offset rides in the LDS instruction.
#define LIG_TAB_PERM(tb, q, j) __builtin_amdgcn_perm((tb), (q), 0x07060500u + (j))

__device__ __forceinline__ uint32_t mul_tab_lds(const uint32_t* tabs, uint32_t tb, uint32_t w) {
    const uint32_t qa = (w << 2) & 0x3C3C3C3Cu, qb = (w >> 2) & 0x3C3C3C3Cu;
    uint32_t r0 = lds_ld32(tabs + 0, LIG_TAB_PERM(tb, qa, 0));
    uint32_t r1 = lds_ld32(tabs + 16, LIG_TAB_PERM(tb, qb, 0));
    uint32_t r2 = lds_ld32(tabs + 32, LIG_TAB_PERM(tb, qa, 1));
    uint32_t r3 = lds_ld32(tabs + 48, LIG_TAB_PERM(tb, qb, 1));
    uint32_t r4 = lds_ld32(tabs + 64, LIG_TAB_PERM(tb, qa, 2));
    uint32_t r5 = lds_ld32(tabs + 80, LIG_TAB_PERM(tb, qb, 2));
    uint32_t r6 = lds_ld32(tabs + 96, LIG_TAB_PERM(tb, qa, 3));
    uint32_t r7 = lds_ld32(tabs + 112, LIG_TAB_PERM(tb, qb, 3));
    return xor3(xor3(r0, r1, r2), xor3(r3, r4, r5), r6 ^ r7);
}

// NM independent butterflies u[i] ^= w[i] * (table at byte offset tb[i])
// with every LDS read issued before the first XOR, so one wave keeps 8*NM
// reads in flight.  Returns the updated u in place.
template <int NM>
__device__ __forceinline__ void mul_tab_lds_n(const uint32_t* tabs, const uint32_t (&tb)[NM], const uint32_t (&w)[NM],
                                              uint32_t (&u)[NM]) {
    uint32_t r[NM][8];
#pragma unroll
    for (int i = 0; i < NM; i++) {
        const uint32_t qa = (w[i] << 2) & 0x3C3C3C3Cu, qb = (w[i] >> 2) & 0x3C3C3C3Cu;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            r[i][2 * j] = LIG_TAB_PERM(tb[i], qa, j);
            r[i][2 * j + 1] = LIG_TAB_PERM(tb[i], qb, j);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < NM; i++)
#pragma unroll
        for (int k = 0; k < 8; k++) r[i][k] = lds_ld32(tabs + 16 * k, r[i][k]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < NM; i++)
        u[i] = xor3(xor3(r[i][0], r[i][1], r[i][2]), xor3(r[i][3], r[i][4], r[i][5]), xor3(r[i][6], r[i][7], u[i]));
}

// --------------------------------------------- GF(2^32) register-radix pass --
// One launch = stages d0..d0+K-1 of quarter q's size-m FFT for every tile.
// A tile is the 2^K rows r(t) = B*2^(ell-d0) + t*S + lo (S = 2^(ell-d0-K))
// of up to 64 columns.  Phase 1 keeps a chain of 2^K1 rows (t = a*2^K2 + g)
// of one column in registers and runs the first K1 stages; one LDS exchange
// regroups the tile into chains of 2^K2 consecutive rows for the last K2
// stages.  All 2^K - 1 twiddle tables of the tile are staged once: the
// twiddle of block B*2^j + t' at stage j is (by GF(2)-linearity in the block
// index) lambda(base_j) + lambda(t'), so each entry is the XOR of the tile's
// base table and a table shared by every tile.
// Global word access at a 32-bit byte offset from a wave-uniform base (the
// base stays in scalar registers, the offset in one vector register).
__device__ __forceinline__ uint32_t gld_off(const uint32_t* base, uint32_t byte_off) {
    return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(base) + byte_off);
}
__device__ __forceinline__ void gst_off(uint32_t* base, uint32_t byte_off, uint32_t v) {
    *reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

// Raw buffer access (gfx9 resource word 3 = 0x00020000): base and the
// wave-uniform offset live in scalar registers, the lane offset in one vector
// register, and accesses past num_bytes read 0 / are dropped.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t num_bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)num_bytes, 0x00020000);
}
__device__ __forceinline__ uint32_t buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, uint32_t v) {
    __builtin_amdgcn_raw_buffer_store_b32(v, r, (int)voff, (int)soff, 0);
}

// Six waves per SIMD: three 512-thread workgroups per CU, the most their LDS
// (48.6 KB each at K = 6) allows.  Offsets inside one quarter are 32-bit
// (the launcher requires a quarter below 2^31 bytes).  The loop-invariant
// values that do not fit in the 80-register budget live in scratch; they
// are reloaded once per tile.
template <int K1, int K2>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(6))) void k_fft_pass32(
    const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, size_t m, size_t n, int ell, int d0, int q_arg,
    int src_shared, int tiles_per_wg, const uint32_t* __restrict__ tw) {
    constexpr int K = K1 + K2, R = 1 << K, NT = R - 1;
    constexpr int L1 = 1 << K1, L2 = 1 << K2;
    __shared__ __attribute__((aligned(256))) uint32_t tabs[NT * 128];
    __shared__ uint32_t xch[K2 > 0 ? R * 64 : 1];
    // blockIdx.y runs over the quarters of one or more codewords laid out
    // one after another ([codeword][quarter][m][n]); quarter qq is quarter
    // qq & 3 of codeword qq >> 2, whose message (when shared) is the
    // (qq >> 2)-th m x n block of src
    const uint32_t qq = q_arg >= 0 ? (uint32_t)q_arg : blockIdx.y;
    const uint32_t q = qq & 3u;
    const int cc = n < 64 ? (int)n : 64;
    const int lgcc = 31 - __clz(cc);
    const uint32_t nchunks = (uint32_t)((n + 63) >> 6);  // a power of two
    const int lgnc = 31 - __clz(nchunks);
    const int sh = ell - d0 - K;
    const uint32_t S = 1u << sh;
    const int tid = threadIdx.x;
    const uint32_t c = tid & (cc - 1);
    const uint32_t g = tid >> lgcc;  // < L2
    const uint32_t rs4 = (S * (uint32_t)n) << 2;  // byte stride between tile rows
    const uint32_t lane4 = (g * (S * (uint32_t)n) + c) << 2;
    const uint32_t qbytes = (uint32_t)(m * n * 4);
    const __amdgpu_buffer_rsrc_t rsrc = buf_rsrc(src + (size_t)(src_shared ? (qq >> 2) : qq) * m * n, qbytes);
    const __amdgpu_buffer_rsrc_t rdst = buf_rsrc(dst + (size_t)qq * m * n, qbytes);
    // tiles of one workgroup share B (consecutive lo), hence their tables
    const uint32_t first_job = blockIdx.x * (uint32_t)tiles_per_wg;
    const uint32_t B = (first_job >> lgnc) >> sh;
    uint32_t zt = 0;  // bit j: phase-1 stage j's first table is zero

    // byte offset of tile tj's first element
    auto tile_off = [&](int tj) -> uint32_t {
        const uint32_t job = first_job + tj;
        const uint32_t tile_id = job >> lgnc;
        const uint32_t c0 = (job & (nchunks - 1)) << 6;
        const uint32_t lo = tile_id & (S - 1);
        const uint32_t rbase = (B << (ell - d0)) + lo;
        return (rbase * (uint32_t)n + c0) << 2;
    };
    // Twiddle tables of the workgroup's tiles: table (j, t') multiplies by
    // the heap twiddle of layer d0+j, block B*2^j + t'.  Work item (table,
    // nibble position k) builds entries v = 0..15 from c_b = lambda *
    // x^(4k+b): entry v = XOR of c_b over the bits of v.
    auto lam_of = [&](int idx) -> uint32_t {
        const int tt = idx >> 3;
        const int j = 31 - __clz(tt + 1);
        const uint32_t tp = (uint32_t)(tt + 1) - (1u << j);
        const int d = d0 + j;
        const size_t first = ((size_t)1 << (d + 2)) - 1;
        return tw[first + ((size_t)q << d) + ((size_t)B << j) + tp];
    };
    // A thread's first twiddle is loaded before tile 0's words, so the table
    // build waits for it alone while the tile's loads are in flight.  Tiles
    // then alternate between the two register sets: the next tile's words
    // are loaded while this tile computes (one set per role, no copies).
    uint32_t va[L1], vb[L1];
    const uint32_t lam0 = tid < NT * 8 ? lam_of(tid) : 0u;
    {
        const uint32_t toff0 = tile_off(0);
#pragma unroll
        for (int a = 0; a < L1; a++) va[a] = buf_ld(rsrc, lane4, toff0 + (uint32_t)(a * L2) * rs4);
    }
    if (tid < NT * 8) gf32_tab_build<true>(tabs + (tid >> 3) * 128, lam0, tid & 7);
    for (int idx = tid + (int)blockDim.x; idx < NT * 8; idx += blockDim.x)
        gf32_tab_build<true>(tabs + (idx >> 3) * 128, lam_of(idx), idx & 7);
    LIG_LDS_BARRIER();
    // Stage j's first table (block B*2^j) is all zero when its twiddle is:
    // entry 1 of nibble 0 is the twiddle itself.  The heap's first twiddle
    // of every layer is zero, so quarter 0's leading tiles skip those
    // multiplies (a quarter of the first stage's work and less after it).
    // Same for every tile of the workgroup.
#pragma unroll
    for (int j = 0; j < K1; j++) zt |= (tabs[((1 << j) - 1) * 128 + 1] == 0u ? 1u : 0u) << j;
    zt = __builtin_amdgcn_readfirstlane(zt);

    auto tile = [&](uint32_t(&v)[L1], uint32_t(&vn)[L1], int tj) {
        const uint32_t toff = tile_off(tj);
        if (tj > 0) LIG_LDS_BARRIER();  // the previous tile's exchange reads are done
        {
            // issued for every tile (a conditional load here makes the
            // compiler wait for all loads in flight): after the last tile
            // the lane offset lies past the quarter, so the buffer loads
            // return zeros without touching memory (offsets stay below 2^32:
            // quarter offsets are below 2^31)
            const uint32_t toffn = tile_off(tj + 1);
            const uint32_t lanen = tj + 1 < tiles_per_wg ? lane4 : 0x80000000u;
#pragma unroll
            for (int a = 0; a < L1; a++) vn[a] = buf_ld(rsrc, lanen, toffn + (uint32_t)(a * L2) * rs4);
        }

        auto stage1 = [&](auto J) {
            constexpr int j = decltype(J)::value, h = L1 >> (j + 1);
            // butterflies a < h sit in the stage's first block (table (j, 0))
            constexpr int NS = L1 / 2 - h;  // butterflies outside the first block
            if ((zt >> j) & 1u) {
#pragma unroll
                for (int a = 0; a < h; a++) v[a + h] ^= v[a];  // u unchanged: twiddle 0
                if constexpr (NS > 0) {
                    uint32_t tbs[NS], ws[NS], us[NS];
                    int b = 0;
#pragma unroll
                    for (int a = 2 * h; a < L1; a++) {
                        if (a & h) continue;
                        tbs[b] = (uint32_t)(((1 << j) - 1 + (a >> (K1 - j))) * 512);
                        us[b] = v[a];
                        ws[b++] = v[a + h];
                    }
                    mul_tab_lds_n<NS>(tabs, tbs, ws, us);
                    b = 0;
#pragma unroll
                    for (int a = 2 * h; a < L1; a++) {
                        if (a & h) continue;
                        v[a] = us[b];
                        v[a + h] ^= us[b++];
                    }
                }
                return;
            }
            uint32_t tbs[L1 / 2], ws[L1 / 2], us[L1 / 2];
            int b = 0;
#pragma unroll
            for (int a = 0; a < L1; a++) {
                if (a & h) continue;
                tbs[b] = (uint32_t)(((1 << j) - 1 + (a >> (K1 - j))) * 512);
                us[b] = v[a];
                ws[b++] = v[a + h];
            }
            mul_tab_lds_n<L1 / 2>(tabs, tbs, ws, us);
            b = 0;
#pragma unroll
            for (int a = 0; a < L1; a++) {
                if (a & h) continue;
                v[a] = us[b];
                v[a + h] ^= us[b++];
            }
        };
        stage1(std::integral_constant<int, 0>{});
        if constexpr (K1 > 1) stage1(std::integral_constant<int, 1>{});
        if constexpr (K1 > 2) stage1(std::integral_constant<int, 2>{});
        static_assert(K1 <= 3, "phase 1 holds at most three stages");

        if constexpr (K2 == 0) {
#pragma unroll
            for (int a = 0; a < L1; a++) buf_st(rdst, lane4, toff + (uint32_t)(a * L2) * rs4, v[a]);
        } else {
#pragma unroll
            for (int a = 0; a < L1; a++) xch[(a * L2 + g) * 64 + c] = v[a];
            LIG_LDS_BARRIER();
            constexpr int NC2 = L1 / L2;
#pragma unroll
            for (int i = 0; i < NC2; i++) {
                const uint32_t hi = g + L2 * i;
#pragma unroll
                for (int e = 0; e < L2; e++) v[i * L2 + e] = xch[(hi * L2 + e) * 64 + c];
            }
#pragma unroll
            for (int j = K1; j < K; j++) {
                const int h = L2 >> (j - K1 + 1);
                uint32_t tbs[L1 / 2], ws[L1 / 2], us[L1 / 2];
                int b = 0;
#pragma unroll
                for (int i = 0; i < NC2; i++) {
                    const uint32_t hi = g + L2 * i;
#pragma unroll
                    for (int e = 0; e < L2; e++) {
                        if (e & h) continue;
                        tbs[b] = (uint32_t)(((1 << j) - 1 + (hi << (j - K1)) + (e >> (K - j))) * 512);
                        us[b] = v[i * L2 + e];
                        ws[b++] = v[i * L2 + e + h];
                    }
                }
                mul_tab_lds_n<L1 / 2>(tabs, tbs, ws, us);
                b = 0;
#pragma unroll
                for (int i = 0; i < NC2; i++) {
#pragma unroll
                    for (int e = 0; e < L2; e++) {
                        if (e & h) continue;
                        v[i * L2 + e] = us[b];
                        v[i * L2 + e + h] ^= us[b++];
                    }
                }
            }
            // row hi*L2+e of the tile sits (hi*L2+e)*S rows from its first
            // row, hi = g + L2*i: lane part g*L2 rows + column c
            const uint32_t lane_st = g * (uint32_t)L2 * rs4 + (c << 2);
#pragma unroll
            for (int i = 0; i < NC2; i++) {
#pragma unroll
                for (int e = 0; e < L2; e++)
                    buf_st(rdst, lane_st, toff + (uint32_t)(L2 * L2 * i + e) * rs4, v[i * L2 + e]);
            }
        }
    };
    // tile 0 first, so that every entry to the loop has the same memory
    // operations in flight (the prefetched words, then the stores), which
    // lets each tile wait for its own words only
    tile(va, vb, 0);
    for (int tj = 1;; tj += 2) {
        if (tj >= tiles_per_wg) break;
        tile(vb, va, tj);
        if (tj + 1 >= tiles_per_wg) break;
        tile(va, vb, tj + 1);
    }
}

static void launch_pass32(int K, dim3 grid, int block, hipStream_t s, const uint32_t* src, uint32_t* dst, size_t m, size_t n,
                          int ell, int d0, int q, int shared, int tpw, const uint32_t* tw) {
#define LIG_P32(A, B_)                                                                                                    \
    hipLaunchKernelGGL((k_fft_pass32<A, B_>), grid, dim3(block), 0, s, src, dst, m, n, ell, d0, q, shared, tpw, tw);     \
    break;
    switch (K) {
        case 1: LIG_P32(1, 0)
        case 2: LIG_P32(1, 1)
        case 3: LIG_P32(2, 1)
        case 4: LIG_P32(2, 2)
        case 5: LIG_P32(3, 2)
        case 6: LIG_P32(3, 3)
        default: break;
    }
#undef LIG_P32
}

// GF(2^32) encode passes, quarter by quarter so that a quarter (m x n x 4
// bytes) stays resident in the last-level cache between its passes.
static void fft_quarters32(const uint32_t* msg, uint32_t* mat, size_t m, size_t n, const uint32_t* tw, hipStream_t s,
                           PassObserver* obs, PassTune tune, int codewords) {
    if (m * n * sizeof(uint32_t) >= ((size_t)1 << 31)) throw std::invalid_argument("FFT quarter of 2 GiB or more");
    if (n == 0 || (n & (n - 1)) != 0) throw std::invalid_argument("column count must be a power of two");
    if (tune.kmax < 0 || tune.kmax == 1 || tune.kmax > 6 || tune.group < 0 || tune.group > 8 || (tune.group & (tune.group - 1)))
        throw std::invalid_argument("GF(2^32) passes: kmax 2..6, group a power of two up to 8");
    int ell = 0;
    while (((size_t)1 << ell) < m) ell++;
    const int KMAX = tune.kmax ? tune.kmax : 6;
    const int TPW = tune.group ? tune.group : 4;
    const int npass = (ell + KMAX - 1) / KMAX;
    // Narrow matrices (n < 64): the 2^v = 64/n consecutive rows of an
    // aligned group are contiguous, so a layer that pairs rows 2^v or more
    // apart is the same layer (same twiddle blocks) of the (m >> v) x 64
    // matrix over the same words.  Passes made only of such layers run on
    // that view: 64-wide tiles, one set of tables per 64 columns.
    int v = 0;
    while ((n << v) < 64 && v < ell) v++;
    // All four quarters per launch (grid.y): one quarter per launch leaves a
    // partly filled last round of workgroups at the end of every launch,
    // which costs more than the quarter's cache residency between passes
    // saves (measured: 495 us against 534 us for the 2^18 x 64 encode).
    int d0 = 0;
    for (int p = 0; p < npass; p++) {
        const int K = ell / npass + (p < ell % npass ? 1 : 0);
        const int K2 = K / 2;
        const bool wide = d0 + K <= ell - v;
        const size_t pm = wide ? m >> v : m, pn = wide ? n << v : n;
        const int pell = wide ? ell - v : ell;
        const size_t nchunks = (pn + 63) / 64;
        const int cc = pn < 64 ? (int)pn : 64;
        const size_t tiles = (size_t)1 << (pell - K);
        const size_t S = (size_t)1 << (pell - d0 - K);
        // tiles sharing B (hence tables) are grouped per workgroup
        size_t jobs = tiles * nchunks;
        int tpw = 1;
        while (tpw < TPW && (size_t)(tpw * 2) <= S * nchunks && jobs / (tpw * 2) >= 512) tpw *= 2;
        const uint32_t* src = (p == 0 && msg) ? msg : mat;
        const int shared = (p == 0 && msg) ? 1 : 0;
        dim3 grid((unsigned)(jobs / tpw), 4u * (unsigned)codewords);
        const size_t elems = m * n * 4 * (size_t)codewords;
        if (obs) obs->before(K, elems);
        launch_pass32(K, grid, cc << K2, s, src, mat, pm, pn, pell, d0, -1, shared, tpw, tw);
        if (obs) obs->after(K, elems);
        d0 += K;
    }
}

// ------------------------------------------------------------- leaf hash ---
// One lane per leaf: BLAKE3 over the row's little-endian bytes.
template <class F>
__global__ void k_leaf_hash(const typename F::T* __restrict__ mat, size_t rows, size_t n, uint32_t* __restrict__ layer0) {
    const size_t words = n * (sizeof(typename F::T) / 4);
    const size_t len = words * 4;
    const size_t nblocks = (len + 63) / 64;
    for (size_t row = blockIdx.x * (size_t)blockDim.x + threadIdx.x; row < rows; row += (size_t)gridDim.x * blockDim.x) {
        const uint32_t* p = reinterpret_cast<const uint32_t*>(mat) + row * words;
        uint32_t m[16];
        uint32_t out[8];
        const size_t lb = (nblocks - 1) * 16;
        const uint32_t last_len = (uint32_t)(len - (nblocks - 1) * 64);
        if (nblocks <= 16) {
            // single chunk: chain the blocks, ROOT on the last one
            uint32_t cv[8];
#pragma unroll
            for (int i = 0; i < 8; i++) cv[i] = b3::iv(i);
            for (size_t b = 0; b + 1 < nblocks; b++) {
                const uint4* q4 = reinterpret_cast<const uint4*>(p + b * 16);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    uint4 v = q4[i];
                    m[4 * i] = v.x;
                    m[4 * i + 1] = v.y;
                    m[4 * i + 2] = v.z;
                    m[4 * i + 3] = v.w;
                }
                b3::compress_cv(cv, m, 0, 64, b == 0 ? b3::CHUNK_START : 0u, cv);
            }
#pragma unroll
            for (int i = 0; i < 16; i++) m[i] = (lb + i < words) ? p[lb + i] : 0u;
            b3::compress_cv(cv, m, 0, last_len, b3::CHUNK_END | b3::ROOT | (nblocks == 1 ? b3::CHUNK_START : 0u), out);
        } else {
            b3::Hasher h;
            h.init();
            for (size_t b = 0; b + 1 < nblocks; b++) {
                const uint4* q4 = reinterpret_cast<const uint4*>(p + b * 16);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    uint4 v = q4[i];
                    m[4 * i] = v.x;
                    m[4 * i + 1] = v.y;
                    m[4 * i + 2] = v.z;
                    m[4 * i + 3] = v.w;
                }
                h.block(m);
            }
#pragma unroll
            for (int i = 0; i < 16; i++) m[i] = (lb + i < words) ? p[lb + i] : 0u;
            h.finish(m, last_len, out);
        }
        uint32_t* o = layer0 + row * 8;
#pragma unroll
        for (int i = 0; i < 8; i++) o[i] = out[i];
    }
}

// Leaves of any byte length (build_merkle_tree over AsRef<[u8]> leaves,
// merkle_tree.rs:4-20): leaf i is bytes[offs[i] .. offs[i+1]), at most
// 256 KiB.  One lane per leaf assembles little-endian words from bytes (a
// leaf may start at any byte) and runs the BLAKE3 chunk chain; the last
// block, zero-padded, may be partial or (for an empty leaf) empty.
__global__ __launch_bounds__(256) void k_leaf_hash_var(const uint8_t* __restrict__ bytes, const uint64_t* __restrict__ offs,
                                                       size_t n_leaves, uint32_t* __restrict__ layer0) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n_leaves; i += (size_t)gridDim.x * blockDim.x) {
        const uint8_t* p = bytes + offs[i];
        const uint64_t len = offs[i + 1] - offs[i];
        const uint64_t nblocks = len == 0 ? 1 : (len + 63) / 64;
        b3::Hasher h;
        h.init();
        uint32_t m[16];
        for (uint64_t b = 0; b < nblocks; b++) {
            const uint64_t base = b * 64;
#pragma unroll
            for (int w = 0; w < 16; w++) {
                uint32_t x = 0;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint64_t at = base + 4 * w + k;
                    if (at < len) x |= (uint32_t)p[at] << (8 * k);
                }
                m[w] = x;
            }
            if (b + 1 < nblocks) h.block(m);
            else {
                uint32_t out[8];
                h.finish(m, (uint32_t)(len - base), out);
                uint4* o = reinterpret_cast<uint4*>(layer0 + i * 8);
                o[0] = make_uint4(out[0], out[1], out[2], out[3]);
                o[1] = make_uint4(out[4], out[5], out[6], out[7]);
            }
        }
    }
}

void launch_leaf_hash_var(const uint8_t* bytes, const uint64_t* offs, size_t n_leaves, uint32_t* layer0, hipStream_t s) {
    if (!n_leaves) return;
    hipLaunchKernelGGL(k_leaf_hash_var, dim3(grid_for(n_leaves, 256)), dim3(256), 0, s, bytes, offs, n_leaves, layer0);
}

// Rows of NB whole 64-byte blocks (one chunk): each lane loads its whole row
// first (all loads in flight together, so each row's cache lines are read
// from HBM once), then chains the NB compressions from registers.  Rows
// come in groups of 2^lg_per, one per tree: the leaf of row r goes to
// layer0 + (r >> lg_per) * tree_stride + (r mod 2^lg_per) * 8.
template <int NB>
__global__ __launch_bounds__(128) void k_leaf_hash_blocks(const uint32_t* __restrict__ mat, size_t rows,
                                                          uint32_t* __restrict__ layer0, int lg_per, size_t tree_stride) {
    const size_t per_mask = lg_per >= 63 ? ~(size_t)0 : ((size_t)1 << lg_per) - 1;
    for (size_t row = blockIdx.x * (size_t)blockDim.x + threadIdx.x; row < rows; row += (size_t)gridDim.x * blockDim.x) {
        const uint4* p = reinterpret_cast<const uint4*>(mat + row * 16 * NB);
        uint4 v[4 * NB];
#pragma unroll
        for (int i = 0; i < 4 * NB; i++) v[i] = p[i];
        uint32_t cv[8];
#pragma unroll
        for (int i = 0; i < 8; i++) cv[i] = b3::iv(i);
#pragma unroll
        for (int b = 0; b < NB; b++) {
            uint32_t m[16];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                m[4 * i] = v[4 * b + i].x;
                m[4 * i + 1] = v[4 * b + i].y;
                m[4 * i + 2] = v[4 * b + i].z;
                m[4 * i + 3] = v[4 * b + i].w;
            }
            const uint32_t fl = (b == 0 ? b3::CHUNK_START : 0u) | (b == NB - 1 ? (b3::CHUNK_END | b3::ROOT) : 0u);
            b3::compress_cv(cv, m, 0, 64, fl, cv);
        }
        const size_t tr = lg_per >= 63 ? 0 : row >> lg_per;
        uint4* o = reinterpret_cast<uint4*>(layer0 + tr * tree_stride + (row & per_mask) * 8);
        o[0] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
        o[1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
    }
}

// Leaves of nch whole 1 KiB chunks (nch a power of two, 2..256): one
// thread per chunk chains its 16 block compressions (chunk counter = chunk
// index) and writes the chunk's chaining value ...
__global__ __launch_bounds__(256) void k_leaf_chunks(const uint32_t* __restrict__ mat, size_t rows, int lgnch,
                                                     uint32_t* __restrict__ cvs) {
    const size_t total = rows << lgnch;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const uint32_t ch = (uint32_t)(t & ((1u << lgnch) - 1));
        const uint4* p = reinterpret_cast<const uint4*>(mat + t * 256);
        uint32_t cv[8];
#pragma unroll
        for (int i = 0; i < 8; i++) cv[i] = b3::iv(i);
        uint4 nx[4];
#pragma unroll
        for (int i = 0; i < 4; i++) nx[i] = p[i];
        for (int b = 0; b < 16; b++) {
            uint32_t m[16];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                m[4 * i] = nx[i].x;
                m[4 * i + 1] = nx[i].y;
                m[4 * i + 2] = nx[i].z;
                m[4 * i + 3] = nx[i].w;
            }
            if (b < 15) {
#pragma unroll
                for (int i = 0; i < 4; i++) nx[i] = p[4 * (b + 1) + i];
            }
            const uint32_t fl = (b == 0 ? b3::CHUNK_START : 0u) | (b == 15 ? b3::CHUNK_END : 0u);
            b3::compress_cv(cv, m, ch, 64, fl, cv);
        }
        uint4* o = reinterpret_cast<uint4*>(cvs + t * 8);
        o[0] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
        o[1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
    }
}

// ... then one thread per leaf combines its chunk values with PARENT nodes
// (a perfect binary tree, ROOT on the top), in place in cvs.
__global__ __launch_bounds__(128) void k_leaf_chunk_tree(uint32_t* __restrict__ cvs, size_t rows, int lgnch,
                                                         uint32_t* __restrict__ layer0) {
    for (size_t row = blockIdx.x * (size_t)blockDim.x + threadIdx.x; row < rows; row += (size_t)gridDim.x * blockDim.x) {
        uint32_t* c = cvs + (row << lgnch) * 8;
        uint32_t out[8];
        for (int cnt = 1 << (lgnch - 1); cnt >= 1; cnt >>= 1) {
            for (int i = 0; i < cnt; i++) {
                uint32_t m[16];
#pragma unroll
                for (int w = 0; w < 16; w++) m[w] = c[16 * i + w];
                uint32_t k[8];
#pragma unroll
                for (int w = 0; w < 8; w++) k[w] = b3::iv(w);
                b3::compress_cv(k, m, 0, 64, b3::PARENT | (cnt == 1 ? b3::ROOT : 0u), out);
#pragma unroll
                for (int w = 0; w < 8; w++) c[8 * i + w] = out[w];
            }
        }
        uint4* o = reinterpret_cast<uint4*>(layer0 + row * 8);
        o[0] = make_uint4(out[0], out[1], out[2], out[3]);
        o[1] = make_uint4(out[4], out[5], out[6], out[7]);
    }
}

size_t leaf_hash_scratch_words(size_t rows, size_t row_bytes) {
    const size_t nch = row_bytes / 1024;
    if (row_bytes % 1024 != 0 || nch < 2 || nch > 256 || (nch & (nch - 1)) != 0) return 0;
    return rows * nch * 8;
}

template <class F>
void launch_leaf_hash(const typename F::T* mat, size_t rows, size_t n, uint32_t* layer0, uint32_t* scratch, hipStream_t s,
                      int trees, size_t tree_stride) {
    const size_t words = n * (sizeof(typename F::T) / 4);
    if (trees != 1) {
        if (trees < 1 || rows % (size_t)trees != 0) throw std::invalid_argument("leaf hash: rows must split evenly over the trees");
        const size_t per = rows / (size_t)trees;
        if (per & (per - 1)) throw std::invalid_argument("leaf hash: rows per tree must be a power of two");
        int lg = 0;
        while (((size_t)1 << lg) < per) lg++;
        if (words == 16 || words == 32 || words == 64) {
            const dim3 grid(grid_for(rows, 128)), block(128);
            const uint32_t* m = reinterpret_cast<const uint32_t*>(mat);
            if (words == 16) hipLaunchKernelGGL(k_leaf_hash_blocks<1>, grid, block, 0, s, m, rows, layer0, lg, tree_stride);
            else if (words == 32) hipLaunchKernelGGL(k_leaf_hash_blocks<2>, grid, block, 0, s, m, rows, layer0, lg, tree_stride);
            else hipLaunchKernelGGL(k_leaf_hash_blocks<4>, grid, block, 0, s, m, rows, layer0, lg, tree_stride);
            return;
        }
        // other leaf shapes: one tree per launch (the scratch is reused in
        // stream order)
        for (int t = 0; t < trees; t++)
            launch_leaf_hash<F>(mat + (size_t)t * per * n, per, n, layer0 + (size_t)t * tree_stride, scratch, s, 1, 0);
        return;
    }
    const dim3 grid(grid_for(rows, 128)), block(128);
    const uint32_t* m = reinterpret_cast<const uint32_t*>(mat);
    if (scratch && leaf_hash_scratch_words(rows, words * 4)) {
        int lgnch = 0;
        while (((size_t)1024 << lgnch) < words * 4) lgnch++;
        hipLaunchKernelGGL(k_leaf_chunks, dim3(grid_for(rows << lgnch, 256)), dim3(256), 0, s, m, rows, lgnch, scratch);
        hipLaunchKernelGGL(k_leaf_chunk_tree, grid, block, 0, s, scratch, rows, lgnch, layer0);
        return;
    }
    switch (words) {
        case 16: hipLaunchKernelGGL(k_leaf_hash_blocks<1>, grid, block, 0, s, m, rows, layer0, 63, (size_t)0); return;
        case 32: hipLaunchKernelGGL(k_leaf_hash_blocks<2>, grid, block, 0, s, m, rows, layer0, 63, (size_t)0); return;
        case 64: hipLaunchKernelGGL(k_leaf_hash_blocks<4>, grid, block, 0, s, m, rows, layer0, 63, (size_t)0); return;
        default: hipLaunchKernelGGL(k_leaf_hash<F>, grid, block, 0, s, mat, rows, n, layer0); return;
    }
}

// ----------------------------------------------------------- Merkle levels --
// One tree level: each thread hashes one pair of children.
__global__ __launch_bounds__(256) void k_merkle_level(uint32_t* __restrict__ tree, size_t in_off, size_t n_out,
                                                      size_t tree_stride) {
    tree += blockIdx.z * tree_stride;  // one tree per blockIdx.z
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n_out; i += (size_t)gridDim.x * blockDim.x) {
        const uint4* src = reinterpret_cast<const uint4*>(tree + (in_off + 2 * i) * 8);
        uint32_t msg[16];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint4 v = src[k];
            msg[4 * k] = v.x;
            msg[4 * k + 1] = v.y;
            msg[4 * k + 2] = v.z;
            msg[4 * k + 3] = v.w;
        }
        uint32_t out[8];
        b3::hash64(msg, out);
        uint4* dst = reinterpret_cast<uint4*>(tree + (in_off + 2 * n_out + i) * 8);
        dst[0] = make_uint4(out[0], out[1], out[2], out[3]);
        dst[1] = make_uint4(out[4], out[5], out[6], out[7]);
    }
}

// ---------------------------------------- four-lane parent compression ---
// Near the root a level has too few parents to fill the GPU, so its time is
// the latency of one compression.  Four lanes of a quad share one parent:
// lane k holds column k of the 4x4 state (words k, 4+k, 8+k, 12+k), runs
// the column-step G_k, then the diagonal-step G_(4+k) after rotating rows
// 1, 2, 3 left by 1, 2, 3 lanes (quad DPP moves), and rotates them back.
// Message word for round r: column step S_r[2k], S_r[2k+1]; diagonal step
// S_r[8+2k], S_r[9+2k], with S_r the BLAKE3 schedule of compress_cv.
namespace b3x4 {
constexpr uint8_t SCHED[7][16] = {{0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15},
                                  {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8},
                                  {3, 4, 10, 12, 13, 2, 7, 14, 6, 5, 9, 0, 11, 15, 8, 1},
                                  {10, 7, 12, 9, 14, 3, 13, 15, 4, 0, 11, 2, 5, 8, 1, 6},
                                  {12, 13, 9, 11, 15, 10, 14, 8, 7, 2, 5, 3, 0, 1, 6, 4},
                                  {9, 14, 11, 5, 8, 12, 15, 1, 13, 3, 0, 10, 2, 6, 4, 7},
                                  {11, 15, 5, 0, 1, 9, 8, 6, 14, 10, 2, 12, 3, 4, 7, 13}};
// word indices of slot s (0, 1: column step; 2, 3: diagonal step) for the
// four lanes, four bits per lane
constexpr uint32_t packed(int r, int s) {
    uint32_t v = 0;
    for (int k = 0; k < 4; k++) v |= (uint32_t)SCHED[r][(s < 2 ? 2 * k + s : 8 + 2 * k + (s - 2))] << (4 * k);
    return v;
}
// lane i of each quad receives lane sel_i: quad_perm(1,2,3,0) = 0x39,
// (2,3,0,1) = 0x4E, (3,0,1,2) = 0x93
template <int CTRL>
__device__ __forceinline__ uint32_t qperm(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
}  // namespace b3x4

// Subtree of depth g per workgroup (blockDim.x / 4 quads): 2^g nodes of the
// layer at in_off become the g layers above them.  Dynamic LDS holds two
// layers of at most 2^g nodes (2^(g+6) bytes).  Each level's nodes go to LDS
// for the next level and to the tree in global memory; the barrier between
// levels does not wait for the global stores.  With a flag (the one
// workgroup of the top launch), once every node is written the kernel stores
// seq there (host memory; system scope, after the nodes and the root).
// With rows, the input layer is a matrix of 64-byte rows read as pairs of
// 32-byte nodes: its first level are the leaves (BLAKE3 of a 64-byte row is
// one block with a parent's flags), written at node 0 of the tree (in_off is
// then minus n_in, modulo 2^64).
__global__ __launch_bounds__(1024) void k_merkle_subtree4(uint32_t* __restrict__ tree, size_t in_off, size_t n_in, int g,
                                                          uint32_t* __restrict__ root_out, uint32_t* __restrict__ flag,
                                                          uint32_t seq, size_t tree_stride,
                                                          const uint32_t* __restrict__ rows, size_t rows_stride) {
    using namespace b3x4;
    // one tree per blockIdx.z (a flag is only given for a single tree)
    tree += blockIdx.z * tree_stride;
    if (rows) rows += blockIdx.z * rows_stride;
    if (root_out) root_out += blockIdx.z * 8;
    // the two layers live at word 0 and word 8 << g; picked by offset (a
    // pointer array indexed at run time would turn the LDS accesses into
    // flat ones)
    extern __shared__ uint32_t sub4_lds[];
    const uint32_t lv_words = 8u << g;
    const int k = threadIdx.x & 3, quad = threadIdx.x >> 2, nquads = blockDim.x >> 2;
    const size_t base = blockIdx.x * ((size_t)1 << g);  // first input node of this workgroup
    // message word index per (round, slot) for this lane
    uint32_t widx[28];
#pragma unroll
    for (int r = 0; r < 7; r++)
#pragma unroll
        for (int s = 0; s < 4; s++) widx[4 * r + s] = (packed(r, s) >> (4 * k)) & 15u;
    // initial state column k: IV[k], IV[4+k], IV[k], (counter lo, hi, 64, flags)[k]
    const uint32_t iv_lo = k == 0 ? b3::iv(0) : k == 1 ? b3::iv(1) : k == 2 ? b3::iv(2) : b3::iv(3);
    const uint32_t iv_hi = k == 0 ? b3::iv(4) : k == 1 ? b3::iv(5) : k == 2 ? b3::iv(6) : b3::iv(7);
    const uint32_t d_init = k == 2 ? 64u : k == 3 ? (uint32_t)(b3::CHUNK_START | b3::CHUNK_END | b3::ROOT) : 0u;
    {
        const uint4* src = reinterpret_cast<const uint4*>(rows ? rows + base * 8 : tree + (in_off + base) * 8);
        uint4* dst = reinterpret_cast<uint4*>(sub4_lds);
        for (int i = threadIdx.x; i < (2 << g); i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    size_t lay_off = in_off, lay_n = n_in;
    for (int l = 1; l <= g; l++) {
        const int cnt = 1 << (g - l);
        const uint32_t* in = sub4_lds + ((l - 1) & 1) * lv_words;
        uint32_t* out = sub4_lds + (l & 1) * lv_words;
        const size_t out_off = lay_off + lay_n;
        const size_t out_base = base >> l;
        for (int p = quad; p < cnt; p += nquads) {
            const uint32_t* m = in + 16 * p;  // left child, then right child
            uint32_t mw[28];
#pragma unroll
            for (int i = 0; i < 28; i++) mw[i] = m[widx[i]];
            uint32_t a = iv_lo, b = iv_hi, c = iv_lo, d = d_init;
#pragma unroll
            for (int r = 0; r < 7; r++) {
                LIG_B3_G(a, b, c, d, mw[4 * r], mw[4 * r + 1])
                b = qperm<0x39>(b);
                c = qperm<0x4E>(c);
                d = qperm<0x93>(d);
                LIG_B3_G(a, b, c, d, mw[4 * r + 2], mw[4 * r + 3])
                b = qperm<0x93>(b);
                c = qperm<0x4E>(c);
                d = qperm<0x39>(d);
            }
            const uint32_t lo = a ^ c, hi = b ^ d;  // words k and 4 + k of the parent
            out[8 * p + k] = lo;
            out[8 * p + 4 + k] = hi;
            uint32_t* dst = tree + (out_off + out_base + p) * 8;
            dst[k] = lo;
            dst[4 + k] = hi;
            if (root_out && cnt == 1 && lay_n == 2) {
                root_out[k] = lo;
                root_out[4 + k] = hi;
            }
        }
        LIG_LDS_BARRIER();
        lay_off = out_off;
        lay_n >>= 1;
    }
    if (flag) {
        __syncthreads();  // every wave's global stores done
        if (threadIdx.x == 0) {
            __threadfence_system();
            __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// Three levels per launch for wide layers: thread t turns nodes 8t..8t+7 of
// the layer at in_off into their 4 parents, 2 grandparents and 1
// great-grandparent (7 compressions, no idle lanes, one launch instead of
// three).  n_out = number of great-grandparents = n_in / 8.
__global__ __launch_bounds__(256) void k_merkle_block3(uint32_t* __restrict__ tree, size_t in_off, size_t n_in,
                                                       size_t tree_stride) {
    tree += blockIdx.z * tree_stride;  // one tree per blockIdx.z
    const size_t n_out = n_in / 8;
    const size_t l1 = in_off + n_in, l2 = l1 + n_in / 2, l3 = l2 + n_in / 4;  // first node of each layer
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n_out; t += (size_t)gridDim.x * blockDim.x) {
        uint32_t c[64];  // 8 children, then reused for the upper nodes
        const uint4* src = reinterpret_cast<const uint4*>(tree + (in_off + 8 * t) * 8);
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const uint4 v = src[k];
            c[4 * k] = v.x;
            c[4 * k + 1] = v.y;
            c[4 * k + 2] = v.z;
            c[4 * k + 3] = v.w;
        }
        // level 1: parent p of children 2p, 2p+1 (words 16p..16p+15) into words 8p..8p+7
#pragma unroll
        for (int p = 0; p < 4; p++) {
            uint32_t out[8];
            b3::hash64(c + 16 * p, out);
#pragma unroll
            for (int w = 0; w < 8; w++) c[8 * p + w] = out[w];
            uint4* d = reinterpret_cast<uint4*>(tree + (l1 + 4 * t + p) * 8);
            d[0] = make_uint4(out[0], out[1], out[2], out[3]);
            d[1] = make_uint4(out[4], out[5], out[6], out[7]);
        }
#pragma unroll
        for (int p = 0; p < 2; p++) {
            uint32_t out[8];
            b3::hash64(c + 16 * p, out);
#pragma unroll
            for (int w = 0; w < 8; w++) c[8 * p + w] = out[w];
            uint4* d = reinterpret_cast<uint4*>(tree + (l2 + 2 * t + p) * 8);
            d[0] = make_uint4(out[0], out[1], out[2], out[3]);
            d[1] = make_uint4(out[4], out[5], out[6], out[7]);
        }
        uint32_t out[8];
        b3::hash64(c, out);
        uint4* d = reinterpret_cast<uint4*>(tree + (l3 + t) * 8);
        d[0] = make_uint4(out[0], out[1], out[2], out[3]);
        d[1] = make_uint4(out[4], out[5], out[6], out[7]);
    }
}

// Wide levels one launch each, one lane per parent (throughput-bound);
// once a level has fewer than MERKLE_WIDE parents, k_merkle_subtree4
// finishes the top with four lanes per parent (latency-bound).
constexpr size_t MERKLE_WIDE = (size_t)1 << 17;
// Those levels go in as few launches as the workgroup depth allows (at most
// SUB4_MAXG levels: two layers of 2^g nodes in LDS), of equal depth, the
// last one a single workgroup; a workgroup of depth g has 2^(g-1) quads
// (one per parent of its first level), at most 1024 threads.  Every level
// costs one compression latency (about 0.75 us with a quad per parent), so
// the first launch's workgroups should not be deeper than the second:
// measured on MI355X for 2^17 nodes, depths (9, 8) 22.2 us against (7, 10)
// 28.7 us; 2^16 nodes (8, 8) 19.6 against (7, 9) 21.7; 2^12 nodes (6, 6)
// 14.2 against (7, 5) 14.5.
constexpr int SUB4_MAXG = 10;

static void merkle_tops(uint32_t* tree, size_t in_off, size_t nodes, hipStream_t s, uint32_t* root_out, uint32_t* flag,
                        uint32_t seq, unsigned z, size_t tree_stride, const uint32_t* rows, size_t rows_stride);

void launch_merkle_levels(uint32_t* tree, size_t n_leaves, hipStream_t s, uint32_t* root_out, uint32_t* flag, uint32_t seq,
                          int trees, size_t tree_stride) {
    if (trees < 1 || trees > 65535 || (trees > 1 && flag)) throw std::invalid_argument("merkle levels: trees 1..65535, no flag for several");
    const unsigned z = (unsigned)trees;
    size_t in_off = 0, nodes = n_leaves;
    while (nodes / 8 >= MERKLE_WIDE) {
        hipLaunchKernelGGL(k_merkle_block3, dim3(grid_for(nodes / 8, 256), 1, z), dim3(256), 0, s, tree, in_off, nodes, tree_stride);
        in_off += nodes + nodes / 2 + nodes / 4;
        nodes /= 8;
    }
    while (nodes / 2 >= MERKLE_WIDE) {
        const size_t n_out = nodes / 2;
        hipLaunchKernelGGL(k_merkle_level, dim3((unsigned)((n_out + 255) / 256), 1, z), dim3(256), 0, s, tree, in_off, n_out,
                           tree_stride);
        in_off += nodes;
        nodes = n_out;
    }
    merkle_tops(tree, in_off, nodes, s, root_out, flag, seq, z, tree_stride, nullptr, 0);
}

// the levels from a layer of fewer than 2 * MERKLE_WIDE nodes up to the root
// (k_merkle_subtree4 launches); with rows, the layer is the matrix's rows
// (see the kernel)
static void merkle_tops(uint32_t* tree, size_t in_off, size_t nodes, hipStream_t s, uint32_t* root_out, uint32_t* flag,
                        uint32_t seq, unsigned z, size_t tree_stride, const uint32_t* rows, size_t rows_stride) {
    int lg = 0;
    while (((size_t)1 << lg) < nodes) lg++;
    const int launches = lg < 4 ? 1 : (lg + SUB4_MAXG - 1) / SUB4_MAXG < 2 ? 2 : (lg + SUB4_MAXG - 1) / SUB4_MAXG;
    for (int li = 0; nodes > 1; li++) {
        const int left = launches - li;  // launches still to make, this one included
        int lgn = 0;
        while (((size_t)1 << lgn) < nodes) lgn++;
        const int g = (lgn + left - 1) / left;  // deeper launches first
        const bool top = g == lgn;
        const size_t groups = nodes >> g;
        const size_t lds = (size_t)64 << g;  // two layers of 2^g nodes
        const unsigned thr = g >= 9 ? 1024u : g >= 5 ? (2u << g) : 64u;
        hipLaunchKernelGGL(k_merkle_subtree4, dim3((unsigned)groups, 1, z), dim3(thr), lds, s, tree, in_off, nodes, g,
                           root_out, top ? flag : nullptr, seq, tree_stride, li == 0 ? rows : nullptr, rows_stride);
        for (int l = 0; l < g; l++) {
            in_off += nodes;
            nodes >>= 1;
        }
    }
}

template <class F>
void launch_merkle_commit(const typename F::T* mat, size_t rows, size_t n, uint32_t* tree, uint32_t* scratch, hipStream_t s,
                          uint32_t* root_out, uint32_t* flag, uint32_t seq, int trees, size_t tree_stride) {
    if (trees < 1 || trees > 65535 || (trees > 1 && flag) || rows % (size_t)trees != 0)
        throw std::invalid_argument("merkle commit: trees 1..65535 splitting the rows evenly, no flag for several");
    const size_t per = rows / (size_t)trees;
    if (n * sizeof(typename F::T) == 64 && per >= 1 && per < MERKLE_WIDE && (per & (per - 1)) == 0) {
        // 64-byte rows: the leaves are the first level of the top launches
        merkle_tops(tree, (size_t)0 - 2 * per, 2 * per, s, root_out, flag, seq, (unsigned)trees, tree_stride,
                    reinterpret_cast<const uint32_t*>(mat), per * 16);
        return;
    }
    launch_leaf_hash<F>(mat, rows, n, tree, scratch, s, trees, tree_stride);
    launch_merkle_levels(tree, per, s, root_out, flag, seq, trees, tree_stride);
}

// One wave: the first extra.n lanes copy extra words, then lane 0 stores
// seq at flag (host memory, system scope) once the kernels before it on the
// stream and those copies are done.
__global__ void k_signal(uint32_t* flag, uint32_t seq, WordCopy extra) {
    if ((int)threadIdx.x < extra.n) extra.dst[threadIdx.x] = extra.src[threadIdx.x];
    if (threadIdx.x == 0 && flag) {
        __threadfence_system();
        __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

void launch_signal(uint32_t* flag, uint32_t seq, hipStream_t s, const WordCopy* extra) {
    const WordCopy e = extra ? *extra : WordCopy{};
    if (e.n > 64) throw std::invalid_argument("launch_signal: at most 64 extra words");
    hipLaunchKernelGGL(k_signal, dim3(1), dim3(64), 0, s, flag, seq, e);
}

__global__ void k_copy_words(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

void launch_copy_words(uint32_t* dst, const uint32_t* src, size_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_copy_words, dim3(grid_for(n, 256, 1024)), dim3(256), 0, s, dst, src, n);
}

// n words of each of count blocks (block j = blockIdx.y at src + j *
// src_stride, to dst + j * dst_stride).
__global__ void k_copy_blocks(uint32_t* __restrict__ dst, size_t dst_stride, const uint32_t* __restrict__ src,
                              size_t src_stride, size_t n) {
    dst += blockIdx.y * dst_stride;
    src += blockIdx.y * src_stride;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

void launch_copy_blocks(uint32_t* dst, size_t dst_stride, const uint32_t* src, size_t src_stride, size_t n, int count,
                        hipStream_t s) {
    if (!n || count < 1) return;
    if (count > 65535) throw std::invalid_argument("launch_copy_blocks: at most 65535 blocks");
    hipLaunchKernelGGL(k_copy_blocks, dim3(grid_for(n, 256, 1024), (unsigned)count), dim3(256), 0, s, dst, dst_stride, src,
                       src_stride, n);
}

__global__ void k_gather_nodes(const uint32_t* __restrict__ tree, const uint64_t* __restrict__ idx, size_t count,
                               uint32_t* __restrict__ out, WordCopy extra) {
    if (blockIdx.x == 0 && (int)threadIdx.x < extra.n) extra.dst[threadIdx.x] = extra.src[threadIdx.x];
    size_t total = count * 8;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x)
        out[i] = tree[idx[i / 8] * 8 + (i % 8)];
}

void launch_gather_nodes(const uint32_t* tree, const uint64_t* node_idx, size_t count, uint32_t* out, hipStream_t s,
                         const WordCopy* extra) {
    const WordCopy e = extra ? *extra : WordCopy{};
    if (e.n > 256) throw std::invalid_argument("launch_gather_nodes: at most 256 extra words");
    if (!count && !e.n) return;
    hipLaunchKernelGGL(k_gather_nodes, dim3(count ? grid_for(count * 8, 256) : 1u), dim3(256), 0, s, tree, node_idx, count,
                       out, e);
}

// ------------------------------------------------------------------ fold ---
template <class F>
__global__ void k_fold(const typename F::T* __restrict__ poly, size_t m, int kk, const typename F::T* __restrict__ gr_tabs,
                       typename F::T* __restrict__ out) {
    using T = typename F::T;
    const int nc = 1 << kk;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x) {
        T acc = Ops<F>::zero();
        for (int j = 0; j < nc; j++) acc ^= Ops<F>::mul_tab(gr_tabs + (size_t)j * F::TAB, poly[(size_t)j * m + i]);
        out[i] = acc;
    }
}

// R consecutive output rows per thread (R = 2: one 8-byte load per column,
// 512 bytes per wave and load instruction).  The thread's 2^kk column values
// are loaded 16 / R columns at a time with buffer loads (lane offset in one
// register, column offset in a scalar register), the next columns issued
// before the current ones are multiplied; the first are in flight while the
// workgroup builds its 2^kk multiply-by-weight tables in LDS from the
// weights (a kernel argument, copied to LDS by one lane).
struct FoldWeights32 {
    uint32_t w[1 << FOLD32_MAXK];
};
template <int R>
__global__ __launch_bounds__(256) void k_fold32(const uint32_t* __restrict__ poly, size_t m, int kk, FoldWeights32 gr,
                                                 uint64_t* __restrict__ zero, int n_zero,
                                                 uint32_t* __restrict__ out, FoldJobs jobs) {
    static_assert(R == 1 || R == 2, "one or two rows per thread");
    constexpr int NC = 1 << FOLD32_MAXK, BATCH = 16 / R;
    __shared__ __attribute__((aligned(256))) uint32_t tabs[NC * 128];
    __shared__ uint32_t wsh[NC];
    // several proves at once: job blockIdx.y folds its own input with its
    // own weights into its own output
    const uint32_t job = blockIdx.y;
    if (jobs.polys) poly = jobs.polys[job];
    out += job * jobs.out_stride;
    if (zero) zero += job * jobs.zero_stride;
    if (blockIdx.x == 0 && (int)threadIdx.x < n_zero) zero[threadIdx.x] = 0ull;
    const int nc = 1 << kk;
    const uint32_t col_bytes = (uint32_t)(m * 4);
    const __amdgpu_buffer_rsrc_t rsrc = buf_rsrc(poly, col_bytes * (uint32_t)nc);
    const size_t i = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) * R;
    const uint32_t roff = (uint32_t)(i * 4);  // rows past m read other columns or 0; never stored
    auto ld = [&](int col, uint32_t (&dst)[R]) {
        if constexpr (R == 1) {
            dst[0] = buf_ld(rsrc, roff, (uint32_t)col * col_bytes);
        } else {
            const auto x = __builtin_amdgcn_raw_buffer_load_b64(rsrc, (int)roff, (int)((uint32_t)col * col_bytes), 0);
            dst[0] = x[0];
            dst[1] = x[1];
        }
    };
    uint32_t v[BATCH][R];
#pragma unroll
    for (int t = 0; t < BATCH; t++) {
        if (t < nc) {
            ld(t, v[t]);
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) v[t][r] = 0u;
        }
    }
    if (jobs.w) {
        if ((int)threadIdx.x < NC) wsh[threadIdx.x] = jobs.w[(size_t)job * NC + threadIdx.x];
    } else if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NC; j++) wsh[j] = gr.w[j];
    }
    LIG_LDS_BARRIER();
    for (int idx = threadIdx.x; idx < nc * 8; idx += blockDim.x) gf32_tab_build<true>(tabs + (idx >> 3) * 128, wsh[idx >> 3], idx & 7);
    LIG_LDS_BARRIER();
    uint32_t acc[R];
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = 0u;
    for (int j0 = 0; j0 < nc; j0 += BATCH) {
        uint32_t w[BATCH][R];
#pragma unroll
        for (int t = 0; t < BATCH; t++)
#pragma unroll
            for (int r = 0; r < R; r++) w[t][r] = v[t][r];
        if (j0 + BATCH < nc) {
#pragma unroll
            for (int t = 0; t < BATCH; t++) {
                if (j0 + BATCH + t < nc) {
                    ld(j0 + BATCH + t, v[t]);
                } else {
#pragma unroll
                    for (int r = 0; r < R; r++) v[t][r] = 0u;
                }
            }
        }
#pragma unroll
        for (int t = 0; t < BATCH; t++)
            if (j0 + t < nc) {
#pragma unroll
                for (int r = 0; r < R; r++) acc[r] ^= mul_tab_lds(tabs, (uint32_t)((j0 + t) * 512), w[t][r]);
            }
    }
#pragma unroll
    for (int r = 0; r < R; r++)
        if (i + r < m) out[i + r] = acc[r];
}

bool launch_fold32_weights(const uint32_t* poly, size_t m, int kk, const uint32_t* gr_host, uint32_t* out, hipStream_t s,
                           uint64_t* zero, int n_zero) {
    if (kk < 0 || kk > FOLD32_MAXK || (m << kk) * 4 >= ((size_t)1 << 31) || (m + 255) / 256 > 0x7fffffffu) return false;
    if (n_zero < 0 || n_zero > 256 || (n_zero && !zero)) throw std::invalid_argument("launch_fold32_weights: zero range");
    FoldWeights32 gr{};
    for (int j = 0; j < (1 << kk); j++) gr.w[j] = gr_host[j];
    // two rows per thread when m is even (a thread's pair never straddles m)
    // and the grid still holds a workgroup per CU or more
    const FoldJobs one{};
    if (m % 2 == 0 && m >= ((size_t)1 << 17))
        hipLaunchKernelGGL(k_fold32<2>, dim3((unsigned)((m + 511) / 512)), dim3(256), 0, s, poly, m, kk, gr, zero, n_zero, out,
                           one);
    else
        hipLaunchKernelGGL(k_fold32<1>, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, poly, m, kk, gr, zero, n_zero, out,
                           one);
    return true;
}

void launch_fold32_jobs(size_t m, int kk, const FoldJobs& jobs, int count, uint32_t* out, uint64_t* zero, int n_zero,
                        hipStream_t s) {
    if (kk < 1 || kk > FOLD32_MAXK || (m << kk) * 4 >= ((size_t)1 << 31) || (m + 255) / 256 > 0x7fffffffu || !jobs.polys ||
        !jobs.w || count < 1 || count > 65535 || n_zero < 0 || n_zero > 64 || (n_zero && !zero))
        throw std::invalid_argument("launch_fold32_jobs: shape");
    const FoldWeights32 none{};
    if (m % 2 == 0 && m >= ((size_t)1 << 17))
        hipLaunchKernelGGL(k_fold32<2>, dim3((unsigned)((m + 511) / 512), (unsigned)count), dim3(256), 0, s,
                           (const uint32_t*)nullptr, m, kk, none, zero, n_zero, out, jobs);
    else
        hipLaunchKernelGGL(k_fold32<1>, dim3((unsigned)((m + 255) / 256), (unsigned)count), dim3(256), 0, s,
                           (const uint32_t*)nullptr, m, kk, none, zero, n_zero, out, jobs);
}

template <class F>
void launch_fold(const typename F::T* poly, size_t m, int kk, const typename F::T* gr_tabs, typename F::T* out, hipStream_t s) {
    hipLaunchKernelGGL(k_fold<F>, dim3(grid_for(m, 256, 4096)), dim3(256), 0, s, poly, m, kk, gr_tabs, out);
}

template <class F>
__global__ void k_gather_rows(const typename F::T* __restrict__ mat, size_t n, const uint64_t* __restrict__ rows, size_t count,
                              typename F::T* __restrict__ out, typename F::T* __restrict__ out2) {
    size_t total = count * n;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const typename F::T v = mat[rows[i / n] * n + (i % n)];
        out[i] = v;
        if (out2) out2[i] = v;
    }
}

template <class F>
void launch_gather_rows(const typename F::T* mat, size_t n, const uint64_t* rows, size_t count, typename F::T* out,
                        typename F::T* out2, hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL(k_gather_rows<F>, dim3(grid_for(count * n, 256)), dim3(256), 0, s, mat, n, rows, count, out, out2);
}

// ---------------------------------------------------------------- induce ---
// basis[X*2^clo + x] = sum_i K[i][X] * P[i][x] with
//   P[i][x] = prod_{t < clo, bit t of x} W_t(q_i),
//   K[i][X] = alpha^i * prod_{t >= clo, bit t-clo of X} W_t(q_i),
// which is the sum over queries of evaluate_scaled_basis_inplace
// (binary_fft.rs:276-301, sumcheck_polys.rs:74-160).
template <class F>
__global__ void k_induce_tensor(const typename F::T* __restrict__ cvals, const typename F::T* __restrict__ ap, int ell, int clo,
                                typename F::T* __restrict__ P, typename F::T* __restrict__ Kt) {
    using T = typename F::T;
    const int i = blockIdx.x;
    const T* c = cvals + (size_t)i * ell;
    const size_t np = (size_t)1 << clo, nk = (size_t)1 << (ell - clo);
    for (size_t x = threadIdx.x; x < np + nk; x += blockDim.x) {
        if (x < np) {
            T v = Ops<F>::one();
            for (int t = 0; t < clo; t++)
                if ((x >> t) & 1) v = Ops<F>::mul(v, c[t]);
            P[(size_t)i * np + x] = v;
        } else {
            size_t X = x - np;
            T v = ap[i];
            for (int t = 0; t < ell - clo; t++)
                if ((X >> t) & 1) v = Ops<F>::mul(v, c[clo + t]);
            Kt[(size_t)i * nk + X] = v;
        }
    }
}

template <class F>
__global__ __launch_bounds__(256) void k_induce_accum(const typename F::T* __restrict__ P, const typename F::T* __restrict__ Kt,
                                                      int nq, int ell, int clo, typename F::T* __restrict__ basis) {
    using T = typename F::T;
    __shared__ T tab[2][F::TAB];
    constexpr int PER = 4;  // x values per thread (np <= 1024)
    const size_t np = (size_t)1 << clo, nk = (size_t)1 << (ell - clo);
    const size_t X = blockIdx.x;
    T acc[PER];
#pragma unroll
    for (int r = 0; r < PER; r++) acc[r] = Ops<F>::zero();
    for (int i = 0; i < nq; i++) {
        T kv = Kt[(size_t)i * nk + X];
        for (int e = threadIdx.x; e < F::TAB; e += blockDim.x) tab[i & 1][e] = Ops<F>::table_entry(kv, e);
        __syncthreads();
        const T* Pi = P + (size_t)i * np;
#pragma unroll
        for (int r = 0; r < PER; r++) {
            size_t x = threadIdx.x + (size_t)r * blockDim.x;
            if (x < np) acc[r] ^= Ops<F>::mul_tab(tab[i & 1], Pi[x]);
        }
    }
#pragma unroll
    for (int r = 0; r < PER; r++) {
        size_t x = threadIdx.x + (size_t)r * blockDim.x;
        if (x < np) basis[X * np + x] = acc[r];
    }
}

// GF(2^32) variant: a workgroup owns one X and holds the multiply tables of
// up to ACC_QMAX queries in LDS (built per nibble position).  Lane
// (s, x) with x < 2^clo accumulates basis entry x over the queries
// i = s, s + QS, s + 2QS, ... in registers (QS = blockDim.x / 2^clo query
// slices, so narrow X rows still fill the workgroup); the QS partial sums of
// an entry are XORed through LDS and the entry is written once (no atomics,
// no zeroing).  More queries than ACC_QMAX are taken in chunks.
//
// With f, the kernel also adds the triple of eval_013_product(f, basis)
// into res (and zeroes res_clear): entry j of the basis pairs with j +- h
// (h = 2^(ell-1)), which another workgroup holds, but the triple is
// bilinear, so the holder of b0 = basis[j] (j < h) adds f0*b0 to s0 and
// a2*(2 b0) to s2, the holder of b1 = basis[j] (j >= h) adds f1*b1 to s1
// and a2*(3 b1) to s2, with a2 = 3 f1 + 2 f0 loaded from f; summed over
// both, s2 gets a2*(3 b1 + 2 b0) as in eval_013_product.
template <class F> __device__ inline void block_xor3(typename F::Acc a0, typename F::Acc a1, typename F::Acc a2,
                                                     typename F::Acc* res);
// 148 query tables (79 KB of LDS with the zero table and the partial
// sums) let two 1024-thread workgroups share a CU (eight waves per SIMD,
// hence at most 64 registers per lane).  The loop is bound by LDS
// throughput: eight table reads per product.
constexpr int ACC_QMAX = 148;
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void k_induce_accum32(
    const uint32_t* __restrict__ P, const uint32_t* __restrict__ Kt, int nq, int ell, int clo, uint32_t* __restrict__ basis,
    const uint32_t* __restrict__ f, uint64_t* res, uint64_t* res_clear, InduceJobs jobs) {
    __shared__ __attribute__((aligned(256))) uint32_t tabs[(ACC_QMAX + 1) * 128];
    __shared__ uint32_t part[1024];
    const size_t np = (size_t)1 << clo, nk = (size_t)1 << (ell - clo);
    {
        // job blockIdx.y of several proves of one shape
        const size_t jb = blockIdx.y;
        P += jb * ((size_t)nq << clo);
        Kt += jb * ((size_t)nq << (ell - clo));
        basis += jb << ell;
        if (f) f += jb * jobs.f;
        if (res) res += jb * jobs.res;
        if (res_clear) res_clear += jb * jobs.res;
    }
    const size_t X = blockIdx.x;
    const int qs = (int)(blockDim.x >> clo);  // query slices (blockDim.x = qs * 2^clo)
    // the slice is uniform over a wave (2^clo >= 64); saying so keeps the
    // loop's counts, table offsets and load offsets in scalar registers
    const int sl = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> clo));
    const size_t x = threadIdx.x & (np - 1);
    const size_t j = X * np + x, h = (nk * np) / 2;
    // the triple's f entries do not depend on the basis: loaded first
    uint32_t fj = 0, fo = 0;
    if (f && sl == 0) {
        fj = f[j];
        fo = X < nk / 2 ? f[j + h] : f[j - h];
    }
    // Lane queries i = sl + k*qs, k < cnt.  P entries are loaded AP loads
    // ahead of their multiplies (the first AP before the tables are built),
    // so the loop waits on a memory latency once, not once per few queries;
    // multiplies go four at a time with all their LDS reads in flight.  Past
    // cnt a load repeats the last valid entry (in range) and its multiply
    // uses the all-zero table after the query tables, so the loads need no
    // branches.
    constexpr int AP = 8;
    constexpr uint32_t ZTAB = ACC_QMAX * 512;  // byte offset of the zero table
    const __amdgpu_buffer_rsrc_t prs = buf_rsrc(P, (uint32_t)((size_t)nq * np * 4));
    if (threadIdx.x < 128) tabs[ACC_QMAX * 128 + threadIdx.x] = 0u;
    uint32_t acc4[4] = {0u, 0u, 0u, 0u};
    for (int q0 = 0; q0 < nq; q0 += ACC_QMAX) {
        const int qn = nq - q0 < ACC_QMAX ? nq - q0 : ACC_QMAX;
        const int cnt = sl < qn ? (qn - sl + qs - 1) / qs : 0;
        // P entry (q0 + sl + k*qs, x): wave-uniform byte offset in a scalar
        // register, the lane's x*4 in a vector one
        const uint32_t pofs = (uint32_t)(((size_t)q0 + (size_t)sl) * np * 4), pst = (uint32_t)(qs * np * 4);
        const uint32_t xo = (uint32_t)(x * 4);
        auto ld = [&](int k) { return buf_ld(prs, xo, pofs + (uint32_t)(k < cnt ? k : cnt - 1) * pst); };
        auto step = [&](const uint32_t(&cur)[AP], uint32_t(&nxt)[AP], int k0) {
#pragma unroll
            for (int e = 0; e < AP; e++) nxt[e] = ld(k0 + AP + e);
#pragma unroll
            for (int g = 0; g < AP / 4; g++) {
                if (k0 + 4 * g < cnt) {
                    uint32_t tb[4], ww[4];
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const int k = k0 + 4 * g + e;
                        tb[e] = k < cnt ? (uint32_t)((sl + k * qs) * 512) : ZTAB;
                        ww[e] = cur[4 * g + e];
                    }
                    mul_tab_lds_n<4>(tabs, tb, ww, acc4);
                }
            }
        };
        uint32_t w[AP], wn[AP];
        if (cnt > 0) {
#pragma unroll
            for (int e = 0; e < AP; e++) w[e] = ld(e);
        }
        if (q0 > 0) LIG_LDS_BARRIER();  // the previous chunk's lookups are done
        for (int idx = threadIdx.x; idx < qn * 8; idx += blockDim.x)
            gf32_tab_build(tabs + (idx >> 3) * 128, Kt[(size_t)(q0 + (idx >> 3)) * nk + X], idx & 7);
        LIG_LDS_BARRIER();
        for (int k0 = 0; k0 < cnt; k0 += 2 * AP) {
            step(w, wn, k0);
            if (k0 + AP < cnt) step(wn, w, k0 + AP);
        }
    }
    uint32_t b = (acc4[0] ^ acc4[1]) ^ (acc4[2] ^ acc4[3]);
    if (qs > 1) {
        part[threadIdx.x] = b;
        LIG_LDS_BARRIER();
        if (sl == 0)
            for (int k = 1; k < qs; k++) b ^= part[k * np + x];
    }
    if (sl == 0) basis[j] = b;
    if (!f) return;
    // this entry's share of the triple (see above); X < nk / 2 is uniform
    uint64_t s0 = 0, s1 = 0, s2 = 0;
    if (blockIdx.x == 0 && threadIdx.x < 3) res_clear[threadIdx.x] = 0;
    if (sl == 0) {
        if (X < nk / 2) {
            const uint32_t f1 = fo;
            const uint32_t a2 = Ops<F32>::mul2(f1) ^ f1 ^ Ops<F32>::mul2(fj);
            Ops<F32>::acc_mul(s0, fj, b);
            Ops<F32>::acc_mul(s2, a2, Ops<F32>::mul2(b));
        } else {
            const uint32_t f0 = fo;
            const uint32_t a2 = Ops<F32>::mul2(fj) ^ fj ^ Ops<F32>::mul2(f0);
            Ops<F32>::acc_mul(s1, fj, b);
            Ops<F32>::acc_mul(s2, a2, Ops<F32>::mul2(b) ^ b);
        }
    }
    block_xor3<F32>(s0, s1, s2, res);
}

// Low bits of a basis index: P holds 2^clo entries per query, K the other
// 2^(ell - clo).  Aims at 512 K rows (one accumulate workgroup each, two
// per compute unit) with 2^6..2^10 entries per row.
int induce_low_bits(int ell) {
    if (ell < 6) return ell;
    const int clo = ell - 9;
    return clo < 6 ? 6 : (clo > 10 ? 10 : clo);
}

// GF(2^32) variant: P and K by doubling, P[x + 2^t] = P[x] * c_t, with
// the per-query tables of every c_t in LDS.  With parts, workgroup i also
// writes query i's part of the enforced sum: (row_i . gr) times ap[i],
// unreduced (the parts XOR to sum_i ap[i] * (row_i . gr)).
constexpr int TENSOR_MAXK = 4096;
__device__ __forceinline__ uint32_t wave_xor32(uint32_t x);
__global__ __launch_bounds__(256) void k_induce_tensor32(const uint32_t* __restrict__ cvals, const uint32_t* __restrict__ ap,
                                                         int ell, int clo, uint32_t* __restrict__ P, uint32_t* __restrict__ Kt,
                                                         const uint32_t* __restrict__ rows, int ncols,
                                                         const uint32_t* __restrict__ gr, uint64_t* __restrict__ parts,
                                                         InduceJobs jobs) {
    __shared__ __attribute__((aligned(256))) uint32_t tabs[PREP_MAXELL * 128];
    __shared__ uint32_t pb[1024];
    __shared__ uint32_t kb[TENSOR_MAXK];
    const int i = blockIdx.x;
    {
        // job blockIdx.y of several proves of one shape (gridDim.x queries each)
        const size_t j = blockIdx.y, nq = gridDim.x;
        cvals += j * nq * ell;
        ap += j * nq;
        P += j * (nq << clo);
        Kt += j * (nq << (ell - clo));
        if (rows) rows += j * nq * ncols;
        if (gr) gr += j * jobs.gr;
        if (parts) parts += j * jobs.parts;
    }
    if (parts && threadIdx.x < 64) {
        uint32_t d = 0u;
        for (int j = threadIdx.x; j < ncols; j += 64) d ^= Ops<F32>::mul(rows[(size_t)i * ncols + j], gr[j]);
        d = wave_xor32(d);
        if (threadIdx.x == 0) {
            uint64_t acc = 0;
            Ops<F32>::acc_mul(acc, d, ap[i]);
            parts[i] = acc;
        }
    }
    const uint32_t* c = cvals + (size_t)i * ell;
    for (int idx = threadIdx.x; idx < ell * 8; idx += blockDim.x) gf32_tab_build(tabs + (idx >> 3) * 128, c[idx >> 3], idx & 7);
    if (threadIdx.x == 0) {
        pb[0] = 1u;
        kb[0] = ap[i];
    }
    __syncthreads();
    for (int t = 0; t < clo; t++) {
        for (int x = threadIdx.x; x < (1 << t); x += blockDim.x) pb[x + (1 << t)] = mul_tab_lds(tabs, (uint32_t)(t * 512), pb[x]);
        __syncthreads();
    }
    const int hk = ell - clo;
    for (int t = 0; t < hk; t++) {
        for (int x = threadIdx.x; x < (1 << t); x += blockDim.x)
            kb[x + (1 << t)] = mul_tab_lds(tabs, (uint32_t)((clo + t) * 512), kb[x]);
        __syncthreads();
    }
    const size_t np = (size_t)1 << clo, nk = (size_t)1 << hk;
    for (size_t x = threadIdx.x; x < np; x += blockDim.x) P[(size_t)i * np + x] = pb[x];
    for (size_t x = threadIdx.x; x < nk; x += blockDim.x) Kt[(size_t)i * nk + x] = kb[x];
}

template <class F>
void launch_induce_accum(const typename F::T* P, const typename F::T* K, int nq, int ell, typename F::T* basis, hipStream_t s,
                         const Eval013In<F>* ev, int count, const InduceJobs& jobs);

template <class F>
int launch_induce(const typename F::T* cvals, const typename F::T* ap, int nq, int ell, typename F::T* P, typename F::T* K,
                  typename F::T* basis, hipStream_t s, const EnforcedIn<F>* e, const Eval013In<F>* ev, int count,
                  const InduceJobs& jobs) {
    const int clo = induce_low_bits(ell);
    size_t nk = (size_t)1 << (ell - clo);
    int n_parts = 0;
    bool fused = false;
    if (count < 1 || count > 65535) throw std::invalid_argument("launch_induce: 1..65535 jobs");
    if constexpr (F::BITS == 32) {
        if (ell <= PREP_MAXELL && nk <= (size_t)TENSOR_MAXK && clo <= 10 && (!e || nq <= ENFORCED_MAX_PARTS)) {
            hipLaunchKernelGGL(k_induce_tensor32, dim3(nq, (unsigned)count), dim3(256), 0, s, cvals, ap, ell, clo, P, K,
                               e ? e->rows : nullptr, e ? e->ncols : 0, e ? e->gr : nullptr, e ? e->parts : nullptr, jobs);
            fused = true;
            if (e) n_parts = nq;
        }
    }
    if (!fused) {
        if (count != 1) throw std::invalid_argument("launch_induce: several jobs need the GF(2^32) kernels");
        if (e) {
            launch_enforced<F>(e->rows, nq, e->ncols, e->gr, ap, e->parts, s);
            n_parts = enforced_parts(nq, e->ncols);
        }
        hipLaunchKernelGGL(k_induce_tensor<F>, dim3(nq), dim3(256), 0, s, cvals, ap, ell, clo, P, K);
    }
    launch_induce_accum<F>(P, K, nq, ell, basis, s, ev, count, jobs);
    return n_parts;
}

// The basis from the tensors (and its 0/1/3 triple with ev->f).
template <class F>
void launch_induce_accum(const typename F::T* P, const typename F::T* K, int nq, int ell, typename F::T* basis, hipStream_t s,
                         const Eval013In<F>* ev, int count, const InduceJobs& jobs) {
    const int clo = induce_low_bits(ell);
    const size_t nk = (size_t)1 << (ell - clo);
    if constexpr (F::BITS == 32) {
        // 1024 lanes per X row: 2^clo entries times 1024 / 2^clo query slices
        // the triple pairs rows X and X + nk/2; P is addressed with 32-bit offsets
        if (clo >= 6 && clo <= 10 && (!ev || nk >= 2) && ((size_t)nq << (clo + 2)) < ((size_t)1 << 31)) {
            hipLaunchKernelGGL(k_induce_accum32, dim3((unsigned)nk, (unsigned)count), dim3(1024), 0, s, P, K, nq, ell, clo, basis,
                               ev ? ev->f : nullptr, ev ? ev->res : nullptr, ev ? ev->res_clear : nullptr, jobs);
            return;
        }
    }
    if (count != 1) throw std::invalid_argument("launch_induce: several jobs need the GF(2^32) kernels");
    hipLaunchKernelGGL(k_induce_accum<F>, dim3((unsigned)nk), dim3(256), 0, s, P, K, nq, ell, clo, basis);
    if (ev) launch_eval013<F>(ev->f, basis, (size_t)1 << ell, ev->res, ev->res_clear, s);
}

template <class F>
__global__ void k_induce_prep(const uint64_t* __restrict__ q, const typename F::T* __restrict__ sks,
                              const typename F::T* __restrict__ alpha, int nq, int ell, typename F::T* __restrict__ cvals,
                              typename F::T* __restrict__ ap) {
    using T = typename F::T;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    T x = Ops<F>::from_u64(q[i]);
    T* c = cvals + (size_t)i * ell;
    c[0] = x;
    for (int t = 1; t < ell; t++) {
        x = Ops<F>::mul(x, x) ^ Ops<F>::mul(sks[t - 1], x);  // next_s (binary_fft.rs:100-107)
        c[t] = x;
    }
    // alpha^i by square-and-multiply
    T r = Ops<F>::one(), b = alpha[0];
    for (int e = i; e; e >>= 1) {
        if (e & 1) r = Ops<F>::mul(r, b);
        b = Ops<F>::mul(b, b);
    }
    ap[i] = r;
}

// GF(2^32) variant, one workgroup: squaring and multiplication by each
// sks[t] and each alpha^(2^k) are table multiplies (squaring is
// GF(2)-linear).  The inputs sit in host memory (query indices, sks, alpha);
// every thread issues its loads first, so the kernel waits for one round
// trip.  With at most 192 queries the fourth wave builds the alpha tables
// and the powers alpha^i while the first three run the next_s chains.
// Workgroups 1.. (if any) gather the opened rows instead: row q[i] of mat
// (n per row) to out and, if given, out2.
__global__ __launch_bounds__(256) void k_induce_prep32(const uint64_t* __restrict__ q, const uint32_t* __restrict__ sks,
                                                       const uint32_t* __restrict__ alpha, int nq, int ell,
                                                       uint32_t* __restrict__ cvals, uint32_t* __restrict__ ap,
                                                       const uint32_t* __restrict__ mat, size_t n, uint32_t* __restrict__ out,
                                                       uint32_t* __restrict__ out2, InduceJobs jobs) {
    // table 0: squaring; 1..ell-1: times sks[t-1]; ell..ell+7: times alpha^(2^k)
    __shared__ __attribute__((aligned(256))) uint32_t tabs[(PREP_MAXELL + 8) * 128];
    const int tid = threadIdx.x;
    {
        // job blockIdx.y of several proves of one shape (shared sks)
        const size_t j = blockIdx.y;
        q += j * jobs.q;
        alpha += j * jobs.alpha;
        cvals += j * (size_t)nq * ell;
        ap += j * (size_t)nq;
        if (mat) mat += j * jobs.mat;
        if (out) out += j * (size_t)nq * n;
        if (out2) out2 += j * jobs.rows2;
    }
    if (blockIdx.x > 0) {
        const size_t total = (size_t)nq * n, stride = (size_t)(gridDim.x - 1) * blockDim.x;
        for (size_t i = (size_t)(blockIdx.x - 1) * blockDim.x + tid; i < total; i += stride) {
            const uint32_t v = mat[q[i / n] * n + (i % n)];
            out[i] = v;
            if (out2) out2[i] = v;
        }
        return;
    }
    const int tbl = tid >> 3, k = tid & 7;  // ell * 8 <= 208 work items
    uint32_t lam = 0u;
    if (tid >= 8 && tid < ell * 8) lam = sks[tbl - 1];
    const uint32_t a0 = alpha[0];
    uint32_t x = tid < nq ? (uint32_t)q[tid] : 0u;
    if (tid < ell * 8) {
        if (tbl == 0)
            gf32_tab_build_sqr(tabs, k);
        else
            gf32_tab_build(tabs + tbl * 128, lam, k);
    }
    __syncthreads();
    // lane j of 64: table ell + j/8, nibble position j%8, of alpha^(2^(j/8))
    auto alpha_tables = [&](int j) {
        uint32_t a = a0;
        for (int t = 0; t < (j >> 3); t++) a = mul_tab_lds(tabs, 0u, a);  // squaring
        gf32_tab_build(tabs + (ell + (j >> 3)) * 128, a, j & 7);
    };
    auto alpha_power = [&](int i) {
        uint32_t r = 1u;
        for (int b = 0; b < 8; b++)
            if ((i >> b) & 1) r = mul_tab_lds(tabs, (uint32_t)((ell + b) * 512), r);
        ap[i] = r;
    };
    const bool spare = nq <= 192;
    if (spare && tid >= 192) {
        alpha_tables(tid - 192);
        // only this wave reads the alpha tables: its own LDS writes done
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_waitcnt(0xC07F);
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
        for (int i = tid - 192; i < nq; i += 64) alpha_power(i);
    }
    if (tid < nq) {
        uint32_t* c = cvals + (size_t)tid * ell;
        c[0] = x;
        for (int t = 1; t < ell; t++) {
            x = mul_tab_lds(tabs, 0u, x) ^ mul_tab_lds(tabs, (uint32_t)(t * 512), x);  // next_s
            c[t] = x;
        }
    }
    if (!spare) {
        __syncthreads();
        if (tid < 64) alpha_tables(tid);
        __syncthreads();
        if (tid < nq) alpha_power(tid);
    }
}

template <class F>
void launch_induce_prep(const uint64_t* q, const typename F::T* sks, const typename F::T* alpha, int nq, int ell,
                        typename F::T* cvals, typename F::T* ap, hipStream_t s, const RowGather<F>* g, int count,
                        const InduceJobs& jobs) {
    if (count < 1 || count > 65535) throw std::invalid_argument("launch_induce_prep: 1..65535 jobs");
    if constexpr (F::BITS == 32) {
        if (nq <= 256 && ell <= PREP_MAXELL) {
            // workgroup 0 the constants, the rest the row gather
            const unsigned ng = g && nq ? (unsigned)grid_for((size_t)nq * g->n, 256) : 0u;
            hipLaunchKernelGGL(k_induce_prep32, dim3(1 + ng, (unsigned)count), dim3(256), 0, s, q, sks, alpha, nq, ell, cvals,
                               ap, g ? g->mat : nullptr, g ? g->n : (size_t)0, g ? g->out : nullptr, g ? g->out2 : nullptr,
                               jobs);
            return;
        }
    }
    if (count != 1) throw std::invalid_argument("launch_induce_prep: several jobs need the GF(2^32) kernel");
    if (g) launch_gather_rows<F>(g->mat, g->n, q, (size_t)nq, g->out, g->out2, s);
    hipLaunchKernelGGL(k_induce_prep<F>, dim3((nq + 63) / 64), dim3(64), 0, s, q, sks, alpha, nq, ell, cvals, ap);
}

// GF(2^32): the query constants and the tensors of one query per
// workgroup (k_induce_prep32 and k_induce_tensor32 in one launch): the
// workgroup reads its query index, sks and alpha (host memory), builds the
// squaring, sks and alpha^(2^b) tables, runs the query's next_s chain on
// wave 0 and alpha^i on wave 1 while wave 2 gathers the opened row (to out
// and out2) and its dot product with gr; then the part of the enforced sum,
// the tables of the chain's values and the P / K tensors by doubling.
__global__ __launch_bounds__(256) void k_induce_pt32(const uint64_t* __restrict__ q, const uint32_t* __restrict__ sks,
                                                     const uint32_t* __restrict__ alpha, int ell, int clo,
                                                     const uint32_t* __restrict__ mat, int n, uint32_t* __restrict__ out,
                                                     uint32_t* __restrict__ out2, const uint32_t* __restrict__ gr,
                                                     uint64_t* __restrict__ parts, uint32_t* __restrict__ P,
                                                     uint32_t* __restrict__ Kt, InduceJobs jobs) {
    // tables 0: squaring, 1..ell-1: times sks[t-1], ell..ell+7: times
    // alpha^(2^b); later 0..ell-1: times the chain's values c_t
    __shared__ __attribute__((aligned(256))) uint32_t tabs[(PREP_MAXELL + 8) * 128];
    __shared__ uint32_t cbuf[PREP_MAXELL];
    __shared__ uint32_t ap_sh;
    __shared__ uint32_t pb[1024];
    __shared__ uint32_t kb[TENSOR_MAXK];
    const int i = blockIdx.x, tid = threadIdx.x, nq = (int)gridDim.x;
    {
        const size_t j = blockIdx.y;
        q += j * jobs.q;
        alpha += j * jobs.alpha;
        mat += j * jobs.mat;
        out += j * (size_t)nq * n;
        out2 += j * jobs.rows2;
        gr += j * jobs.gr;
        parts += j * jobs.parts;
        P += j * ((size_t)nq << clo);
        Kt += j * ((size_t)nq << (ell - clo));
    }
    const uint32_t qi = (uint32_t)q[i];
    const uint32_t a0 = alpha[0];
    for (int task = tid; task < (ell + 8) * 8; task += blockDim.x) {
        const int tbl = task >> 3, k = task & 7;
        if (tbl == 0) {
            gf32_tab_build_sqr(tabs, k);
        } else if (tbl < ell) {
            gf32_tab_build(tabs + tbl * 128, sks[tbl - 1], k);
        } else {
            uint32_t a = a0;
            for (int b = 0; b < tbl - ell; b++) a = Ops<F32>::mul(a, a);
            gf32_tab_build(tabs + tbl * 128, a, k);
        }
    }
    // wave 2: the opened row (query index qi) and its dot product with gr
    uint32_t dot = 0u;
    if (tid >= 128 && tid < 192) {
        for (int c = tid - 128; c < n; c += 64) {
            const uint32_t v = mat[(size_t)qi * n + c];
            out[(size_t)i * n + c] = v;
            if (out2) out2[(size_t)i * n + c] = v;
            if (parts) dot ^= Ops<F32>::mul(v, gr[c]);
        }
        dot = wave_xor32(dot);
    }
    __syncthreads();
    if (tid < 64) {
        // next_s (binary_fft.rs:100-107): c_t = c_{t-1}^2 + sks[t-1] * c_{t-1}
        uint32_t x = qi;
        if (tid == 0) cbuf[0] = x;
        for (int t = 1; t < ell; t++) {
            x = mul_tab_lds(tabs, 0u, x) ^ mul_tab_lds(tabs, (uint32_t)(t * 512), x);
            if (tid == 0) cbuf[t] = x;
        }
    } else if (tid < 128) {
        uint32_t r = 1u;
        for (int b = 0; b < 8; b++)
            if ((i >> b) & 1) r = mul_tab_lds(tabs, (uint32_t)((ell + b) * 512), r);
        if (tid == 64) ap_sh = r;
    }
    __syncthreads();
    const uint32_t api = ap_sh;
    if (parts && tid == 128) {
        uint64_t acc = 0;
        Ops<F32>::acc_mul(acc, dot, api);
        parts[i] = acc;
    }
    for (int task = tid; task < ell * 8; task += blockDim.x) gf32_tab_build(tabs + (task >> 3) * 128, cbuf[task >> 3], task & 7);
    if (tid == 0) {
        pb[0] = 1u;
        kb[0] = api;
    }
    __syncthreads();
    for (int t = 0; t < clo; t++) {
        for (int x = tid; x < (1 << t); x += blockDim.x) pb[x + (1 << t)] = mul_tab_lds(tabs, (uint32_t)(t * 512), pb[x]);
        __syncthreads();
    }
    const int hk = ell - clo;
    for (int t = 0; t < hk; t++) {
        for (int x = tid; x < (1 << t); x += blockDim.x) kb[x + (1 << t)] = mul_tab_lds(tabs, (uint32_t)((clo + t) * 512), kb[x]);
        __syncthreads();
    }
    const size_t np = (size_t)1 << clo, nk = (size_t)1 << hk;
    for (size_t x = tid; x < np; x += blockDim.x) P[(size_t)i * np + x] = pb[x];
    for (size_t x = tid; x < nk; x += blockDim.x) Kt[(size_t)i * nk + x] = kb[x];
}

// The whole induction of one step (sumcheck_polys.rs:74-160): query
// constants, opened rows, tensors, basis and its triple with f.
template <class F>
int launch_induce_all(const uint64_t* q, const typename F::T* sks, const typename F::T* alpha, int nq, int ell,
                      typename F::T* cvals, typename F::T* ap, const RowGather<F>& g, typename F::T* P, typename F::T* K,
                      typename F::T* basis, hipStream_t s, const EnforcedIn<F>& e, const Eval013In<F>* ev, int count,
                      const InduceJobs& jobs) {
    if constexpr (F::BITS == 32) {
        const int clo = induce_low_bits(ell);
        const size_t nk = (size_t)1 << (ell - clo);
        if (nq >= 1 && nq <= 256 && nq <= ENFORCED_MAX_PARTS && ell <= PREP_MAXELL && nk <= (size_t)TENSOR_MAXK &&
            clo <= 10 && g.n >= 1 && g.n <= 4096 && e.rows == g.out && e.ncols == (int)g.n && count >= 1 && count <= 65535) {
            hipLaunchKernelGGL(k_induce_pt32, dim3((unsigned)nq, (unsigned)count), dim3(256), 0, s, q, sks, alpha, ell, clo,
                               g.mat, (int)g.n, g.out, g.out2, e.gr, e.parts, P, K, jobs);
            launch_induce_accum<F>(P, K, nq, ell, basis, s, ev, count, jobs);
            return nq;
        }
    }
    launch_induce_prep<F>(q, sks, alpha, nq, ell, cvals, ap, s, &g, count, jobs);
    return launch_induce<F>(cvals, ap, nq, ell, P, K, basis, s, &e, ev, count, jobs);
}

// Partial enforced sums: workgroup g XOR-accumulates the products of its
// share of the (row, column) pairs and writes parts[g]; the transcript step
// that checks the sum adds the parts.
template <class F>
__global__ __launch_bounds__(256) void k_enforced(const typename F::T* __restrict__ rows, int nrows, int ncols,
                                                  const typename F::T* __restrict__ gr, const typename F::T* __restrict__ ap,
                                                  typename F::Acc* __restrict__ parts) {
    using T = typename F::T;
    using A = typename F::Acc;
    __shared__ A red[256];
    const int lg = 31 - __clz(ncols);
    A acc = Ops<F>::acc_zero();
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nrows * ncols; p += gridDim.x * blockDim.x) {
        const T prod = Ops<F>::mul(rows[p], gr[p & (ncols - 1)]);
        Ops<F>::acc_mul(acc, prod, ap[p >> lg]);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] ^= red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = red[0];
}

int enforced_parts(int nrows, int ncols) {
    const int g = (nrows * ncols + 255) / 256;
    return g < 1 ? 1 : (g > ENFORCED_MAX_PARTS ? ENFORCED_MAX_PARTS : g);
}

template <class F>
void launch_enforced(const typename F::T* rows, int nrows, int ncols, const typename F::T* gr, const typename F::T* ap,
                     typename F::Acc* parts, hipStream_t s) {
    hipLaunchKernelGGL(k_enforced<F>, dim3(enforced_parts(nrows, ncols)), dim3(256), 0, s, rows, nrows, ncols, gr, ap, parts);
}

// -------------------------------------------------------------- sumcheck ---
// XOR over the 64 lanes of a wave (every lane gets the total; all 64 lanes
// active): DPP within each row of 16 (swap neighbours, swap pairs, mirror
// halves of 8, mirror the row), then the four row totals read from lanes
// 0, 16, 32 and 48.
__device__ __forceinline__ uint32_t wave_xor32(uint32_t x) {
    x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, true);   // quad_perm [1,0,3,2]
    x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, true);   // quad_perm [2,3,0,1]
    x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x141, 0xF, 0xF, true);  // row_half_mirror
    x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x140, 0xF, 0xF, true);  // row_mirror
    return (uint32_t)(__builtin_amdgcn_readlane((int)x, 0) ^ __builtin_amdgcn_readlane((int)x, 16) ^
                      __builtin_amdgcn_readlane((int)x, 32) ^ __builtin_amdgcn_readlane((int)x, 48));
}
__device__ __forceinline__ uint64_t wave_xor(uint64_t x) {
    return (uint64_t)wave_xor32((uint32_t)x) | ((uint64_t)wave_xor32((uint32_t)(x >> 32)) << 32);
}
__device__ __forceinline__ u128v wave_xor(u128v x) { return {wave_xor(x.lo), wave_xor(x.hi)}; }

template <class F> __device__ inline void acc_atomic_xor(typename F::Acc* dst, typename F::Acc v);
template <> __device__ inline void acc_atomic_xor<F32>(uint64_t* dst, uint64_t v) {
    if (v) atomicXor(reinterpret_cast<unsigned long long*>(dst), (unsigned long long)v);
}
template <> __device__ inline void acc_atomic_xor<F128>(u128v* dst, u128v v) {
    if (v.lo) atomicXor(reinterpret_cast<unsigned long long*>(&dst->lo), (unsigned long long)v.lo);
    if (v.hi) atomicXor(reinterpret_cast<unsigned long long*>(&dst->hi), (unsigned long long)v.hi);
}

// XOR-reduces three accumulators across the workgroup and folds them into res.
// XOR of three accumulators over the workgroup (a multiple of 64 threads,
// at most 1024), added into res with atomics: wave shuffles, then one
// partial per wave through LDS.
template <class F> __device__ inline void block_xor3(typename F::Acc a0, typename F::Acc a1, typename F::Acc a2, typename F::Acc* res) {
    using A = typename F::Acc;
    __shared__ A red[3][16];
    a0 = wave_xor(a0);
    a1 = wave_xor(a1);
    a2 = wave_xor(a2);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = a0;
        red[1][w] = a1;
        red[2][w] = a2;
    }
    LIG_LDS_BARRIER();
    if (threadIdx.x < 3) {
        A t = red[threadIdx.x][0];
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) t ^= red[threadIdx.x][i];
        acc_atomic_xor<F>(&res[threadIdx.x], t);
    }
}

template <class F> __global__ void k_zero_acc(typename F::Acc* res, int count) {
    using A = typename F::Acc;
    if ((int)threadIdx.x < count) res[threadIdx.x] = A{};
}

template <class F>
__global__ __launch_bounds__(256) void k_eval013(const typename F::T* __restrict__ f, const typename F::T* __restrict__ g,
                                                 size_t len, typename F::Acc* res, typename F::Acc* res_clear) {
    using T = typename F::T;
    using A = typename F::Acc;
    if (blockIdx.x == 0 && threadIdx.x < 3) res_clear[threadIdx.x] = A{};
    const size_t h = len / 2;
    A s0 = Ops<F>::acc_zero(), s1 = Ops<F>::acc_zero(), s2 = Ops<F>::acc_zero();
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < h; i += (size_t)gridDim.x * blockDim.x) {
        T a0 = f[i], a1 = f[i + h], b0 = g[i], b1 = g[i + h];
        Ops<F>::acc_mul(s0, a0, b0);
        Ops<F>::acc_mul(s1, a1, b1);
        T a2 = Ops<F>::mul2(a1) ^ a1 ^ Ops<F>::mul2(a0);  // 3*f1 + (1+3)*f0
        T b2 = Ops<F>::mul2(b1) ^ b1 ^ Ops<F>::mul2(b0);
        Ops<F>::acc_mul(s2, a2, b2);
    }
    block_xor3<F>(s0, s1, s2, res);
}

template <class F>
void launch_eval013(const typename F::T* f, const typename F::T* g, size_t len, typename F::Acc* res,
                    typename F::Acc* res_clear, hipStream_t s) {
    hipLaunchKernelGGL(k_eval013<F>, dim3(grid_for(len / 2, 256, 1024)), dim3(256), 0, s, f, g, len, res, res_clear);
}

template <class F> void launch_zero_acc(typename F::Acc* res, int count, hipStream_t s) {
    if (count < 1 || count > 64) throw std::invalid_argument("launch_zero_acc: 1..64 accumulators");
    hipLaunchKernelGGL(k_zero_acc<F>, dim3(1), dim3(64), 0, s, res, count);
}

template <class F, bool STEP>
__global__ __launch_bounds__(256) void k_sumcheck_fold(const typename F::T* __restrict__ f, const typename F::T* __restrict__ b,
                                                       size_t len, const typename F::T* __restrict__ rp, FoldStep<F> step,
                                                       typename F::T* __restrict__ f2, typename F::T* __restrict__ b2,
                                                       typename F::Acc* res, typename F::Acc* res_clear, SumcheckJobs jobs) {
    using T = typename F::T;
    using A = typename F::Acc;
    {
        // job blockIdx.y of several proves of one shape: its own polys,
        // accumulators, transcript state and logs
        const size_t j = blockIdx.y;
        f += j * jobs.f;
        b += j * jobs.b;
        f2 += j * jobs.f2;
        b2 += j * jobs.b2;
        res += j * jobs.acc;
        res_clear += j * jobs.acc;
        if constexpr (STEP) {
            if (step.st_in) step.st_in += j * jobs.st;
            step.st_out += j * jobs.st;
            step.res_in += j * jobs.acc;
            if (step.enf) step.enf += j * jobs.acc;
            if (step.err) step.err += j * jobs.log_words;
            if (step.bnew) step.bnew += j * jobs.bnew;
            step.tlog += j * jobs.log_elems;
            step.rlog += j * jobs.log_elems;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < 3) res_clear[threadIdx.x] = A{};
    const size_t L2 = len / 2, h2 = L2 / 2;
    const size_t i0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
    // the first element group is loaded while the challenge (and its table)
    // is produced
    T fa{}, fb{}, fc{}, fd{}, ba{}, bb{}, bc{}, bd{};
    if (i0 < h2) {
        fa = f[i0], fb = f[i0 + L2], fc = f[i0 + h2], fd = f[i0 + h2 + L2];
        ba = b[i0], bb = b[i0 + L2], bc = b[i0 + h2], bd = b[i0 + h2 + L2];
    }
    __shared__ T tab[F::TAB];
    __shared__ T tab_beta[STEP ? F::TAB : 1];
    T r;
    bool glue = false;
    if constexpr (STEP) {
        // the preceding transcript step, computed by wave 0 of every workgroup
        __shared__ T r_sh[2];
        if (step.lead && threadIdx.x >= 64 && threadIdx.x < 128) {
            // a section's leading step: wave 1, idle until the challenge is
            // out, runs one compression on dummy words first, so the shared
            // compression code is in this CU's instruction cache by the time
            // wave 0 (still loading the state and triple) reaches it
            const dfs::Sha8 d = dfs::compress_shared(threadIdx.x, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                                      0, 0, 0);
            asm volatile("" ::"v"(d.v[0]));
        }
        if (step.enf && blockIdx.x == 0 && threadIdx.x >= 128 && threadIdx.x < 192) {
            // the enforced-sum check (ligerito.rs:35,52) on wave 2 of the
            // first workgroup, off the transcript's path: its loads would
            // otherwise hold up the compressions of every workgroup
            A e = Ops<F>::acc_zero();
            for (int i = threadIdx.x - 128; i < step.n_enf; i += 64) e ^= step.enf[i];
            e = wave_xor(e);
            const T t0 = Ops<F>::acc_reduce(dfs::uniform_acc(step.res_in[0]));
            const T t1 = Ops<F>::acc_reduce(dfs::uniform_acc(step.res_in[1]));
            if (threadIdx.x == 128 && !Ops<F>::eq(t0 ^ t1, Ops<F>::acc_reduce(dfs::uniform_acc(e)))) *step.err = 1;
        }
        if (threadIdx.x < 64) {
            dfs::State st;
            if (step.st_in)
                st.load(step.st_in);
            else
                st.load_value(step.init);
            T t[3];
#pragma unroll
            for (int k = 0; k < 3; k++) t[k] = Ops<F>::acc_reduce(dfs::uniform_acc(step.res_in[k]));
#pragma unroll
            for (int k = 0; k < 3; k++)
#pragma unroll
                for (int i = 0; i < Ops<F>::WORDS; i++) st.word_le(Ops<F>::word(t[k], i));
            uint32_t o[4];
            st.squeeze(o);
            const T rv0 = Ops<F>::from_squeeze(o);
            T rv1 = rv0;
            if (step.n_squeeze == 2) {
                st.squeeze(o);
                rv1 = Ops<F>::from_squeeze(o);
            }
            if (threadIdx.x == 0) {
                r_sh[0] = rv0;
                r_sh[1] = rv1;
                if (blockIdx.x == 0) {
                    st.store(step.st_out);
                    for (int k = 0; k < 3; k++) step.tlog[k] = t[k];
                    step.rlog[0] = rv0;
                    if (step.n_squeeze == 2) step.rlog[1] = rv1;
                }
            }
        }
        LIG_LDS_BARRIER();
        r = r_sh[1];
        glue = step.bnew != nullptr;
        if (glue) {
            const T beta = r_sh[0];
            for (int e = threadIdx.x; e < F::TAB; e += blockDim.x) tab_beta[e] = Ops<F>::table_entry(beta, e);
        }
    } else {
        r = *rp;
    }
    for (int e = threadIdx.x; e < F::TAB; e += blockDim.x) tab[e] = Ops<F>::table_entry(r, e);
    LIG_LDS_BARRIER();
    A s0 = Ops<F>::acc_zero(), s1 = Ops<F>::acc_zero(), s2 = Ops<F>::acc_zero();
    if (h2 == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            T b0 = b[0], b1 = b[1];
            if constexpr (STEP) {
                if (glue) {
                    b0 ^= Ops<F>::mul_tab(tab_beta, step.bnew[0]);
                    b1 ^= Ops<F>::mul_tab(tab_beta, step.bnew[1]);
                }
            }
            f2[0] = f[0] ^ Ops<F>::mul_tab(tab, f[0] ^ f[1]);
            b2[0] = b0 ^ Ops<F>::mul_tab(tab, b0 ^ b1);
        }
    } else {
        for (size_t i = i0; i < h2; i += stride) {
            if (i != i0) {
                fa = f[i], fb = f[i + L2], fc = f[i + h2], fd = f[i + h2 + L2];
                ba = b[i], bb = b[i + L2], bc = b[i + h2], bd = b[i + h2 + L2];
            }
            if constexpr (STEP) {
                if (glue) {
                    const T* bn = step.bnew;
                    ba ^= Ops<F>::mul_tab(tab_beta, bn[i]);
                    bb ^= Ops<F>::mul_tab(tab_beta, bn[i + L2]);
                    bc ^= Ops<F>::mul_tab(tab_beta, bn[i + h2]);
                    bd ^= Ops<F>::mul_tab(tab_beta, bn[i + h2 + L2]);
                }
            }
            // (1+r)*e[i] + r*e[i+L2] = e[i] + r*(e[i] + e[i+L2])  (multilinear_poly.rs:71-73)
            T f0 = fa ^ Ops<F>::mul_tab(tab, fa ^ fb);
            T f1 = fc ^ Ops<F>::mul_tab(tab, fc ^ fd);
            T g0 = ba ^ Ops<F>::mul_tab(tab, ba ^ bb);
            T g1 = bc ^ Ops<F>::mul_tab(tab, bc ^ bd);
            f2[i] = f0;
            f2[i + h2] = f1;
            b2[i] = g0;
            b2[i + h2] = g1;
            Ops<F>::acc_mul(s0, f0, g0);
            Ops<F>::acc_mul(s1, f1, g1);
            T fx = Ops<F>::mul2(f1) ^ f1 ^ Ops<F>::mul2(f0);
            T gx = Ops<F>::mul2(g1) ^ g1 ^ Ops<F>::mul2(g0);
            Ops<F>::acc_mul(s2, fx, gx);
        }
    }
    block_xor3<F>(s0, s1, s2, res);
}

template <class F>
void launch_sumcheck_fold_step(const typename F::T* f, const typename F::T* b, size_t len, const FoldStep<F>& step,
                               typename F::T* f2, typename F::T* b2, typename F::Acc* res, typename F::Acc* res_clear,
                               hipStream_t s, int count, const SumcheckJobs& jobs) {
    if (step.res_in == res || step.res_in == res_clear || step.st_in == step.st_out)
        throw std::invalid_argument("fold step: the triple or state read must not be written by the same kernel");
    if (step.n_squeeze < 1 || step.n_squeeze > 2 || (step.bnew && step.n_squeeze != 2))
        throw std::invalid_argument("fold step: one challenge, or two with glue");
    if ((!step.lead && !step.st_in) || count < 1 || count > 65535 || (count > 1 && !step.st_in))
        throw std::invalid_argument("fold step: a state to start from (per job for several jobs), 1..65535 jobs");
    size_t h2 = len / 4;
    hipLaunchKernelGGL((k_sumcheck_fold<F, true>), dim3(grid_for(h2 ? h2 : 1, 256, 1024), (unsigned)count), dim3(256), 0, s, f,
                       b, len, (const typename F::T*)nullptr, step, f2, b2, res, res_clear, jobs);
}

// -------------------------------------------------------- microbenchmarks ---
// Peak rates the rooflines are priced against.  Every lane runs NB
// independent chains; one step of a chain is one product (or one VALU
// instruction), so a rate is lanes x chains x steps over the kernel time.
// The chain values feed one conditional store, so no step can be dropped.
constexpr int BENCH_CHAINS = 8, BENCH_BLOCKS = 4096, BENCH_THREADS = 256;

__device__ __forceinline__ uint32_t bench_mix(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    return x ^ (x >> 16);
}

// GF(2^32) products by wave-uniform constants through the FFT passes'
// nibble tables in LDS (mul_tab_lds_n: 8 table reads and the address and
// XOR instructions of one product), each followed by the butterfly's XOR.
__global__ __launch_bounds__(BENCH_THREADS) void k_bench_mul_tab(uint32_t seed, int steps, uint32_t* out) {
    constexpr int NB = BENCH_CHAINS;
    __shared__ __attribute__((aligned(256))) uint32_t tabs[NB * 128];
    for (int idx = threadIdx.x; idx < NB * 8; idx += blockDim.x)
        gf32_tab_build(tabs + (idx >> 3) * 128, bench_mix(seed + (uint32_t)(idx >> 3)), idx & 7);
    __syncthreads();
    uint32_t tb[NB], w[NB], u[NB];
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int i = 0; i < NB; i++) {
        tb[i] = (uint32_t)i * 512u;
        w[i] = bench_mix(id * NB + i + seed);
        u[i] = bench_mix(w[i]);
    }
    for (int s = 0; s < steps; s++) {
        mul_tab_lds_n<NB>(tabs, tb, w, u);
#pragma unroll
        for (int i = 0; i < NB; i++) w[i] ^= u[i];
    }
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < NB; i++) x ^= u[i] ^ w[i];
    if (x == seed) out[0] = x;
}

// General GF(2^32) products (both operands per lane): the carry-less
// product from integer multiplies and its reduction (Ops<F32>::mul).
__global__ __launch_bounds__(BENCH_THREADS) void k_bench_mul_clmul(uint32_t seed, int steps, uint32_t* out) {
    constexpr int NB = BENCH_CHAINS;
    uint32_t a[NB], b[NB];
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int i = 0; i < NB; i++) {
        a[i] = bench_mix(id * NB + i + seed);
        b[i] = bench_mix(a[i] + 1u);
    }
    for (int s = 0; s < steps; s++)
#pragma unroll
        for (int i = 0; i < NB; i++) a[i] = Ops<F32>::mul(a[i], b[i]) ^ b[(i + 1) % NB];
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < NB; i++) x ^= a[i];
    if (x == seed) out[0] = x;
}

// 32-bit integer VALU: each step is one three-input bitwise instruction
// (v_bitop3_b32) per chain, majority and select functions so that no chain
// simplifies across steps.
__global__ __launch_bounds__(BENCH_THREADS) void k_bench_valu(uint32_t seed, int steps, uint32_t* out) {
    constexpr int NB = 16;
    uint32_t x[NB];
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int i = 0; i < NB; i++) x[i] = bench_mix(id * NB + i + seed);
    for (int s = 0; s < steps; s++) {
#pragma unroll
        for (int i = 0; i < NB; i += 2) {
            x[i] = __builtin_amdgcn_bitop3_b32(x[i], x[(i + 5) % NB], x[(i + 11) % NB], 0xCA);
            x[i + 1] = __builtin_amdgcn_bitop3_b32(x[i + 1], x[(i + 6) % NB], x[(i + 12) % NB], 0xE8);
        }
    }
    uint32_t r = 0;
#pragma unroll
    for (int i = 0; i < NB; i++) r ^= x[i];
    if (r == seed) out[0] = r;
}

// A formulation with fewer LDS bytes per product: 5-bit windows of w (seven
// positions, 32-entry tables of 128 bytes, 28 bytes read per product instead
// of 32) at two instructions per address (window extract, scaled add).
__global__ __launch_bounds__(BENCH_THREADS) void k_bench_mul_tab5(uint32_t seed, int steps, uint32_t* out) {
    constexpr int NB = BENCH_CHAINS, NP = 7;
    __shared__ uint32_t tabs[NB * NP * 32];
    for (int idx = threadIdx.x; idx < NB * NP; idx += blockDim.x) {
        const int t = idx / NP, k = idx % NP;
        uint32_t c[5];
        c[0] = gf32_reduce64((uint64_t)bench_mix(seed + (uint32_t)t) << (5 * k));
        for (int b = 1; b < 5; b++) c[b] = gf32_dbl(c[b - 1]);
        for (int v = 0; v < 32; v++) {
            uint32_t e = 0;
            for (int b = 0; b < 5; b++)
                if ((v >> b) & 1) e ^= c[b];
            tabs[(t * NP + k) * 32 + v] = e;
        }
    }
    __syncthreads();
    uint32_t w[NB], u[NB];
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int i = 0; i < NB; i++) {
        w[i] = bench_mix(id * NB + i + seed);
        u[i] = bench_mix(w[i]);
    }
    for (int s = 0; s < steps; s++) {
        uint32_t r[NB][NP];
#pragma unroll
        for (int i = 0; i < NB; i++)
#pragma unroll
            for (int k = 0; k < NP; k++) r[i][k] = tabs[(i * NP + k) * 32 + ((w[i] >> (5 * k)) & 31u)];
#pragma unroll
        for (int i = 0; i < NB; i++) {
            u[i] = xor3(xor3(r[i][0], r[i][1], r[i][2]), xor3(r[i][3], r[i][4], r[i][5]), r[i][6] ^ u[i]);
            w[i] ^= u[i];
        }
    }
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < NB; i++) x ^= u[i] ^ w[i];
    if (x == seed) out[0] = x;
}

// LDS read bandwidth with the table pattern of the GF(2^128) passes: 64
// lanes read 16-byte entries of one 256-byte group (ds_read_b128), 16
// distinct entries per instruction; the rate is bytes read per second.
__global__ __launch_bounds__(BENCH_THREADS) void k_bench_lds128(uint32_t seed, int steps, uint32_t* out) {
    constexpr int NB = BENCH_CHAINS;
    __shared__ uint4 t[NB * 16];
    for (int idx = threadIdx.x; idx < NB * 16; idx += blockDim.x)
        t[idx] = make_uint4(bench_mix(idx + seed), bench_mix(idx + 1000u), bench_mix(idx + 2000u), bench_mix(idx + 3000u));
    __syncthreads();
    uint32_t sel = bench_mix(blockIdx.x * blockDim.x + threadIdx.x + seed);
    uint4 acc = make_uint4(0u, 0u, 0u, 0u);
    for (int s = 0; s < steps; s++) {
        uint4 r[NB];
#pragma unroll
        for (int i = 0; i < NB; i++) r[i] = t[i * 16 + ((sel >> (4 * i)) & 15u)];
#pragma unroll
        for (int i = 0; i < NB; i++) {
            acc.x ^= r[i].x;
            acc.y ^= r[i].y;
            acc.z ^= r[i].z;
            acc.w ^= r[i].w;
        }
        sel = (sel << 1) | (acc.x & 1u);  // the next addresses depend on the reads
    }
    if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == seed) out[0] = acc.x;
}

// GF(2^32) products by a wave-uniform constant without tables: its 32
// multiples c_j = lam * x^j sit in scalar registers and each set bit j of w
// adds c_j, one sign-extended bit extract and one three-input AND-XOR per
// bit (64 VALU instructions per product, no LDS).
__global__ __launch_bounds__(BENCH_THREADS) void k_bench_mul_mask(uint32_t seed, int steps, uint32_t* out) {
    constexpr int NB = BENCH_CHAINS;
    uint32_t c[32];
    c[0] = bench_mix(seed);
#pragma unroll
    for (int j = 1; j < 32; j++) c[j] = gf32_dbl(c[j - 1]);
    uint32_t w[NB];
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int i = 0; i < NB; i++) w[i] = bench_mix(id * NB + i + seed);
    for (int s = 0; s < steps; s++) {
#pragma unroll
        for (int i = 0; i < NB; i++) {
            uint32_t u = 0;
#pragma unroll
            for (int j = 0; j < 32; j++) {
                const uint32_t m = (uint32_t)((int32_t)(w[i] << (31 - j)) >> 31);  // bit j of w, sign-extended
                u = __builtin_amdgcn_bitop3_b32(u, c[j], m, 0x78);                    // u ^ (c_j & m)
            }
            w[i] ^= u;
        }
    }
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < NB; i++) x ^= w[i];
    if (x == seed) out[0] = x;
}

// GF(2^8) products by a constant through log/exp tables in LDS (the
// subfield step of a tower representation GF(((2^8)^2)^2), whose GF(2^32)
// product takes nine of them by Karatsuba): per byte a log read, an add of
// the constant's log, an exp read (510-entry table, no reduction) and a
// zero select; four bytes per word.  The rate counts GF(2^8) products.
__global__ __launch_bounds__(BENCH_THREADS) void k_bench_gf8_logexp(uint32_t seed, int steps, uint32_t* out) {
    constexpr int NB = BENCH_CHAINS;
    __shared__ uint8_t lg[256];
    __shared__ uint8_t ex[512];
    if (threadIdx.x == 0) {
        // GF(2^8) mod x^8 + x^4 + x^3 + x^2 + 1 (0x11D), where x (= 2) is primitive
        uint32_t v = 1;
        for (int e = 0; e < 255; e++) {
            ex[e] = (uint8_t)v;
            ex[e + 255] = (uint8_t)v;
            lg[v] = (uint8_t)e;
            v <<= 1;
            if (v & 0x100u) v ^= 0x11Du;
        }
        ex[510] = ex[511] = 0;
        lg[0] = 0;
    }
    __syncthreads();
    const uint32_t lc = bench_mix(seed) % 255u;  // log of the constant
    uint32_t w[NB];
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
    for (int i = 0; i < NB; i++) w[i] = bench_mix(id * NB + i + seed);
    for (int s = 0; s < steps; s++) {
#pragma unroll
        for (int i = 0; i < NB; i++) {
            uint32_t r = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t byte = (w[i] >> (8 * k)) & 255u;
                const uint32_t p = ex[lg[byte] + lc];
                r |= (byte ? p : 0u) << (8 * k);
            }
            w[i] ^= r;
        }
    }
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < NB; i++) x ^= w[i];
    if (x == seed) out[0] = x;
}

double bench_rate(int which, uint32_t* scratch_word, hipStream_t s) {
    static const int steps_of[7] = {1024, 256, 4096, 1024, 1024, 128, 512};
    static const double per_step_of[7] = {BENCH_CHAINS, BENCH_CHAINS, 16.0, BENCH_CHAINS, BENCH_CHAINS * 16.0,
                                          BENCH_CHAINS, BENCH_CHAINS * 4.0};
    if (which < 0 || which > 6) throw std::invalid_argument("unknown microbenchmark");
    const int steps = steps_of[which];
    const double per_step = per_step_of[which];
    auto launch = [&](int st) {
        const dim3 g(BENCH_BLOCKS), b(BENCH_THREADS);
        switch (which) {
            case 0: hipLaunchKernelGGL(k_bench_mul_tab, g, b, 0, s, 12345u, st, scratch_word); break;
            case 1: hipLaunchKernelGGL(k_bench_mul_clmul, g, b, 0, s, 12345u, st, scratch_word); break;
            case 2: hipLaunchKernelGGL(k_bench_valu, g, b, 0, s, 12345u, st, scratch_word); break;
            case 3: hipLaunchKernelGGL(k_bench_mul_tab5, g, b, 0, s, 12345u, st, scratch_word); break;
            case 4: hipLaunchKernelGGL(k_bench_lds128, g, b, 0, s, 12345u, st, scratch_word); break;
            case 5: hipLaunchKernelGGL(k_bench_mul_mask, g, b, 0, s, 12345u, st, scratch_word); break;
            default: hipLaunchKernelGGL(k_bench_gf8_logexp, g, b, 0, s, 12345u, st, scratch_word); break;
        }
    };
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) throw std::runtime_error("hipEventCreate failed");
    launch(steps / 8);  // warm-up (code load, clocks)
    float best = 0.f;
    for (int rep = 0; rep < 5; rep++) {
        (void)hipEventRecord(a, s);
        launch(steps);
        (void)hipEventRecord(b, s);
        if (hipEventSynchronize(b) != hipSuccess) throw std::runtime_error("microbenchmark failed");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a, b);
        if (rep == 0 || ms < best) best = ms;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (hipGetLastError() != hipSuccess) throw std::runtime_error("microbenchmark launch failed");
    const double work = (double)BENCH_BLOCKS * BENCH_THREADS * per_step * steps;
    return work / (best * 1e-3);
}

// ------------------------------------------------------- instantiations ---
#define LIG_INST(F)                                                                                                         \
    template void launch_build_tables<F>(const F::T*, size_t, F::T*, hipStream_t);                                          \
    template void launch_transpose_scale<F>(const F::T*, size_t, size_t, const F::T*, const F::T*, const F::T*, F::T*,      \
                                            hipStream_t, const F::T* const*, int);                                          \
    template void launch_scale_rows<F>(F::T*, size_t, size_t, const F::T*, hipStream_t);                                    \
    template void launch_transpose<F>(const F::T*, size_t, size_t, F::T*, hipStream_t);                                     \
    template void launch_fft_stage<F>(F::T*, size_t, size_t, int, const F::T*, int, hipStream_t);                           \
    template void launch_fft_quarters<F>(const F::T*, F::T*, size_t, size_t, const F::T*, const F::T*, hipStream_t,         \
                                         PassObserver*, PassTune, int);                                                      \
    template void launch_leaf_hash<F>(const F::T*, size_t, size_t, uint32_t*, uint32_t*, hipStream_t, int, size_t);                   \
    template void launch_merkle_commit<F>(const F::T*, size_t, size_t, uint32_t*, uint32_t*, hipStream_t, uint32_t*, uint32_t*,     \
                                          uint32_t, int, size_t);                                                                  \
    template void launch_fold<F>(const F::T*, size_t, int, const F::T*, F::T*, hipStream_t);                                \
    template void launch_gather_rows<F>(const F::T*, size_t, const uint64_t*, size_t, F::T*, F::T*, hipStream_t);                  \
    template int launch_induce<F>(const F::T*, const F::T*, int, int, F::T*, F::T*, F::T*, hipStream_t,                     \
                                  const EnforcedIn<F>*, const Eval013In<F>*, int, const InduceJobs&);                       \
    template void launch_eval013<F>(const F::T*, const F::T*, size_t, F::Acc*, F::Acc*, hipStream_t);                       \
    template void launch_zero_acc<F>(F::Acc*, int, hipStream_t);                                                                 \
    template void launch_sumcheck_fold_step<F>(const F::T*, const F::T*, size_t, const FoldStep<F>&, F::T*, F::T*, F::Acc*, \
                                               F::Acc*, hipStream_t, int, const SumcheckJobs&);                             \
    template void launch_induce_prep<F>(const uint64_t*, const F::T*, const F::T*, int, int, F::T*, F::T*, hipStream_t,    \
                                        const RowGather<F>*, int, const InduceJobs&);                                       \
    template int launch_induce_all<F>(const uint64_t*, const F::T*, const F::T*, int, int, F::T*, F::T*,                  \
                                      const RowGather<F>&, F::T*, F::T*, F::T*, hipStream_t, const EnforcedIn<F>&,         \
                                      const Eval013In<F>*, int, const InduceJobs&);                                        \
    template void launch_enforced<F>(const F::T*, int, int, const F::T*, const F::T*, F::Acc*, hipStream_t);

LIG_INST(F32)
LIG_INST(F128)

}  // namespace lig
