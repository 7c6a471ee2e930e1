// Launchers for the Ligerito hot-path kernels (kernels.hip).  All pointers
// are device pointers; every launcher is asynchronous on `s`.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include "gf.hpp"

namespace lig {

// Element layouts shared by host and device code.  Device operations on
// them live in fields.hpp (Ops<F>).
struct F32 {
    using T = uint32_t;
    using Acc = uint64_t;                 // unreduced carry-less products
    static constexpr int BITS = 32;
    static constexpr int TAB = 128;        // multiply table: 8 nibble positions x 16 values
    static constexpr int TAB_STRIDE = 144; // LDS stride: consecutive tables start 16 banks apart
};
struct F128 {
    using T = u128v;
    using Acc = u128v;                     // products are reduced before accumulation
    static constexpr int BITS = 128;
    static constexpr int TAB = 512;        // 32 nibble positions x 16 values
    static constexpr int TAB_STRIDE = 512;
};

// Device copy of a Fiat–Shamir transcript (host Transcript: SHA-256 state,
// pending block as big-endian words zero past blen, byte count, squeeze
// counter), plus a flag device-side checks set.
struct DevTranscript {
    uint32_t h[8];
    uint32_t buf[16];
    uint32_t blen, total_lo, total_hi, counter;
    uint32_t err;
    uint32_t pad[3];
};

// FFT tile shape per field: up to 2^K rows x C columns per pass.
template <class F> struct PassShape;

// ---------------------------------------------------------------- tables --
// tabs[i * F::TAB + e] = multiply-by-lams[i] table entry e.
template <class F> void launch_build_tables(const typename F::T* lams, size_t count, typename F::T* tabs, hipStream_t s);

// ------------------------------------------------------------ Reed–Solomon --
// msg[i*n + c] = poly[c*m + i] * pis[i]  (poly is column-major as in
// ligero.rs:11-27; msg is row-major).  Both fields read the plain pis (m
// elements) and build one table per message row of a tile in LDS; narrow
// matrices (GF(2^32): fewer than 4 columns or 64 rows, GF(2^128): fewer
// than 64 columns) use pis[i] = pis[i & 63] * pis[i & ~63] with lo_tabs
// (min(m,64) tables) and hi_tabs (max(m/64,1) tables) instead.
// With inputs > 1 (GF(2^32), m >= 64, n >= 4): that many inputs at once,
// polys a device array of their pointers, msg receiving their messages one
// after another ([input][m][n]); poly is then unused.
template <class F>
void launch_transpose_scale(const typename F::T* poly, size_t m, size_t n, const typename F::T* pis,
                            const typename F::T* lo_tabs, const typename F::T* hi_tabs, typename F::T* msg, hipStream_t s,
                            const typename F::T* const* polys = nullptr, int inputs = 1);
// buf[i*n + c] *= pis[i] for i < rows (general multiply).
template <class F> void launch_scale_rows(typename F::T* buf, size_t rows, size_t n, const typename F::T* pis, hipStream_t s);
// dst[c*rows + r] = src[r*cols + c] (src row-major rows x cols; dst != src).
template <class F> void launch_transpose(const typename F::T* src, size_t rows, size_t cols, typename F::T* dst, hipStream_t s);
// Static device scratch of the entry points that allocate nothing:
// NOALLOC_SCRATCH_BYTES per device.  noalloc_scratch_ptr returns its device
// address on `device`, which must be current (looked up once per device).
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;
void* noalloc_scratch_ptr(int device);

// One additive-FFT stage on a row-major [N][n] buffer with general multiply
// (binary_fft.rs:180-206 butterflies); inverse selects the IFFT butterfly.
template <class F>
void launch_fft_stage(typename F::T* buf, size_t N, size_t n, int stage, const typename F::T* tw, int inverse, hipStream_t s);
// Sees each GF(2^32) FFT pass launch of launch_fft_quarters: log2 of the
// pass radix and the elements it reads (and writes), before and after it.
struct PassObserver {
    virtual void before(int radix_log, size_t elems) = 0;
    virtual void after(int radix_log, size_t elems) = 0;

protected:
    ~PassObserver() = default;
};

// Stages 2..log2(4m)-1 of the length-4m FFT on the four quarters of the
// row-major [4m][n] matrix `mat`.  If msg != nullptr, the first pass reads
// its input rows from msg (all four quarters start equal to msg), otherwise
// from mat itself.  tw holds the heap-ordered twiddles; tw_tabs one
// multiply table per twiddle, read only where the GF(2^32) passes do not
// apply (GF(2^128), or m == 1), and may be null otherwise.
// Launch shape of the FFT passes (no effect on the result): kmax = most
// butterfly stages per pass (GF(2^32): 2..6, GF(2^128): 1..3), group = most
// tiles (GF(2^32)) or 64-column jobs (GF(2^128)) per workgroup, a power of
// two (GF(2^32): 1..8, GF(2^128): 1..16).  0 takes the default.
struct PassTune {
    int kmax = 0;
    int group = 0;
};
// With codewords > 1 (GF(2^32), m >= 2): that many codewords of the same
// plan at once, mat holding them one after another ([codeword][4m][n]) and
// msg (if given) their messages ([codeword][m][n]).
template <class F>
void launch_fft_quarters(const typename F::T* msg, typename F::T* mat, size_t m, size_t n, const typename F::T* tw,
                         const typename F::T* tw_tabs, hipStream_t s, PassObserver* obs = nullptr, PassTune tune = {},
                         int codewords = 1);

// -------------------------------------------------------------- Merkle ----
// layer0[row*8 .. +8] = BLAKE3(row bytes) for every row of the [rows][n] matrix.
// Leaves of whole 1 KiB chunks (a power of two of them, 2..256) hash chunk-
// parallel when scratch is given: leaf_hash_scratch_words(rows, row bytes)
// words (0 when the leaf shape does not qualify).
// With trees > 1 the rows are that many equal matrices one after another,
// the leaves of matrix t going to layer0 + t * tree_stride (words); scratch
// (if the leaf shape needs it) is sized for one matrix and reused.
size_t leaf_hash_scratch_words(size_t rows, size_t row_bytes);
template <class F>
void launch_leaf_hash(const typename F::T* mat, size_t rows, size_t n, uint32_t* layer0, uint32_t* scratch, hipStream_t s,
                      int trees = 1, size_t tree_stride = 0);
// layer0[i*8 .. +8] = BLAKE3(bytes[offs[i] .. offs[i+1])) for leaves of any
// length up to 256 KiB (offs: n_leaves + 1 non-decreasing byte offsets).
void launch_leaf_hash_var(const uint8_t* bytes, const uint64_t* offs, size_t n_leaves, uint32_t* layer0, hipStream_t s);
// Builds layers 1.. in place after layer 0 (contiguous node layout,
// merkle_tree.rs:22-60).  n_leaves must be a power of two, at least 2.
// If root_out is given, the root's 8 words are also written there (for
// instance to mapped host memory); if flag is given, the last launch stores
// seq there after every node and the root (for a host polling mapped memory).
// With trees > 1: that many trees of n_leaves leaves, tree_stride words
// apart, their roots to root_out + 8t (if given); no flag (launch_signal
// after them instead).
void launch_merkle_levels(uint32_t* tree, size_t n_leaves, hipStream_t s, uint32_t* root_out = nullptr,
                          uint32_t* flag = nullptr, uint32_t seq = 0, int trees = 1, size_t tree_stride = 0);
// The whole tree of a [rows][n] matrix (launch_leaf_hash, then
// launch_merkle_levels with the same root_out, flag, trees and tree_stride;
// rows per tree a power of two).  Rows of 64 bytes in trees of fewer than
// 2^17 leaves skip the separate leaf launch: a 64-byte leaf is hashed like a
// parent, so the first top launch reads the rows as its input layer.
template <class F>
void launch_merkle_commit(const typename F::T* mat, size_t rows, size_t n, uint32_t* tree, uint32_t* scratch, hipStream_t s,
                          uint32_t* root_out = nullptr, uint32_t* flag = nullptr, uint32_t seq = 0, int trees = 1,
                          size_t tree_stride = 0);
// A few words copied by the first threads of a launch (for instance a
// section's last triple, as raw accumulators, to mapped host memory).
struct WordCopy {
    const uint32_t* src = nullptr;
    uint32_t* dst = nullptr;
    int n = 0;
};
// Stores seq at flag (mapped host memory, if flag is given) after the
// stream's earlier kernels and the extra copy (at most 64 words).
void launch_signal(uint32_t* flag, uint32_t seq, hipStream_t s, const WordCopy* extra = nullptr);
// out[i] = node at byte offset offs[i] of the tree (32 bytes each).
// With extra, the first threads also copy those words (at most 256).
void launch_gather_nodes(const uint32_t* tree, const uint64_t* node_idx, size_t count, uint32_t* out, hipStream_t s,
                         const WordCopy* extra = nullptr);

// dst[i] = src[i] for i < n words (dst may be mapped pinned host memory).
void launch_copy_words(uint32_t* dst, const uint32_t* src, size_t n, hipStream_t s);
// n words from each of count blocks src + j * src_stride to dst + j * dst_stride
void launch_copy_blocks(uint32_t* dst, size_t dst_stride, const uint32_t* src, size_t src_stride, size_t n, int count,
                        hipStream_t s);

// Microbenchmarks (best of five timed launches, synchronous): which = 0
// GF(2^32) products per second by wave-uniform constants with nibble tables
// in LDS (the FFT passes' multiply), 1 general GF(2^32) products per second
// (carry-less product from integer multiplies), 2 32-bit VALU instructions
// per second per lane (lane-ops/s), 3 products per second with 5-bit-window
// tables, 4 LDS bytes read per second with the GF(2^128) table pattern.
// scratch_word: one device word.
double bench_rate(int which, uint32_t* scratch_word, hipStream_t s);

// ------------------------------------------------------------ folds -------
// out[i] = sum_j gr[j] * poly[j*m + i], j < 2^kk (partial_eval over the top
// kk variables, multilinear_poly.rs:66-87).  gr_tabs: 2^kk tables.
template <class F>
void launch_fold(const typename F::T* poly, size_t m, int kk, const typename F::T* gr_tabs, typename F::T* out, hipStream_t s);
// The same for GF(2^32) with the weights themselves (host memory, passed as
// a kernel argument); the kernel builds its tables.  Returns false (nothing
// launched) unless kk <= FOLD32_MAXK and the 2^kk columns span < 2 GiB.
constexpr int FOLD32_MAXK = 6;
// It also zeroes zero[0..n_zero) (at most 256), if given.
bool launch_fold32_weights(const uint32_t* poly, size_t m, int kk, const uint32_t* gr_host, uint32_t* out, hipStream_t s,
                           uint64_t* zero = nullptr, int n_zero = 0);
// The same fold for `count` proves of one shape in one launch (job j =
// blockIdx.y): input polys[j], the 2^kk weights at w + j * 64, output at
// out + j * out_stride, accumulators to zero at zero + j * zero_stride.
struct FoldJobs {
    const uint32_t* const* polys = nullptr;
    const uint32_t* w = nullptr;
    size_t out_stride = 0;
    size_t zero_stride = 0;
};
void launch_fold32_jobs(size_t m, int kk, const FoldJobs& jobs, int count, uint32_t* out, uint64_t* zero, int n_zero,
                        hipStream_t s);
// out (and out2, if given) receive the count rows listed in rows.
template <class F>
void launch_gather_rows(const typename F::T* mat, size_t n, const uint64_t* rows, size_t count, typename F::T* out,
                        typename F::T* out2, hipStream_t s);

// ------------------------------------------------------------ induce ------
// The basis index splits into clo low bits (P tables, nq << clo elements)
// and ell - clo high bits (K tables, nq << (ell - clo) elements).
int induce_low_bits(int ell);
// Inputs of the enforced sum of sumcheck_polys.rs:249-256: the opened rows
// (nq x ncols), the row weights gr (ncols) and the parts buffer.
template <class F> struct EnforcedIn {
    const typename F::T* rows = nullptr;
    int ncols = 0;
    const typename F::T* gr = nullptr;
    typename F::Acc* parts = nullptr;
};
// The triple of eval_013_product(f, basis) (as launch_eval013: added into
// res, res_clear zeroed), f of the basis length.
template <class F> struct Eval013In {
    const typename F::T* f = nullptr;
    typename F::Acc* res = nullptr;
    typename F::Acc* res_clear = nullptr;
};
// cvals[i*ell + t] = W_t(q_i); ap[i] = alpha^i.  Produces basis[2^ell].
// With e, also the enforced sum as parts whose XOR is the sum (unreduced):
// one per query, from the same launch, where the GF(2^32) kernel applies,
// else enforced_parts(nq, ncols) from launch_enforced.  With ev, also the
// 0/1/3 triple of (f, basis), inside the accumulation where the GF(2^32)
// kernel applies, else by launch_eval013.  Returns the number of enforced
// parts written (0 without e).
// Several proves of one shape in one launch (job j = blockIdx.y): the
// strides, in elements of each buffer, between the jobs' inputs and outputs
// that are not laid out by the kernels themselves (those intermediate
// buffers hold the jobs one after another: query constants nq * ell,
// alpha powers nq, P nq * 2^clo, K nq * 2^(ell - clo), basis 2^ell, rows
// nq * n).  Strides 0 with one job are the single prove.
struct InduceJobs {
    size_t q = 0;       // query indices
    size_t alpha = 0;   // alpha
    size_t mat = 0;     // the committed matrix the rows are gathered from
    size_t rows2 = 0;   // second copy of the gathered rows
    size_t gr = 0;      // row-combination weights
    size_t parts = 0;   // enforced-sum parts
    size_t f = 0;       // the folded polynomial of the triple
    size_t res = 0;     // triple accumulators (res and res_clear)
};
template <class F>
int launch_induce(const typename F::T* cvals, const typename F::T* ap, int nq, int ell, typename F::T* P, typename F::T* K,
                  typename F::T* basis, hipStream_t s, const EnforcedIn<F>* e = nullptr, const Eval013In<F>* ev = nullptr,
                  int count = 1, const InduceJobs& jobs = InduceJobs());

// The opened rows of a step: out (and out2, if given) receive the rows of
// mat (row-major, n per row) listed in the query indices.
template <class F> struct RowGather {
    const typename F::T* mat = nullptr;
    size_t n = 0;
    typename F::T* out = nullptr;
    typename F::T* out2 = nullptr;
};
// cvals[i*ell + t] = W_t(q_i) (the sks_at_x recursion, binary_fft.rs:289-292)
// and ap[i] = alpha^i for the nq queries.  With g, also the row gather of
// the same queries (one launch for both where the GF(2^32) kernel applies).
template <class F>
void launch_induce_prep(const uint64_t* q, const typename F::T* sks, const typename F::T* alpha, int nq, int ell,
                        typename F::T* cvals, typename F::T* ap, hipStream_t s, const RowGather<F>* g = nullptr,
                        int count = 1, const InduceJobs& jobs = InduceJobs());
// The enforced sum of sumcheck_polys.rs:249-256, sum_i ap[i] * (row_i . gr)
// (unreduced), as enforced_parts(nrows, ncols) parts whose XOR is the sum.
// The whole induction (launch_induce_prep then launch_induce, or for
// GF(2^32) one launch for the query constants, opened rows, enforced-sum
// parts and tensors, then the accumulation): the opened rows go to g.out
// (and g.out2), which e.rows must name.  Returns the number of parts.
template <class F>
int launch_induce_all(const uint64_t* q, const typename F::T* sks, const typename F::T* alpha, int nq, int ell,
                      typename F::T* cvals, typename F::T* ap, const RowGather<F>& g, typename F::T* P, typename F::T* K,
                      typename F::T* basis, hipStream_t s, const EnforcedIn<F>& e, const Eval013In<F>* ev, int count = 1,
                      const InduceJobs& jobs = InduceJobs());
constexpr int ENFORCED_MAX_PARTS = 256;
int enforced_parts(int nrows, int ncols);
template <class F>
void launch_enforced(const typename F::T* rows, int nrows, int ncols, const typename F::T* gr, const typename F::T* ap,
                     typename F::Acc* parts, hipStream_t s);

// ------------------------------------------------------------ sumcheck ----
// res[0..3) = (sum f0*g0, sum f1*g1, sum f2*g2) unreduced (eval_013_product).
// res_clear (3 accumulators) is zeroed for the next accumulation; res must
// have been zeroed by the previous call (or launch_zero_acc, which zeroes
// count consecutive accumulators).
template <class F>
void launch_eval013(const typename F::T* f, const typename F::T* g, size_t len, typename F::Acc* res,
                    typename F::Acc* res_clear, hipStream_t s);
template <class F> void launch_zero_acc(typename F::Acc* res, int count, hipStream_t s);
// Folds f and b by the round's challenge (length len -> len/2, ligerito.rs:
// 119-127) into f2, b2 and adds the triple of the folded pair into res
// (res_clear zeroed for the next round).  The transcript step that precedes
// the fold runs in the fold kernel itself:
// every workgroup takes the state at st_in (or, if st_in is null, the state
// init, passed by value), absorbs the triple at res_in (reduced) and
// squeezes n_squeeze challenges; the fold's r is the last one.  Workgroup 0
// also stores the new state at st_out (!= st_in), logs the triple to
// tlog[0..3) and the challenges to rlog[0..n_squeeze), and, if enf is given,
// sets *err to 1 when t0 + t1 differs from the enforced sum (the XOR of the
// n_enf parts at enf, ligerito.rs:35,52).  With bnew, the fold reads
// b + beta * bnew instead of b (glue, ligerito.rs:59-67), beta being the
// first of two challenges.  res_in must differ from res and res_clear (all
// workgroups read it while the kernel writes those).
template <class F> struct FoldStep {
    int lead;  // a section's leading step (enforced-sum check, one or two challenges)
    const DevTranscript* st_in;
    DevTranscript init;
    DevTranscript* st_out;
    const typename F::Acc* res_in;
    const typename F::Acc* enf;
    int n_enf;
    uint32_t* err;
    int n_squeeze;
    const typename F::T* bnew;
    typename F::T* tlog;
    typename F::T* rlog;
};
// Strides between the jobs of a fold step launched for several proves of
// one shape (job j = blockIdx.y), in elements, accumulators, transcript
// states and log words; zero for a single prove.
struct SumcheckJobs {
    size_t f = 0, b = 0, f2 = 0, b2 = 0;  // polys
    size_t bnew = 0;                      // the basis glued in by a leading step
    size_t acc = 0;                       // res, res_clear, the triple read and the enforced-sum parts
    size_t st = 0;                        // transcript states in and out
    size_t log_elems = 0;                 // triple and challenge logs
    size_t log_words = 0;                 // check flag
};
template <class F>
void launch_sumcheck_fold_step(const typename F::T* f, const typename F::T* b, size_t len, const FoldStep<F>& step,
                               typename F::T* f2, typename F::T* b2, typename F::Acc* res, typename F::Acc* res_clear,
                               hipStream_t s, int count = 1, const SumcheckJobs& jobs = SumcheckJobs());
}  // namespace lig
