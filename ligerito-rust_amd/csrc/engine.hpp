// Device engine of the Ligerito library: context with a caching device
// allocator, Reed–Solomon plans, device-resident witnesses, the GPU prover
// and the proof object.
#pragma once
#include <algorithm>
#include <array>
#include <atomic>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "hostmath.hpp"
#include "kernels.hpp"

namespace lig {

// Device-memory cache: freed blocks are kept per exact size and reused, so
// repeated proofs of one configuration allocate nothing after the first.
// The byte counters let a server charge a pool's whole device footprint
// (blocks handed out plus blocks cached) against its memory budget.
struct Pool {
    std::multimap<size_t, void*> free_;
    std::mutex mu;
    std::atomic<size_t> live{0}, cached{0};
    void* get(size_t bytes);
    void put(void* p, size_t bytes);
    // returns every cached block to the device (blocks in use stay)
    void drain();
    size_t held() const { return live.load() + cached.load(); }
    ~Pool() { drain(); }
};

// Phases timed with HIP events when profiling is enabled.
// PH_PASS times each FFT pass launch of the first commit of a prove alone.
enum Phase { PH_ENCODE0 = 0, PH_ENCODE, PH_MERKLE, PH_FOLD, PH_INDUCE, PH_SUMCHECK, PH_OPEN, PH_PASS, PH_COUNT };
const char* phase_name(int p);

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Pool pool;
    bool profiling = false;
    struct Mark {
        int phase;
        hipEvent_t a, b;
    };
    std::vector<Mark> marks;
    double phase_ms[PH_COUNT] = {};
    uint64_t phase_n[PH_COUNT] = {};
    // Pinned, device-mapped, coherent host staging (grown on demand;
    // contents are only valid until the next call that uses it).  Kernels
    // write results the host needs straight into it (through pin_dev).
    uint8_t* pin = nullptr;
    uint8_t* pin_dev = nullptr;
    size_t pin_cap = 0;
    uint8_t* pinned(size_t bytes) {
        if (bytes > pin_cap) {
            sync();
            if (side) wait(side);
            if (pin) (void)hipHostFree(pin);
            pin = pin_dev = nullptr;
            size_t cap = std::max<size_t>(bytes, (size_t)1 << 20);
            LIG_HIP(hipHostMalloc((void**)&pin, cap, hipHostMallocMapped | hipHostMallocCoherent));
            LIG_HIP(hipHostGetDevicePointer((void**)&pin_dev, pin, 0));
            memset(pin, 0, 256);  // the completion flag starts below any sequence number
            pin_cap = cap;
        }
        return pin;
    }
    // Waits for the stream.  With spin_wait the host polls the stream (a
    // blocking wait wakes the thread several microseconds after the last
    // kernel ends, which the prover's host steps then add to every sync);
    // contexts that share the host with many others (batch and serve
    // workers) block instead.
    bool spin_wait = true;
    void wait(hipStream_t st) {
        if (spin_wait) {
            for (;;) {
                const hipError_t e = hipStreamQuery(st);
                if (e == hipSuccess) return;
                if (e != hipErrorNotReady) LIG_HIP(e);
                __builtin_ia32_pause();
            }
        }
        LIG_HIP(hipStreamSynchronize(st));
    }
    void sync() { wait(stream); }
    // Completion flag: a kernel stores a sequence number at byte PIN_FLAG of
    // the pinned buffer once the results the host needs are written (in
    // mapped memory, system scope).  Polling it sees the results several
    // microseconds before the stream reports the kernel complete.
    static constexpr size_t PIN_FLAG = 128;
    uint32_t flag_seq = 0;
    uint32_t* flag_dev() { return reinterpret_cast<uint32_t*>(pin_dev + PIN_FLAG); }
    uint32_t next_flag() { return ++flag_seq; }
    void wait_flag(uint32_t seq) {
        const uint32_t* f = reinterpret_cast<const uint32_t*>(pin + PIN_FLAG);
        if (!spin_wait) {
            sync();
            LIG_REQUIRE(__atomic_load_n(f, __ATOMIC_ACQUIRE) == seq, LIG_E_ASSERT, "stream done without its completion flag");
            return;
        }
        // every 1024 polls the stream is asked too, so that an error or a
        // stream that ends without the flag stops the wait
        for (uint32_t it = 1;; it++) {
            if (__atomic_load_n(f, __ATOMIC_ACQUIRE) == seq) return;
            if ((it & 1023) == 0) {
                const hipError_t e = hipStreamQuery(stream);
                if (e == hipSuccess) {
                    LIG_REQUIRE(__atomic_load_n(f, __ATOMIC_ACQUIRE) == seq, LIG_E_ASSERT,
                                "stream done without its completion flag");
                    return;
                }
                if (e != hipErrorNotReady) LIG_HIP(e);
            }
            __builtin_ia32_pause();
        }
    }
    // A second stream (default priority) for device-to-host copies that
    // overlap the work on the context stream; created on first use, ended
    // by release_side().
    hipStream_t side = nullptr;
    hipStream_t side_stream() {
        if (!side) LIG_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        return side;
    }
    // Orders the side stream's copies after work on the context stream.
    hipEvent_t side_ev = nullptr;
    hipEvent_t side_event() {
        if (!side_ev) LIG_HIP(hipEventCreateWithFlags(&side_ev, hipEventDisableTiming));
        return side_ev;
    }
    void release_side() {
        if (side) {
            (void)hipStreamSynchronize(side);
            (void)hipStreamDestroy(side);
            side = nullptr;
        }
        if (side_ev) {
            (void)hipEventDestroy(side_ev);
            side_ev = nullptr;
        }
    }
    // Resolves recorded event pairs into per-phase totals.
    void collect();
    // Copies bytes from pageable host memory to device memory, ordered on
    // the context stream; the host buffer may change once it returns.
    // Large copies go through a ring of pinned staging slots (allocated on
    // first use) filled by several host threads while the DMA engine moves
    // earlier slots.
    void upload(void* dst, const void* src, size_t bytes);
    static constexpr int UP_SLOTS = 4, UP_THREADS = 4;
    static constexpr size_t UP_SLOT_BYTES = (size_t)4 << 20;
    uint8_t* up_ring = nullptr;
    hipEvent_t up_ev[UP_SLOTS] = {};
    // The owner synchronizes and destroys the stream first (no copy from the
    // staging ring or kernel reading pinned memory is then in flight).
    ~Ctx() {
        if (up_ring) (void)hipHostFree(up_ring);
        for (hipEvent_t e : up_ev)
            if (e) (void)hipEventDestroy(e);
        if (pin) (void)hipHostFree(pin);
    }
};

// Records a HIP event pair on the context stream around a scope.
struct PhaseTimer {
    Ctx* ctx;
    int phase;
    hipEvent_t a = nullptr, b = nullptr;
    PhaseTimer(Ctx* c, int p) : ctx(c), phase(p) {
        if (!ctx->profiling) return;
        LIG_HIP(hipEventCreate(&a));
        LIG_HIP(hipEventCreate(&b));
        LIG_HIP(hipEventRecord(a, ctx->stream));
    }
    ~PhaseTimer() {
        if (!a) return;
        if (hipEventRecord(b, ctx->stream) == hipSuccess) ctx->marks.push_back({phase, a, b});
    }
};

// Pool-backed device buffer (or, made by view(), a view of memory owned
// elsewhere, which release() leaves alone).
template <class T> struct PBuf {
    Pool* pool = nullptr;
    T* p = nullptr;
    size_t n = 0;
    PBuf() = default;
    PBuf(Pool* pl, size_t count) : pool(pl), n(count) {
        p = count ? (T*)pool->get(count * sizeof(T)) : nullptr;
    }
    PBuf(const PBuf&) = delete;
    PBuf& operator=(const PBuf&) = delete;
    PBuf(PBuf&& o) noexcept : pool(o.pool), p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    PBuf& operator=(PBuf&& o) noexcept {
        if (this != &o) {
            release();
            pool = o.pool; p = o.p; n = o.n;
            o.p = nullptr; o.n = 0;
        }
        return *this;
    }
    ~PBuf() { release(); }
    static PBuf view(T* ptr, size_t count) {
        PBuf b;
        b.p = ptr;
        b.n = count;
        return b;
    }
    void release() {
        if (p && pool) pool->put(p, n * sizeof(T));
        p = nullptr;
        n = 0;
    }
};

// Host field H paired with its device traits.
template <class H> struct DevField;
template <> struct DevField<H32> { using F = F32; };
template <> struct DevField<H128> { using F = F128; };

// An unreduced device accumulator as a field element.
template <class H> typename H::T acc_reduce_host(const typename DevField<H>::F::Acc& a);
template <> inline uint32_t acc_reduce_host<H32>(const uint64_t& a) { return gf32_reduce64(a); }
template <> inline u128v acc_reduce_host<H128>(const u128v& a) { return a; }

// ReedSolomonEncoding<F> (reed_solomon.rs:7-64) with device tables.
template <class H> struct RsPlan {
    using T = typename H::T;
    Ctx* ctx = nullptr;
    int log_msg = 0, log_block = 0;
    Vec<H> tw, pis;
    DevBuf<T> d_tw, d_tw_tabs, d_pis, d_pis_lo, d_pis_hi;
    size_t m() const { return (size_t)1 << log_msg; }
    size_t M() const { return (size_t)1 << log_block; }
    void init(Ctx* c, size_t msg_len, size_t block_len);
    // encode_non_systematic on n columns, column-major, in place (device),
    // asynchronous on the context stream.  For n > 1 the kernels work on a
    // row-major copy in scratch (M*n elements; from the pool if null).
    void encode_non_systematic_cols(T* d_cols, size_t n, T* scratch = nullptr);
    // encode (systematic) on one host message.
    void encode_systematic(const T* msg, T* out);
    // Launch shape of the FFT passes per column count (no effect on the
    // result): chosen at the first encode of that many columns by timing
    // the candidate shapes on this GPU (autotune), unless set_tune fixed it
    // or autotuning is off (LIG_AUTOTUNE=0 in the environment) or the
    // matrix is small (the defaults then).
    std::map<size_t, PassTune> tunes;
    PassTune tune_for(size_t n);
    void set_tune(size_t n, PassTune t) { tunes[n] = t; }
    // device memory held by the plan's tables
    size_t device_bytes() const {
        return (d_tw.n + d_tw_tabs.n + d_pis.n + d_pis_lo.n + d_pis_hi.n) * sizeof(T);
    }
};

// RecursiveLigeroWitness<F> (data_structures.rs:32-40), device-resident:
// the encoded matrix row-major [M][n] and the Merkle tree, all layers back
// to back.
template <class H> struct Witness {
    using T = typename H::T;
    Ctx* ctx = nullptr;
    size_t m = 0, n = 0, M = 0;
    PBuf<T> mat;
    PBuf<uint32_t> tree;
    uint8_t root[32];
};

// ligero_commit (ligero.rs:70-95) on a device poly (column-major m x n).
template <class H> void commit(RsPlan<H>& rs, const typename H::T* d_poly, size_t m, size_t n, Witness<H>& w, int encode_phase = PH_ENCODE);
// ligero_commit of count device polys of one plan at once (GF(2^32) only:
// H = H32; m >= 64, n >= 4): every launch covers all of them.  out[i] views buffers that keep
// owns (from the plan's context pool); keep must outlive the witnesses and
// every kernel that reads them.  Returns after the roots are on the host.
template <class H>
void commit_group(RsPlan<H>& rs, const typename H::T* const* d_polys, int count, size_t m, size_t n,
                  std::vector<Witness<H>>& out, std::shared_ptr<void>& keep);

// The dimensions of a ProverConfig (data_structures.rs:9-21); its
// VerifierConfig (:23-30) is the same shape in log form (initial_dim =
// log2 initial_m, log_dims[i] = log2 dims_m[i]).
struct Shape {
    int recursive_steps = 0;
    size_t initial_m = 0, initial_n = 0;
    int initial_k = 0;
    std::vector<size_t> dims_m, dims_n;
    std::vector<int> ks;
    // The generated configuration for 2^L coefficients, L in [16, 27]:
    // hardcoded_config_20 / _24 (config.rs:6-47, 69-112) at L = 20 / 24,
    // the same rules elsewhere.
    static Shape generated(int L);
    // From a VerifierConfig's fields (config.rs:49-67, 114-134).
    static Shape from_logs(int recursive_steps, int initial_dim, const int* log_dims, int initial_k, const int* ks);
    // The invariants prover() and verifier() rely on (LIG_E_INVALID with
    // the broken one otherwise), and this library's size limits.
    void validate(int field_bytes) const;
    size_t num_coeffs() const { return initial_m * initial_n; }
};

// ProverConfig<F> (data_structures.rs:9-21).
template <class H> struct Config {
    Ctx* ctx = nullptr;
    int log_n = 0;  // log2 of the coefficient count
    int recursive_steps = 0;
    size_t initial_m = 0, initial_n = 0;
    int initial_k = 0;
    std::vector<size_t> dims_m, dims_n;
    std::vector<int> ks;
    std::unique_ptr<RsPlan<H>> initial_rs;
    std::vector<std::unique_ptr<RsPlan<H>>> rs;
    // eval_sk_at_vks(2^ell) for every ell a prove's sumcheck inductions use
    std::vector<Vec<H>> sks_by_ell;
    void init(Ctx* c, int L) { init(c, Shape::generated(L)); }
    void init(Ctx* c, const Shape& sh);
    Shape shape() const {
        Shape sh;
        sh.recursive_steps = recursive_steps;
        sh.initial_m = initial_m;
        sh.initial_n = initial_n;
        sh.initial_k = initial_k;
        sh.dims_m = dims_m;
        sh.dims_n = dims_n;
        sh.ks = ks;
        return sh;
    }
    size_t device_bytes() const {
        size_t b = initial_rs ? initial_rs->device_bytes() : 0;
        for (const auto& r : rs) b += r->device_bytes();
        return b;
    }
};

// RecursiveLigeroProof / FinalLigeroProof parts (data_structures.rs:42-66).
template <class H> struct Rlp {
    size_t n_rows = 0, n_cols = 0;
    std::vector<typename H::T> rows;
    std::vector<uint8_t> nodes;  // 32 bytes each
};

// FinalizedLigeritoProof<F> (data_structures.rs:91-103).
template <class H> struct Proof {
    std::array<uint8_t, 32> initial_root;
    Rlp<H> initial;
    std::vector<std::array<uint8_t, 32>> rec_roots;
    std::vector<Rlp<H>> rec_proofs;
    std::vector<typename H::T> yr;
    Rlp<H> final_;
    std::vector<typename H::T> tr;  // 3 per round
};

// prover (prover.rs:15-199) on a device-resident poly.
template <class H> void prove(Config<H>& cfg, const typename H::T* d_poly, uint32_t fs_seed, Proof<H>& out);
// prove() of count device polys of one GF(2^32) configuration with one
// recursive step (H = H32), in lockstep: every kernel launch covers all of
// them.  proofs[j] is prove(cfg, d_polys[j], seeds[j]).
constexpr int GANG_MAX = 256;
template <class H>
void prove_gang(Config<H>& cfg, const typename H::T* const* d_polys, const uint32_t* seeds, int count, Proof<H>* proofs);
// verifier (verifier.rs:10-164) on the host for a configuration's shape:
// 1 accept, 0 reject, <0 at the reference's panic sites; with dev, the
// basis inductions (verifier.rs:54-61,148-155) run on that context's GPU.
template <class H> int verify(const Shape& sh, const Proof<H>& p, uint32_t fs_seed, Ctx* dev = nullptr);
template <class H> std::vector<uint8_t> serialize(const Proof<H>& p);
template <class H> bool deserialize(const uint8_t* b, size_t len, Proof<H>& p);
template <class H> size_t proof_size(const Proof<H>& p);


}  // namespace lig
