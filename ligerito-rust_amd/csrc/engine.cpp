// Device engine: plans, commit, GPU prover.
#include "engine.hpp"

#include <algorithm>
#include <cstdlib>
#include <exception>
#include <thread>
#include <tuple>

namespace lig {

// ------------------------------------------------------------------- pool --
void* Pool::get(size_t bytes) {
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = free_.find(bytes);
        if (it != free_.end()) {
            void* p = it->second;
            free_.erase(it);
            cached -= bytes;
            live += bytes;
            return p;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        drain();  // retry once with the cache returned to the device
        LIG_HIP(hipMalloc(&p, bytes));
    }
    live += bytes;
    return p;
}

void Pool::put(void* p, size_t bytes) {
    std::lock_guard<std::mutex> g(mu);
    free_.emplace(bytes, p);
    live -= bytes;
    cached += bytes;
}

void Pool::drain() {
    std::lock_guard<std::mutex> g(mu);
    for (auto& kv : free_) (void)hipFree(kv.second);
    free_.clear();
    cached = 0;
}

const char* phase_name(int p) {
    static const char* names[PH_COUNT] = {"encode0", "encode", "merkle", "fold", "induce", "sumcheck", "open", "pass"};
    return (p >= 0 && p < PH_COUNT) ? names[p] : "";
}

void Ctx::collect() {
    if (marks.empty()) return;
    sync();
    for (auto& mk : marks) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, mk.a, mk.b) == hipSuccess) {
            phase_ms[mk.phase] += ms;
            phase_n[mk.phase] += 1;
        }
        (void)hipEventDestroy(mk.a);
        (void)hipEventDestroy(mk.b);
    }
    marks.clear();
}

// A plain hipMemcpyAsync from pageable memory moves about 30 GB/s on the
// MI355X host (the runtime stages it); from pinned memory the DMA engine
// moves 57 GB/s, and four host threads copy pageable into pinned memory at
// 65-100 GB/s.  So chunk c of a large copy goes to slot c % UP_SLOTS of a
// pinned ring, filled by thread c % UP_THREADS (UP_SLOTS is a multiple of
// UP_THREADS, so a slot always belongs to one thread), which then queues
// the slot's DMA and records the slot's event; a slot is refilled once its
// event, recorded after its previous DMA, has completed.  Like every use of
// a context, one host thread at a time.  Measured with a 2^24 prove from a
// host-resident input: 3.29 ms per call with the plain copy, 2.62 ms with
// four 4 MiB slots (2.69 with eight 2 MiB slots, 2.65 with four 8 MiB).
void Ctx::upload(void* dst, const void* src, size_t bytes) {
    static_assert(UP_SLOTS % UP_THREADS == 0, "a slot belongs to one thread");
    bool pinned_src = false;
    if (bytes >= 4 * UP_SLOT_BYTES) {
        // pinned (or device) memory goes straight to the DMA engine
        hipPointerAttribute_t at;
        if (hipPointerGetAttributes(&at, src) == hipSuccess)
            pinned_src = at.type == hipMemoryTypeHost || at.type == hipMemoryTypeDevice;
        else
            (void)hipGetLastError();  // pageable memory is not known to the runtime
    }
    if (bytes < 4 * UP_SLOT_BYTES || pinned_src) {
        LIG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, stream));
        return;
    }
    if (!up_ring) {
        LIG_HIP(hipHostMalloc((void**)&up_ring, UP_SLOTS * UP_SLOT_BYTES, hipHostMallocDefault));
        for (auto& e : up_ev) LIG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    const size_t chunks = (bytes + UP_SLOT_BYTES - 1) / UP_SLOT_BYTES;
    std::exception_ptr err;
    std::mutex err_mu;
    auto body = [&](int t) {
        try {
            for (size_t c = (size_t)t; c < chunks; c += UP_THREADS) {
                const int slot = (int)(c % UP_SLOTS);
                uint8_t* stage = up_ring + (size_t)slot * UP_SLOT_BYTES;
                const size_t at = c * UP_SLOT_BYTES, len = std::min(UP_SLOT_BYTES, bytes - at);
                // the slot's previous DMA (of this copy, or of an earlier
                // one; a never recorded event counts as complete)
                LIG_HIP(hipEventSynchronize(up_ev[slot]));
                memcpy(stage, (const uint8_t*)src + at, len);
                LIG_HIP(hipMemcpyAsync((uint8_t*)dst + at, stage, len, hipMemcpyHostToDevice, stream));
                LIG_HIP(hipEventRecord(up_ev[slot], stream));
            }
        } catch (...) {
            std::lock_guard<std::mutex> g(err_mu);
            if (!err) err = std::current_exception();
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < UP_THREADS; t++) th.emplace_back(body, t);
    body(0);
    for (auto& x : th) x.join();
    if (err) std::rethrow_exception(err);
}

// ---------------------------------------------------------------- RS plan --
template <class H> void RsPlan<H>::init(Ctx* c, size_t msg_len, size_t block_len) {
    using F = typename DevField<H>::F;
    LIG_REQUIRE(is_pow2(msg_len) && is_pow2(block_len) && msg_len < block_len, LIG_E_INVALID,
                "ReedSolomonEncoding::new: lengths must be powers of two with message < block");
    ctx = c;
    log_msg = ilog2(msg_len);
    log_block = ilog2(block_len);
    tw = compute_twiddles<H>(log_block);
    Vec<H> sks = eval_sk_at_vks<H>(msg_len);
    pis = compute_pis<H>(msg_len, sks);
    hipStream_t s = ctx->stream;
    const size_t M = block_len, m = msg_len;
    d_tw.alloc(M - 1);
    LIG_HIP(hipMemcpyAsync(d_tw.p, tw.data(), (M - 1) * sizeof(T), hipMemcpyHostToDevice, s));
    // per-twiddle multiply tables: only the generic passes (m == 1) read
    // them; the register passes build theirs from d_tw in LDS
    if (m == 1) {
        d_tw_tabs.alloc((M - 1) * F::TAB);
        launch_build_tables<F>(d_tw.p, M - 1, d_tw_tabs.p, s);
    }
    d_pis.alloc(m);
    LIG_HIP(hipMemcpyAsync(d_pis.p, pis.data(), m * sizeof(T), hipMemcpyHostToDevice, s));
    size_t nlo = std::min<size_t>(m, 64), nhi = std::max<size_t>(m / 64, 1);
    d_pis_lo.alloc(nlo * F::TAB);
    launch_build_tables<F>(d_pis.p, nlo, d_pis_lo.p, s);
    Vec<H> hi(nhi);
    for (size_t h = 0; h < nhi; h++) hi[h] = pis[h * 64 < m ? h * 64 : 0];
    DevBuf<T> d_hi(nhi);
    LIG_HIP(hipMemcpyAsync(d_hi.p, hi.data(), nhi * sizeof(T), hipMemcpyHostToDevice, s));
    d_pis_hi.alloc(nhi * F::TAB);
    launch_build_tables<F>(d_hi.p, nhi, d_pis_hi.p, s);
    LIG_HIP(hipGetLastError());
    ctx->sync();
}

// reed_solomon.rs:104-113 over n columns: pis scaling, then the full FFT
// (including the first two stages, which are replication only when the
// tail beyond m is zero).
template <class H> void RsPlan<H>::encode_non_systematic_cols(T* d_cols, size_t n, T* scratch) {
    using F = typename DevField<H>::F;
    LIG_REQUIRE(is_pow2(n), LIG_E_INVALID, "encode_non_systematic: the column count must be a power of two");
    hipStream_t s = ctx->stream;
    const size_t M_ = M(), m_ = m();
    // the FFT kernels work on the row-major [M][n] matrix; one column is
    // that matrix already
    PBuf<T> own;
    T* rm = d_cols;
    if (n > 1) {
        if (!scratch) {
            own = PBuf<T>(&ctx->pool, M_ * n);
            scratch = own.p;
        }
        rm = scratch;
        launch_transpose<F>(d_cols, n, M_, rm, s);  // column-major [n][M] -> [M][n]
    }
    launch_scale_rows<F>(rm, m_, n, d_pis.p, s);
    // two leading stages with the general butterfly, then the quarter passes
    int stages = log_block < 2 ? log_block : 2;
    for (int st = 0; st < stages; st++) launch_fft_stage<F>(rm, M_, n, st, d_tw.p, 0, s);
    // a shape already tuned by a commit of n columns; never tunes here (the
    // no-allocation entry point runs this too)
    const auto kt = tunes.find(n);
    const PassTune tune = kt == tunes.end() ? PassTune{} : kt->second;
    if (log_block >= 2) launch_fft_quarters<F>(nullptr, rm, M_ / 4, n, d_tw.p, d_tw_tabs.p, s, nullptr, tune);
    if (n > 1) launch_transpose<F>(rm, M_, n, d_cols, s);
    LIG_HIP(hipGetLastError());
    // the pooled buffer returns to the pool after the kernels that use it
    // (stream order); the caller's scratch stays busy until the stream
    // reaches this point
}

// ---------------------------------------------------------------- autotune --
static bool autotune_enabled() {
    static const bool on = [] {
        const char* e = getenv("LIG_AUTOTUNE");
        return !(e && e[0] == '0');
    }();
    return on;
}

// Choices shared by every plan of the process, per (device, field, message
// length, columns): a shape is timed once, by the first plan that encodes
// it (batch and serve workers each own plans of the same shapes).
namespace {
std::mutex g_tune_mu;
std::map<std::tuple<int, int, size_t, size_t>, PassTune> g_tunes;
}  // namespace

// Times every candidate launch shape of the FFT passes for n columns (one
// warm-up and three timed encodes of scratch buffers each, best of three)
// and keeps the fastest.  The candidates differ only in how the passes are
// launched, never in the codeword.
template <class H> PassTune RsPlan<H>::tune_for(size_t n) {
    using F = typename DevField<H>::F;
    auto it = tunes.find(n);
    if (it != tunes.end()) return it->second;
    const size_t m_ = m();
    const auto key = std::make_tuple(ctx->device, (int)F::BITS, m_, n);
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto gt = g_tunes.find(key);
        if (gt != g_tunes.end()) return tunes[n] = gt->second;
    }
    PassTune best{};
    // small encodes take the defaults: their passes are latency-bound and
    // timing them costs more than any shape saves
    if (!autotune_enabled() || m_ < 64 || m_ * n < ((size_t)1 << 18)) {
        tunes[n] = best;
        return best;
    }
    std::vector<PassTune> cands;
    if constexpr (F::BITS == 32) {
        for (int k : {6, 5, 4})
            for (int g : {1, 2, 4, 8}) cands.push_back({k, g});
    } else {
        for (int k : {3, 2})
            for (int g : {4, 8, 16}) cands.push_back({k, g});
    }
    hipStream_t s = ctx->stream;
    PBuf<T> msg(&ctx->pool, m_ * n), mat(&ctx->pool, 4 * m_ * n);
    hipEvent_t a, b;
    LIG_HIP(hipEventCreate(&a));
    LIG_HIP(hipEventCreate(&b));
    float best_ms = 0.f;
    try {
        for (const PassTune& c : cands) {
            launch_fft_quarters<F>(msg.p, mat.p, m_, n, d_tw.p, d_tw_tabs.p, s, nullptr, c);
            float t_best = 0.f;
            for (int rep = 0; rep < 3; rep++) {
                LIG_HIP(hipEventRecord(a, s));
                launch_fft_quarters<F>(msg.p, mat.p, m_, n, d_tw.p, d_tw_tabs.p, s, nullptr, c);
                LIG_HIP(hipEventRecord(b, s));
                LIG_HIP(hipEventSynchronize(b));
                float ms = 0.f;
                LIG_HIP(hipEventElapsedTime(&ms, a, b));
                if (rep == 0 || ms < t_best) t_best = ms;
            }
            if (best_ms == 0.f || t_best < best_ms) {
                best_ms = t_best;
                best = c;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        throw;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    LIG_HIP(hipGetLastError());
    tunes[n] = best;
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tunes.emplace(key, best);
    return best;
}

// reed_solomon.rs:19-41
template <class H> static Vec<H> short_from_long(const Vec<H>& l, int n, int k) {
    Vec<H> s(((size_t)1 << k) - 1);
    size_t jump = (size_t)1 << (n - k);
    s[0] = l[jump - 1];
    size_t idx = 2;
    for (int i = 1; i < k; i++) {
        jump *= 2;
        size_t take = (size_t)1 << i;
        for (size_t j = 0; j < take; j++) s[idx + j - 1] = l[jump + j - 1];
        idx += take;
    }
    return s;
}

// reed_solomon.rs:90-102: IFFT of the message with the short twiddles, then
// the FFT of the zero-padded block.
template <class H> void RsPlan<H>::encode_systematic(const T* msg, T* out) {
    using F = typename DevField<H>::F;
    hipStream_t s = ctx->stream;
    const size_t M_ = M(), m_ = m();
    Vec<H> stw = short_from_long<H>(tw, log_block, log_msg);
    PBuf<T> buf(&ctx->pool, M_), dstw(&ctx->pool, std::max<size_t>(stw.size(), 1));
    LIG_HIP(hipMemsetAsync(buf.p, 0, M_ * sizeof(T), s));
    LIG_HIP(hipMemcpyAsync(buf.p, msg, m_ * sizeof(T), hipMemcpyHostToDevice, s));
    if (!stw.empty()) LIG_HIP(hipMemcpyAsync(dstw.p, stw.data(), stw.size() * sizeof(T), hipMemcpyHostToDevice, s));
    // binary_fft.rs:161-178: the recursive IFFT applies the deepest stage first
    for (int st = log_msg - 1; st >= 0; st--) launch_fft_stage<F>(buf.p, m_, 1, st, dstw.p, 1, s);
    for (int st = 0; st < log_block; st++) launch_fft_stage<F>(buf.p, M_, 1, st, d_tw.p, 0, s);
    LIG_HIP(hipGetLastError());
    LIG_HIP(hipMemcpyAsync(out, buf.p, M_ * sizeof(T), hipMemcpyDeviceToHost, s));
    ctx->sync();
}

// ----------------------------------------------------------------- commit --
template <class H>
void commit(RsPlan<H>& rs, const typename H::T* d_poly, size_t m, size_t n, Witness<H>& w, int encode_phase) {
    using F = typename DevField<H>::F;
    using T = typename H::T;
    LIG_REQUIRE(rs.m() == m && rs.M() == 4 * m, LIG_E_INVALID, "ligero_commit: plan must encode m -> 4m");
    LIG_REQUIRE(is_pow2(n), LIG_E_INVALID, "ligero_commit: the column count n must be a power of two");
    LIG_REQUIRE(n * sizeof(T) <= 256 * 1024, LIG_E_INVALID, "ligero_commit: leaves above 256 KiB are not supported");
    Ctx* ctx = rs.ctx;
    hipStream_t s = ctx->stream;
    const size_t M = 4 * m;
    const PassTune tune = rs.tune_for(n);  // times the candidates on the first commit of n columns
    w.ctx = ctx;
    w.m = m;
    w.n = n;
    w.M = M;
    uint32_t seq = 0;
    {
        PBuf<T> msg(&ctx->pool, m * n);
        w.mat = PBuf<T>(&ctx->pool, M * n);
        // times the FFT pass launches of the first commit one by one
        struct Pass64Timer final : PassObserver {
            Ctx* ctx;
            hipEvent_t a = nullptr;
            explicit Pass64Timer(Ctx* c) : ctx(c) {}
            void before(int radix_log, size_t) override {
                LIG_HIP(hipEventCreate(&a));
                LIG_HIP(hipEventRecord(a, ctx->stream));
            }
            void after(int radix_log, size_t) override {
                if (!a) return;
                hipEvent_t b = nullptr;
                LIG_HIP(hipEventCreate(&b));
                LIG_HIP(hipEventRecord(b, ctx->stream));
                ctx->marks.push_back({PH_PASS, a, b});
                a = nullptr;
            }
        } p64(ctx);
        const bool time_passes = ctx->profiling && encode_phase == PH_ENCODE0;
        PhaseTimer pt(ctx, encode_phase);
        launch_transpose_scale<F>(d_poly, m, n, rs.d_pis.p, rs.d_pis_lo.p, rs.d_pis_hi.p, msg.p, s);
        launch_fft_quarters<F>(msg.p, w.mat.p, m, n, rs.d_tw.p, rs.d_tw_tabs.p, s, time_passes ? &p64 : nullptr, tune);
        // msg returns to the pool; stream order keeps the FFT's reads first
    }
    w.tree = PBuf<uint32_t>(&ctx->pool, (2 * M - 1) * 8);
    {
        PhaseTimer pt(ctx, PH_MERKLE);
        PBuf<uint32_t> scratch(&ctx->pool, leaf_hash_scratch_words(M, n * sizeof(T)));
        uint8_t* pin = ctx->pinned(256);
        (void)pin;
        seq = ctx->next_flag();
        launch_merkle_commit<F>(w.mat.p, M, n, w.tree.p, scratch.n ? scratch.p : nullptr, s, (uint32_t*)ctx->pin_dev,
                                ctx->flag_dev(), seq);
    }
    LIG_HIP(hipGetLastError());
    ctx->wait_flag(seq);  // the last tree kernel wrote the root to pinned offset 0, then the flag
    memcpy(w.root, ctx->pin, 32);
}

// The first commits of several polys of one shape, each kernel launched once
// for all of them: messages [count][m][n] from a table of input pointers,
// codewords [count][4m][n] (the FFT passes see 4*count quarters), trees of
// the standard layout tree_stride words apart, roots to pinned memory.
template <class H>
void commit_group(RsPlan<H>& rs, const typename H::T* const* d_polys, int count, size_t m, size_t n,
                  std::vector<Witness<H>>& out, std::shared_ptr<void>& keep) {
    using F = typename DevField<H>::F;
    using T = typename H::T;
    LIG_REQUIRE(F::BITS == 32, LIG_E_INVALID, "grouped commits are GF(2^32) only");
    LIG_REQUIRE(count >= 1 && count <= 4096, LIG_E_INVALID, "grouped commits: 1..4096 polys");
    LIG_REQUIRE(rs.m() == m && rs.M() == 4 * m && m >= 64 && n >= 4 && is_pow2(n), LIG_E_INVALID,
                "grouped commits: plan m -> 4m, m >= 64, n >= 4 a power of two");
    Ctx* ctx = rs.ctx;
    hipStream_t s = ctx->stream;
    const size_t M = 4 * m, tstride = (2 * M - 1) * 8;
    const PassTune tune = rs.tune_for(n);
    struct Group {
        PBuf<T> mat;
        PBuf<uint32_t> tree;
    };
    std::shared_ptr<Group> g = std::make_shared<Group>();
    g->mat = PBuf<T>(&ctx->pool, (size_t)count * M * n);
    g->tree = PBuf<uint32_t>(&ctx->pool, (size_t)count * tstride);
    // pinned: the pointer table at PIN_PTRS, the roots after it
    constexpr size_t PIN_PTRS = 4096;
    uint8_t* pin = ctx->pinned(PIN_PTRS + (size_t)count * (8 + 32));
    memcpy(pin + PIN_PTRS, d_polys, (size_t)count * sizeof(const T*));
    uint32_t* roots_dev = reinterpret_cast<uint32_t*>(ctx->pin_dev + PIN_PTRS + (size_t)count * 8);
    PBuf<const T*> d_ptrs(&ctx->pool, (size_t)count);
    LIG_HIP(hipMemcpyAsync(d_ptrs.p, pin + PIN_PTRS, (size_t)count * sizeof(const T*), hipMemcpyHostToDevice, s));
    {
        PBuf<T> msg(&ctx->pool, (size_t)count * m * n);
        launch_transpose_scale<F>(d_polys[0], m, n, rs.d_pis.p, rs.d_pis_lo.p, rs.d_pis_hi.p, msg.p, s, d_ptrs.p, count);
        launch_fft_quarters<F>(msg.p, g->mat.p, m, n, rs.d_tw.p, rs.d_tw_tabs.p, s, nullptr, tune, count);
        PBuf<uint32_t> scratch(&ctx->pool, leaf_hash_scratch_words(M, n * sizeof(T)));
        launch_merkle_commit<F>(g->mat.p, (size_t)count * M, n, g->tree.p, scratch.n ? scratch.p : nullptr, s, roots_dev, nullptr,
                                0, count, tstride);
        const uint32_t seq = ctx->next_flag();
        launch_signal(ctx->flag_dev(), seq, s);
        LIG_HIP(hipGetLastError());
        ctx->wait_flag(seq);  // the roots are in pinned memory; msg and scratch return to the pool
    }
    out.clear();
    out.resize((size_t)count);
    for (int i = 0; i < count; i++) {
        Witness<H>& w = out[(size_t)i];
        w.ctx = ctx;
        w.m = m;
        w.n = n;
        w.M = M;
        w.mat = PBuf<T>::view(g->mat.p + (size_t)i * M * n, M * n);
        w.tree = PBuf<uint32_t>::view(g->tree.p + (size_t)i * tstride, tstride);
        memcpy(w.root, ctx->pin + PIN_PTRS + (size_t)count * 8 + (size_t)i * 32, 32);
    }
    keep = g;
}

// ----------------------------------------------------------------- config --
Shape Shape::generated(int L) {
    LIG_REQUIRE(L >= 16 && L <= 27, LIG_E_INVALID, "prover config: log_n must be in [16, 27]");
    Shape sh;
    // config.rs:19-28 / 79-90: one recursive step up to 2^23 coefficients
    // (cfg20), two at 2^24 (cfg24), one more per further 2^4
    sh.recursive_steps = L <= 23 ? 1 : 2 + (L - 24) / 4;
    sh.initial_k = 6;
    sh.initial_m = (size_t)1 << (L - 6);
    sh.initial_n = 64;
    for (int i = 0; i < sh.recursive_steps; i++) {
        sh.dims_m.push_back((size_t)1 << (L - 6 - 4 * (i + 1)));
        sh.dims_n.push_back(16);
        sh.ks.push_back(4);
    }
    return sh;
}

Shape Shape::from_logs(int recursive_steps, int initial_dim, const int* log_dims, int initial_k, const int* ks) {
    LIG_REQUIRE(recursive_steps >= 1 && recursive_steps <= 8, LIG_E_INVALID,
                "recursive_steps must be in [1, 8] (prover.rs:198 panics at 0)");
    LIG_REQUIRE(initial_dim >= 0 && initial_dim < 40 && initial_k >= 0 && initial_k < 40, LIG_E_INVALID,
                "configuration dimension out of range");
    Shape sh;
    sh.recursive_steps = recursive_steps;
    sh.initial_m = (size_t)1 << initial_dim;
    sh.initial_k = initial_k;
    sh.initial_n = (size_t)1 << initial_k;
    for (int i = 0; i < recursive_steps; i++) {
        LIG_REQUIRE(log_dims[i] >= 0 && log_dims[i] < 40 && ks[i] >= 0 && ks[i] < 40, LIG_E_INVALID,
                    "configuration dimension out of range");
        sh.dims_m.push_back((size_t)1 << log_dims[i]);
        sh.dims_n.push_back((size_t)1 << ks[i]);
        sh.ks.push_back(ks[i]);
    }
    return sh;
}

void Shape::validate(int field_bytes) const {
    auto fail = [](const char* what) { throw Error(LIG_E_INVALID, std::string("prover config: ") + what); };
    if (recursive_steps < 1 || recursive_steps > 8) fail("recursive_steps must be in [1, 8] (prover.rs:198 panics at 0)");
    if ((int)dims_m.size() != recursive_steps || (int)dims_n.size() != recursive_steps || (int)ks.size() != recursive_steps)
        fail("dims and ks need one entry per recursive step");
    if (!is_pow2(initial_m) || !is_pow2(initial_n)) fail("dimensions must be powers of two (ReedSolomonEncoding::new)");
    // partial_eval must reduce the polynomial to initial_dims.0 rows
    // (prover.rs:50-53); evaluate_lagrange_basis needs a challenge (utils.rs:9)
    if (initial_k < 1 || initial_n != ((size_t)1 << initial_k)) fail("initial_k must be log2(initial_dims.1) >= 1");
    if (initial_m != dims_m[0] * dims_n[0]) fail("initial_dims.0 must equal dims[0].0 * dims[0].1");
    for (int i = 0; i < recursive_steps; i++) {
        if (!is_pow2(dims_m[i]) || !is_pow2(dims_n[i])) fail("dimensions must be powers of two (ReedSolomonEncoding::new)");
        // each step's rounds fold the column variables away (prover.rs:97-102)
        if (ks[i] < 1 || dims_n[i] != ((size_t)1 << ks[i])) fail("ks[i] must be log2(dims[i].1) >= 1");
        if (i + 1 < recursive_steps && dims_m[i] != dims_m[i + 1] * dims_n[i + 1])
            fail("dims[i].0 must equal dims[i+1].0 * dims[i+1].1");
        // get_distinct_queries(4m, 148) never ends below 148 rows (emulated_fs.rs:65)
        if (4 * dims_m[i] < 148) fail("every encoded matrix needs at least 148 rows (4 * dims[i].0 >= 148)");
        if (dims_n[i] * (size_t)field_bytes > 4096) fail("rows above 4 KiB are not supported");
    }
    if (4 * initial_m < 148) fail("every encoded matrix needs at least 148 rows (4 * initial_dims.0 >= 148)");
    if (initial_n * (size_t)field_bytes > 4096) fail("rows above 4 KiB are not supported");
    // device limits of this library: 2^27 coefficients, FFT quarters below
    // 2 GiB, the final message (yr) within its 2 MiB staging region
    if (num_coeffs() > ((size_t)1 << 27)) fail("at most 2^27 coefficients are supported");
    if (initial_m * initial_n * (size_t)field_bytes >= ((size_t)2 << 30)) fail("FFT quarters must stay below 2 GiB");
    if (dims_m.back() * (size_t)field_bytes > ((size_t)2 << 20)) fail("the final message must stay within 2 MiB");
}

template <class H> void Config<H>::init(Ctx* c, const Shape& sh) {
    sh.validate(H::BYTES);
    ctx = c;
    log_n = ilog2(sh.num_coeffs());
    recursive_steps = sh.recursive_steps;
    initial_k = sh.initial_k;
    initial_m = sh.initial_m;
    initial_n = sh.initial_n;
    dims_m = sh.dims_m;
    dims_n = sh.dims_n;
    ks = sh.ks;
    initial_rs.reset(new RsPlan<H>());
    initial_rs->init(c, initial_m, 4 * initial_m);
    rs.clear();
    for (int i = 0; i < recursive_steps; i++) {
        rs.emplace_back(new RsPlan<H>());
        rs.back()->init(c, dims_m[i], 4 * dims_m[i]);
    }
    // the inductions run over the first fold's 2^ell entries (ell =
    // log2 initial_m) and over each non-final step's folded message
    // (log2 dims_m[i], prover.rs:145-195)
    sks_by_ell.assign(ilog2(initial_m) + 1, Vec<H>());
    sks_by_ell[ilog2(initial_m)] = eval_sk_at_vks<H>(initial_m);
    for (int i = 0; i + 1 < recursive_steps; i++) sks_by_ell[ilog2(dims_m[i])] = eval_sk_at_vks<H>(dims_m[i]);
}

// ----------------------------------------------------------------- prover --
namespace {


// Openings (prover.rs:72-76 / 114-125, merkle_tree.rs:98-111) whose copies
// to pinned host memory are in flight on the context stream.  Pinned layout
// from the region's offset: query indices | opened rows | Merkle node
// indices | Merkle nodes, the node regions sized for the most nodes the
// queries can need (one per query and tree level).
template <class H> struct PendingOpen {
    Rlp<H>* out = nullptr;
    // query indices in pinned memory, read by the kernels through the
    // mapping (valid until the stream is next synchronized)
    const uint64_t* d_q = nullptr;
    PBuf<typename H::T> d_rows;
    // the row gather, launched with the induce constants
    RowGather<typename DevField<H>::F> gather;
    size_t n_nodes = 0, max_nodes = 0, nodes_off = 0;
    uint8_t* h_rows = nullptr;
    uint8_t* h_nodes = nullptr;
};

constexpr size_t pin_round(size_t b) { return (b + 255) & ~(size_t)255; }

// The opened rows: query indices to pinned memory and the row gather set up
// (induce_launch launches it with the induce constants; its output feeds the
// enforced sum).  open_launch_nodes adds the Merkle proof.
template <class H>
void open_launch_rows(Witness<H>& w, const std::vector<size_t>& q, Rlp<H>& out, PendingOpen<H>& po, size_t pin_off,
                      size_t pin_end) {
    using T = typename H::T;
    Ctx* ctx = w.ctx;
    po.out = &out;
    const size_t q_bytes = pin_round(q.size() * 8), rows_bytes = pin_round(q.size() * w.n * sizeof(T));
    po.max_nodes = q.size() * (size_t)ilog2(w.M);
    po.nodes_off = pin_off + q_bytes + rows_bytes;
    const size_t end = po.nodes_off + pin_round(po.max_nodes * 8) + po.max_nodes * 32;
    LIG_REQUIRE(end <= pin_end, LIG_E_INVALID, "opening exceeds its staging region");
    uint8_t* pin = ctx->pinned(end);
    uint64_t* h_idx = (uint64_t*)(pin + pin_off);
    for (size_t i = 0; i < q.size(); i++) h_idx[i] = q[i];
    po.d_q = reinterpret_cast<const uint64_t*>(ctx->pin_dev + pin_off);
    po.d_rows = PBuf<T>(&ctx->pool, q.size() * w.n);
    po.h_rows = pin + pin_off + q_bytes;
    // the gather writes the proof's rows straight to pinned memory too
    po.gather.mat = w.mat.p;
    po.gather.n = w.n;
    po.gather.out = po.d_rows.p;
    po.gather.out2 = (T*)(ctx->pin_dev + pin_off + q_bytes);
    out.n_rows = q.size();
    out.n_cols = w.n;
}

// The Merkle proof of the opened rows (merkle_tree.rs:98-111): node indices
// on the host, the gather of the nodes into pinned memory on the stream.
template <class H>
void open_launch_nodes(Witness<H>& w, const std::vector<size_t>& q, PendingOpen<H>& po, const WordCopy* extra = nullptr) {
    Ctx* ctx = w.ctx;
    std::vector<uint64_t> nodes = merkle_proof_nodes(w.M, q);
    LIG_REQUIRE(nodes.size() <= po.max_nodes, LIG_E_ASSERT, "Merkle proof needs more nodes than queries x depth");
    po.n_nodes = nodes.size();
    const size_t nodes_at = po.nodes_off + pin_round(po.max_nodes * 8);
    memcpy(ctx->pin + po.nodes_off, nodes.data(), nodes.size() * 8);
    po.h_nodes = ctx->pin + nodes_at;
    PhaseTimer pt(ctx, PH_OPEN);
    launch_gather_nodes(w.tree.p, reinterpret_cast<const uint64_t*>(ctx->pin_dev + po.nodes_off), nodes.size(),
                        (uint32_t*)(ctx->pin_dev + nodes_at), ctx->stream, extra);
}

// An opening read on the host from a copy of the witness: the matrix
// (M rows of n elements) and the Merkle tree (2M - 1 nodes of 32 bytes,
// leaves first), as the gathers above read them on the device.
template <class H>
void open_host(const uint8_t* mat, const uint8_t* tree, size_t M, size_t n, const std::vector<size_t>& q, Rlp<H>& out) {
    using T = typename H::T;
    out.n_rows = q.size();
    out.n_cols = n;
    out.rows.resize(q.size() * n);
    for (size_t i = 0; i < q.size(); i++) {
        LIG_REQUIRE(q[i] < M, LIG_E_ASSERT, "query index beyond the witness");
        memcpy(out.rows.data() + i * n, mat + q[i] * n * sizeof(T), n * sizeof(T));
    }
    const std::vector<uint64_t> nodes = merkle_proof_nodes(M, q);
    out.nodes.resize(nodes.size() * 32);
    for (size_t i = 0; i < nodes.size(); i++) {
        LIG_REQUIRE(nodes[i] < 2 * M - 1, LIG_E_ASSERT, "Merkle node index beyond the tree");
        memcpy(out.nodes.data() + i * 32, tree + nodes[i] * 32, 32);
    }
}

// Call after the stream has been synchronized.
template <class H> void open_finish(PendingOpen<H>& po) {
    Rlp<H>& out = *po.out;
    out.rows.resize(out.n_rows * out.n_cols);
    memcpy(out.rows.data(), po.h_rows, out.rows.size() * sizeof(typename H::T));
    out.nodes.assign(po.h_nodes, po.h_nodes + po.n_nodes * 32);
}

// The host transcript in the device layout: pending bytes as big-endian
// words, zero past the pending length.
void export_transcript(const Transcript& fs, DevTranscript& d) {
    LIG_REQUIRE(fs.hasher.blen % 4 == 0, LIG_E_ASSERT, "transcript length is not a whole number of words");
    memset(&d, 0, sizeof(d));
    memcpy(d.h, fs.hasher.h, sizeof(d.h));
    for (uint32_t i = 0; i < fs.hasher.blen; i++) d.buf[i >> 2] |= (uint32_t)fs.hasher.buf[i] << (24 - 8 * (i & 3));
    d.blen = fs.hasher.blen;
    d.total_lo = (uint32_t)fs.hasher.total;
    d.total_hi = (uint32_t)(fs.hasher.total >> 32);
    d.counter = fs.counter;
}

// induce_sumcheck_poly (sumcheck_polys.rs:74-160) entirely on the device:
// basis polynomial into d_basis, enforced sum (unreduced) as *n_enf parts
// at d_enf.
template <class H>
void induce_launch(Ctx* ctx, int ell, const Vec<H>& sks, const PendingOpen<H>& po, size_t n_cols, const Vec<H>& gr,
                   const std::vector<size_t>& q, typename H::T alpha, PBuf<typename H::T>& d_basis,
                   typename DevField<H>::F::Acc* d_enf, int* n_enf, size_t pin_off, const Eval013In<typename DevField<H>::F>& ev) {
    using F = typename DevField<H>::F;
    using T = typename H::T;
    hipStream_t s = ctx->stream;
    const size_t nq = q.size();
    LIG_REQUIRE(gr.size() == n_cols, LIG_E_ASSERT, "opened row length != 2^k");
    LIG_REQUIRE(sks.size() == (size_t)ell + 1, LIG_E_ASSERT, "subspace constants of another size");
    // sks | gr | alpha in pinned memory, read by the kernels through the
    // mapping (valid until the stream is next synchronized)
    const size_t nvals = sks.size() + gr.size() + 1;
    LIG_REQUIRE(nvals * sizeof(T) <= 16384 - 256, LIG_E_INVALID, "induce constants exceed the staging region");
    T* h = (T*)ctx->pinned(pin_off + nvals * sizeof(T));
    h = (T*)((uint8_t*)h + pin_off);
    memcpy(h, sks.data(), sks.size() * sizeof(T));
    memcpy(h + sks.size(), gr.data(), gr.size() * sizeof(T));
    h[nvals - 1] = alpha;
    const T* d_vals = (const T*)(ctx->pin_dev + pin_off);
    const int clo = induce_low_bits(ell);
    PBuf<T> d_c(&ctx->pool, nq * ell), d_ap(&ctx->pool, nq);
    PBuf<T> d_P(&ctx->pool, nq << clo), d_K(&ctx->pool, nq << (ell - clo));
    d_basis = PBuf<T>(&ctx->pool, (size_t)1 << ell);
    {
        PhaseTimer pt(ctx, PH_INDUCE);
        EnforcedIn<F> e;
        e.rows = po.d_rows.p;
        e.ncols = (int)n_cols;
        e.gr = d_vals + sks.size();
        e.parts = d_enf;
        LIG_REQUIRE(nq <= (size_t)ENFORCED_MAX_PARTS, LIG_E_INVALID, "more queries than enforced-sum parts");
        *n_enf = launch_induce_all<F>(po.d_q, d_vals, d_vals + nvals - 1, (int)nq, ell, d_c.p, d_ap.p, po.gather, d_P.p,
                                      d_K.p, d_basis.p, s, e, &ev);
    }
    LIG_HIP(hipGetLastError());
}

}  // namespace

template <class H>
void prove(Config<H>& cfg, const typename H::T* d_poly, uint32_t fs_seed, Proof<H>& proof) {
    using F = typename DevField<H>::F;
    using T = typename H::T;
    using A = typename F::Acc;
    Ctx* ctx = cfg.ctx;
    hipStream_t s = ctx->stream;
    const size_t S = 148;  // prover.rs:30
    Transcript fs(fs_seed);

    proof = Proof<H>();
    // Pinned staging regions (disjoint, so nothing in flight touches a
    // region the host is refilling): [0,256) commit roots, [256,16K) uploads
    // of induce constants and fold weights, [16K,32K) device transcript
    // logs, [32K,1M) and [1M,2M) openings (alternating: the host copies one
    // step's opening out while the next step's gathers fill the other),
    // [2M,4M) yr, from 4M the last witness (matrix, then Merkle tree), whose
    // openings the host reads there.
    constexpr size_t PIN_UP = 256, PIN_FSDN = 16384, PIN_YR = (size_t)2 << 20, PIN_FINAL = (size_t)4 << 20;
    const size_t pin_open[3] = {32768, (size_t)1 << 20, (size_t)2 << 20};
    int open_slot = 0;
    const size_t final_M = 4 * cfg.dims_m.back(), final_mat_bytes = final_M * cfg.dims_n.back() * sizeof(T);
    ctx->pinned(PIN_FINAL + final_mat_bytes + (2 * final_M - 1) * 32);
    Witness<H> w0;
    commit<H>(*cfg.initial_rs, d_poly, cfg.initial_m, cfg.initial_n, w0, PH_ENCODE0);
    memcpy(proof.initial_root.data(), w0.root, 32);
    fs.absorb(w0.root, 32);

    Vec<H> pe0(cfg.initial_k);
    for (auto& r : pe0) r = fs.get_field<H>();
    Vec<H> gr0 = lagrange_basis<H>(pe0);

    // Sumcheck sections run on the device transcript: the host hands its
    // transcript over after the section's host-side steps, a chain of
    // kernels absorbs each triple and squeezes each challenge, and after the
    // next synchronisation the host replays the section from the logged
    // triples (checking that it squeezes the same challenges) to continue
    // from the identical state.
    int maxk = 0;
    for (int k : cfg.ks) maxk = std::max(maxk, k);
    // two transcript-state slots: a fold that runs the preceding step reads
    // one and writes the other (its workgroups read while workgroup 0 writes)
    PBuf<DevTranscript> d_fs(&ctx->pool, 2);
    int fsi = 0;
    // three triple accumulators in rotation: a triple kernel reads the latest
    // triple (res_cur, as the transcript input of a fold step), adds into
    // res_next (clear) and clears res_free; then the roles rotate, so
    // res_cur again holds the latest triple and res_next is clear; plus the
    // enforced sum
    PBuf<A> d_acc(&ctx->pool, 9 + ENFORCED_MAX_PARTS);
    A* res_cur = d_acc.p;
    A* res_next = d_acc.p + 3;
    A* res_free = d_acc.p + 6;
    A* d_enf = d_acc.p + 9;
    auto rotate = [&]() {
        A* t = res_cur;
        res_cur = res_next;
        res_next = res_free;
        res_free = t;
    };
    int n_enf = 0;
    // res_next and res_free are adjacent; the fold below zeroes them (or a
    // launch of its own), ahead of the recursive commit rather than after the
    // synchronisation that ends it

    // partial_eval over the top initial_k variables == row combination with gr
    const size_t N = cfg.initial_m * cfg.initial_n;
    size_t flen = N >> cfg.initial_k;
    PBuf<T> d_f(&ctx->pool, flen), d_f2(&ctx->pool, flen / 2);
    bool folded = false;
    if constexpr (F::BITS == 32) {
        // the weights travel as a kernel argument
        PhaseTimer pt(ctx, PH_FOLD);
        folded = launch_fold32_weights(reinterpret_cast<const uint32_t*>(d_poly), flen, cfg.initial_k,
                                       reinterpret_cast<const uint32_t*>(gr0.data()), reinterpret_cast<uint32_t*>(d_f.p), s,
                                       reinterpret_cast<uint64_t*>(res_next), 6);
        LIG_HIP(hipGetLastError());
    }
    if (!folded) {
        T* h = (T*)(ctx->pinned(PIN_UP + gr0.size() * sizeof(T)) + PIN_UP);
        memcpy(h, gr0.data(), gr0.size() * sizeof(T));
        PBuf<T> d_gr(&ctx->pool, gr0.size()), d_grt(&ctx->pool, gr0.size() * F::TAB);
        LIG_HIP(hipMemcpyAsync(d_gr.p, h, gr0.size() * sizeof(T), hipMemcpyHostToDevice, s));
        PhaseTimer pt(ctx, PH_FOLD);
        launch_build_tables<F>(d_gr.p, gr0.size(), d_grt.p, s);
        launch_fold<F>(d_poly, flen, cfg.initial_k, d_grt.p, d_f.p, s);
        launch_zero_acc<F>(res_next, 6, s);
        LIG_HIP(hipGetLastError());
    }


    // The last witness is copied to pinned memory on the side stream while
    // the last sumcheck section runs; the final openings then need no GPU
    // round trip.  The copies are issued only once the host has seen an
    // event recorded after the witness's commit complete: then every store
    // of the commit's kernels is visible to them by stream semantics, with
    // no device-side wait between the streams (a wait of the side stream
    // on the context stream's event held up the hardware queue the side
    // stream shares with other contexts' streams).
    const T* final_mat = nullptr;
    const uint32_t* final_tree = nullptr;
    auto stage_final = [&](const Witness<H>& w) {
        LIG_REQUIRE(w.M == final_M && w.M * w.n * sizeof(T) == final_mat_bytes, LIG_E_ASSERT, "last witness shape");
        LIG_HIP(hipEventRecord(ctx->side_event(), s));
        final_mat = w.mat.p;
        final_tree = w.tree.p;
    };
    // issues the copies if the commit is known complete (or, with block,
    // once it is)
    auto issue_final = [&](bool block) {
        if (!final_mat) return;
        if (!block) {
            const hipError_t e = hipEventQuery(ctx->side_event());
            if (e == hipErrorNotReady) return;
            LIG_HIP(e);
        } else {
            LIG_HIP(hipEventSynchronize(ctx->side_event()));
        }
        hipStream_t ss = ctx->side_stream();
        LIG_HIP(hipMemcpyAsync(ctx->pin + PIN_FINAL, final_mat, final_mat_bytes, hipMemcpyDeviceToHost, ss));
        LIG_HIP(hipMemcpyAsync(ctx->pin + PIN_FINAL + final_mat_bytes, final_tree, (2 * final_M - 1) * 32,
                               hipMemcpyDeviceToHost, ss));
        final_mat = nullptr;
    };

    Witness<H> wprev;
    commit<H>(*cfg.rs[0], d_f.p, cfg.dims_m[0], cfg.dims_n[0], wprev);  // synchronizes
    if (cfg.recursive_steps == 1) stage_final(wprev);
    proof.rec_roots.emplace_back();
    memcpy(proof.rec_roots.back().data(), wprev.root, 32);
    fs.absorb(wprev.root, 32);

    std::vector<size_t> q = fs.get_distinct_queries(w0.M, S);
    T alpha = fs.get_field<H>();
    int nv = ilog2(flen);

    PBuf<T> d_b, d_b2(&ctx->pool, flen / 2);

    struct Section {
        int n0 = 0, k = 0;  // squeezes of the leading step, rounds
    } sec;
    // Section logs go straight to pinned memory: triples, then squeezed
    // challenges, then the check flag.
    const size_t n_tlog = 3 * (size_t)(maxk + 1), n_r = (size_t)maxk + 2;
    uint8_t* h_fsdn = ctx->pin + PIN_FSDN;
    T* dev_tlog = (T*)(ctx->pin_dev + PIN_FSDN);
    T* dev_rlog = dev_tlog + n_tlog;
    uint32_t* dev_err = (uint32_t*)(dev_rlog + n_r);
    // a section's last triple, as raw accumulators (the host reduces them)
    constexpr size_t LAST_OFF = 8192;
    LIG_REQUIRE((n_tlog + n_r) * sizeof(T) + 4 <= LAST_OFF, LIG_E_INVALID, "section logs exceed their staging region");
    const A* h_last = (const A*)(ctx->pin + PIN_FSDN + LAST_OFF);
    A* dev_last = (A*)(ctx->pin_dev + PIN_FSDN + LAST_OFF);
    // Launches: k fold rounds, each running the transcript step before it:
    // round 0 the section's leading step (absorb the triple in res_cur from
    // the host's exported state, check it against the enforced sum, squeeze
    // n0 challenges, glue b + beta * bnew when n0 == 2), round j the step
    // after round j-1 (absorb its triple, squeeze round j's challenge); then
    // a last step absorbs the final triple.
    auto launch_section = [&](int n0, int k, PBuf<T>* d_bnew) {
        DevTranscript init;
        export_transcript(fs, init);
        *(uint32_t*)(h_fsdn + (n_tlog + n_r) * sizeof(T)) = 0;
        PhaseTimer pt(ctx, PH_SUMCHECK);
        for (int j = 0; j < k; j++) {
            FoldStep<F> step{};
            step.st_out = d_fs.p + (fsi ^ 1);
            step.res_in = res_cur;
            if (j == 0) {
                // the leading step: from the host's state, with the
                // enforced-sum check, n0 challenges and (n0 == 2) the glue
                step.lead = 1;
                step.st_in = nullptr;
                step.init = init;
                step.enf = d_enf;
                step.n_enf = n_enf;
                step.err = dev_err;
                step.n_squeeze = n0;
                step.bnew = d_bnew ? d_bnew->p : nullptr;
                step.tlog = dev_tlog;
                step.rlog = dev_rlog;
            } else {
                // the step after round j-1 (absorb its triple, squeeze r)
                step.st_in = d_fs.p + fsi;
                step.n_squeeze = 1;
                step.tlog = dev_tlog + 3 * j;
                step.rlog = dev_rlog + n0 + (j - 1);
            }
            launch_sumcheck_fold_step<F>(d_f.p, d_b.p, flen, step, d_f2.p, d_b2.p, res_next, res_free, s);
            fsi ^= 1;
            rotate();
            std::swap(d_f, d_f2);
            std::swap(d_b, d_b2);
            flen /= 2;
        }
        LIG_HIP(hipGetLastError());
        sec.n0 = n0;
        sec.k = k;
    };
    // The section's last triple needs no transcript step on the device (the
    // next section starts from the host's exported transcript, and the host
    // replays this one): its raw accumulators ride with the node gather
    // launched right after the section.
    auto last_triple_copy = [&]() {
        WordCopy c;
        c.src = reinterpret_cast<const uint32_t*>(res_cur);
        c.dst = reinterpret_cast<uint32_t*>(dev_last);
        c.n = (int)(3 * sizeof(A) / 4);
        return c;
    };
    Vec<H> fold_rs;  // challenges of the section's fold rounds
    // after the stream has been synchronized
    auto replay_section = [&](const char* what) {
        const T* tl = (const T*)h_fsdn;
        const T* rr = tl + n_tlog;
        LIG_REQUIRE(*(const uint32_t*)(rr + n_r) == 0, LIG_E_ASSERT, what);
        int ri = 0;
        fold_rs.clear();
        auto squeeze_check = [&]() {
            T v = fs.get_field<H>();
            LIG_REQUIRE(H::eq(v, rr[ri]), LIG_E_ASSERT, "device transcript diverged from the host transcript");
            if (ri >= sec.n0 - 1) fold_rs.push_back(v);
            ri++;
        };
        proof.tr.insert(proof.tr.end(), tl, tl + 3);
        fs.absorb_elems<H>(tl, 3);
        for (int i = 0; i < sec.n0; i++) squeeze_check();
        for (int j = 0; j < sec.k; j++) {
            const T* t3 = tl + 3 * (j + 1);
            T last3[3];
            if (j + 1 == sec.k) {
                for (int i = 0; i < 3; i++) last3[i] = acc_reduce_host<H>(h_last[i]);
                t3 = last3;
            }
            proof.tr.insert(proof.tr.end(), t3, t3 + 3);
            fs.absorb_elems<H>(t3, 3);
            if (j + 1 < sec.k) squeeze_check();
        }
    };

    // the GPU starts on the rows, induce and the sumcheck section before the
    // host lists the Merkle proof nodes
    std::unique_ptr<PendingOpen<H>> pending(new PendingOpen<H>());
    // pending openings point into rec_proofs until they are copied out
    proof.rec_proofs.reserve(cfg.recursive_steps);
    open_launch_rows<H>(w0, q, proof.initial, *pending, pin_open[0], pin_open[1]);
    {
        // the basis, and with it the 0/1/3 triple of (f, basis)
        Eval013In<F> ev;
        ev.f = d_f.p;
        ev.res = res_next;
        ev.res_clear = res_free;
        induce_launch<H>(ctx, nv, cfg.sks_by_ell.at(nv), *pending, w0.n, gr0, q, alpha, d_b, d_enf, &n_enf, PIN_UP, ev);
        rotate();
    }
    // SumcheckProverInstance::new (ligerito.rs:32-41) and the rounds of the
    // first step (prover.rs:96-102)
    launch_section(1, cfg.ks[0], nullptr);
    {
        const WordCopy lc = last_triple_copy();
        open_launch_nodes<H>(w0, q, *pending, &lc);
    }
    issue_final(false);
    w0 = Witness<H>();  // the first witness is released once its kernels ran

    for (int i = 0; i < cfg.recursive_steps; i++) {
        if (i == cfg.recursive_steps - 1) {
            // final step (prover.rs:104-142): yr, queries, openings
            issue_final(true);  // the last commit completed long ago
            T* hy = (T*)(ctx->pin + PIN_YR);
            LIG_REQUIRE(PIN_YR + flen * sizeof(T) <= ((size_t)4 << 20), LIG_E_INVALID, "yr exceeds the staging region");
            // written by a kernel straight into the mapped pinned buffer
            launch_copy_words((uint32_t*)(ctx->pin_dev + PIN_YR), (const uint32_t*)d_f.p, flen * (sizeof(T) / 4), s);
            const uint32_t seq = ctx->next_flag();
            launch_signal(ctx->flag_dev(), seq, s);
            LIG_HIP(hipGetLastError());
            ctx->wait_flag(seq);  // yr and the section's logs are in pinned memory
            replay_section("sumcheck: s0 + s1 != enforced sum (ligerito.rs:35,52)");
            proof.yr.assign(hy, hy + flen);
            fs.absorb_elems<H>(proof.yr.data(), flen);
            q = fs.get_distinct_queries(wprev.M, S);
            open_finish<H>(*pending);
            ctx->wait(ctx->side_stream());  // the last witness copy, long done
            open_host<H>(ctx->pin + PIN_FINAL, ctx->pin + PIN_FINAL + final_mat_bytes, wprev.M, wprev.n, q, proof.final_);
            return;
        }
        Witness<H> wnext;
        commit<H>(*cfg.rs[i + 1], d_f.p, cfg.dims_m[i + 1], cfg.dims_n[i + 1], wnext);  // synchronizes
        if (i + 1 == cfg.recursive_steps - 1) stage_final(wnext);
        replay_section("sumcheck: s0 + s1 != enforced sum (ligerito.rs:35,52)");
        proof.rec_roots.emplace_back();
        memcpy(proof.rec_roots.back().data(), wnext.root, 32);
        fs.absorb(wnext.root, 32);

        q = fs.get_distinct_queries(wprev.M, S);
        T a = fs.get_field<H>();
        nv = ilog2(flen);
        proof.rec_proofs.emplace_back();
        PBuf<T> d_bnew;
        std::unique_ptr<PendingOpen<H>> next(new PendingOpen<H>());
        open_slot ^= 1;
        open_launch_rows<H>(wprev, q, proof.rec_proofs.back(), *next, pin_open[open_slot], pin_open[open_slot + 1]);
        Vec<H> gr = lagrange_basis<H>(fold_rs);
        // introduce_new (ligerito.rs:50-57): the new basis and the triple of
        // f with it alone (from the same launches); then glue
        // (ligerito.rs:59-67) with the next challenge: the summed basis
        // absorbs beta * b_new (folds and the 0/1/3 evaluations are linear in
        // the basis polys); then the rounds of step i+1
        {
            Eval013In<F> ev;
            ev.f = d_f.p;
            ev.res = res_next;
            ev.res_clear = res_free;
            induce_launch<H>(ctx, nv, cfg.sks_by_ell.at(nv), *next, wprev.n, gr, q, a, d_bnew, d_enf, &n_enf, PIN_UP, ev);
            rotate();
        }
        launch_section(2, cfg.ks[i + 1], &d_bnew);
        {
            const WordCopy lc = last_triple_copy();
            open_launch_nodes<H>(wprev, q, *next, &lc);
        }
        issue_final(false);
        open_finish<H>(*pending);  // the previous opening, while this section runs
        pending = std::move(next);
        wprev = std::move(wnext);
    }
    throw Error(LIG_E_ASSERT, "Prover completed without entering final step (prover.rs:198)");
}

// ------------------------------------------------------------ gang prove --
// `count` proves of one GF(2^32) configuration in lockstep: each kernel is
// launched once for all of them (job j = blockIdx.y, or one table of
// pointers), and at each synchronisation the host runs every job's
// transcript steps.  A small prove is a chain of short latency-bound
// kernels (transcript steps, Merkle tops, small folds) that leave most of
// the GPU idle, and a device runs few kernels at once (one per hardware
// queue); here a chain serves the whole gang.  Each job follows prove()'s
// transcript schedule (prover.rs:29-195), so its proof bytes are the same.
//
// Witness levels: 0 is the first commit, l = i + 1 the commit of step i;
// level l is opened at queries over its 4 m rows (level 0 as the proof's
// first opening, levels 1..R-1 as its recursive openings, level R as the
// final one).  Pinned staging, from GANG_PIN (below it: the flag and
// commit_group's region), per job at fixed strides: fold input pointer,
// fold weights, query indices, induce constants, the section's initial
// transcript state and logs, yr, and per level the opened rows, the Merkle
// node indices and the nodes (sized by the queries' worst case).
template <class H>
void prove_gang(Config<H>& cfg, const typename H::T* const* d_polys, const uint32_t* seeds, int count, Proof<H>* proofs) {
    static_assert(std::is_same<H, H32>::value, "gang proves are GF(2^32)");
    using F = F32;
    using T = uint32_t;
    using A = uint64_t;
    const int R = cfg.recursive_steps;
    LIG_REQUIRE(count >= 1 && count <= GANG_MAX, LIG_E_INVALID, "gang prove: 1..GANG_MAX proves");
    LIG_REQUIRE(R >= 1 && cfg.initial_k >= 1 && cfg.initial_k <= FOLD32_MAXK && cfg.initial_n <= 64, LIG_E_INVALID,
                "gang prove: recursive steps, fold width");
    int maxk = 0;
    for (int k : cfg.ks) maxk = std::max(maxk, k);
    LIG_REQUIRE(maxk >= 1 && maxk <= 8, LIG_E_INVALID, "gang prove: fold rounds per step");
    for (int j = 0; j < count; j++) LIG_REQUIRE(d_polys[j], LIG_E_INVALID, "NULL polynomial");
    Ctx* ctx = cfg.ctx;
    hipStream_t s = ctx->stream;
    const size_t S = 148;  // prover.rs:30
    const size_t G = (size_t)count;
    std::vector<size_t> Mw(R + 1), nw(R + 1), depth(R + 1);
    Mw[0] = 4 * cfg.initial_m;
    nw[0] = cfg.initial_n;
    for (int i = 0; i < R; i++) {
        Mw[i + 1] = 4 * cfg.dims_m[i];
        nw[i + 1] = cfg.dims_n[i];
    }
    for (int l = 0; l <= R; l++) depth[l] = (size_t)ilog2(Mw[l]);
    const size_t flen0 = cfg.initial_m;  // N >> initial_k
    const size_t yr_len = cfg.dims_m.back();

    // pinned layout (bytes)
    constexpr size_t GANG_PIN = (size_t)64 << 10;
    constexpr size_t CONST_W = 80;  // per job: up to 64 row weights | alpha at word 64
    constexpr size_t LOG_W = 64;    // per job: triples | challenges | check flag; raw last triple at LAST_W
    constexpr size_t LAST_W = 40;
    const size_t n_tlog = 3 * (size_t)(maxk + 1), n_r = (size_t)maxk + 2;
    LIG_REQUIRE(n_tlog + n_r + 1 <= LAST_W, LIG_E_INVALID, "gang prove: section logs exceed their region");
    for (int l = 0; l < R; l++) LIG_REQUIRE(nw[l] <= 64, LIG_E_INVALID, "gang prove: rows of at most 64 elements");
    size_t off = GANG_PIN;
    auto region = [&](size_t bytes) {
        const size_t at = off;
        off = pin_round(off + bytes);
        return at;
    };
    const size_t P_PTR = region(G * 8), P_W = region(G * 64 * 4), P_Q = region(G * S * 8),
                 P_SKS = region(((size_t)ilog2(flen0) + 1) * 4), P_CONST = region(G * CONST_W * 4),
                 P_INIT = region(G * 2 * sizeof(DevTranscript)), P_LOG = region(G * LOG_W * 4),
                 P_YR = region(G * yr_len * 4), P_FQ = region(G * S * 8);
    std::vector<size_t> P_ROWS(R + 1), P_NIDX(R + 1), P_NODES(R + 1);
    for (int l = 0; l <= R; l++) {
        P_ROWS[l] = region(G * S * nw[l] * 4);
        P_NIDX[l] = region(G * S * depth[l] * 8);
        P_NODES[l] = region(G * S * depth[l] * 32);
    }
    uint8_t* pin = ctx->pinned(off);
    uint8_t* pdev = ctx->pin_dev;
    T* h_const = (T*)(pin + P_CONST);
    T* d_tlog = (T*)(pdev + P_LOG);

    // witnesses per level (each a view into its gang buffers, kept to the end)
    std::vector<std::vector<Witness<H>>> wl(R + 1);
    std::vector<std::shared_ptr<void>> keep(R + 1);
    commit_group<H>(*cfg.initial_rs, d_polys, count, cfg.initial_m, cfg.initial_n, wl[0], keep[0]);

    std::vector<Transcript> fs;
    fs.reserve(G);
    std::vector<Vec<H>> gr(G);
    T* h_w = (T*)(pin + P_W);
    for (size_t j = 0; j < G; j++) {
        proofs[j] = Proof<H>();
        fs.emplace_back(seeds[j]);
        memcpy(proofs[j].initial_root.data(), wl[0][j].root, 32);
        fs[j].absorb(wl[0][j].root, 32);
        Vec<H> pe0(cfg.initial_k);
        for (auto& r : pe0) r = fs[j].get_field<H>();
        gr[j] = lagrange_basis<H>(pe0);
        memcpy(h_w + j * 64, gr[j].data(), gr[j].size() * 4);
        ((const T**)(pin + P_PTR))[j] = d_polys[j];
    }

    // accumulators per job: triples in rotation at 0, 3, 6, enforced-sum
    // parts from 9
    constexpr size_t ACC = 9 + ENFORCED_MAX_PARTS;
    PBuf<A> d_acc(&ctx->pool, G * ACC);
    size_t rc = 0, rn = 3, rf = 6;  // res_cur, res_next, res_free
    auto rotate = [&]() {
        const size_t t = rc;
        rc = rn;
        rn = rf;
        rf = t;
    };
    // partial_eval of every job (zeroing res_next and res_free)
    PBuf<T> d_f(&ctx->pool, G * flen0), d_f2(&ctx->pool, G * flen0 / 2);
    PBuf<T> d_b(&ctx->pool, G * flen0), d_b2(&ctx->pool, G * flen0 / 2);
    size_t sf = flen0, sf2 = flen0 / 2, sb = flen0, sb2 = flen0 / 2, flen = flen0;
    {
        FoldJobs fj;
        fj.polys = (const T* const*)(pdev + P_PTR);
        fj.w = (const T*)(pdev + P_W);
        fj.out_stride = flen0;
        fj.zero_stride = ACC;
        launch_fold32_jobs(flen0, cfg.initial_k, fj, count, d_f.p, d_acc.p + rn, 6, s);
        LIG_HIP(hipGetLastError());
    }
    // the commit of level l >= 1: every job's folded polynomial (synchronizes)
    auto commit_level = [&](int l) {
        std::vector<const T*> fp(G);
        for (size_t j = 0; j < G; j++) fp[j] = d_f.p + j * sf;
        commit_group<H>(*cfg.rs[l - 1], fp.data(), count, cfg.dims_m[l - 1], cfg.dims_n[l - 1], wl[l], keep[l]);
    };
    commit_level(1);

    // after a commit of level l + 1: per job, absorb its root, draw level l's
    // queries and alpha, list level l's Merkle proof nodes, and export the
    // transcript the next section starts from; then the induction over
    // level l's opened rows (the basis into d_basis)
    std::vector<size_t> nodes_at(G + 1, 0);
    auto open_and_induce = [&](int l, PBuf<T>& d_basis) {
        uint64_t* h_q = (uint64_t*)(pin + P_Q);
        uint64_t* h_nidx = (uint64_t*)(pin + P_NIDX[l]);
        const int ell = ilog2(flen);
        const Vec<H>& sks = cfg.sks_by_ell.at(ell);
        memcpy(pin + P_SKS, sks.data(), sks.size() * 4);
        nodes_at.assign(G + 1, 0);
        for (size_t j = 0; j < G; j++) {
            proofs[j].rec_roots.emplace_back();
            memcpy(proofs[j].rec_roots.back().data(), wl[l + 1][j].root, 32);
            fs[j].absorb(wl[l + 1][j].root, 32);
            const std::vector<size_t> q = fs[j].get_distinct_queries(Mw[l], S);
            const T alpha = fs[j].get_field<H>();
            LIG_REQUIRE(q.size() == S && gr[j].size() == nw[l], LIG_E_ASSERT, "queries or opened row length");
            for (size_t i = 0; i < S; i++) h_q[j * S + i] = q[i];
            memcpy(h_const + j * CONST_W, gr[j].data(), gr[j].size() * 4);
            h_const[j * CONST_W + 64] = alpha;
            const std::vector<uint64_t> nd = merkle_proof_nodes(Mw[l], q);
            LIG_REQUIRE(nd.size() <= S * depth[l], LIG_E_ASSERT, "Merkle proof needs more nodes than queries x depth");
            nodes_at[j + 1] = nodes_at[j] + nd.size();
            for (size_t i = 0; i < nd.size(); i++) h_nidx[nodes_at[j] + i] = nd[i] + j * (2 * Mw[l] - 1);
            export_transcript(fs[j], *((DevTranscript*)(pin + P_INIT) + 2 * j));
            ((uint32_t*)(pin + P_LOG))[j * LOG_W + n_tlog + n_r] = 0;  // check flag
        }
        const int clo = induce_low_bits(ell);
        PBuf<T> d_c(&ctx->pool, G * S * ell), d_ap(&ctx->pool, G * S);
        PBuf<T> d_P(&ctx->pool, G * (S << clo)), d_K(&ctx->pool, G * (S << (ell - clo)));
        PBuf<T> d_rows(&ctx->pool, G * S * nw[l]);
        d_basis = PBuf<T>(&ctx->pool, G * flen);
        InduceJobs ij;
        ij.q = S;
        ij.alpha = CONST_W;
        ij.mat = Mw[l] * nw[l];
        ij.rows2 = S * nw[l];
        ij.gr = CONST_W;
        ij.parts = ACC;
        ij.f = sf;
        ij.res = ACC;
        RowGather<F> g;
        g.mat = wl[l][0].mat.p;  // the level's gang matrices, Mw[l] x nw[l] apart
        g.n = nw[l];
        g.out = d_rows.p;
        g.out2 = (T*)(pdev + P_ROWS[l]);
        const T* d_const = (const T*)(pdev + P_CONST);
        EnforcedIn<F> e;
        e.rows = d_rows.p;
        e.ncols = (int)nw[l];
        e.gr = d_const;
        e.parts = d_acc.p + 9;
        Eval013In<F> ev;
        ev.f = d_f.p;
        ev.res = d_acc.p + rn;
        ev.res_clear = d_acc.p + rf;
        const int n_parts = launch_induce_all<F>((const uint64_t*)(pdev + P_Q), (const T*)(pdev + P_SKS), d_const + 64, (int)S,
                                                 ell, d_c.p, d_ap.p, g, d_P.p, d_K.p, d_basis.p, s, e, &ev, count, ij);
        LIG_REQUIRE(n_parts == (int)S, LIG_E_ASSERT, "gang prove: one enforced-sum part per query");
        LIG_HIP(hipGetLastError());
        rotate();
    };

    // the sumcheck section of a step: k fold rounds, each running the
    // preceding transcript step per job on the device (the leading step with
    // n0 challenges and, with bnew, the glue)
    PBuf<DevTranscript> d_fs(&ctx->pool, 2 * G);
    int fsi = 0;
    auto section = [&](int n0, int k, const T* bnew, size_t bnew_stride) {
        for (int r = 0; r < k; r++) {
            FoldStep<F> step{};
            step.st_out = d_fs.p + (fsi ^ 1);
            step.res_in = d_acc.p + rc;
            if (r == 0) {
                step.lead = 1;
                step.st_in = (const DevTranscript*)(pdev + P_INIT);
                step.enf = d_acc.p + 9;
                step.n_enf = (int)S;
                step.err = (uint32_t*)(d_tlog + n_tlog + n_r);
                step.n_squeeze = n0;
                step.bnew = bnew;
                step.tlog = d_tlog;
                step.rlog = d_tlog + n_tlog;
            } else {
                step.st_in = d_fs.p + fsi;
                step.n_squeeze = 1;
                step.tlog = d_tlog + 3 * r;
                step.rlog = d_tlog + n_tlog + n0 + (r - 1);
            }
            SumcheckJobs sj;
            sj.f = sf;
            sj.b = sb;
            sj.f2 = sf2;
            sj.b2 = sb2;
            sj.bnew = bnew_stride;
            sj.acc = ACC;
            sj.st = 2;
            sj.log_elems = LOG_W;
            sj.log_words = LOG_W;
            launch_sumcheck_fold_step<F>(d_f.p, d_b.p, flen, step, d_f2.p, d_b2.p, d_acc.p + rn, d_acc.p + rf, s, count, sj);
            fsi ^= 1;
            rotate();
            std::swap(d_f, d_f2);
            std::swap(d_b, d_b2);
            std::swap(sf, sf2);
            std::swap(sb, sb2);
            flen /= 2;
        }
        // the last triple (raw accumulators) to each job's log
        launch_copy_blocks((uint32_t*)(d_tlog + LAST_W), LOG_W, (const uint32_t*)(d_acc.p + rc), ACC * 2, 6, count, s);
    };
    // after a synchronisation: per job, replay the section from its logs
    // (checking the device's challenges), keeping its fold challenges as the
    // next row weights, and copy level l's opening into the proof
    auto replay_and_open = [&](int n0, int k, int l) {
        for (size_t j = 0; j < G; j++) {
            Proof<H>& pr = proofs[j];
            const T* tl = (const T*)(pin + P_LOG) + j * LOG_W;
            const T* rr = tl + n_tlog;
            LIG_REQUIRE(*(const uint32_t*)(rr + n_r) == 0, LIG_E_ASSERT,
                        "sumcheck: s0 + s1 != enforced sum (ligerito.rs:35,52)");
            int ri = 0;
            Vec<H> fold_rs;
            auto squeeze_check = [&]() {
                const T v = fs[j].get_field<H>();
                LIG_REQUIRE(H::eq(v, rr[ri]), LIG_E_ASSERT, "device transcript diverged from the host transcript");
                if (ri >= n0 - 1) fold_rs.push_back(v);
                ri++;
            };
            pr.tr.insert(pr.tr.end(), tl, tl + 3);
            fs[j].absorb_elems<H>(tl, 3);
            for (int i = 0; i < n0; i++) squeeze_check();
            const A* last = (const A*)(tl + LAST_W);
            for (int r = 0; r < k; r++) {
                const T* t3 = tl + 3 * (r + 1);
                T last3[3];
                if (r + 1 == k) {
                    for (int i = 0; i < 3; i++) last3[i] = acc_reduce_host<H>(last[i]);
                    t3 = last3;
                }
                pr.tr.insert(pr.tr.end(), t3, t3 + 3);
                fs[j].absorb_elems<H>(t3, 3);
                if (r + 1 < k) squeeze_check();
            }
            gr[j] = lagrange_basis<H>(fold_rs);
            Rlp<H>* o = &pr.initial;
            if (l > 0) {
                pr.rec_proofs.emplace_back();
                o = &pr.rec_proofs.back();
            }
            o->n_rows = S;
            o->n_cols = nw[l];
            const T* rows = (const T*)(pin + P_ROWS[l]) + j * S * nw[l];
            o->rows.assign(rows, rows + S * nw[l]);
            const uint8_t* nodes = pin + P_NODES[l] + nodes_at[j] * 32;
            o->nodes.assign(nodes, nodes + (nodes_at[j + 1] - nodes_at[j]) * 32);
        }
    };

    // step 0: level 0's opening and induction, the first section
    open_and_induce(0, d_b);
    section(1, cfg.ks[0], nullptr, 0);
    launch_gather_nodes(wl[0][0].tree.p, (const uint64_t*)(pdev + P_NIDX[0]), nodes_at[G], (uint32_t*)(pdev + P_NODES[0]), s);
    for (int i = 0; i < R; i++) {
        const int n0 = i == 0 ? 1 : 2;
        if (i == R - 1) {
            // final step (prover.rs:104-142): yr, queries, the last level's opening
            launch_copy_blocks((uint32_t*)(pdev + P_YR), yr_len, d_f.p, sf, yr_len, count, s);
            {
                const uint32_t seq = ctx->next_flag();
                launch_signal(ctx->flag_dev(), seq, s);
                LIG_HIP(hipGetLastError());
                ctx->wait_flag(seq);
            }
            LIG_REQUIRE(flen == yr_len, LIG_E_ASSERT, "final message length");
            replay_and_open(n0, cfg.ks[i], i);
            uint64_t* h_fq = (uint64_t*)(pin + P_FQ);
            uint64_t* h_nidx = (uint64_t*)(pin + P_NIDX[R]);
            nodes_at.assign(G + 1, 0);
            for (size_t j = 0; j < G; j++) {
                const T* yr = (const T*)(pin + P_YR) + j * yr_len;
                proofs[j].yr.assign(yr, yr + yr_len);
                fs[j].absorb_elems<H>(proofs[j].yr.data(), yr_len);
                const std::vector<size_t> q = fs[j].get_distinct_queries(Mw[R], S);
                for (size_t k = 0; k < S; k++) h_fq[j * S + k] = q[k] + j * Mw[R];
                const std::vector<uint64_t> nd = merkle_proof_nodes(Mw[R], q);
                LIG_REQUIRE(nd.size() <= S * depth[R], LIG_E_ASSERT, "Merkle proof needs more nodes than queries x depth");
                nodes_at[j + 1] = nodes_at[j] + nd.size();
                for (size_t k = 0; k < nd.size(); k++) h_nidx[nodes_at[j] + k] = nd[k] + j * (2 * Mw[R] - 1);
            }
            launch_gather_rows<F>(wl[R][0].mat.p, nw[R], (const uint64_t*)(pdev + P_FQ), G * S, (T*)(pdev + P_ROWS[R]),
                                  nullptr, s);
            launch_gather_nodes(wl[R][0].tree.p, (const uint64_t*)(pdev + P_NIDX[R]), nodes_at[G],
                                (uint32_t*)(pdev + P_NODES[R]), s);
            {
                const uint32_t seq = ctx->next_flag();
                launch_signal(ctx->flag_dev(), seq, s);
                LIG_HIP(hipGetLastError());
                ctx->wait_flag(seq);
            }
            for (size_t j = 0; j < G; j++) {
                Rlp<H>& fo = proofs[j].final_;
                fo.n_rows = S;
                fo.n_cols = nw[R];
                const T* rows = (const T*)(pin + P_ROWS[R]) + j * S * nw[R];
                fo.rows.assign(rows, rows + S * nw[R]);
                const uint8_t* nodes = pin + P_NODES[R] + nodes_at[j] * 32;
                fo.nodes.assign(nodes, nodes + (nodes_at[j + 1] - nodes_at[j]) * 32);
            }
            return;
        }
        // step i + 1 (prover.rs:144-195): commit the folded polynomial, replay
        // section i, open level i + 1, induce and glue, the next section
        commit_level(i + 2);
        replay_and_open(n0, cfg.ks[i], i);
        PBuf<T> d_bnew;
        open_and_induce(i + 1, d_bnew);
        const size_t bnew_stride = flen;
        section(2, cfg.ks[i + 1], d_bnew.p, bnew_stride);
        launch_gather_nodes(wl[i + 1][0].tree.p, (const uint64_t*)(pdev + P_NIDX[i + 1]), nodes_at[G],
                            (uint32_t*)(pdev + P_NODES[i + 1]), s);
    }
    throw Error(LIG_E_ASSERT, "Prover completed without entering final step (prover.rs:198)");
}

template struct RsPlan<H32>;
template struct RsPlan<H128>;
template struct Config<H32>;
template struct Config<H128>;
template void commit<H32>(RsPlan<H32>&, const uint32_t*, size_t, size_t, Witness<H32>&, int);
template void commit<H128>(RsPlan<H128>&, const u128v*, size_t, size_t, Witness<H128>&, int);
template void prove<H32>(Config<H32>&, const uint32_t*, uint32_t, Proof<H32>&);
template void prove<H128>(Config<H128>&, const u128v*, uint32_t, Proof<H128>&);
template void commit_group<H32>(RsPlan<H32>&, const uint32_t* const*, int, size_t, size_t, std::vector<Witness<H32>>&,
                                std::shared_ptr<void>&);
template void prove_gang<H32>(Config<H32>&, const uint32_t* const*, const uint32_t*, int, Proof<H32>*);

}  // namespace lig
