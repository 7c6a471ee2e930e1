// extern "C" boundary (include/ligerito.h).  Every entry point catches
// library exceptions and returns a status; nothing unwinds across the ABI.
#include <array>
#include <atomic>
#include <exception>

#include "batch.hpp"
#include "engine.hpp"

using namespace lig;

// A context is reference counted: the caller's handle holds one reference
// and every handle made from the context (plans, witnesses, prover
// configurations) holds another, since they keep pointers into its stream
// and memory pool.  lig_ctx_destroy drops the caller's reference; the
// context is torn down when the last reference goes.
struct lig_ctx {
    Ctx c;
    std::atomic<int> refs{1};
};

namespace {
lig_ctx* ctx_retain(lig_ctx* ctx) {
    ctx->refs.fetch_add(1, std::memory_order_relaxed);
    return ctx;
}

void ctx_release(lig_ctx* ctx) {
    if (!ctx || ctx->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    (void)hipSetDevice(ctx->c.device);
    (void)hipStreamSynchronize(ctx->c.stream);
    ctx->c.release_side();
    try {
        ctx->c.collect();
    } catch (...) {
    }
    ctx->c.pool.drain();
    (void)hipStreamDestroy(ctx->c.stream);
    ctx->c.stream = nullptr;
    delete ctx;
}

// Holds one reference to a context for the lifetime of a child handle.
struct CtxRef {
    lig_ctx* ctx = nullptr;
    CtxRef() = default;
    explicit CtxRef(lig_ctx* c) : ctx(ctx_retain(c)) {}
    CtxRef(const CtxRef&) = delete;
    CtxRef& operator=(const CtxRef&) = delete;
    ~CtxRef() { ctx_release(ctx); }
};
}  // namespace

// Members are destroyed in reverse order: the device objects go before the
// context reference that keeps their pool alive.
struct lig_rs {
    int bits;
    CtxRef ref;
    lig_ctx* ctx;
    std::unique_ptr<RsPlan<H32>> p32;
    std::unique_ptr<RsPlan<H128>> p128;
    explicit lig_rs(lig_ctx* c) : ref(c), ctx(c) {}
};

struct lig_witness {
    int bits;
    CtxRef ref;
    Witness<H32> w32;
    Witness<H128> w128;
    explicit lig_witness(lig_ctx* c) : ref(c) {}
};

struct lig_fs {
    Transcript t;
    explicit lig_fs(uint32_t s) : t(s) {}
};

struct lig_prover_config {
    int bits;
    CtxRef ref;
    Config<H32> c32;
    Config<H128> c128;
    explicit lig_prover_config(lig_ctx* c) : ref(c) {}
};

struct lig_proof {
    int bits;
    Proof<H32> p32;
    Proof<H128> p128;
};

struct lig_batch {
    int bits;
    std::unique_ptr<Batch<H32>> b32;
    std::unique_ptr<Batch<H128>> b128;
};

struct lig_server {
    int bits;
    std::unique_ptr<Server<H32>> s32;
    std::unique_ptr<Server<H128>> s128;
};

namespace {
template <class Fn> int guarded(Fn&& fn) {
    try {
        set_error("");
        return fn();
    } catch (const Error& e) {
        set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return LIG_E_OOM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return LIG_E_ASSERT;
    }
}

int check_bits(int bits) {
    LIG_REQUIRE(bits == 32 || bits == 128, LIG_E_INVALID, "field_bits must be 32 or 128");
    return 0;
}

size_t elem_bytes(int bits) { return (size_t)bits / 8; }

// Upload a host buffer into a pooled device buffer.
template <class T> PBuf<T> upload(Ctx& c, const void* host, size_t count) {
    PBuf<T> b(&c.pool, count);
    c.upload(b.p, host, count * sizeof(T));
    return b;
}
}  // namespace

extern "C" {

const char* lig_last_error(void) { return last_error(); }

int lig_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int lig_ctx_create(int device, lig_ctx** out) {
    return guarded([&] {
        LIG_REQUIRE(out, LIG_E_INVALID, "out is NULL");
        int n = lig_device_count();
        LIG_REQUIRE(n > 0, LIG_E_NO_DEVICE, "no HIP device visible");
        LIG_REQUIRE(device >= 0 && device < n, LIG_E_INVALID, "device index out of range");
        LIG_HIP(hipSetDevice(device));
        std::unique_ptr<lig_ctx> c(new lig_ctx());
        c->c.device = device;
        LIG_HIP(hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking));
        // set-up of the no-allocation entry point's scratch on this device
        // (its address looked up once, with the device current); a failure
        // here is reported by that entry point, which looks again
        try {
            (void)noalloc_scratch_ptr(device);
        } catch (const std::exception&) {
        }
        *out = c.release();
        return LIG_OK;
    });
}

void lig_ctx_destroy(lig_ctx* ctx) { ctx_release(ctx); }

int lig_ctx_profile(lig_ctx* ctx, int enable) {
    return guarded([&] {
        LIG_REQUIRE(ctx, LIG_E_INVALID, "ctx is NULL");
        ctx->c.collect();
        ctx->c.profiling = enable != 0;
        for (int i = 0; i < PH_COUNT; i++) {
            ctx->c.phase_ms[i] = 0;
            ctx->c.phase_n[i] = 0;
        }
        return LIG_OK;
    });
}

int lig_ctx_profile_read(lig_ctx* ctx, const char* phase, double* total_ms, uint64_t* count) {
    return guarded([&] {
        LIG_REQUIRE(ctx && phase && total_ms && count, LIG_E_INVALID, "NULL argument");
        ctx->c.collect();
        for (int i = 0; i < PH_COUNT; i++)
            if (strcmp(phase, phase_name(i)) == 0) {
                *total_ms = ctx->c.phase_ms[i];
                *count = ctx->c.phase_n[i];
                return LIG_OK;
            }
        throw Error(LIG_E_INVALID, std::string("unknown phase ") + phase);
    });
}

int lig_bench_rate(lig_ctx* ctx, int which, double* rate) {
    return guarded([&] {
        LIG_REQUIRE(ctx && rate, LIG_E_INVALID, "NULL argument");
        LIG_REQUIRE(which >= 0 && which <= 6, LIG_E_INVALID, "unknown microbenchmark");
        LIG_HIP(hipSetDevice(ctx->c.device));
        PBuf<uint32_t> word(&ctx->c.pool, 1);
        try {
            *rate = bench_rate(which, word.p, ctx->c.stream);
        } catch (const std::exception& e) {
            throw Error(LIG_E_HIP, e.what());
        }
        return LIG_OK;
    });
}

void* lig_ctx_stream(lig_ctx* ctx) { return ctx ? (void*)ctx->c.stream : nullptr; }

int lig_ctx_synchronize(lig_ctx* ctx) {
    return guarded([&] {
        LIG_REQUIRE(ctx, LIG_E_INVALID, "ctx is NULL");
        ctx->c.sync();
        return LIG_OK;
    });
}

// ------------------------------------------------------------ Reed–Solomon --
int lig_rs_new(lig_ctx* ctx, int field_bits, size_t message_length, size_t block_length, lig_rs** out) {
    return guarded([&] {
        LIG_REQUIRE(ctx && out, LIG_E_INVALID, "NULL argument");
        check_bits(field_bits);
        LIG_HIP(hipSetDevice(ctx->c.device));
        std::unique_ptr<lig_rs> r(new lig_rs(ctx));
        r->bits = field_bits;
        if (field_bits == 32) {
            r->p32.reset(new RsPlan<H32>());
            r->p32->init(&ctx->c, message_length, block_length);
        } else {
            r->p128.reset(new RsPlan<H128>());
            r->p128->init(&ctx->c, message_length, block_length);
        }
        *out = r.release();
        return LIG_OK;
    });
}

void lig_rs_free(lig_rs* rs) { delete rs; }

size_t lig_rs_message_length(const lig_rs* rs) { return !rs ? 0 : rs->bits == 32 ? rs->p32->m() : rs->p128->m(); }
size_t lig_rs_block_length(const lig_rs* rs) { return !rs ? 0 : rs->bits == 32 ? rs->p32->M() : rs->p128->M(); }

int lig_rs_twiddles(const lig_rs* rs, void* out) {
    return guarded([&] {
        LIG_REQUIRE(rs && out, LIG_E_INVALID, "NULL argument");
        if (rs->bits == 32) memcpy(out, rs->p32->tw.data(), rs->p32->tw.size() * 4);
        else memcpy(out, rs->p128->tw.data(), rs->p128->tw.size() * 16);
        return LIG_OK;
    });
}

int lig_rs_set_tuning(lig_rs* rs, size_t n_cols, int kmax, int group) {
    return guarded([&] {
        LIG_REQUIRE(rs && n_cols, LIG_E_INVALID, "bad argument");
        const bool ok32 = kmax == 0 || (kmax >= 2 && kmax <= 6), ok128 = kmax >= 0 && kmax <= 3;
        const int gmax = rs->bits == 32 ? 8 : 16;
        LIG_REQUIRE((rs->bits == 32 ? ok32 : ok128) && group >= 0 && group <= gmax && (group & (group - 1)) == 0,
                    LIG_E_INVALID, "tuning: kmax 2..6 (GF(2^32)) or 1..3 (GF(2^128)), group a power of two up to 8 / 16");
        PassTune t;
        t.kmax = kmax;
        t.group = group;
        if (rs->bits == 32) rs->p32->set_tune(n_cols, t);
        else rs->p128->set_tune(n_cols, t);
        return LIG_OK;
    });
}

int lig_rs_autotune(lig_rs* rs, size_t n_cols, int* kmax, int* group) {
    return guarded([&] {
        LIG_REQUIRE(rs && n_cols && kmax && group, LIG_E_INVALID, "bad argument");
        LIG_REQUIRE(is_pow2(n_cols), LIG_E_INVALID, "the column count must be a power of two");
        LIG_HIP(hipSetDevice(rs->ctx->c.device));
        const PassTune t = rs->bits == 32 ? rs->p32->tune_for(n_cols) : rs->p128->tune_for(n_cols);
        *kmax = t.kmax;
        *group = t.group;
        return LIG_OK;
    });
}

int lig_rs_encode(lig_rs* rs, const void* message, void* out) {
    return guarded([&] {
        LIG_REQUIRE(rs && message && out, LIG_E_INVALID, "NULL argument");
        LIG_HIP(hipSetDevice(rs->ctx->c.device));
        if (rs->bits == 32) rs->p32->encode_systematic((const uint32_t*)message, (uint32_t*)out);
        else rs->p128->encode_systematic((const u128v*)message, (u128v*)out);
        return LIG_OK;
    });
}

int lig_rs_encode_non_systematic_cols_dev(lig_rs* rs, void* d_columns, size_t n_cols) {
    return guarded([&] {
        LIG_REQUIRE(rs && d_columns && n_cols, LIG_E_INVALID, "bad argument");
        LIG_HIP(hipSetDevice(rs->ctx->c.device));
        if (rs->bits == 32) rs->p32->encode_non_systematic_cols((uint32_t*)d_columns, n_cols);
        else rs->p128->encode_non_systematic_cols((u128v*)d_columns, n_cols);
        rs->ctx->c.sync();
        return LIG_OK;
    });
}

int lig_rs_encode_non_systematic_cols_dev_noalloc(lig_rs* rs, void* d_columns, size_t n_cols) {
    return guarded([&] {
        LIG_REQUIRE(rs && d_columns && n_cols, LIG_E_INVALID, "bad argument");
        const size_t row_bytes = lig_rs_block_length(rs) * elem_bytes(rs->bits);
        LIG_REQUIRE(n_cols <= NOALLOC_SCRATCH_BYTES / row_bytes, LIG_E_INVALID,
                    "encode_non_systematic (no allocation): block_length * n_cols elements exceed the 64 MiB scratch");
        Ctx& c = rs->ctx->c;
        LIG_HIP(hipSetDevice(c.device));
        void* scratch = nullptr;
        try {
            scratch = noalloc_scratch_ptr(c.device);
        } catch (const std::exception& e) {
            throw Error(LIG_E_HIP, e.what());
        }
        if (rs->bits == 32) rs->p32->encode_non_systematic_cols((uint32_t*)d_columns, n_cols, (uint32_t*)scratch);
        else rs->p128->encode_non_systematic_cols((u128v*)d_columns, n_cols, (u128v*)scratch);
        c.sync();
        return LIG_OK;
    });
}

int lig_rs_encode_non_systematic(lig_rs* rs, void* message) {
    return guarded([&] {
        LIG_REQUIRE(rs && message, LIG_E_INVALID, "NULL argument");
        Ctx& c = rs->ctx->c;
        LIG_HIP(hipSetDevice(c.device));
        size_t M = lig_rs_block_length(rs), eb = elem_bytes(rs->bits);
        PBuf<uint8_t> d(&c.pool, M * eb);
        LIG_HIP(hipMemcpyAsync(d.p, message, M * eb, hipMemcpyHostToDevice, c.stream));
        if (rs->bits == 32) rs->p32->encode_non_systematic_cols((uint32_t*)d.p, 1);
        else rs->p128->encode_non_systematic_cols((u128v*)d.p, 1);
        LIG_HIP(hipMemcpyAsync(message, d.p, M * eb, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        return LIG_OK;
    });
}

// ------------------------------------------------------------------ Merkle --
int lig_build_merkle_tree(lig_ctx* ctx, const void* leaves, size_t n_leaves, size_t leaf_len, uint8_t* layers_out) {
    return guarded([&] {
        LIG_REQUIRE(ctx && leaves && layers_out, LIG_E_INVALID, "NULL argument");
        LIG_REQUIRE(is_pow2(n_leaves), LIG_E_INVALID, "build_merkle_tree: leaf count must be a power of two");
        LIG_REQUIRE(leaf_len % 4 == 0 && leaf_len > 0 && leaf_len <= 256 * 1024, LIG_E_INVALID,
                    "leaf_len must be a positive multiple of 4 up to 256 KiB");
        Ctx& c = ctx->c;
        LIG_HIP(hipSetDevice(c.device));
        // a leaf is a row of leaf_len/4 32-bit words
        PBuf<uint32_t> d = upload<uint32_t>(c, leaves, n_leaves * leaf_len / 4);
        PBuf<uint32_t> tree(&c.pool, (2 * n_leaves - 1) * 8);
        PBuf<uint32_t> scratch(&c.pool, leaf_hash_scratch_words(n_leaves, leaf_len));
        launch_leaf_hash<F32>(d.p, n_leaves, leaf_len / 4, tree.p, scratch.n ? scratch.p : nullptr, c.stream);
        launch_merkle_levels(tree.p, n_leaves, c.stream);
        LIG_HIP(hipGetLastError());
        LIG_HIP(hipMemcpyAsync(layers_out, tree.p, (2 * n_leaves - 1) * 32, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        return LIG_OK;
    });
}

static void check_offsets(const size_t* offsets, size_t n_leaves) {
    for (size_t i = 0; i < n_leaves; i++) {
        LIG_REQUIRE(offsets[i + 1] >= offsets[i], LIG_E_INVALID, "leaf offsets must not decrease");
        LIG_REQUIRE(offsets[i + 1] - offsets[i] <= 256 * 1024, LIG_E_INVALID, "leaves above 256 KiB are not supported");
    }
}

int lig_build_merkle_tree_var(lig_ctx* ctx, const void* bytes, const size_t* offsets, size_t n_leaves, uint8_t* layers_out) {
    return guarded([&] {
        LIG_REQUIRE(ctx && offsets && layers_out, LIG_E_INVALID, "NULL argument");
        // merkle_tree.rs:26 asserts a power-of-two leaf count
        LIG_REQUIRE(is_pow2(n_leaves), LIG_E_ASSERT, "build_merkle_tree: leaf count must be a power of two");
        check_offsets(offsets, n_leaves);
        const size_t total = offsets[n_leaves] - offsets[0];
        LIG_REQUIRE(bytes || total == 0, LIG_E_INVALID, "NULL leaf bytes");
        Ctx& c = ctx->c;
        LIG_HIP(hipSetDevice(c.device));
        std::vector<uint64_t> offs(n_leaves + 1);
        for (size_t i = 0; i <= n_leaves; i++) offs[i] = offsets[i] - offsets[0];
        PBuf<uint8_t> d_bytes(&c.pool, std::max<size_t>(total, 1));
        if (total)
            LIG_HIP(hipMemcpyAsync(d_bytes.p, (const uint8_t*)bytes + offsets[0], total, hipMemcpyHostToDevice, c.stream));
        PBuf<uint64_t> d_offs = upload<uint64_t>(c, offs.data(), offs.size());
        PBuf<uint32_t> tree(&c.pool, (2 * n_leaves - 1) * 8);
        launch_leaf_hash_var(d_bytes.p, d_offs.p, n_leaves, tree.p, c.stream);
        if (n_leaves > 1) launch_merkle_levels(tree.p, n_leaves, c.stream);
        LIG_HIP(hipGetLastError());
        LIG_HIP(hipMemcpyAsync(layers_out, tree.p, (2 * n_leaves - 1) * 32, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        return LIG_OK;
    });
}

int lig_merkle_verify_var(const uint8_t root[32], const uint8_t* proof, size_t n_nodes, size_t depth, const void* bytes,
                          const size_t* offsets, const size_t* indices, size_t n_queries) {
    int r = guarded([&] {
        LIG_REQUIRE(root && offsets && indices && n_queries, LIG_E_INVALID, "bad argument");
        LIG_REQUIRE(bytes || offsets[n_queries] == offsets[0], LIG_E_INVALID, "NULL leaf bytes");
        check_offsets(offsets, n_queries);
        std::vector<std::array<uint8_t, 32>> lh(n_queries);
        for (size_t i = 0; i < n_queries; i++)
            blake3_bytes((const uint8_t*)bytes + offsets[i], offsets[i + 1] - offsets[i], lh[i].data());
        std::vector<size_t> q(indices, indices + n_queries);
        return merkle_verify_hashes(root, proof, n_nodes, depth, lh, q) ? 1 : 0;
    });
    return r < 0 ? 0 : r;
}

int lig_merkle_prove(const uint8_t* layers, size_t n_leaves, const size_t* queries, size_t n_queries, uint8_t* proof_out,
                     size_t cap_nodes, size_t* n_nodes) {
    return guarded([&] {
        LIG_REQUIRE(layers && queries && n_nodes && n_queries, LIG_E_INVALID, "bad argument");
        LIG_REQUIRE(is_pow2(n_leaves), LIG_E_INVALID, "leaf count must be a power of two");
        std::vector<size_t> q(queries, queries + n_queries);
        for (size_t i = 0; i < q.size(); i++) {
            LIG_REQUIRE(q[i] < n_leaves, LIG_E_INVALID, "query out of range");
            LIG_REQUIRE(i == 0 || q[i] > q[i - 1], LIG_E_INVALID, "queries must be sorted and distinct");
        }
        std::vector<uint64_t> nodes = merkle_proof_nodes(n_leaves, q);
        *n_nodes = nodes.size();
        if (!proof_out) return LIG_OK;
        LIG_REQUIRE(nodes.size() <= cap_nodes, LIG_E_CAPACITY, "proof buffer too small");
        for (size_t i = 0; i < nodes.size(); i++) memcpy(proof_out + 32 * i, layers + 32 * nodes[i], 32);
        return LIG_OK;
    });
}

int lig_merkle_verify(const uint8_t root[32], const uint8_t* proof, size_t n_nodes, size_t depth, const void* leaves,
                      size_t leaf_len, const size_t* indices, size_t n_queries) {
    int r = guarded([&] {
        LIG_REQUIRE(root && leaves && indices && n_queries, LIG_E_INVALID, "bad argument");
        std::vector<std::array<uint8_t, 32>> lh(n_queries);
        for (size_t i = 0; i < n_queries; i++) blake3_bytes((const uint8_t*)leaves + i * leaf_len, leaf_len, lh[i].data());
        std::vector<size_t> q(indices, indices + n_queries);
        return merkle_verify_hashes(root, proof, n_nodes, depth, lh, q) ? 1 : 0;
    });
    return r < 0 ? 0 : r;
}

// -------------------------------------------------------------- transcript --
int lig_fs_new(uint32_t seed, lig_fs** out) {
    return guarded([&] {
        LIG_REQUIRE(out, LIG_E_INVALID, "out is NULL");
        *out = new lig_fs(seed);
        return LIG_OK;
    });
}
void lig_fs_free(lig_fs* fs) { delete fs; }
int lig_fs_absorb(lig_fs* fs, const void* bytes, size_t len) {
    return guarded([&] {
        LIG_REQUIRE(fs && (bytes || !len), LIG_E_INVALID, "bad argument");
        fs->t.absorb(bytes, len);
        return LIG_OK;
    });
}
int lig_fs_squeeze(lig_fs* fs, uint8_t out16[16]) {
    return guarded([&] {
        LIG_REQUIRE(fs && out16, LIG_E_INVALID, "bad argument");
        uint64_t lo, hi;
        fs->t.squeeze(&lo, &hi);
        memcpy(out16, &lo, 8);
        memcpy(out16 + 8, &hi, 8);
        return LIG_OK;
    });
}
int lig_fs_get_field(lig_fs* fs, int field_bits, void* out) {
    return guarded([&] {
        LIG_REQUIRE(fs && out, LIG_E_INVALID, "bad argument");
        check_bits(field_bits);
        if (field_bits == 32) {
            uint32_t v = fs->t.get_field<H32>();
            memcpy(out, &v, 4);
        } else {
            u128v v = fs->t.get_field<H128>();
            H128::to_bytes(v, (uint8_t*)out);
        }
        return LIG_OK;
    });
}
size_t lig_fs_get_query(lig_fs* fs, size_t n) { return (fs && n) ? fs->t.get_query(n) : 0; }
int lig_fs_get_distinct_queries(lig_fs* fs, size_t n, size_t s, size_t* out) {
    return guarded([&] {
        LIG_REQUIRE(fs && out, LIG_E_INVALID, "bad argument");
        std::vector<size_t> q = fs->t.get_distinct_queries(n, s);
        memcpy(out, q.data(), s * sizeof(size_t));
        return LIG_OK;
    });
}

// ----------------------------------------------------------- ligero_commit --
static int commit_impl(lig_rs* rs, const void* poly, size_t m, size_t n, bool on_device, lig_witness** out) {
    return guarded([&] {
        LIG_REQUIRE(rs && poly && out, LIG_E_INVALID, "NULL argument");
        Ctx& c = rs->ctx->c;
        LIG_HIP(hipSetDevice(c.device));
        std::unique_ptr<lig_witness> w(new lig_witness(rs->ctx));
        w->bits = rs->bits;
        if (rs->bits == 32) {
            PBuf<uint32_t> tmp;
            const uint32_t* d = (const uint32_t*)poly;
            if (!on_device) {
                tmp = upload<uint32_t>(c, poly, m * n);
                d = tmp.p;
            }
            commit<H32>(*rs->p32, d, m, n, w->w32);
        } else {
            PBuf<u128v> tmp;
            const u128v* d = (const u128v*)poly;
            if (!on_device) {
                tmp = upload<u128v>(c, poly, m * n);
                d = tmp.p;
            }
            commit<H128>(*rs->p128, d, m, n, w->w128);
        }
        *out = w.release();
        return LIG_OK;
    });
}

int lig_ligero_commit(lig_rs* rs, const void* poly, size_t m, size_t n, lig_witness** out) {
    return commit_impl(rs, poly, m, n, false, out);
}
int lig_ligero_commit_dev(lig_rs* rs, const void* d_poly, size_t m, size_t n, lig_witness** out) {
    return commit_impl(rs, d_poly, m, n, true, out);
}

int lig_witness_root(const lig_witness* w, uint8_t root[32]) {
    return guarded([&] {
        LIG_REQUIRE(w && root, LIG_E_INVALID, "NULL argument");
        memcpy(root, w->bits == 32 ? w->w32.root : w->w128.root, 32);
        return LIG_OK;
    });
}
size_t lig_witness_num_rows(const lig_witness* w) { return !w ? 0 : w->bits == 32 ? w->w32.M : w->w128.M; }
size_t lig_witness_num_cols(const lig_witness* w) { return !w ? 0 : w->bits == 32 ? w->w32.n : w->w128.n; }

}  // extern "C"

template <class H> static void flat_mat_of(const Witness<H>& w, void* out) {
    using T = typename H::T;
    std::vector<T> rm(w.M * w.n);
    LIG_HIP(hipMemcpyAsync(rm.data(), w.mat.p, rm.size() * sizeof(T), hipMemcpyDeviceToHost, w.ctx->stream));
    LIG_HIP(hipStreamSynchronize(w.ctx->stream));
    T* o = (T*)out;
    for (size_t j = 0; j < w.n; j++)
        for (size_t i = 0; i < w.M; i++) o[j * w.M + i] = rm[i * w.n + j];
}

extern "C" {

int lig_witness_flat_mat(const lig_witness* w, void* out) {
    return guarded([&] {
        LIG_REQUIRE(w && out, LIG_E_INVALID, "NULL argument");
        if (w->bits == 32) flat_mat_of<H32>(w->w32, out);
        else flat_mat_of<H128>(w->w128, out);
        return LIG_OK;
    });
}

int lig_witness_layers(const lig_witness* w, uint8_t* out) {
    return guarded([&] {
        LIG_REQUIRE(w && out, LIG_E_INVALID, "NULL argument");
        const uint32_t* t = w->bits == 32 ? w->w32.tree.p : w->w128.tree.p;
        size_t M = lig_witness_num_rows(w);
        hipStream_t s = w->bits == 32 ? w->w32.ctx->stream : w->w128.ctx->stream;
        LIG_HIP(hipMemcpyAsync(out, t, (2 * M - 1) * 32, hipMemcpyDeviceToHost, s));
        LIG_HIP(hipStreamSynchronize(s));
        return LIG_OK;
    });
}

int lig_witness_extract_row(const lig_witness* w, size_t row, void* out) {
    return guarded([&] {
        LIG_REQUIRE(w && out, LIG_E_INVALID, "NULL argument");
        LIG_REQUIRE(row < lig_witness_num_rows(w), LIG_E_INVALID, "row out of range");
        size_t n = lig_witness_num_cols(w), eb = elem_bytes(w->bits);
        const void* src = w->bits == 32 ? (const void*)(w->w32.mat.p + row * n) : (const void*)(w->w128.mat.p + row * n);
        hipStream_t s = w->bits == 32 ? w->w32.ctx->stream : w->w128.ctx->stream;
        LIG_HIP(hipMemcpyAsync(out, src, n * eb, hipMemcpyDeviceToHost, s));
        LIG_HIP(hipStreamSynchronize(s));
        return LIG_OK;
    });
}

void lig_witness_free(lig_witness* w) { delete w; }

// ------------------------------------------------------------------ config --
static int config_impl(lig_ctx* ctx, int field_bits, const Shape& sh, lig_prover_config** out) {
    LIG_REQUIRE(ctx && out, LIG_E_INVALID, "NULL argument");
    check_bits(field_bits);
    sh.validate(field_bits / 8);
    LIG_HIP(hipSetDevice(ctx->c.device));
    std::unique_ptr<lig_prover_config> c(new lig_prover_config(ctx));
    c->bits = field_bits;
    if (field_bits == 32) c->c32.init(&ctx->c, sh);
    else c->c128.init(&ctx->c, sh);
    *out = c.release();
    return LIG_OK;
}

int lig_prover_config_new(lig_ctx* ctx, int field_bits, int log_n, lig_prover_config** out) {
    return guarded([&] { return config_impl(ctx, field_bits, Shape::generated(log_n), out); });
}

int lig_prover_config_from_dims(lig_ctx* ctx, int field_bits, size_t recursive_steps, const size_t initial_dims[2],
                                size_t initial_k, const size_t* dims, const size_t* ks, lig_prover_config** out) {
    return guarded([&] {
        LIG_REQUIRE(initial_dims && dims && ks, LIG_E_INVALID, "NULL argument");
        LIG_REQUIRE(recursive_steps >= 1 && recursive_steps <= 8, LIG_E_INVALID,
                    "recursive_steps must be in [1, 8] (prover.rs:198 panics at 0)");
        LIG_REQUIRE(initial_k < 40, LIG_E_INVALID, "configuration dimension out of range");
        Shape sh;
        sh.recursive_steps = (int)recursive_steps;
        sh.initial_m = initial_dims[0];
        sh.initial_n = initial_dims[1];
        sh.initial_k = (int)initial_k;
        for (size_t i = 0; i < recursive_steps; i++) {
            LIG_REQUIRE(ks[i] < 40, LIG_E_INVALID, "configuration dimension out of range");
            sh.dims_m.push_back(dims[2 * i]);
            sh.dims_n.push_back(dims[2 * i + 1]);
            sh.ks.push_back((int)ks[i]);
        }
        return config_impl(ctx, field_bits, sh, out);
    });
}

static void write_shape(const Shape& sh, size_t* recursive_steps, size_t initial_dims[2], size_t* initial_k, size_t* dims,
                        size_t* ks, size_t cap_steps) {
    LIG_REQUIRE(recursive_steps && initial_dims && initial_k, LIG_E_INVALID, "NULL argument");
    *recursive_steps = (size_t)sh.recursive_steps;
    initial_dims[0] = sh.initial_m;
    initial_dims[1] = sh.initial_n;
    *initial_k = (size_t)sh.initial_k;
    if (!dims || !ks) return;
    LIG_REQUIRE((size_t)sh.recursive_steps <= cap_steps, LIG_E_CAPACITY, "more recursive steps than cap_steps");
    for (int i = 0; i < sh.recursive_steps; i++) {
        dims[2 * i] = sh.dims_m[i];
        dims[2 * i + 1] = sh.dims_n[i];
        ks[i] = (size_t)sh.ks[i];
    }
}

int lig_prover_config_shape(int log_n, size_t* recursive_steps, size_t initial_dims[2], size_t* initial_k, size_t* dims,
                            size_t* ks, size_t cap_steps) {
    return guarded([&] {
        write_shape(Shape::generated(log_n), recursive_steps, initial_dims, initial_k, dims, ks, cap_steps);
        return LIG_OK;
    });
}

int lig_prover_config_dims(const lig_prover_config* cfg, size_t* recursive_steps, size_t initial_dims[2],
                           size_t* initial_k, size_t* dims, size_t* ks, size_t cap_steps) {
    return guarded([&] {
        LIG_REQUIRE(cfg, LIG_E_INVALID, "NULL argument");
        write_shape(cfg->bits == 32 ? cfg->c32.shape() : cfg->c128.shape(), recursive_steps, initial_dims, initial_k,
                    dims, ks, cap_steps);
        return LIG_OK;
    });
}
void lig_prover_config_free(lig_prover_config* cfg) { delete cfg; }

// ------------------------------------------------------------------- prove --
static int prover_impl(lig_prover_config* cfg, const void* poly, uint32_t fs_seed, bool on_device, lig_proof** out) {
    return guarded([&] {
        LIG_REQUIRE(cfg && poly && out, LIG_E_INVALID, "NULL argument");
        std::unique_ptr<lig_proof> p(new lig_proof());
        p->bits = cfg->bits;
        if (cfg->bits == 32) {
            Ctx& c = *cfg->c32.ctx;
            LIG_HIP(hipSetDevice(c.device));
            PBuf<uint32_t> tmp;
            const uint32_t* d = (const uint32_t*)poly;
            if (!on_device) {
                tmp = upload<uint32_t>(c, poly, (size_t)1 << cfg->c32.log_n);
                d = tmp.p;
            }
            prove<H32>(cfg->c32, d, fs_seed, p->p32);
        } else {
            Ctx& c = *cfg->c128.ctx;
            LIG_HIP(hipSetDevice(c.device));
            PBuf<u128v> tmp;
            const u128v* d = (const u128v*)poly;
            if (!on_device) {
                tmp = upload<u128v>(c, poly, (size_t)1 << cfg->c128.log_n);
                d = tmp.p;
            }
            prove<H128>(cfg->c128, d, fs_seed, p->p128);
        }
        *out = p.release();
        return LIG_OK;
    });
}

int lig_prover(lig_prover_config* cfg, const void* poly, uint32_t fs_seed, lig_proof** out) {
    return prover_impl(cfg, poly, fs_seed, false, out);
}
int lig_prover_dev(lig_prover_config* cfg, const void* d_poly, uint32_t fs_seed, lig_proof** out) {
    return prover_impl(cfg, d_poly, fs_seed, true, out);
}

int lig_proof_serialize(const lig_proof* p, uint8_t* buf, size_t cap, size_t* len) {
    return guarded([&] {
        LIG_REQUIRE(p && len, LIG_E_INVALID, "NULL argument");
        std::vector<uint8_t> b = p->bits == 32 ? serialize<H32>(p->p32) : serialize<H128>(p->p128);
        *len = b.size();
        if (!buf) return LIG_OK;
        LIG_REQUIRE(cap >= b.size(), LIG_E_CAPACITY, "buffer too small");
        memcpy(buf, b.data(), b.size());
        return LIG_OK;
    });
}

int lig_proof_deserialize(int field_bits, const uint8_t* buf, size_t len, lig_proof** out) {
    return guarded([&] {
        LIG_REQUIRE(buf && out, LIG_E_INVALID, "NULL argument");
        check_bits(field_bits);
        std::unique_ptr<lig_proof> p(new lig_proof());
        p->bits = field_bits;
        bool ok = field_bits == 32 ? deserialize<H32>(buf, len, p->p32) : deserialize<H128>(buf, len, p->p128);
        LIG_REQUIRE(ok, LIG_E_MALFORMED, "malformed proof bytes");
        *out = p.release();
        return LIG_OK;
    });
}

size_t lig_proof_size(const lig_proof* p) { return !p ? 0 : p->bits == 32 ? proof_size<H32>(p->p32) : proof_size<H128>(p->p128); }
void lig_proof_free(lig_proof* p) { delete p; }

static int verify_impl(lig_ctx* ctx, int field_bits, const Shape& sh, const lig_proof* p, uint32_t fs_seed) {
    LIG_REQUIRE(p, LIG_E_INVALID, "NULL proof");
    LIG_REQUIRE(p->bits == field_bits, LIG_E_INVALID, "proof field does not match field_bits");
    Ctx* dev = nullptr;
    if (ctx) {
        LIG_HIP(hipSetDevice(ctx->c.device));
        dev = &ctx->c;
    }
    return field_bits == 32 ? verify<H32>(sh, p->p32, fs_seed, dev) : verify<H128>(sh, p->p128, fs_seed, dev);
}

static Shape verifier_shape(size_t recursive_steps, size_t initial_dim, const size_t* log_dims, size_t initial_k,
                            const size_t* ks) {
    LIG_REQUIRE(log_dims && ks, LIG_E_INVALID, "NULL argument");
    LIG_REQUIRE(recursive_steps >= 1 && recursive_steps <= 8, LIG_E_INVALID,
                "recursive_steps must be in [1, 8] (prover.rs:198 panics at 0)");
    int ld[8], kk[8];
    for (size_t i = 0; i < recursive_steps; i++) {
        LIG_REQUIRE(log_dims[i] < 40 && ks[i] < 40, LIG_E_INVALID, "configuration dimension out of range");
        ld[i] = (int)log_dims[i];
        kk[i] = (int)ks[i];
    }
    LIG_REQUIRE(initial_dim < 40 && initial_k < 40, LIG_E_INVALID, "configuration dimension out of range");
    return Shape::from_logs((int)recursive_steps, (int)initial_dim, ld, (int)initial_k, kk);
}

int lig_verifier(int field_bits, int log_n, const lig_proof* p, uint32_t fs_seed) {
    return guarded([&] { return verify_impl(nullptr, field_bits, Shape::generated(log_n), p, fs_seed); });
}

int lig_verifier_dev(lig_ctx* ctx, int field_bits, int log_n, const lig_proof* p, uint32_t fs_seed) {
    return guarded([&] {
        LIG_REQUIRE(ctx, LIG_E_INVALID, "NULL argument");
        return verify_impl(ctx, field_bits, Shape::generated(log_n), p, fs_seed);
    });
}

int lig_verifier_with_config(int field_bits, size_t recursive_steps, size_t initial_dim, const size_t* log_dims,
                             size_t initial_k, const size_t* ks, const lig_proof* p, uint32_t fs_seed) {
    return guarded([&] {
        return verify_impl(nullptr, field_bits, verifier_shape(recursive_steps, initial_dim, log_dims, initial_k, ks), p,
                           fs_seed);
    });
}

int lig_verifier_with_config_dev(lig_ctx* ctx, int field_bits, size_t recursive_steps, size_t initial_dim,
                                 const size_t* log_dims, size_t initial_k, const size_t* ks, const lig_proof* p,
                                 uint32_t fs_seed) {
    return guarded([&] {
        LIG_REQUIRE(ctx, LIG_E_INVALID, "NULL argument");
        return verify_impl(ctx, field_bits, verifier_shape(recursive_steps, initial_dim, log_dims, initial_k, ks), p,
                           fs_seed);
    });
}

// ------------------------------------------------------------------ batch --
int lig_batch_create(int device, int field_bits, int n_streams, lig_batch** out) {
    return guarded([&] {
        LIG_REQUIRE(out, LIG_E_INVALID, "out is NULL");
        check_bits(field_bits);
        int n = lig_device_count();
        LIG_REQUIRE(n > 0, LIG_E_NO_DEVICE, "no HIP device visible");
        LIG_REQUIRE(device >= 0 && device < n, LIG_E_INVALID, "device index out of range");
        std::unique_ptr<lig_batch> b(new lig_batch());
        b->bits = field_bits;
        if (field_bits == 32) b->b32.reset(new Batch<H32>(device, n_streams));
        else b->b128.reset(new Batch<H128>(device, n_streams));
        *out = b.release();
        return LIG_OK;
    });
}

void lig_batch_destroy(lig_batch* b) { delete b; }

int lig_prove_batch(lig_batch* b, size_t n_jobs, const int* log_n, const void* const* polys, int polys_on_device,
                    const uint32_t* fs_seeds, lig_proof** out) {
    return guarded([&] {
        LIG_REQUIRE(b && out && (n_jobs == 0 || (log_n && polys && fs_seeds)), LIG_E_INVALID, "NULL argument");
        for (size_t i = 0; i < n_jobs; i++) out[i] = nullptr;
        std::vector<std::unique_ptr<lig_proof>> res(n_jobs);
        if (b->bits == 32) {
            std::vector<Proof<H32>> ps;
            b->b32->run(n_jobs, log_n, polys, polys_on_device != 0, fs_seeds, ps);
            for (size_t i = 0; i < n_jobs; i++) {
                res[i].reset(new lig_proof());
                res[i]->bits = 32;
                res[i]->p32 = std::move(ps[i]);
            }
        } else {
            std::vector<Proof<H128>> ps;
            b->b128->run(n_jobs, log_n, polys, polys_on_device != 0, fs_seeds, ps);
            for (size_t i = 0; i < n_jobs; i++) {
                res[i].reset(new lig_proof());
                res[i]->bits = 128;
                res[i]->p128 = std::move(ps[i]);
            }
        }
        for (size_t i = 0; i < n_jobs; i++) out[i] = res[i].release();
        return LIG_OK;
    });
}

// ---------------------------------------------------------------- serving --
int lig_serve_start(int device, int field_bits, int n_slots, uint64_t hbm_budget, lig_server** out) {
    return guarded([&] {
        LIG_REQUIRE(out, LIG_E_INVALID, "out is NULL");
        check_bits(field_bits);
        int n = lig_device_count();
        LIG_REQUIRE(n > 0, LIG_E_NO_DEVICE, "no HIP device visible");
        LIG_REQUIRE(device >= 0 && device < n, LIG_E_INVALID, "device index out of range");
        LIG_REQUIRE(hbm_budget > 0, LIG_E_INVALID, "hbm_budget must be positive");
        std::unique_ptr<lig_server> s(new lig_server());
        s->bits = field_bits;
        if (field_bits == 32) s->s32.reset(new Server<H32>(device, n_slots, hbm_budget));
        else s->s128.reset(new Server<H128>(device, n_slots, hbm_budget));
        *out = s.release();
        return LIG_OK;
    });
}

int lig_serve_submit(lig_server* s, int log_n, const void* poly, uint32_t fs_seed, uint64_t* ticket) {
    return guarded([&] {
        LIG_REQUIRE(s && ticket, LIG_E_INVALID, "NULL argument");
        *ticket = s->bits == 32 ? s->s32->submit(log_n, poly, fs_seed) : s->s128->submit(log_n, poly, fs_seed);
        return LIG_OK;
    });
}

static int serve_take(lig_server* s, uint64_t ticket, bool wait, lig_proof** out, double* latency_ms) {
    return guarded([&] {
        LIG_REQUIRE(s && out, LIG_E_INVALID, "NULL argument");
        *out = nullptr;
        std::unique_ptr<lig_proof> p(new lig_proof());
        p->bits = s->bits;
        std::string msg;
        const int r = s->bits == 32 ? s->s32->take(ticket, wait, p->p32, latency_ms, msg)
                                    : s->s128->take(ticket, wait, p->p128, latency_ms, msg);
        if (r < 0) throw Error(r, msg);
        if (r == 1) *out = p.release();
        return r;
    });
}

int lig_serve_poll(lig_server* s, uint64_t ticket, lig_proof** out, double* latency_ms) {
    return serve_take(s, ticket, false, out, latency_ms);
}

int lig_serve_wait(lig_server* s, uint64_t ticket, lig_proof** out, double* latency_ms) {
    return serve_take(s, ticket, true, out, latency_ms);
}

int lig_serve_warm(lig_server* s, const int* log_n, size_t count, size_t* warmed) {
    return guarded([&] {
        LIG_REQUIRE(s, LIG_E_INVALID, "NULL server");
        const size_t w = s->bits == 32 ? s->s32->warm(log_n, count) : s->s128->warm(log_n, count);
        if (warmed) *warmed = w;
        return LIG_OK;
    });
}

void lig_serve_stop(lig_server* s) { delete s; }

}  // extern "C"
