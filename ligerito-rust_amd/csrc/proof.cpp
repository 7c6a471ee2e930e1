// Proof encoding, ProofSize and the host verifier.
#include "engine.hpp"

namespace lig {

// -------------------------------------------------------------- encoding --
// Little-endian, FinalizedLigeritoProof field order (data_structures.rs:91-103):
//   "LGRT" | u32 version=1 | u32 field_bytes | initial root[32]
//   RLP(initial) | u64 n + roots | u64 n + RLPs | u64 len + yr | RLP(final)
//   | u64 n + n*(e0,e1,e2)
//   RLP = u64 n_rows | u64 n_cols | rows | u64 n_nodes | nodes*32
namespace {
struct Writer {
    std::vector<uint8_t> b;
    void u32(uint32_t v) { for (int i = 0; i < 4; i++) b.push_back((uint8_t)(v >> (8 * i))); }
    void u64(uint64_t v) { for (int i = 0; i < 8; i++) b.push_back((uint8_t)(v >> (8 * i))); }
    void raw(const void* p, size_t n) { b.insert(b.end(), (const uint8_t*)p, (const uint8_t*)p + n); }
    template <class H> void elems(const std::vector<typename H::T>& v) {
        size_t o = b.size();
        b.resize(o + v.size() * H::BYTES);
        for (size_t i = 0; i < v.size(); i++) H::to_bytes(v[i], &b[o + i * H::BYTES]);
    }
    template <class H> void rlp(const Rlp<H>& r) {
        u64(r.n_rows);
        u64(r.n_cols);
        elems<H>(r.rows);
        u64(r.nodes.size() / 32);
        raw(r.nodes.data(), r.nodes.size());
    }
};

struct Reader {
    const uint8_t *p, *e;
    bool bad = false;
    bool need(size_t n) {
        if ((size_t)(e - p) < n) bad = true;
        return !bad;
    }
    uint64_t u64() {
        if (!need(8)) return 0;
        uint64_t v = 0;
        for (int i = 7; i >= 0; i--) v = (v << 8) | p[i];
        p += 8;
        return v;
    }
    template <class H> void elems(std::vector<typename H::T>& v, uint64_t n) {
        if (n > (uint64_t)(e - p) / H::BYTES) { bad = true; return; }
        v.resize(n);
        for (uint64_t i = 0; i < n; i++, p += H::BYTES) v[i] = H::from_bytes(p);
    }
    template <class H> void rlp(Rlp<H>& r) {
        r.n_rows = u64();
        r.n_cols = u64();
        if (bad || r.n_rows > (1u << 24) || r.n_cols > (1u << 24)) { bad = true; return; }
        elems<H>(r.rows, (uint64_t)r.n_rows * r.n_cols);
        uint64_t nn = u64();
        if (bad || nn > (uint64_t)(e - p) / 32) { bad = true; return; }
        r.nodes.assign(p, p + nn * 32);
        p += nn * 32;
    }
};
}  // namespace

template <class H> std::vector<uint8_t> serialize(const Proof<H>& pr) {
    Writer w;
    w.raw("LGRT", 4);
    w.u32(1);
    w.u32(H::BYTES);
    w.raw(pr.initial_root.data(), 32);
    w.rlp<H>(pr.initial);
    w.u64(pr.rec_roots.size());
    for (auto& r : pr.rec_roots) w.raw(r.data(), 32);
    w.u64(pr.rec_proofs.size());
    for (auto& r : pr.rec_proofs) w.rlp<H>(r);
    w.u64(pr.yr.size());
    w.elems<H>(pr.yr);
    w.rlp<H>(pr.final_);
    w.u64(pr.tr.size() / 3);
    w.elems<H>(pr.tr);
    return w.b;
}

template <class H> bool deserialize(const uint8_t* b, size_t len, Proof<H>& pr) {
    pr = Proof<H>();
    if (len < 44 || memcmp(b, "LGRT", 4) != 0 || b[4] != 1 || b[5] || b[6] || b[7] || b[8] != H::BYTES) return false;
    Reader r{b + 12, b + len};
    memcpy(pr.initial_root.data(), b + 12, 32);
    r.p += 32;
    r.rlp<H>(pr.initial);
    uint64_t n = r.u64();
    if (r.bad || n > 64 || !r.need(32 * n)) return false;
    pr.rec_roots.resize(n);
    for (auto& x : pr.rec_roots) {
        memcpy(x.data(), r.p, 32);
        r.p += 32;
    }
    n = r.u64();
    if (r.bad || n > 64) return false;
    pr.rec_proofs.resize(n);
    for (auto& x : pr.rec_proofs) r.rlp<H>(x);
    n = r.u64();
    r.elems<H>(pr.yr, n);
    r.rlp<H>(pr.final_);
    n = r.u64();
    if (r.bad || n > (uint64_t)(r.e - r.p) / (3 * H::BYTES)) return false;
    r.elems<H>(pr.tr, 3 * n);
    return !r.bad && r.p == r.e;
}

// ProofSize (data_structures.rs:105-194)
template <class H> size_t proof_size(const Proof<H>& p) {
    auto rlp = [](const Rlp<H>& r) { return r.rows.size() * sizeof(typename H::T) + r.nodes.size(); };
    size_t s = 32 + rlp(p.initial) + 32 * p.rec_roots.size();
    for (auto& r : p.rec_proofs) s += rlp(r);
    s += p.yr.size() * sizeof(typename H::T) + rlp(p.final_);
    s += (p.tr.size() / 3) * 3 * sizeof(typename H::T);
    return s;
}

// -------------------------------------------------------------- verifier --
namespace {
template <class H> bool check_openings(const uint8_t* root, const Rlp<H>& r, int depth, const std::vector<size_t>& q) {
    if (r.n_rows != q.size() || r.rows.size() != r.n_rows * r.n_cols) return false;
    std::vector<std::array<uint8_t, 32>> leaves(q.size());
    std::vector<uint8_t> bytes(r.n_cols * H::BYTES);
    for (size_t i = 0; i < q.size(); i++) {
        for (size_t j = 0; j < r.n_cols; j++) H::to_bytes(r.rows[i * r.n_cols + j], &bytes[j * H::BYTES]);
        blake3_bytes(bytes.data(), bytes.size(), leaves[i].data());
    }
    return merkle_verify_hashes(root, r.nodes.data(), r.nodes.size() / 32, (size_t)depth, leaves, q);
}

// ligero.rs:97-131
template <class H> bool ligero_verify(const std::vector<size_t>& q, const Rlp<H>& r, const Vec<H>& yr, const Vec<H>& ch) {
    Vec<H> gr = lagrange_basis<H>(ch);
    if (gr.size() != r.n_cols || !is_pow2(yr.size())) return false;
    Vec<H> sks = eval_sk_at_vks<H>(yr.size());
    Vec<H> basis(yr.size());
    for (size_t i = 0; i < q.size(); i++) {
        typename H::T dot = H::zero(), e = H::zero();
        for (size_t j = 0; j < r.n_cols; j++) dot ^= H::mul(r.rows[i * r.n_cols + j], gr[j]);
        scaled_basis<H>(sks, H::from_int(q[i]), H::one(), basis);
        for (size_t j = 0; j < yr.size(); j++) e ^= H::mul(yr[j], basis[j]);
        if (!H::eq(e, dot)) return false;
    }
    return true;
}
}  // namespace

namespace {
// induce_sumcheck_poly (sumcheck_polys.rs:74-160) for the verifier on the
// GPU: the prover's induce kernels on the opened rows, the basis polynomial
// and the enforced-sum parts copied back.
template <class H>
void induce_device(Ctx* ctx, int n, const Vec<H>& sks, const std::vector<typename H::T>& rows, size_t n_rows, size_t n_cols,
                   const Vec<H>& v_ch, const std::vector<size_t>& queries, typename H::T alpha, Vec<H>& basis_out,
                   typename H::T& enforced) {
    using F = typename DevField<H>::F;
    using T = typename H::T;
    using A = typename F::Acc;
    LIG_REQUIRE(((size_t)1 << v_ch.size()) == n_cols, LIG_E_ASSERT, "opened row length != 2^k");
    LIG_REQUIRE(rows.size() == n_rows * n_cols && queries.size() == n_rows, LIG_E_ASSERT, "opening shape");
    Vec<H> gr = lagrange_basis<H>(v_ch);
    hipStream_t s = ctx->stream;
    std::vector<uint64_t> q64(queries.begin(), queries.end());
    std::vector<T> vals(sks.begin(), sks.end());
    vals.insert(vals.end(), gr.begin(), gr.end());
    vals.push_back(alpha);
    const int clo = induce_low_bits(n);
    PBuf<uint64_t> d_q(&ctx->pool, n_rows);
    PBuf<T> d_rows(&ctx->pool, rows.size()), d_vals(&ctx->pool, vals.size());
    PBuf<T> d_c(&ctx->pool, n_rows * n), d_ap(&ctx->pool, n_rows);
    PBuf<T> d_P(&ctx->pool, n_rows << clo), d_K(&ctx->pool, n_rows << (n - clo)), d_basis(&ctx->pool, (size_t)1 << n);
    PBuf<A> d_parts(&ctx->pool, ENFORCED_MAX_PARTS);
    LIG_HIP(hipMemcpyAsync(d_q.p, q64.data(), n_rows * 8, hipMemcpyHostToDevice, s));
    LIG_HIP(hipMemcpyAsync(d_rows.p, rows.data(), rows.size() * sizeof(T), hipMemcpyHostToDevice, s));
    LIG_HIP(hipMemcpyAsync(d_vals.p, vals.data(), vals.size() * sizeof(T), hipMemcpyHostToDevice, s));
    launch_induce_prep<F>(d_q.p, d_vals.p, d_vals.p + vals.size() - 1, (int)n_rows, n, d_c.p, d_ap.p, s);
    launch_enforced<F>(d_rows.p, (int)n_rows, (int)n_cols, d_vals.p + sks.size(), d_ap.p, d_parts.p, s);
    launch_induce<F>(d_c.p, d_ap.p, (int)n_rows, n, d_P.p, d_K.p, d_basis.p, s);
    LIG_HIP(hipGetLastError());
    basis_out.resize((size_t)1 << n);
    std::vector<A> parts(ENFORCED_MAX_PARTS);
    const int np = enforced_parts((int)n_rows, (int)n_cols);
    LIG_HIP(hipMemcpyAsync(basis_out.data(), d_basis.p, basis_out.size() * sizeof(T), hipMemcpyDeviceToHost, s));
    LIG_HIP(hipMemcpyAsync(parts.data(), d_parts.p, np * sizeof(A), hipMemcpyDeviceToHost, s));
    ctx->sync();
    A acc{};
    for (int i = 0; i < np; i++) acc ^= parts[i];
    enforced = acc_reduce_host<H>(acc);
}
}  // namespace

template <class H> int verify(const Shape& sh, const Proof<H>& p, uint32_t fs_seed, Ctx* dev) {
    using T = typename H::T;
    auto induce = [&](int n, const Vec<H>& sks, const std::vector<T>& rows, size_t n_rows, size_t n_cols, const Vec<H>& v_ch,
                      const std::vector<size_t>& queries, T alpha, Vec<H>& basis_out, T& enforced) {
        if (dev) induce_device<H>(dev, n, sks, rows, n_rows, n_cols, v_ch, queries, alpha, basis_out, enforced);
        else induce_host<H>(n, sks, rows, n_rows, n_cols, v_ch, queries, alpha, basis_out, enforced);
    };
    sh.validate(H::BYTES);
    // VerifierConfig (data_structures.rs:23-30) of the shape
    const int R = sh.recursive_steps, initial_dim = ilog2(sh.initial_m), initial_k = sh.initial_k, log_inv_rate = 2;
    std::vector<int> log_dims, ks;
    for (int i = 0; i < R; i++) {
        log_dims.push_back(ilog2(sh.dims_m[i]));
        ks.push_back(sh.ks[i]);
    }
    const size_t S = 148;
    if ((int)p.rec_roots.size() < R || p.tr.size() < 3) return LIG_E_MALFORMED;
    Transcript fs(fs_seed);
    fs.absorb(p.initial_root.data(), 32);
    Vec<H> pe0(initial_k);
    for (auto& r : pe0) r = fs.get_field<H>();
    fs.absorb(p.rec_roots[0].data(), 32);
    int depth = initial_dim + log_inv_rate;
    std::vector<size_t> q = fs.get_distinct_queries((size_t)1 << depth, S);
    if (p.initial.n_cols != ((size_t)1 << initial_k)) return LIG_E_MALFORMED;
    if (!check_openings<H>(p.initial_root.data(), p.initial, depth, q)) return LIG_E_MERKLE;
    T alpha = fs.get_field<H>();
    Vec<H> sks = eval_sk_at_vks<H>((size_t)1 << initial_dim);
    Vec<H> b1;
    T h1;
    induce(initial_dim, sks, p.initial.rows, S, p.initial.n_cols, pe0, q, alpha, b1, h1);

    // SumcheckVerifierInstance (ligerito.rs:143-296)
    std::vector<Vec<H>> basis{b1};
    Vec<H> sep{H::one()};
    Vec<H> ris;
    size_t trr = 0;
    auto read_tr = [&](T g[3]) {
        LIG_REQUIRE(trr * 3 + 3 <= p.tr.size(), LIG_E_ASSERT, "sumcheck transcript exhausted (ligerito.rs:188)");
        for (int i = 0; i < 3; i++) g[i] = p.tr[3 * trr + i];
        trr++;
    };
    T g[3];
    read_tr(g);
    if (!H::eq(g[0] ^ g[1], h1)) return LIG_E_ASSERT;
    Quad<H> running = quadratic_from_evals<H>(g[0], g[1], g[2]);
    fs.absorb_elems<H>(g, 3);
    for (int i = 0; i < R; i++) {
        Vec<H> rs;
        for (int j = 0; j < ks[i]; j++) {
            T ri = fs.get_field<H>();
            ris.push_back(ri);
            T sum = running.eval(ri);
            read_tr(g);
            if (!H::eq(g[0] ^ g[1], sum)) return LIG_E_ASSERT;
            running = quadratic_from_evals<H>(g[0], g[1], g[2]);
            fs.absorb_elems<H>(g, 3);
            rs.push_back(ri);
        }
        const uint8_t* root = p.rec_roots[i].data();
        if (i == R - 1) {
            fs.absorb_elems<H>(p.yr.data(), p.yr.size());
            depth = log_dims[i] + log_inv_rate;
            q = fs.get_distinct_queries((size_t)1 << depth, S);
            if (p.final_.n_cols != ((size_t)1 << ks[i]) || !is_pow2(p.yr.size()) || p.yr.size() < 2) return LIG_E_MALFORMED;
            if (!check_openings<H>(root, p.final_, depth, q)) return LIG_E_MERKLE;
            if (!ligero_verify<H>(q, p.final_, p.yr, rs)) return LIG_E_LIGERO;
            T final_r = fs.get_field<H>();
            Vec<H> fpe = partial_eval<H>(p.yr, &final_r, 1);
            int k = ilog2(fpe.size());
            // verify_partial (ligerito.rs:282-295)
            T sum = running.eval(final_r);
            ris.push_back(final_r);
            if (basis[0].size() >> ris.size() != fpe.size()) return 0;
            Vec<H> acc = partial_eval<H>(basis[0], ris.data(), ris.size());
            for (size_t b = 1; b < basis.size(); b++) {
                int nvb = ilog2(basis[b].size());
                size_t cnt = (size_t)(nvb - k);
                Vec<H> bi = partial_eval<H>(basis[b], ris.data() + (ris.size() - cnt), cnt);
                if (bi.size() != acc.size()) return 0;
                for (size_t j = 0; j < acc.size(); j++) acc[j] ^= H::mul(sep[b], bi[j]);
            }
            T tot = H::zero();
            for (size_t j = 0; j < fpe.size(); j++) tot ^= H::mul(fpe[j], acc[j]);
            return H::eq(tot, sum) ? 1 : 0;
        }
        fs.absorb(p.rec_roots[i + 1].data(), 32);
        depth = log_dims[i] + log_inv_rate;
        if ((size_t)i >= p.rec_proofs.size()) return LIG_E_MALFORMED;
        const Rlp<H>& lp = p.rec_proofs[i];
        q = fs.get_distinct_queries((size_t)1 << depth, S);
        if (lp.n_cols != ((size_t)1 << ks[i])) return LIG_E_MALFORMED;
        if (!check_openings<H>(root, lp, depth, q)) return LIG_E_MERKLE;
        T a = fs.get_field<H>();
        sks = eval_sk_at_vks<H>((size_t)1 << log_dims[i]);
        Vec<H> bn;
        T h;
        induce(log_dims[i], sks, lp.rows, S, lp.n_cols, rs, q, a, bn, h);
        read_tr(g);
        if (!H::eq(g[0] ^ g[1], h)) return LIG_E_ASSERT;
        basis.push_back(bn);
        Quad<H> to_glue = quadratic_from_evals<H>(g[0], g[1], g[2]);
        fs.absorb_elems<H>(g, 3);
        T beta = fs.get_field<H>();
        sep.push_back(beta);
        running = fold_quadratic<H>(running, to_glue, beta);
    }
    return LIG_E_ASSERT;  // verifier.rs:163
}

template std::vector<uint8_t> serialize<H32>(const Proof<H32>&);
template std::vector<uint8_t> serialize<H128>(const Proof<H128>&);
template bool deserialize<H32>(const uint8_t*, size_t, Proof<H32>&);
template bool deserialize<H128>(const uint8_t*, size_t, Proof<H128>&);
template size_t proof_size<H32>(const Proof<H32>&);
template size_t proof_size<H128>(const Proof<H128>&);
template int verify<H32>(const Shape&, const Proof<H32>&, uint32_t, Ctx*);
template int verify<H128>(const Shape&, const Proof<H128>&, uint32_t, Ctx*);

}  // namespace lig
