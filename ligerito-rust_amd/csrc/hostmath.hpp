// Host-side field algorithms of the reference that run once per plan or per
// transcript step: twiddles, subspace-polynomial values, pis, Lagrange
// tensor, small partial evaluations and the sumcheck quadratic helpers.
#pragma once
#include <vector>

#include "host.hpp"

namespace lig {

template <class H> using Vec = std::vector<typename H::T>;

// binary_fft.rs:100-107
template <class H> typename H::T next_s(typename H::T s, typename H::T r) { return H::mul(s, s) ^ H::mul(r, s); }

inline int ilog2(size_t n) {
    int l = 0;
    while (((size_t)1 << (l + 1)) <= n) l++;
    return l;
}
inline bool is_pow2(size_t n) { return n && !(n & (n - 1)); }

// binary_fft.rs:30-58 with beta = 0.
template <class H> Vec<H> compute_twiddles(int log_n) {
    using T = typename H::T;
    size_t layer_size = (size_t)1 << (log_n - 1);
    Vec<H> tw(((size_t)1 << log_n) - 1);
    Vec<H> layer(layer_size);
    size_t write_at = layer_size;
    for (size_t i = 0; i < layer_size; i++) layer[i] = H::from_int((uint64_t)i << 1);
    T s_prev = H::one();
    for (size_t i = 0; i < layer_size; i++) tw[write_at - 1 + i] = layer[i];
    for (int it = 0; it < log_n - 1; it++) {
        write_at >>= 1;
        layer_size = write_at;
        T s_root = next_s<H>(layer[1] ^ layer[0], s_prev);
        for (size_t w = 0; w < layer_size; w++) layer[w] = next_s<H>(layer[2 * w], s_prev);
        s_prev = s_root;
        T s_inv = h_inverse<H>(s_prev);
        for (size_t i = 0; i < layer_size; i++) tw[write_at + i - 1] = H::mul(s_inv, layer[i]);
    }
    return tw;
}

// binary_fft.rs:217-246
template <class H> Vec<H> eval_sk_at_vks(size_t n) {
    int ns = ilog2(n);
    Vec<H> sk;
    sk.push_back(H::one());
    Vec<H> layer(ns + 1);
    for (int i = 1; i <= ns; i++) layer[i - 1] = H::from_int((uint64_t)1 << i);
    int cur = ns;
    for (int i = 1; i <= ns; i++) {
        for (int j = 1; j <= cur; j++) {
            if (j == 1)
                sk.push_back(H::mul(layer[0], layer[0]) ^ H::mul(sk[i - 1], layer[0]));
            else
                layer[j - 2] = H::mul(layer[j - 1], layer[j - 1]) ^ H::mul(sk[i - 1], layer[j - 1]);
        }
        cur--;
    }
    return sk;
}

// binary_fft.rs:303-322
template <class H> Vec<H> compute_pis(size_t len, const Vec<H>& sks) {
    Vec<H> pis(len);
    pis[0] = H::one();
    for (size_t i = 2; i <= sks.size(); i++) {
        size_t cur = (size_t)1 << (i - 2);
        for (size_t j = 1; j <= cur; j++) pis[j + cur - 1] = H::mul(sks[i - 2], pis[j - 1]);
    }
    return pis;
}

// sks_at_x[t] = W_t(x) for t < ell (binary_fft.rs:289-292).
template <class H> Vec<H> sks_at(typename H::T x, const Vec<H>& sks_vks, int ell) {
    Vec<H> s(ell > 0 ? ell : 1);
    s[0] = x;
    for (int i = 2; i <= ell; i++) s[i - 1] = next_s<H>(s[i - 2], sks_vks[i - 2]);
    return s;
}

// binary_fft.rs:276-301
template <class H> void scaled_basis(const Vec<H>& sks_vks, typename H::T x, typename H::T alpha, Vec<H>& basis) {
    int ns = ilog2(basis.size());
    Vec<H> s = sks_at<H>(x, sks_vks, ns);
    basis[0] = alpha;
    for (int i = 1; i <= ns; i++) {
        size_t cur = (size_t)1 << (i - 1);
        for (size_t j = 1; j <= cur; j++) basis[j + cur - 1] = H::mul(s[i - 1], basis[j - 1]);
    }
}

// utils.rs:4-27 (rs[0] is the most significant index bit)
template <class H> Vec<H> lagrange_basis(const Vec<H>& rs) {
    Vec<H> cur = {H::one() ^ rs[0], rs[0]};
    for (size_t i = 1; i < rs.size(); i++) {
        Vec<H> nxt(cur.size() * 2);
        typename H::T rp1 = H::one() ^ rs[i];
        for (size_t j = 0; j < cur.size(); j++) {
            nxt[2 * j] = H::mul(cur[j], rp1);
            nxt[2 * j + 1] = H::mul(cur[j], rs[i]);
        }
        cur.swap(nxt);
    }
    return cur;
}

// multilinear_poly.rs:66-87
template <class H> Vec<H> partial_eval(const Vec<H>& e, const typename H::T* rs, size_t k) {
    size_t n = e.size() / 2;
    Vec<H> pe(n);
    for (size_t i = 0; i < n; i++) pe[i] = H::mul(H::one() ^ rs[0], e[i]) ^ H::mul(rs[0], e[i + n]);
    for (size_t i = 1; i < k; i++) {
        n /= 2;
        for (size_t j = 0; j < n; j++) pe[j] = H::mul(H::one() ^ rs[i], pe[j]) ^ H::mul(rs[i], pe[j + n]);
    }
    pe.resize(n);
    return pe;
}

// sumcheck_polys.rs:22-72 on the host (verifier side).
template <class H>
void induce_host(int n, const Vec<H>& sks_vks, const std::vector<typename H::T>& rows, size_t n_rows, size_t n_cols,
                 const Vec<H>& v_ch, const std::vector<size_t>& queries, typename H::T alpha, Vec<H>& basis_out,
                 typename H::T& enforced) {
    LIG_REQUIRE(((size_t)1 << v_ch.size()) == n_cols, LIG_E_ASSERT, "opened row length != 2^k");
    Vec<H> gr = lagrange_basis<H>(v_ch);
    basis_out.assign((size_t)1 << n, H::zero());
    Vec<H> local((size_t)1 << n);
    enforced = H::zero();
    typename H::T ap = H::one();
    for (size_t i = 0; i < n_rows; i++) {
        typename H::T dot = H::zero();
        for (size_t j = 0; j < n_cols; j++) dot ^= H::mul(rows[i * n_cols + j], gr[j]);
        enforced ^= H::mul(dot, ap);
        scaled_basis<H>(sks_vks, H::from_int(queries[i]), ap, local);
        for (size_t j = 0; j < local.size(); j++) basis_out[j] ^= local[j];
        ap = H::mul(ap, alpha);
    }
}

// sumcheck.rs:50-94
template <class H> struct Quad {
    typename H::T a, b, c;
    typename H::T eval(typename H::T r) const { return H::mul(H::mul(a, r), r) ^ H::mul(b, r) ^ c; }
};
template <class H> Quad<H> quadratic_from_evals(typename H::T at0, typename H::T at1, typename H::T atx) {
    typename H::T x = H::from_int(3);
    typename H::T num = atx ^ at0 ^ H::mul(x, at1 ^ at0);
    typename H::T den = H::mul(x, x) ^ x;
    Quad<H> q;
    q.a = H::mul(num, h_inverse<H>(den));
    q.b = at1 ^ at0 ^ q.a;
    q.c = at0;
    return q;
}
template <class H> Quad<H> fold_quadratic(const Quad<H>& p1, const Quad<H>& p2, typename H::T alpha) {
    return {p1.a ^ H::mul(alpha, p2.a), p1.b ^ H::mul(alpha, p2.b), p1.c ^ H::mul(alpha, p2.c)};
}

}  // namespace lig
