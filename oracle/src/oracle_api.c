/* ORACLE — test infrastructure only.
 *
 * Instantiates the C restatement of the reference (gf_tmpl.h) for
 * GF(2^16), GF(2^32) and GF(2^128) (binary_field.rs:176-189) and exposes a
 * flat C API for the test-suite (ctypes) and the CPU-baseline leg of the
 * benchmark.  Elements cross this API as little-endian byte strings of
 * bits/8 bytes each (BinaryField::to_bytes, binary_field.rs:161-163).
 * Nothing in the product library links or calls this file.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include <time.h>

#include "blake3.h"
#include "fs.h"
#include "merkle.h"
#include "sha256.h"

#define FT uint16_t
#define FBITS 16
#define FPOLY 0x2Du
#define FN(x) gf16_##x
#include "gf_tmpl.h"
#undef FT
#undef FBITS
#undef FPOLY
#undef FN

#define FT uint32_t
#define FBITS 32
#define FPOLY 0x8299u
#define FN(x) gf32_##x
#include "gf_tmpl.h"
#undef FT
#undef FBITS
#undef FPOLY
#undef FN

#define FT unsigned __int128
#define FBITS 128
#define FPOLY 0x87u
#define FN(x) gf128_##x
#include "gf_tmpl.h"
#undef FT
#undef FBITS
#undef FPOLY
#undef FN

#define EXPORT __attribute__((visibility("default")))

/* Thread count of the OpenMP loops (column encode, query induce); the
 * single-core CPU baseline sets it to 1. */
EXPORT void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
EXPORT int oracle_get_threads(void) { return omp_get_max_threads(); }

/* Dispatch helper: body is instantiated for each field with T and P set. */
#define DISPATCH(bits, BODY16, BODY32, BODY128) \
    switch (bits) {                              \
    case 16: BODY16; break;                      \
    case 32: BODY32; break;                      \
    case 128: BODY128; break;                    \
    default: return -100;                        \
    }

/* ---- element (de)serialisation helpers -------------------------------- */
#define LOADV(T, P, src, n, dst)                                         \
    T *dst = (T *)calloc((size_t)(n) != 0 ? (size_t)(n) : 1, sizeof(T)); \
    for (size_t _i = 0; _i < (size_t)(n); _i++) dst[_i] = P##_from_le((const uint8_t *)(src) + _i * sizeof(T));
#define STOREV(T, P, src, n, dst) \
    for (size_t _i = 0; _i < (size_t)(n); _i++) P##_to_le((src)[_i], (uint8_t *)(dst) + _i * sizeof(T));

EXPORT int oracle_gf_mul_batch(int bits, const uint8_t *a, const uint8_t *b, uint8_t *out, size_t n) {
#define BODY(T, P)                                            \
    {                                                         \
        LOADV(T, P, a, n, va);                                \
        LOADV(T, P, b, n, vb);                                \
        for (size_t i = 0; i < n; i++) va[i] = P##_mul(va[i], vb[i]); \
        STOREV(T, P, va, n, out);                             \
        free(va); free(vb);                                   \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return 0;
}

EXPORT int oracle_gf_inverse(int bits, const uint8_t *a, uint8_t *out) {
    int rc = 0;
#define BODY(T, P)                         \
    {                                      \
        T x = P##_from_le(a), r = 0;       \
        rc = P##_inverse(x, &r);           \
        P##_to_le(r, out);                 \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return rc;
}

EXPORT int oracle_gf_pow(int bits, const uint8_t *a, const uint8_t *e, uint8_t *out) {
#define BODY(T, P) { P##_to_le(P##_pow(P##_from_le(a), P##_from_le(e)), out); }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return 0;
}

EXPORT int oracle_compute_twiddles(int bits, int log_n, uint8_t *out) {
    size_t n = ((size_t)1 << log_n) - 1;
#define BODY(T, P)                                   \
    {                                                \
        T *tw = (T *)malloc(n * sizeof(T));          \
        P##_compute_twiddles(log_n, tw);             \
        STOREV(T, P, tw, n, out);                    \
        free(tw);                                    \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return 0;
}

EXPORT int oracle_fft(int bits, uint8_t *v, size_t n, const uint8_t *tw, int inverse) {
    if (n < 2) return -1;
#define BODY(T, P)                                  \
    {                                               \
        LOADV(T, P, v, n, vv);                      \
        LOADV(T, P, tw, n - 1, tt);                 \
        if (inverse) P##_ifft(vv, n, tt); else P##_fft(vv, n, tt); \
        STOREV(T, P, vv, n, v);                     \
        free(vv); free(tt);                         \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return 0;
}

EXPORT int oracle_eval_sk_at_vks(int bits, size_t n, uint8_t *out) {
    int ns = 0;
    while (((size_t)1 << (ns + 1)) <= n) ns++;
#define BODY(T, P)                                        \
    {                                                     \
        T *sk = (T *)malloc((size_t)(ns + 1) * sizeof(T)); \
        P##_eval_sk_at_vks(n, sk);                        \
        STOREV(T, P, sk, ns + 1, out);                    \
        free(sk);                                         \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return ns + 1;
}

EXPORT int oracle_evaluate_basis(int bits, size_t basis_len, const uint8_t *sks, int n_sks,
                                 const uint8_t *x, uint8_t *out) {
#define BODY(T, P)                                            \
    {                                                         \
        LOADV(T, P, sks, n_sks, s);                           \
        T *b = (T *)malloc(basis_len * sizeof(T));            \
        P##_evaluate_basis(basis_len, s, P##_from_le(x), b);  \
        STOREV(T, P, b, basis_len, out);                      \
        free(s); free(b);                                     \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return 0;
}

EXPORT int oracle_compute_pis(int bits, size_t pis_len, const uint8_t *sks, int n_sks, uint8_t *out) {
    if (pis_len != ((size_t)1 << (n_sks - 1))) return -1; /* binary_fft.rs:309 */
#define BODY(T, P)                                    \
    {                                                 \
        LOADV(T, P, sks, n_sks, s);                   \
        T *p = (T *)malloc(pis_len * sizeof(T));      \
        P##_compute_pis(pis_len, s, n_sks, p);        \
        STOREV(T, P, p, pis_len, out);                \
        free(s); free(p);                             \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return 0;
}

/* reed_solomon.rs:90-102 (systematic) and :104-113 (non-systematic, in place
 * over a block-length buffer). */
EXPORT int oracle_rs_encode(int bits, size_t msg_len, size_t block_len, const uint8_t *msg, uint8_t *out,
                            int non_systematic) {
    int rc = 0;
#define BODY(T, P)                                                   \
    {                                                                \
        P##_rs rs;                                                   \
        if (P##_rs_new(&rs, msg_len, block_len)) return -1;          \
        if (non_systematic) {                                        \
            LOADV(T, P, out, block_len, v);                          \
            P##_rs_encode_non_systematic(&rs, v);                    \
            STOREV(T, P, v, block_len, out);                         \
            free(v);                                                 \
        } else {                                                     \
            LOADV(T, P, msg, msg_len, m);                            \
            T *o = (T *)malloc(block_len * sizeof(T));               \
            P##_rs_encode(&rs, m, o);                                \
            STOREV(T, P, o, block_len, out);                         \
            free(m); free(o);                                        \
        }                                                            \
        P##_rs_free(&rs);                                            \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return rc;
}

EXPORT int oracle_lagrange_basis(int bits, const uint8_t *rs, int k, uint8_t *out) {
#define BODY(T, P)                                               \
    {                                                            \
        LOADV(T, P, rs, k, r);                                   \
        T *o = (T *)malloc(((size_t)1 << k) * sizeof(T));        \
        P##_lagrange_basis(r, k, o);                             \
        STOREV(T, P, o, (size_t)1 << k, out);                    \
        free(r); free(o);                                        \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return 0;
}

EXPORT int oracle_partial_eval(int bits, const uint8_t *e, size_t len, const uint8_t *rs, int k, uint8_t *out) {
#define BODY(T, P)                                        \
    {                                                     \
        LOADV(T, P, e, len, ev);                          \
        LOADV(T, P, rs, k, r);                            \
        T *o = (T *)malloc((len >> k) * sizeof(T));       \
        P##_partial_eval(ev, len, r, k, o);               \
        STOREV(T, P, o, len >> k, out);                   \
        free(ev); free(r); free(o);                       \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return 0;
}

EXPORT int oracle_eval_013_product(int bits, const uint8_t *f, const uint8_t *g, size_t len, uint8_t *out) {
#define BODY(T, P)                                 \
    {                                              \
        LOADV(T, P, f, len, fv);                   \
        LOADV(T, P, g, len, gv);                   \
        T o[3];                                    \
        P##_eval_013_product(fv, gv, len, o);      \
        STOREV(T, P, o, 3, out);                   \
        free(fv); free(gv);                        \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return 0;
}

/* sumcheck_polys.rs:22-72 */
EXPORT int oracle_induce(int bits, int n, const uint8_t *rows, size_t n_rows, size_t n_cols,
                         const uint8_t *v_ch, int k, const size_t *queries, const uint8_t *alpha,
                         uint8_t *basis_out, uint8_t *enforced_out) {
    int rc = 0;
#define BODY(T, P)                                                                      \
    {                                                                                   \
        T *sks = (T *)malloc((size_t)(n + 1) * sizeof(T));                              \
        P##_eval_sk_at_vks((size_t)1 << n, sks);                                        \
        LOADV(T, P, rows, n_rows * n_cols, rv);                                         \
        LOADV(T, P, v_ch, k, vc);                                                       \
        T *b = (T *)malloc(((size_t)1 << n) * sizeof(T));                               \
        T enf = 0;                                                                      \
        rc = P##_induce_sumcheck_poly(n, sks, n + 1, rv, n_rows, n_cols, vc, k, queries, \
                                      P##_from_le(alpha), b, &enf);                     \
        STOREV(T, P, b, (size_t)1 << n, basis_out);                                     \
        P##_to_le(enf, enforced_out);                                                   \
        free(sks); free(rv); free(vc); free(b);                                         \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return rc;
}

/* ---- hashes / transcript / Merkle --------------------------------------- */

EXPORT void oracle_blake3(const uint8_t *in, size_t len, uint8_t *out) { ob3_hash(in, len, out); }

/* Incremental hashing with update() pieces of the given sizes. */
EXPORT void oracle_blake3_pieces(const uint8_t *in, const size_t *pieces, size_t n_pieces, uint8_t *out) {
    ob3_hasher h;
    ob3_hasher_init(&h);
    for (size_t i = 0; i < n_pieces; i++) {
        ob3_hasher_update(&h, in, pieces[i]);
        in += pieces[i];
    }
    ob3_hasher_finalize(&h, out);
}

EXPORT void oracle_sha256(const uint8_t *in, size_t len, uint8_t *out) {
    osha256_ctx c;
    osha256_init(&c);
    osha256_update(&c, in, len);
    osha256_final_copy(&c, out);
}

EXPORT void *oracle_fs_new(uint32_t seed) {
    ofs *f = (ofs *)malloc(sizeof(ofs));
    ofs_new(f, seed);
    return f;
}
EXPORT void oracle_fs_free(void *f) { free(f); }
EXPORT void oracle_fs_absorb(void *f, const uint8_t *b, size_t n) { ofs_absorb((ofs *)f, b, n); }
EXPORT void oracle_fs_squeeze(void *f, uint8_t out[16]) {
    ou128 v = ofs_squeeze((ofs *)f);
    for (int i = 0; i < 16; i++) out[i] = (uint8_t)(v >> (8 * i));
}
EXPORT size_t oracle_fs_get_query(void *f, size_t n) { return ofs_get_query((ofs *)f, n); }
EXPORT int oracle_fs_get_distinct_queries(void *f, size_t n, size_t s, size_t *out) {
    return ofs_get_distinct_queries((ofs *)f, n, s, out);
}
EXPORT uint32_t oracle_fs_counter(void *f) { return ((ofs *)f)->counter; }

/* Builds the tree and copies every layer, leaf hashes first, into out
 * (32 * (2n - 1) bytes).  Returns the layer count. */
EXPORT int oracle_merkle_build(const uint8_t *bytes, const size_t *offs, size_t n, uint8_t *out, int hash_kind) {
    omerkle t;
    if (omerkle_build(&t, bytes, offs, n, hash_kind)) return -1;
    uint8_t *p = out;
    for (int d = 0; d < t.n_layers; d++) {
        size_t nodes = n >> d;
        memcpy(p, t.layers[d], nodes * 32);
        p += nodes * 32;
    }
    int nl = t.n_layers;
    omerkle_free(&t);
    return nl;
}

EXPORT long oracle_merkle_prove(const uint8_t *bytes, const size_t *offs, size_t n, const size_t *q, size_t nq,
                                uint8_t *out, size_t cap, int hash_kind) {
    omerkle t;
    if (omerkle_build(&t, bytes, offs, n, hash_kind)) return -2;
    long r = omerkle_prove(&t, q, nq, out, cap);
    omerkle_free(&t);
    return r;
}

EXPORT int oracle_merkle_verify(const uint8_t *root, const uint8_t *proof, size_t n_proof, int depth,
                                const uint8_t *bytes, const size_t *offs, const size_t *idx, size_t nq,
                                int hash_kind) {
    return omerkle_verify(root, proof, n_proof, depth, bytes, offs, idx, nq, hash_kind);
}

/* ---- ligero_commit ------------------------------------------------------- */

/* ligero.rs:70-95.  poly has m*n elements.  Outputs: root (32 B); if
 * flat_out is non-NULL the column-major encoded matrix (4m*n elements); if
 * layers_out is non-NULL every Merkle layer, leaves first. */
EXPORT int oracle_ligero_commit(int bits, const uint8_t *poly, size_t m, size_t n, uint8_t *root,
                                uint8_t *flat_out, uint8_t *layers_out) {
    int rc = 0;
#define BODY(T, P)                                                               \
    {                                                                            \
        P##_rs rs;                                                               \
        if (P##_rs_new(&rs, m, 4 * m)) return -1;                                \
        LOADV(T, P, poly, m *n, pv);                                             \
        P##_witness w;                                                           \
        rc = P##_ligero_commit(pv, m, n, &rs, &w);                               \
        if (rc == 0) {                                                           \
            memcpy(root, w.tree.layers[w.tree.n_layers - 1], 32);                \
            if (flat_out) STOREV(T, P, w.flat_mat, 4 * m * n, flat_out);         \
            if (layers_out) {                                                    \
                uint8_t *lp = layers_out;                                        \
                for (int d = 0; d < w.tree.n_layers; d++) {                      \
                    memcpy(lp, w.tree.layers[d], (4 * m >> d) * 32);             \
                    lp += (4 * m >> d) * 32;                                     \
                }                                                                \
            }                                                                    \
            P##_witness_free(&w);                                                \
        }                                                                        \
        free(pv);                                                                \
        P##_rs_free(&rs);                                                        \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return rc;
}

/* ---- prover / verifier --------------------------------------------------- */

/* prover.rs:15-199 on the generated configuration for 2^L coefficients
 * (L = 20/24 are hardcoded_config_20/24).  If out is NULL only *len is set.
 * The proof is the canonical serialisation documented in gf_tmpl.h. */
EXPORT int oracle_prove(int bits, int L, const uint8_t *poly, uint32_t fs_seed, uint8_t *out, size_t cap,
                        size_t *len) {
    int rc = 0;
#define BODY(T, P)                                                   \
    {                                                                \
        P##_config cfg;                                              \
        if (P##_config_new(&cfg, L)) return -10;                     \
        LOADV(T, P, poly, (size_t)1 << L, pv);                       \
        P##_proof pr;                                                \
        rc = P##_prover(&cfg, pv, fs_seed, &pr);                     \
        if (rc == 0) {                                               \
            size_t sz = P##_proof_serialized_size(&pr);              \
            *len = sz;                                               \
            if (out) {                                               \
                if (cap < sz) rc = -11;                              \
                else P##_proof_serialize(&pr, out);                  \
            }                                                        \
            P##_proof_free(&pr);                                     \
        }                                                            \
        free(pv);                                                    \
        P##_config_free(&cfg);                                       \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return rc;
}

/* The generated configuration's shape for 2^L (the rule of gf32_config_new):
 * dims[2i], dims[2i+1] = (rows, cols) of recursive step i.  Returns the
 * number of recursive steps, or -10 for an unsupported L. */
EXPORT int oracle_config_shape(int L, size_t *initial_m, size_t *initial_n, int *initial_k, size_t *dims, int *ks) {
    gf32_config cfg;
    if (gf32_config_new(&cfg, L)) return -10;
    *initial_m = cfg.initial_m;
    *initial_n = cfg.initial_n;
    *initial_k = cfg.initial_k;
    for (int i = 0; i < cfg.recursive_steps; i++) {
        dims[2 * i] = cfg.dims_m[i];
        dims[2 * i + 1] = cfg.dims_n[i];
        ks[i] = cfg.ks[i];
    }
    int r = cfg.recursive_steps;
    gf32_config_free(&cfg);
    return r;
}

/* prover.rs:15-199 on a configuration given by its dimensions (ProverConfig,
 * data_structures.rs:9-21); dims as in oracle_config_shape.  The input has
 * initial_m * initial_n elements. */
EXPORT int oracle_prove_dims(int bits, int R, size_t initial_m, size_t initial_n, int initial_k, const size_t *dims,
                             const int *ks, const uint8_t *poly, uint32_t fs_seed, uint8_t *out, size_t cap,
                             size_t *len) {
    int rc = 0;
    if (R < 1 || R > 8) return -10;
    size_t dm[8], dn[8];
    for (int i = 0; i < R; i++) {
        dm[i] = dims[2 * i];
        dn[i] = dims[2 * i + 1];
    }
#define BODY(T, P)                                                                 \
    {                                                                              \
        P##_config cfg;                                                            \
        if (P##_config_from_dims(&cfg, R, initial_m, initial_n, initial_k, dm, dn, ks)) return -10; \
        LOADV(T, P, poly, initial_m * initial_n, pv);                              \
        P##_proof pr;                                                              \
        rc = P##_prover(&cfg, pv, fs_seed, &pr);                                   \
        if (rc == 0) {                                                             \
            size_t sz = P##_proof_serialized_size(&pr);                            \
            *len = sz;                                                             \
            if (out) {                                                             \
                if (cap < sz) rc = -11;                                            \
                else P##_proof_serialize(&pr, out);                                \
            }                                                                      \
            P##_proof_free(&pr);                                                   \
        }                                                                          \
        free(pv);                                                                  \
        P##_config_free(&cfg);                                                     \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return rc;
}

/* verifier.rs:10-164 on an explicit VerifierConfig (data_structures.rs:
 * 23-30). */
EXPORT int oracle_verify_dims(int bits, int R, int initial_dim, const int *log_dims, int initial_k, const int *ks,
                              const uint8_t *proof, size_t len, uint32_t fs_seed) {
    int rc = 0;
    if (R < 1 || R > 8) return -10;
#define BODY(T, P)                                                               \
    {                                                                            \
        P##_proof pr;                                                            \
        if (P##_proof_deserialize(proof, len, &pr)) return -20;                  \
        rc = P##_verifier(R, initial_dim, log_dims, initial_k, ks, &pr, fs_seed); \
        P##_proof_free(&pr);                                                     \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return rc;
}

/* The timing the reference's main prints (main.rs:29-49): the configuration
 * (twiddles, pis) is built before the clock starts and the proof bytes are
 * not serialized; *prove_s is the wall time of prover() alone. */
EXPORT int oracle_prove_timed(int bits, int L, const uint8_t *poly, uint32_t fs_seed, double *prove_s) {
    int rc = 0;
#define BODY(T, P)                                                              \
    {                                                                           \
        P##_config cfg;                                                         \
        if (P##_config_new(&cfg, L)) return -10;                                \
        LOADV(T, P, poly, (size_t)1 << L, pv);                                  \
        P##_proof pr;                                                           \
        struct timespec t0, t1;                                                 \
        clock_gettime(CLOCK_MONOTONIC, &t0);                                    \
        rc = P##_prover(&cfg, pv, fs_seed, &pr);                                \
        clock_gettime(CLOCK_MONOTONIC, &t1);                                    \
        *prove_s = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec); \
        if (rc == 0) P##_proof_free(&pr);                                       \
        free(pv);                                                               \
        P##_config_free(&cfg);                                                  \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return rc;
}

/* verifier.rs:10-164 on the generated verifier configuration for 2^L
 * (config.rs:49-67, 114-134 for L = 20/24).  Returns 1 accept, 0 reject,
 * negative at the reference's panic sites or -20 for undecodable bytes. */
EXPORT int oracle_verify(int bits, int L, const uint8_t *proof, size_t len, uint32_t fs_seed) {
    int rc = 0;
    if (L < 16 || L > 27) return -10;
    int R = L <= 23 ? 1 : 2 + (L - 24) / 4;
    int log_dims[8], ks[8];
    for (int i = 0; i < R; i++) {
        log_dims[i] = L - 6 - 4 * (i + 1);
        ks[i] = 4;
    }
#define BODY(T, P)                                                         \
    {                                                                      \
        P##_proof pr;                                                      \
        if (P##_proof_deserialize(proof, len, &pr)) return -20;            \
        rc = P##_verifier(R, L - 6, log_dims, 6, ks, &pr, fs_seed);        \
        P##_proof_free(&pr);                                               \
    }
    DISPATCH(bits, BODY(uint16_t, gf16), BODY(uint32_t, gf32), BODY(unsigned __int128, gf128));
#undef BODY
    return rc;
}

/* SplitMix64 input generator used by every workload (documented seed ->
 * element mapping): element i of a GF(2^32) vector is the low 32 bits of
 * draw i; a GF(2^128) element takes two draws (low word first). */
EXPORT void oracle_splitmix_fill(uint64_t seed, int bits, size_t n, uint8_t *out) {
    uint64_t s = seed;
    size_t eb = (size_t)bits / 8;
    size_t words = eb <= 8 ? 1 : 2;
    for (size_t i = 0; i < n; i++) {
        for (size_t w = 0; w < words; w++) {
            uint64_t z = (s += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            size_t nb = eb <= 8 ? eb : 8;
            for (size_t b = 0; b < nb; b++) out[i * eb + w * 8 + b] = (uint8_t)(z >> (8 * b));
        }
    }
}
