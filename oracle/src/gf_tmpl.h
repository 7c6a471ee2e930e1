/* ORACLE — test infrastructure only.
 *
 * C restatement of the reference's field-generic code, "instantiated" once per
 * field by including this file with:
 *   FT     element storage type (uint16_t, uint32_t, unsigned __int128)
 *   FBITS  field width in bits
 *   FPOLY  irreducible polynomial without the x^FBITS term
 *   FN(x)  name prefix macro (e.g. gf32_##x)
 *
 * Every function cites the reference file:line it follows.  The arithmetic
 * deliberately keeps the reference's bit-serial multiply so that the CPU
 * baseline times the reference algorithm.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define FBYTES (FBITS / 8)

/* ---------------------------------------------------------------- field --- */

/* binary_field.rs:83-105.  The conditional XORs are written as masks (the
 * select a compiler emits for the reference's `if`), the loop and the
 * operation count are the reference's. */
static void FN(poly_multiply_gf2)(FT a, FT b, FT *hi, FT *lo) {
    FT rh = 0, rl = 0, th = 0, tl = a, tb = b;
    while (tb != 0) {
        FT m = (FT)0 - (FT)(tb & 1);
        rl ^= tl & m;
        rh ^= th & m;
        th = (FT)((FT)(th << 1) | (FT)(tl >> (FBITS - 1)));
        tl = (FT)(tl << 1);
        tb >>= 1;
    }
    *hi = rh;
    *lo = rl;
}

/* binary_field.rs:52-81 */
static FT FN(mod_irreducible)(FT hi, FT lo) {
    const FT irr = (FT)FPOLY;
    for (int i = 2 * FBITS - 1; i >= FBITS; i--) {
        int sh = i - FBITS;
        FT m = (FT)0 - (FT)((hi >> sh) & 1);
        if (sh == 0) {
            lo ^= irr & m;
        } else {
            hi ^= (FT)(irr >> (FBITS - sh)) & m;
            lo ^= (FT)(irr << sh) & m;
        }
    }
    return lo;
}

/* binary_field.rs:132-140 */
FT FN(mul)(FT a, FT b) {
    FT hi, lo;
    FN(poly_multiply_gf2)(a, b, &hi, &lo);
    return FN(mod_irreducible)(hi, lo);
}

/* binary_field.rs:21-38 */
FT FN(pow)(FT a, FT e) {
    if (e == 0) return (FT)1;
    FT r = 1, base = a;
    while (e > 0) {
        if (e & 1) r = FN(mul)(r, base);
        base = FN(mul)(base, base);
        e >>= 1;
    }
    return r;
}

/* binary_field.rs:40-50; returns -1 for zero ("Cannot compute inverse of zero"). */
int FN(inverse)(FT a, FT *out) {
    if (a == 0) return -1;
    FT e = (FT)0;
    e = (FT)(~e) - 1; /* 2^FBITS - 2 */
    *out = FN(pow)(a, e);
    return 0;
}

static FT FN(inv_or_die)(FT a) {
    FT r = 0;
    if (FN(inverse)(a, &r) != 0) abort();
    return r;
}

static void FN(to_le)(FT v, uint8_t *out) {
    for (int i = 0; i < FBYTES; i++) out[i] = (uint8_t)(v >> (8 * i));
}

static FT FN(from_le)(const uint8_t *in) {
    FT v = 0;
    for (int i = FBYTES - 1; i >= 0; i--) v = (FT)((v << 8) | in[i]);
    return v;
}

/* ------------------------------------------------------------------ FFT --- */

/* binary_fft.rs:100-107 */
static FT FN(next_s)(FT s_prev, FT s_prev_at_root) {
    return FN(mul)(s_prev, s_prev) ^ FN(mul)(s_prev_at_root, s_prev);
}

/* binary_fft.rs:30-58 with beta = 0 (layer_0 :60-72, layer_i :74-89). */
void FN(compute_twiddles)(int log_n, FT *tw) {
    size_t layer_size = (size_t)1 << (log_n - 1);
    FT *layer = (FT *)malloc(layer_size * sizeof(FT));
    size_t write_at = layer_size;
    for (size_t i = 0; i < layer_size; i++) layer[i] = (FT)(i << 1); /* layer_0 */
    FT s_prev_at_root = 1;
    memcpy(tw + (write_at - 1), layer, layer_size * sizeof(FT));
    for (int it = 0; it < log_n - 1; it++) {
        write_at >>= 1;
        layer_size = write_at;
        /* layer_i: compute_s_at_root on the previous layer, then next_s in place */
        FT s_at_root = FN(next_s)(layer[1] ^ layer[0], s_prev_at_root);
        for (size_t w = 0; w < layer_size; w++) layer[w] = FN(next_s)(layer[2 * w], s_prev_at_root);
        s_prev_at_root = s_at_root;
        FT s_inv = FN(inv_or_die)(s_prev_at_root);
        for (size_t i = 0; i < layer_size; i++) tw[write_at + i - 1] = FN(mul)(s_inv, layer[i]);
    }
    free(layer);
}

/* binary_fft.rs:180-192 */
static void FN(fft_mul)(FT *v, size_t n, FT lambda) {
    size_t h = n / 2;
    FT *u = v, *w = v + h;
    for (size_t i = 0; i < h; i++) {
        u[i] = u[i] ^ FN(mul)(lambda, w[i]);
        w[i] = w[i] ^ u[i];
    }
}

/* binary_fft.rs:194-206 */
static void FN(ifft_mul)(FT *v, size_t n, FT lambda) {
    size_t h = n / 2;
    FT *u = v, *w = v + h;
    for (size_t i = 0; i < h; i++) {
        w[i] = w[i] ^ u[i];
        u[i] = u[i] ^ FN(mul)(lambda, w[i]);
    }
}

/* binary_fft.rs:143-159 (the rayon split of :109-141 computes the same thing) */
static void FN(fft_twiddles)(FT *v, size_t n, const FT *tw, size_t idx) {
    if (n == 1) return;
    FN(fft_mul)(v, n, tw[idx - 1]);
    FN(fft_twiddles)(v, n / 2, tw, idx * 2);
    FN(fft_twiddles)(v + n / 2, n / 2, tw, idx * 2 + 1);
}

/* binary_fft.rs:161-178 */
static void FN(ifft_twiddles)(FT *v, size_t n, const FT *tw, size_t idx) {
    if (n == 1) return;
    FN(ifft_twiddles)(v, n / 2, tw, idx * 2);
    FN(ifft_twiddles)(v + n / 2, n / 2, tw, idx * 2 + 1);
    FN(ifft_mul)(v, n, tw[idx - 1]);
}

/* binary_fft.rs:4-17 */
void FN(fft)(FT *v, size_t n, const FT *tw) { FN(fft_twiddles)(v, n, tw, 1); }
/* binary_fft.rs:19-28 */
void FN(ifft)(FT *v, size_t n, const FT *tw) { FN(ifft_twiddles)(v, n, tw, 1); }

static int FN(ilog2)(size_t n) {
    int l = 0;
    while (((size_t)1 << (l + 1)) <= n) l++;
    return l;
}

/* binary_fft.rs:217-246; out has log2(n)+1 entries. */
void FN(eval_sk_at_vks)(size_t n, FT *sk) {
    int ns = FN(ilog2)(n);
    FT *layer = (FT *)malloc((size_t)(ns + 1) * sizeof(FT));
    sk[0] = 1;
    for (int i = 1; i <= ns; i++) layer[i - 1] = (FT)((FT)1 << i);
    int cur_len = ns;
    for (int i = 1; i <= ns; i++) {
        for (int j = 1; j <= cur_len; j++) {
            if (j == 1)
                sk[i] = FN(mul)(layer[0], layer[0]) ^ FN(mul)(sk[i - 1], layer[0]);
            else
                layer[j - 2] = FN(mul)(layer[j - 1], layer[j - 1]) ^ FN(mul)(sk[i - 1], layer[j - 1]);
        }
        cur_len--;
    }
    free(layer);
}

/* binary_fft.rs:276-301 (evaluate_basis :248-274 is the alpha = 1 case). */
void FN(evaluate_scaled_basis_inplace)(FT *sks_x, FT *basis, size_t basis_len, const FT *sks_vks,
                                       FT x, FT alpha) {
    int ns = FN(ilog2)(basis_len);
    sks_x[0] = x;
    for (int i = 2; i <= ns; i++) sks_x[i - 1] = FN(next_s)(sks_x[i - 2], sks_vks[i - 2]);
    basis[0] = alpha;
    for (int i = 1; i <= ns; i++) {
        size_t cur = (size_t)1 << (i - 1);
        for (size_t j = 1; j <= cur; j++) basis[j + cur - 1] = FN(mul)(sks_x[i - 1], basis[j - 1]);
    }
}

/* binary_fft.rs:248-274 */
void FN(evaluate_basis)(size_t basis_len, const FT *sks_vks, FT x, FT *basis) {
    FT sks_x[130];
    FN(evaluate_scaled_basis_inplace)(sks_x, basis, basis_len, sks_vks, x, (FT)1);
}

/* binary_fft.rs:303-322; n_sks = len(sks_vks), pis_len = 2^(n_sks-1). */
void FN(compute_pis)(size_t pis_len, const FT *sks, int n_sks, FT *pis) {
    (void)pis_len;
    pis[0] = 1;
    for (int i = 2; i <= n_sks; i++) {
        size_t cur = (size_t)1 << (i - 2);
        for (size_t j = 1; j <= cur; j++) pis[j + cur - 1] = FN(mul)(sks[i - 2], pis[j - 1]);
    }
}

/* ------------------------------------------------------- Reed–Solomon --- */

/* reed_solomon.rs:7-17 */
typedef struct {
    int log_msg, log_block;
    FT *tw;  /* 2^log_block - 1 */
    FT *pis; /* 2^log_msg */
} FN(rs);

/* reed_solomon.rs:49-64 */
int FN(rs_new)(FN(rs) *rs, size_t msg_len, size_t block_len) {
    if (msg_len == 0 || (msg_len & (msg_len - 1)) || block_len == 0 || (block_len & (block_len - 1)) ||
        msg_len >= block_len)
        return -1;
    rs->log_msg = FN(ilog2)(msg_len);
    rs->log_block = FN(ilog2)(block_len);
    rs->tw = (FT *)malloc((block_len - 1) * sizeof(FT));
    FN(compute_twiddles)(rs->log_block, rs->tw);
    FT *sks = (FT *)malloc((size_t)(rs->log_msg + 1) * sizeof(FT));
    FN(eval_sk_at_vks)(msg_len, sks);
    rs->pis = (FT *)malloc(msg_len * sizeof(FT));
    FN(compute_pis)(msg_len, sks, rs->log_msg + 1, rs->pis);
    free(sks);
    return 0;
}

void FN(rs_free)(FN(rs) *rs) {
    free(rs->tw);
    free(rs->pis);
    rs->tw = rs->pis = NULL;
}

/* reed_solomon.rs:19-41 */
static void FN(short_from_long_tw)(const FT *l_tw, int n, int k, FT *s_tw) {
    size_t jump = (size_t)1 << (n - k);
    s_tw[0] = l_tw[jump - 1];
    size_t idx = 2;
    for (int i = 1; i < k; i++) {
        jump *= 2;
        size_t take = (size_t)1 << i;
        for (size_t j = 0; j < take; j++) s_tw[idx + j - 1] = l_tw[jump + j - 1];
        idx += take;
    }
}

/* reed_solomon.rs:90-102: systematic encode; out has block length. */
void FN(rs_encode)(const FN(rs) *rs, const FT *msg, FT *out) {
    size_t m = (size_t)1 << rs->log_msg, M = (size_t)1 << rs->log_block;
    memset(out, 0, M * sizeof(FT));
    memcpy(out, msg, m * sizeof(FT));
    FT *s_tw = (FT *)malloc(m * sizeof(FT));
    FN(short_from_long_tw)(rs->tw, rs->log_block, rs->log_msg, s_tw);
    FN(ifft)(out, m, s_tw);
    FN(fft)(out, M, rs->tw);
    free(s_tw);
}

/* reed_solomon.rs:104-113: message has block length, first m entries set. */
void FN(rs_encode_non_systematic)(const FN(rs) *rs, FT *msg) {
    size_t m = (size_t)1 << rs->log_msg, M = (size_t)1 << rs->log_block;
    for (size_t i = 0; i < m; i++) msg[i] = FN(mul)(msg[i], rs->pis[i]);
    FN(fft)(msg, M, rs->tw);
}

/* ------------------------------------------------ multilinear / utils --- */

/* utils.rs:4-27: 2^k entries, rs[0] is the most significant index bit. */
void FN(lagrange_basis)(const FT *rs, int k, FT *out) {
    FT *cur = (FT *)malloc(((size_t)1 << k) * sizeof(FT));
    FT *nxt = (FT *)malloc(((size_t)1 << k) * sizeof(FT));
    cur[0] = 1 ^ rs[0];
    cur[1] = rs[0];
    size_t len = 2;
    for (int i = 1; i < k; i++) {
        FT rp1 = 1 ^ rs[i];
        for (size_t j = 0; j < len; j++) {
            nxt[2 * j] = FN(mul)(cur[j], rp1);
            nxt[2 * j + 1] = FN(mul)(cur[j], rs[i]);
        }
        FT *t = cur; cur = nxt; nxt = t;
        len *= 2;
    }
    memcpy(out, cur, len * sizeof(FT));
    free(cur);
    free(nxt);
}

/* multilinear_poly.rs:66-87; writes len >> k values. */
void FN(partial_eval)(const FT *e, size_t len, const FT *rs, int k, FT *out) {
    size_t n = len / 2;
    FT *pe = (FT *)malloc(n * sizeof(FT));
#pragma omp parallel for if (n > 4096)
    for (size_t i = 0; i < n; i++) pe[i] = FN(mul)(1 ^ rs[0], e[i]) ^ FN(mul)(rs[0], e[i + n]);
    for (int i = 1; i < k; i++) {
        n /= 2;
#pragma omp parallel for if (n > 4096)
        for (size_t j = 0; j < n; j++) pe[j] = FN(mul)(1 ^ rs[i], pe[j]) ^ FN(mul)(rs[i], pe[j + n]);
    }
    memcpy(out, pe, n * sizeof(FT));
    free(pe);
}

/* multilinear_poly.rs:90-131 */
void FN(eval_013_product)(const FT *f, const FT *g, size_t len, FT out[3]) {
    size_t h = len / 2;
    FT s0 = 0, s1 = 0, s2 = 0, three = 3, four = 1 ^ 3;
    for (size_t i = 0; i < h; i++) s0 ^= FN(mul)(f[i], g[i]);
    for (size_t i = 0; i < h; i++) s1 ^= FN(mul)(f[h + i], g[h + i]);
    for (size_t i = 0; i < h; i++) {
        FT f2 = FN(mul)(f[h + i], three) ^ FN(mul)(four, f[i]);
        FT g2 = FN(mul)(g[h + i], three) ^ FN(mul)(four, g[i]);
        s2 ^= FN(mul)(f2, g2);
    }
    out[0] = s0; out[1] = s1; out[2] = s2;
}

/* sumcheck_polys.rs:6-20 */
static void FN(alpha_powers)(FT alpha, size_t n, FT *out) {
    out[0] = 1;
    for (size_t i = 1; i < n; i++) out[i] = FN(mul)(out[i - 1], alpha);
}

/* sumcheck_polys.rs:22-72 (the rayon fold/reduce of :74-160 XORs the same
 * terms, so it produces the same result).  rows is n_rows x n_cols with
 * n_cols == 2^k.  basis_out has 2^n entries. */
int FN(induce_sumcheck_poly)(int n, const FT *sks_vks, int n_sks, const FT *rows, size_t n_rows,
                             size_t n_cols, const FT *v_ch, int k, const size_t *queries, FT alpha,
                             FT *basis_out, FT *enforced_out) {
    if (((size_t)1 << k) != n_cols) return -1; /* sumcheck_polys.rs:168 */
    FT *gr = (FT *)malloc(n_cols * sizeof(FT));
    FN(lagrange_basis)(v_ch, k, gr);
    FT *ap = (FT *)malloc((n_rows ? n_rows : 1) * sizeof(FT));
    FN(alpha_powers)(alpha, n_rows, ap);
    size_t N = (size_t)1 << n;
    memset(basis_out, 0, N * sizeof(FT));
    FT enforced = 0;
    /* sumcheck_polys.rs:233-288: per-worker accumulators, XOR-reduced. */
#pragma omp parallel
    {
        FT *local = (FT *)malloc(N * sizeof(FT));
        FT *acc = (FT *)calloc(N, sizeof(FT));
        FT *sks_x = (FT *)malloc((size_t)(n_sks + 1) * sizeof(FT));
        FT enf_local = 0;
#pragma omp for schedule(dynamic)
        for (size_t i = 0; i < n_rows; i++) {
            const FT *row = rows + i * n_cols;
            FT dot = 0;
            for (size_t j = 0; j < n_cols; j++) dot ^= FN(mul)(row[j], gr[j]);
            enf_local ^= FN(mul)(dot, ap[i]);
            FN(evaluate_scaled_basis_inplace)(sks_x, local, N, sks_vks, (FT)queries[i], ap[i]);
            for (size_t j = 0; j < N; j++) acc[j] ^= local[j];
        }
#pragma omp critical
        {
            for (size_t j = 0; j < N; j++) basis_out[j] ^= acc[j];
            enforced ^= enf_local;
        }
        free(local); free(acc); free(sks_x);
    }
    *enforced_out = enforced;
    free(gr); free(ap);
    return 0;
}

/* ------------------------------------------------------- sumcheck.rs --- */

typedef struct { FT a, b, c; } FN(quad);

/* sumcheck.rs:67-81 (x = 3) */
static FN(quad) FN(quadratic_from_evals)(FT at0, FT at1, FT atx) {
    FT x = 3;
    FT num = atx ^ at0 ^ FN(mul)(x, at1 ^ at0);
    FT den = FN(mul)(x, x) ^ x;
    FN(quad) q;
    q.a = FN(mul)(num, FN(inv_or_die)(den));
    q.b = at1 ^ at0 ^ q.a;
    q.c = at0;
    return q;
}

/* sumcheck.rs:60-62 */
static FT FN(eval_quadratic)(FN(quad) p, FT r) {
    return FN(mul)(FN(mul)(p.a, r), r) ^ FN(mul)(p.b, r) ^ p.c;
}

/* sumcheck.rs:83-94 */
static FN(quad) FN(fold_quadratic)(FN(quad) p1, FN(quad) p2, FT alpha) {
    FN(quad) q;
    q.a = p1.a ^ FN(mul)(alpha, p2.a);
    q.b = p1.b ^ FN(mul)(alpha, p2.b);
    q.c = p1.c ^ FN(mul)(alpha, p2.c);
    return q;
}

/* ---------------------------------------------------------- ligero.rs --- */

/* data_structures.rs:32-40 */
typedef struct {
    FT *flat_mat; /* column-major: flat_mat[j*num_rows + i] */
    omerkle tree;
    size_t num_rows, num_cols;
} FN(witness);

/* ligero.rs:70-95 (poly2flatmat :11-27, encode_cols :29-38, extract_leaves :53-68) */
int FN(ligero_commit)(const FT *poly, size_t m, size_t n, const FN(rs) *rs, FN(witness) *w) {
    const size_t inv_rate = 4;
    size_t mt = m * inv_rate;
    if (((size_t)1 << rs->log_msg) != m || ((size_t)1 << rs->log_block) != mt) return -1;
    FT *mat = (FT *)calloc(mt * n, sizeof(FT));
    for (size_t j = 0; j < n; j++)
        for (size_t i = 0; i < m; i++) mat[j * mt + i] = poly[j * m + i];
#pragma omp parallel for schedule(dynamic)
    for (size_t j = 0; j < n; j++) FN(rs_encode_non_systematic)(rs, mat + j * mt);
    size_t leaf_len = n * FBYTES;
    uint8_t *leaves = (uint8_t *)malloc(mt * leaf_len);
#pragma omp parallel for
    for (size_t r = 0; r < mt; r++)
        for (size_t j = 0; j < n; j++) FN(to_le)(mat[j * mt + r], leaves + r * leaf_len + j * FBYTES);
    int rc = omerkle_build_fixed(&w->tree, leaves, leaf_len, mt);
    free(leaves);
    if (rc) { free(mat); return rc; }
    w->flat_mat = mat;
    w->num_rows = mt;
    w->num_cols = n;
    return 0;
}

void FN(witness_free)(FN(witness) *w) {
    free(w->flat_mat);
    w->flat_mat = NULL;
    omerkle_free(&w->tree);
}

/* ligero.rs:40-51 */
static void FN(extract_row)(const FN(witness) *w, size_t r, FT *out) {
    for (size_t j = 0; j < w->num_cols; j++) out[j] = w->flat_mat[j * w->num_rows + r];
}

/* ligero.rs:97-131; returns 0 when every opened row checks, -1 otherwise. */
int FN(ligero_verify)(const size_t *queries, const FT *rows, size_t n_rows, size_t n_cols,
                      const FT *yr, size_t yr_len, const FT *ch, int k) {
    FT *gr = (FT *)malloc(((size_t)1 << k) * sizeof(FT));
    FN(lagrange_basis)(ch, k, gr);
    int n = FN(ilog2)(yr_len);
    FT *sks = (FT *)malloc((size_t)(n + 1) * sizeof(FT));
    FN(eval_sk_at_vks)(yr_len, sks);
    FT *basis = (FT *)malloc(yr_len * sizeof(FT));
    FT *sks_x = (FT *)malloc((size_t)(n + 1) * sizeof(FT));
    int ok = 0;
    for (size_t i = 0; i < n_rows && ok == 0; i++) {
        FT dot = 0;
        for (size_t j = 0; j < n_cols; j++) dot ^= FN(mul)(rows[i * n_cols + j], gr[j]);
        FN(evaluate_scaled_basis_inplace)(sks_x, basis, yr_len, sks, (FT)queries[i], (FT)1);
        FT e = 0;
        for (size_t j = 0; j < yr_len; j++) e ^= FN(mul)(yr[j], basis[j]);
        if (e != dot) ok = -1;
    }
    free(gr); free(sks); free(basis); free(sks_x);
    return ok;
}

/* ----------------------------------------------------------- configs --- */

/* data_structures.rs:9-30 */
typedef struct {
    int recursive_steps;
    size_t initial_m, initial_n;
    size_t dims_m[8], dims_n[8];
    int initial_k;
    int ks[8];
    FN(rs) initial_rs;
    FN(rs) rs[8];
} FN(config);

/* Generated configuration for 2^L coefficients.  L = 20 and L = 24 give
 * exactly hardcoded_config_20 / _24 (config.rs:6-47, 69-112); other sizes
 * follow the same shape rules.  Returns -1 for unsupported L. */
int FN(config_new)(FN(config) *c, int L) {
    if (L < 16 || L > 27) return -1;
    memset(c, 0, sizeof(*c));
    c->recursive_steps = L <= 23 ? 1 : 2 + (L - 24) / 4;
    c->initial_k = 6;
    c->initial_m = (size_t)1 << (L - 6);
    c->initial_n = (size_t)1 << 6;
    for (int i = 0; i < c->recursive_steps; i++) {
        c->dims_m[i] = (size_t)1 << (L - 6 - 4 * (i + 1));
        c->dims_n[i] = 16;
        c->ks[i] = 4;
    }
    FN(rs_new)(&c->initial_rs, c->initial_m, c->initial_m * 4);
    for (int i = 0; i < c->recursive_steps; i++) FN(rs_new)(&c->rs[i], c->dims_m[i], c->dims_m[i] * 4);
    return 0;
}

/* A configuration with explicit dimensions, the fields of ProverConfig
 * (data_structures.rs:9-21) as hardcoded_config_20/24 fill them
 * (config.rs:19-46).  No checks here: the caller passes a consistent shape. */
int FN(config_from_dims)(FN(config) *c, int recursive_steps, size_t initial_m, size_t initial_n, int initial_k,
                         const size_t *dims_m, const size_t *dims_n, const int *ks) {
    if (recursive_steps < 1 || recursive_steps > 8) return -1;
    memset(c, 0, sizeof(*c));
    c->recursive_steps = recursive_steps;
    c->initial_m = initial_m;
    c->initial_n = initial_n;
    c->initial_k = initial_k;
    for (int i = 0; i < recursive_steps; i++) {
        c->dims_m[i] = dims_m[i];
        c->dims_n[i] = dims_n[i];
        c->ks[i] = ks[i];
    }
    FN(rs_new)(&c->initial_rs, c->initial_m, c->initial_m * 4);
    for (int i = 0; i < c->recursive_steps; i++) FN(rs_new)(&c->rs[i], c->dims_m[i], c->dims_m[i] * 4);
    return 0;
}

void FN(config_free)(FN(config) *c) {
    FN(rs_free)(&c->initial_rs);
    for (int i = 0; i < c->recursive_steps; i++) FN(rs_free)(&c->rs[i]);
}

/* -------------------------------------------------------------- proof --- */

/* data_structures.rs:42-103 */
typedef struct {
    size_t n_rows, n_cols;
    FT *rows;
    size_t n_nodes;
    uint8_t *nodes; /* 32 bytes each */
} FN(rlp);

typedef struct {
    uint8_t initial_root[32];
    FN(rlp) initial_proof;
    size_t n_rec_cm;
    uint8_t rec_roots[8][32];
    size_t n_rec_proofs;
    FN(rlp) rec_proofs[8];
    size_t yr_len;
    FT *yr;
    FN(rlp) final_proof;
    size_t tr_len;
    FT *tr; /* 3 per entry */
} FN(proof);

static void FN(rlp_free)(FN(rlp) *p) {
    free(p->rows);
    free(p->nodes);
    p->rows = NULL;
    p->nodes = NULL;
}

void FN(proof_free)(FN(proof) *p) {
    FN(rlp_free)(&p->initial_proof);
    for (size_t i = 0; i < p->n_rec_proofs; i++) FN(rlp_free)(&p->rec_proofs[i]);
    FN(rlp_free)(&p->final_proof);
    free(p->yr);
    free(p->tr);
    memset(p, 0, sizeof(*p));
}

/* Openings for sorted queries (prover.rs:72-76, 114-125). */
static int FN(open_rows)(const FN(witness) *w, const size_t *q, size_t s, FN(rlp) *out) {
    out->n_rows = s;
    out->n_cols = w->num_cols;
    out->rows = (FT *)malloc(s * w->num_cols * sizeof(FT));
    for (size_t i = 0; i < s; i++) FN(extract_row)(w, q[i], out->rows + i * w->num_cols);
    size_t cap = s * (size_t)w->tree.n_layers + 1;
    out->nodes = (uint8_t *)malloc(cap * 32);
    long np = omerkle_prove(&w->tree, q, s, out->nodes, cap);
    if (np < 0) return -1;
    out->n_nodes = (size_t)np;
    return 0;
}

/* ---------------------------------------------- sumcheck prover state --- */

/* ligerito.rs:7-17 (the list of basis polys; all share f's current size). */
typedef struct {
    FT *f;
    size_t f_len;
    FT *basis[16];
    int n_basis;
    FT *glue; /* to_be_glued */
    FT *tr;
    size_t tr_len;
} FN(scp);

static void FN(scp_push)(FN(scp) *p, const FT t[3]) {
    p->tr = (FT *)realloc(p->tr, (p->tr_len + 1) * 3 * sizeof(FT));
    memcpy(p->tr + 3 * p->tr_len, t, 3 * sizeof(FT));
    p->tr_len++;
}

/* ligerito.rs:69-117 with x = 3 */
static void FN(scp_eval_01x)(const FN(scp) *p, FT out[3]) {
    size_t h = p->f_len / 2;
    FT x = 3, xp1 = 1 ^ 3;
    FT *b0 = (FT *)malloc(h * sizeof(FT)), *b1 = (FT *)malloc(h * sizeof(FT)),
       *b2 = (FT *)malloc(h * sizeof(FT));
    for (size_t i = 0; i < h; i++) {
        b0[i] = p->basis[0][i];
        b1[i] = p->basis[0][h + i];
        b2[i] = FN(mul)(x, b1[i]) ^ FN(mul)(xp1, b0[i]);
    }
    for (int k = 1; k < p->n_basis; k++) {
        for (size_t i = 0; i < h; i++) {
            FT lo = p->basis[k][i], hi = p->basis[k][h + i];
            b0[i] ^= lo;
            b1[i] ^= hi;
            b2[i] ^= FN(mul)(x, hi) ^ FN(mul)(xp1, lo);
        }
    }
    FT s0 = 0, s1 = 0, s2 = 0;
    for (size_t i = 0; i < h; i++) {
        FT f0 = p->f[i], f1 = p->f[h + i];
        FT f2 = FN(mul)(x, f1) ^ FN(mul)(xp1, f0);
        s0 ^= FN(mul)(f0, b0[i]);
        s1 ^= FN(mul)(f1, b1[i]);
        s2 ^= FN(mul)(f2, b2[i]);
    }
    free(b0); free(b1); free(b2);
    out[0] = s0; out[1] = s1; out[2] = s2;
}

/* ligerito.rs:119-127 */
static void FN(scp_fold)(FN(scp) *p, FT r, FT out[3]) {
    size_t nl = p->f_len / 2;
    FT *nf = (FT *)malloc(nl * sizeof(FT));
    FN(partial_eval)(p->f, p->f_len, &r, 1, nf);
    free(p->f);
    p->f = nf;
    for (int k = 0; k < p->n_basis; k++) {
        FT *nb = (FT *)malloc(nl * sizeof(FT));
        FN(partial_eval)(p->basis[k], p->f_len, &r, 1, nb);
        free(p->basis[k]);
        p->basis[k] = nb;
    }
    p->f_len = nl;
    FN(scp_eval_01x)(p, out);
    FN(scp_push)(p, out);
}

static void FN(scp_free)(FN(scp) *p) {
    free(p->f);
    for (int k = 0; k < p->n_basis; k++) free(p->basis[k]);
    free(p->glue);
    free(p->tr);
    memset(p, 0, sizeof(*p));
}

/* ------------------------------------------------------------- prover --- */

#define ORACLE_S 148

static FT FN(get_field)(ofs *fs) { return (FT)ofs_squeeze(fs); } /* emulated_fs.rs:34-54 */

static void FN(absorb_elems)(ofs *fs, const FT *e, size_t n) {
    uint8_t *b = (uint8_t *)malloc(n * FBYTES + 1);
    for (size_t i = 0; i < n; i++) FN(to_le)(e[i], b + i * FBYTES);
    ofs_absorb(fs, b, n * FBYTES);
    free(b);
}

/* prover.rs:15-199.  poly has 2^L entries matching cfg; fs_seed is 1234 in
 * the reference (prover.rs:29) and a parameter here so batched jobs can run
 * their own transcripts.  Returns 0 on success, negative at the reference's
 * panic sites. */
int FN(prover)(const FN(config) *cfg, const FT *poly, uint32_t fs_seed, FN(proof) *proof) {
    memset(proof, 0, sizeof(*proof));
    ofs fs;
    ofs_new(&fs, fs_seed);
    const size_t s = ORACLE_S;
    int rc = 0;

    FN(witness) w0, w1, wprev, wnext;
    if (FN(ligero_commit)(poly, cfg->initial_m, cfg->initial_n, &cfg->initial_rs, &w0)) return -1;
    memcpy(proof->initial_root, w0.tree.layers[w0.tree.n_layers - 1], 32);
    ofs_absorb(&fs, proof->initial_root, 32);

    FT pe0[16];
    for (int i = 0; i < cfg->initial_k; i++) pe0[i] = FN(get_field)(&fs);

    size_t N = cfg->initial_m * cfg->initial_n;
    size_t flen = N >> cfg->initial_k;
    FT *f = (FT *)malloc(flen * sizeof(FT));
    FN(partial_eval)(poly, N, pe0, cfg->initial_k, f);

    if (FN(ligero_commit)(f, cfg->dims_m[0], cfg->dims_n[0], &cfg->rs[0], &w1)) return -1;
    memcpy(proof->rec_roots[0], w1.tree.layers[w1.tree.n_layers - 1], 32);
    proof->n_rec_cm = 1;
    ofs_absorb(&fs, proof->rec_roots[0], 32);

    size_t *q = (size_t *)malloc(s * sizeof(size_t));
    if (ofs_get_distinct_queries(&fs, w0.num_rows, s, q)) return -2;
    FT alpha = FN(get_field)(&fs);

    int nv = FN(ilog2)(flen);
    FT *sks = (FT *)malloc((size_t)(nv + 1) * sizeof(FT));
    FN(eval_sk_at_vks)((size_t)1 << nv, sks);

    if (FN(open_rows)(&w0, q, s, &proof->initial_proof)) return -3;
    FT *basis = (FT *)malloc(flen * sizeof(FT));
    FT enforced;
    if (FN(induce_sumcheck_poly)(nv, sks, nv + 1, proof->initial_proof.rows, s, w0.num_cols, pe0,
                                 cfg->initial_k, q, alpha, basis, &enforced))
        return -4;
    FN(witness_free)(&w0);

    FN(scp) sp;
    memset(&sp, 0, sizeof(sp));
    sp.f = f;
    sp.f_len = flen;
    FT t[3];
    FN(eval_013_product)(sp.f, basis, flen, t); /* ligerito.rs:32-48 */
    if ((t[0] ^ t[1]) != enforced) return -5;
    FN(scp_push)(&sp, t);
    sp.basis[0] = basis;
    sp.n_basis = 1;
    FN(absorb_elems)(&fs, t, 3);

    wprev = w1;
    for (int i = 0; i < cfg->recursive_steps; i++) {
        FT rs[16];
        for (int j = 0; j < cfg->ks[i]; j++) {
            FT ri = FN(get_field)(&fs);
            FN(scp_fold)(&sp, ri, t);
            FN(absorb_elems)(&fs, t, 3);
            rs[j] = ri;
        }
        if (i == cfg->recursive_steps - 1) {
            FN(absorb_elems)(&fs, sp.f, sp.f_len);
            if (ofs_get_distinct_queries(&fs, wprev.num_rows, s, q)) { rc = -2; break; }
            if (FN(open_rows)(&wprev, q, s, &proof->final_proof)) { rc = -3; break; }
            proof->yr_len = sp.f_len;
            proof->yr = (FT *)malloc(sp.f_len * sizeof(FT));
            memcpy(proof->yr, sp.f, sp.f_len * sizeof(FT));
            proof->tr_len = sp.tr_len;
            proof->tr = (FT *)malloc(sp.tr_len * 3 * sizeof(FT));
            memcpy(proof->tr, sp.tr, sp.tr_len * 3 * sizeof(FT));
            FN(witness_free)(&wprev);
            FN(scp_free)(&sp);
            free(q); free(sks);
            return 0;
        }
        if (FN(ligero_commit)(sp.f, cfg->dims_m[i + 1], cfg->dims_n[i + 1], &cfg->rs[i + 1], &wnext))
            return -1;
        memcpy(proof->rec_roots[i + 1], wnext.tree.layers[wnext.tree.n_layers - 1], 32);
        proof->n_rec_cm++;
        ofs_absorb(&fs, proof->rec_roots[i + 1], 32);

        if (ofs_get_distinct_queries(&fs, wprev.num_rows, s, q)) { rc = -2; break; }
        FT a = FN(get_field)(&fs);
        nv = FN(ilog2)(sp.f_len);
        sks = (FT *)realloc(sks, (size_t)(nv + 1) * sizeof(FT));
        FN(eval_sk_at_vks)((size_t)1 << nv, sks);
        FN(rlp) *rp = &proof->rec_proofs[proof->n_rec_proofs++];
        if (FN(open_rows)(&wprev, q, s, rp)) { rc = -3; break; }
        FT *bnew = (FT *)malloc(sp.f_len * sizeof(FT));
        FT enf;
        if (FN(induce_sumcheck_poly)(nv, sks, nv + 1, rp->rows, s, wprev.num_cols, rs, cfg->ks[i], q, a,
                                     bnew, &enf)) { rc = -4; break; }
        /* ligerito.rs:50-57 introduce_new */
        FN(eval_013_product)(sp.f, bnew, sp.f_len, t);
        if ((t[0] ^ t[1]) != enf) { rc = -5; break; }
        FN(scp_push)(&sp, t);
        FN(absorb_elems)(&fs, t, 3);
        /* ligerito.rs:59-67 glue */
        FT beta = FN(get_field)(&fs);
        for (size_t j = 0; j < sp.f_len; j++) bnew[j] = FN(mul)(bnew[j], beta);
        sp.basis[sp.n_basis++] = bnew;
        FN(witness_free)(&wprev);
        wprev = wnext;
    }
    if (rc == 0) rc = -6; /* prover.rs:198 */
    FN(witness_free)(&wprev);
    FN(scp_free)(&sp);
    free(q); free(sks);
    return rc;
}

/* ----------------------------------------------------------- verifier --- */

/* Partial evaluation used by the verifier instance (multilinear_poly.rs:66-87). */
static FT *FN(pe_alloc)(const FT *e, size_t len, const FT *rs, int k) {
    FT *o = (FT *)malloc((len >> k) * sizeof(FT));
    FN(partial_eval)(e, len, rs, k, o);
    return o;
}

static int FN(rows_to_bytes_verify)(const uint8_t root[32], const FN(rlp) *p, int depth,
                                    const size_t *q, size_t s) {
    if (p->n_rows != s) return 0;
    size_t leaf_len = p->n_cols * FBYTES;
    uint8_t *b = (uint8_t *)malloc(s * leaf_len + 1);
    size_t *offs = (size_t *)malloc((s + 1) * sizeof(size_t));
    for (size_t i = 0; i < s; i++) {
        offs[i] = i * leaf_len;
        for (size_t j = 0; j < p->n_cols; j++) FN(to_le)(p->rows[i * p->n_cols + j], b + i * leaf_len + j * FBYTES);
    }
    offs[s] = s * leaf_len;
    int ok = omerkle_verify(root, p->nodes, p->n_nodes, depth, b, offs, q, s, OMERKLE_BLAKE3);
    free(b);
    free(offs);
    return ok;
}

/* verifier.rs:10-164 with the SumcheckVerifierInstance of ligerito.rs:143-296.
 * Returns 1 (accept), 0 (final sumcheck check false), or negative where the
 * reference panics: -1 Merkle opening, -2 Ligero check, -3 sumcheck
 * consistency assert, -4 malformed proof. */
int FN(verifier)(int recursive_steps, int initial_dim, const int *log_dims, int initial_k,
                 const int *ks, const FN(proof) *proof, uint32_t fs_seed) {
    ofs fs;
    ofs_new(&fs, fs_seed);
    const size_t s = ORACLE_S;
    const int log_inv_rate = 2;
    int rc = -4;
    if (proof->n_rec_cm < (size_t)recursive_steps || proof->tr_len < 1) return -4;
    ofs_absorb(&fs, proof->initial_root, 32);
    FT pe0[16];
    for (int i = 0; i < initial_k; i++) pe0[i] = FN(get_field)(&fs);
    ofs_absorb(&fs, proof->rec_roots[0], 32);
    int depth = initial_dim + log_inv_rate;
    size_t *q = (size_t *)malloc(s * sizeof(size_t));
    if (ofs_get_distinct_queries(&fs, (size_t)1 << depth, s, q)) { free(q); return -4; }
    if (proof->initial_proof.n_cols != ((size_t)1 << initial_k)) { free(q); return -4; }
    if (!FN(rows_to_bytes_verify)(proof->initial_root, &proof->initial_proof, depth, q, s)) {
        free(q);
        return -1;
    }
    FT alpha = FN(get_field)(&fs);
    FT *sks = (FT *)malloc((size_t)(initial_dim + 1) * sizeof(FT));
    FN(eval_sk_at_vks)((size_t)1 << initial_dim, sks);
    size_t blen = (size_t)1 << initial_dim;
    FT *b1 = (FT *)malloc(blen * sizeof(FT));
    FT h1;
    FN(induce_sumcheck_poly)(initial_dim, sks, initial_dim + 1, proof->initial_proof.rows, s,
                             proof->initial_proof.n_cols, pe0, initial_k, q, alpha, b1, &h1);

    /* verifier instance state */
    FT *bp[16]; size_t bl[16]; int nb = 0;
    FT sep[16];
    FT ris[64]; int nris = 0;
    size_t trr = 0;
    FN(quad) running, to_glue;
    bp[nb] = b1; bl[nb] = blen; sep[nb] = 1; nb++;
    FT sum = h1;
    const FT *tr = proof->tr;
#define READ_TR(g0, g1, g2)                                   \
    do {                                                      \
        if (trr >= proof->tr_len) { rc = -4; goto done; }     \
        g0 = tr[3 * trr]; g1 = tr[3 * trr + 1]; g2 = tr[3 * trr + 2]; \
        trr++;                                                \
    } while (0)
    {
        FT g0, g1, g2;
        READ_TR(g0, g1, g2);
        if ((g0 ^ g1) != sum) { rc = -3; goto done; }
        running = FN(quadratic_from_evals)(g0, g1, g2);
        FT gb[3] = {g0, g1, g2};
        FN(absorb_elems)(&fs, gb, 3);
    }
    for (int i = 0; i < recursive_steps; i++) {
        FT rs[16];
        for (int j = 0; j < ks[i]; j++) {
            FT ri = FN(get_field)(&fs);
            /* ligerito.rs:193-203 fold */
            ris[nris++] = ri;
            sum = FN(eval_quadratic)(running, ri);
            FT g0, g1, g2;
            READ_TR(g0, g1, g2);
            if ((g0 ^ g1) != sum) { rc = -3; goto done; }
            running = FN(quadratic_from_evals)(g0, g1, g2);
            FT gb[3] = {g0, g1, g2};
            FN(absorb_elems)(&fs, gb, 3);
            rs[j] = ri;
        }
        const uint8_t *root = proof->rec_roots[i];
        if (i == recursive_steps - 1) {
            FN(absorb_elems)(&fs, proof->yr, proof->yr_len);
            depth = log_dims[i] + log_inv_rate;
            if (ofs_get_distinct_queries(&fs, (size_t)1 << depth, s, q)) { rc = -4; goto done; }
            if (proof->final_proof.n_cols != ((size_t)1 << ks[i])) { rc = -4; goto done; }
            if (!FN(rows_to_bytes_verify)(root, &proof->final_proof, depth, q, s)) { rc = -1; goto done; }
            if (FN(ligero_verify)(q, proof->final_proof.rows, s, proof->final_proof.n_cols, proof->yr,
                                  proof->yr_len, rs, ks[i])) { rc = -2; goto done; }
            FT final_r = FN(get_field)(&fs);
            FT *fpe = FN(pe_alloc)(proof->yr, proof->yr_len, &final_r, 1);
            size_t fl = proof->yr_len / 2;
            int k = FN(ilog2)(fl);
            /* ligerito.rs:282-295 verify_partial */
            sum = FN(eval_quadratic)(running, final_r);
            ris[nris++] = final_r;
            FT *acc = FN(pe_alloc)(bp[0], bl[0], ris, nris);
            size_t al = bl[0] >> nris;
            for (int b = 1; b < nb; b++) {
                int nvb = FN(ilog2)(bl[b]);
                int cnt = nvb - k;
                FT *bi = FN(pe_alloc)(bp[b], bl[b], ris + (nris - cnt), cnt);
                for (size_t j = 0; j < al; j++) acc[j] ^= FN(mul)(sep[b], bi[j]);
                free(bi);
            }
            FT tot = 0;
            if (al == fl)
                for (size_t j = 0; j < fl; j++) tot ^= FN(mul)(fpe[j], acc[j]);
            rc = (al == fl && tot == sum) ? 1 : 0;
            free(fpe);
            free(acc);
            goto done;
        }
        ofs_absorb(&fs, proof->rec_roots[i + 1], 32);
        depth = log_dims[i] + log_inv_rate;
        if ((size_t)i >= proof->n_rec_proofs) { rc = -4; goto done; }
        const FN(rlp) *lp = &proof->rec_proofs[i];
        if (ofs_get_distinct_queries(&fs, (size_t)1 << depth, s, q)) { rc = -4; goto done; }
        if (lp->n_cols != ((size_t)1 << ks[i])) { rc = -4; goto done; }
        if (!FN(rows_to_bytes_verify)(root, lp, depth, q, s)) { rc = -1; goto done; }
        FT a = FN(get_field)(&fs);
        sks = (FT *)realloc(sks, (size_t)(log_dims[i] + 1) * sizeof(FT));
        FN(eval_sk_at_vks)((size_t)1 << log_dims[i], sks);
        size_t nbl = (size_t)1 << log_dims[i];
        FT *bn = (FT *)malloc(nbl * sizeof(FT));
        FT h;
        FN(induce_sumcheck_poly)(log_dims[i], sks, log_dims[i] + 1, lp->rows, s, lp->n_cols, rs, ks[i], q, a,
                                 bn, &h);
        /* ligerito.rs:206-215 introduce_new */
        FT g0, g1, g2;
        READ_TR(g0, g1, g2);
        if ((g0 ^ g1) != h) { free(bn); rc = -3; goto done; }
        bp[nb] = bn; bl[nb] = nbl;
        to_glue = FN(quadratic_from_evals)(g0, g1, g2);
        FT gb[3] = {g0, g1, g2};
        FN(absorb_elems)(&fs, gb, 3);
        FT beta = FN(get_field)(&fs);
        /* ligerito.rs:217-227 glue */
        sep[nb] = beta;
        nb++;
        running = FN(fold_quadratic)(running, to_glue, beta);
    }
    rc = -4; /* verifier.rs:163 */
done:
#undef READ_TR
    for (int b = 0; b < nb; b++) free(bp[b]);
    free(sks);
    free(q);
    return rc;
}

/* ------------------------------------------------------ serialization --- */
/* Canonical proof encoding (the reference has none): little-endian,
 * FinalizedLigeritoProof field order (data_structures.rs:91-103):
 *   "LGRT" | u32 version=1 | u32 field_bytes
 *   initial root[32]
 *   RLP(initial_ligero_proof)
 *   u64 n | n x root[32]                       recursive_commitments
 *   u64 n | n x RLP                            recursive_proofs
 *   u64 len | len x elem | RLP-rows | nodes    final_ligero_proof (yr, rows, merkle)
 *   u64 n | n x (e0,e1,e2)                     sumcheck_transcript
 *   RLP = u64 n_rows | u64 n_cols | rows | u64 n_nodes | nodes x 32 */

static uint8_t *FN(put_u64)(uint8_t *p, uint64_t v) {
    for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (8 * i));
    return p + 8;
}

static uint8_t *FN(put_rlp)(uint8_t *p, const FN(rlp) *r) {
    p = FN(put_u64)(p, r->n_rows);
    p = FN(put_u64)(p, r->n_cols);
    for (size_t i = 0; i < r->n_rows * r->n_cols; i++, p += FBYTES) FN(to_le)(r->rows[i], p);
    p = FN(put_u64)(p, r->n_nodes);
    memcpy(p, r->nodes, r->n_nodes * 32);
    return p + r->n_nodes * 32;
}

static size_t FN(rlp_size)(const FN(rlp) *r) { return 24 + r->n_rows * r->n_cols * FBYTES + r->n_nodes * 32; }

size_t FN(proof_serialized_size)(const FN(proof) *p) {
    size_t n = 12 + 32 + FN(rlp_size)(&p->initial_proof) + 8 + 32 * p->n_rec_cm + 8;
    for (size_t i = 0; i < p->n_rec_proofs; i++) n += FN(rlp_size)(&p->rec_proofs[i]);
    n += 8 + p->yr_len * FBYTES + FN(rlp_size)(&p->final_proof);
    n += 8 + p->tr_len * 3 * FBYTES;
    return n;
}

void FN(proof_serialize)(const FN(proof) *p, uint8_t *out) {
    uint8_t *q = out;
    memcpy(q, "LGRT", 4); q += 4;
    q[0] = 1; q[1] = q[2] = q[3] = 0; q += 4;
    q[0] = FBYTES; q[1] = q[2] = q[3] = 0; q += 4;
    memcpy(q, p->initial_root, 32); q += 32;
    q = FN(put_rlp)(q, &p->initial_proof);
    q = FN(put_u64)(q, p->n_rec_cm);
    for (size_t i = 0; i < p->n_rec_cm; i++, q += 32) memcpy(q, p->rec_roots[i], 32);
    q = FN(put_u64)(q, p->n_rec_proofs);
    for (size_t i = 0; i < p->n_rec_proofs; i++) q = FN(put_rlp)(q, &p->rec_proofs[i]);
    q = FN(put_u64)(q, p->yr_len);
    for (size_t i = 0; i < p->yr_len; i++, q += FBYTES) FN(to_le)(p->yr[i], q);
    q = FN(put_rlp)(q, &p->final_proof);
    q = FN(put_u64)(q, p->tr_len);
    for (size_t i = 0; i < 3 * p->tr_len; i++, q += FBYTES) FN(to_le)(p->tr[i], q);
}

typedef struct { const uint8_t *p, *end; int bad; } FN(rd);

static uint64_t FN(get_u64)(FN(rd) *r) {
    if (r->end - r->p < 8) { r->bad = 1; return 0; }
    uint64_t v = 0;
    for (int i = 7; i >= 0; i--) v = (v << 8) | r->p[i];
    r->p += 8;
    return v;
}

static void FN(get_rlp)(FN(rd) *r, FN(rlp) *o) {
    o->n_rows = FN(get_u64)(r);
    o->n_cols = FN(get_u64)(r);
    size_t ne = o->n_rows * o->n_cols;
    if (r->bad || o->n_rows > (1u << 24) || o->n_cols > (1u << 24) ||
        (size_t)(r->end - r->p) < ne * FBYTES) { r->bad = 1; return; }
    o->rows = (FT *)malloc((ne ? ne : 1) * sizeof(FT));
    for (size_t i = 0; i < ne; i++, r->p += FBYTES) o->rows[i] = FN(from_le)(r->p);
    o->n_nodes = FN(get_u64)(r);
    if (r->bad || (size_t)(r->end - r->p) / 32 < o->n_nodes) { r->bad = 1; return; }
    o->nodes = (uint8_t *)malloc(o->n_nodes * 32 + 1);
    memcpy(o->nodes, r->p, o->n_nodes * 32);
    r->p += o->n_nodes * 32;
}

/* Inverse of proof_serialize; returns 0 or -1 for malformed input. */
int FN(proof_deserialize)(const uint8_t *in, size_t len, FN(proof) *p) {
    memset(p, 0, sizeof(*p));
    FN(rd) r = {in, in + len, 0};
    if (len < 44 || memcmp(in, "LGRT", 4) != 0 || in[4] != 1 || in[8] != FBYTES) return -1;
    r.p += 12;
    memcpy(p->initial_root, r.p, 32); r.p += 32;
    FN(get_rlp)(&r, &p->initial_proof);
    p->n_rec_cm = FN(get_u64)(&r);
    if (r.bad || p->n_rec_cm > 8 || (size_t)(r.end - r.p) < 32 * p->n_rec_cm) goto bad;
    for (size_t i = 0; i < p->n_rec_cm; i++, r.p += 32) memcpy(p->rec_roots[i], r.p, 32);
    p->n_rec_proofs = FN(get_u64)(&r);
    if (r.bad || p->n_rec_proofs > 8) goto bad;
    for (size_t i = 0; i < p->n_rec_proofs && !r.bad; i++) FN(get_rlp)(&r, &p->rec_proofs[i]);
    p->yr_len = FN(get_u64)(&r);
    if (r.bad || (size_t)(r.end - r.p) / FBYTES < p->yr_len) goto bad;
    p->yr = (FT *)malloc((p->yr_len ? p->yr_len : 1) * sizeof(FT));
    for (size_t i = 0; i < p->yr_len; i++, r.p += FBYTES) p->yr[i] = FN(from_le)(r.p);
    FN(get_rlp)(&r, &p->final_proof);
    p->tr_len = FN(get_u64)(&r);
    if (r.bad || (size_t)(r.end - r.p) / (3 * FBYTES) < p->tr_len) goto bad;
    p->tr = (FT *)malloc((p->tr_len ? p->tr_len : 1) * 3 * sizeof(FT));
    for (size_t i = 0; i < 3 * p->tr_len; i++, r.p += FBYTES) p->tr[i] = FN(from_le)(r.p);
    if (r.p != r.end) goto bad;
    return 0;
bad:
    FN(proof_free)(p);
    return -1;
}

#undef FBYTES
