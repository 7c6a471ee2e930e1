/* ORACLE — test infrastructure only.  Restates src/merkle_tree.rs. */
#include "merkle.h"
#include "blake3.h"
#include "sha256.h"
#include <stdlib.h>
#include <string.h>

static int is_pow2(size_t n) { return n != 0 && (n & (n - 1)) == 0; }

static void hash_bytes(int kind, const uint8_t *in, size_t len, uint8_t out[32]) {
    if (kind == OMERKLE_SHA256) {
        osha256_ctx c;
        osha256_init(&c);
        osha256_update(&c, in, len);
        osha256_final_copy(&c, out);
    } else {
        ob3_hash(in, len, out);
    }
}

static int build_from_hashes(omerkle *t, uint8_t *lh, size_t n, int kind) {
    int depth = 0;
    while (((size_t)1 << depth) < n) depth++;
    t->hash_kind = kind;
    t->n_layers = depth + 1;
    t->n_leaves = n;
    t->layers = (uint8_t **)calloc((size_t)t->n_layers, sizeof(uint8_t *));
    t->layers[0] = lh;
    size_t cur = n;
    for (int d = 1; d <= depth; d++) {
        size_t next = (cur + 1) / 2; /* merkle_tree.rs:39 */
        uint8_t *nl = (uint8_t *)malloc(next * 32);
        for (size_t i = 0; i < next; i++) /* merkle_tree.rs:42-52: hash(left32 || right32) */
            hash_bytes(kind, t->layers[d - 1] + i * 64, 64, nl + i * 32);
        t->layers[d] = nl;
        cur = next;
    }
    return 0;
}

int omerkle_build(omerkle *t, const uint8_t *bytes, const size_t *offs, size_t n, int kind) {
    if (!is_pow2(n)) return -1; /* merkle_tree.rs:26 */
    uint8_t *lh = (uint8_t *)malloc(n * 32);
    for (size_t i = 0; i < n; i++) hash_bytes(kind, bytes + offs[i], offs[i + 1] - offs[i], lh + i * 32);
    return build_from_hashes(t, lh, n, kind);
}

int omerkle_build_fixed(omerkle *t, const uint8_t *bytes, size_t leaf_len, size_t n) {
    if (!is_pow2(n)) return -1;
    uint8_t *lh = (uint8_t *)malloc(n * 32);
    for (size_t i = 0; i < n; i++) ob3_hash(bytes + i * leaf_len, leaf_len, lh + i * 32);
    return build_from_hashes(t, lh, n, OMERKLE_BLAKE3);
}

void omerkle_free(omerkle *t) {
    for (int d = 0; d < t->n_layers; d++) free(t->layers[d]);
    free(t->layers);
    t->layers = NULL;
    t->n_layers = 0;
}

/* merkle_tree.rs:62-96 */
static size_t ith_layer(const uint8_t *layer, size_t qlen, size_t *q, uint8_t *out, size_t cap,
                        size_t *np, int *overflow) {
    size_t next = 0, i = 0;
    while (i < qlen) {
        size_t query = q[i];
        size_t sib = query ^ 1;
        next++;
        q[next - 1] = query >> 1;
        int push = 0, step = 1;
        if (i == qlen - 1) {
            push = 1;
            step = 0;
        } else if (query % 2 != 0) {
            push = 1;
        } else if (q[i + 1] != sib) {
            push = 1;
        } else {
            step = 2;
        }
        if (push) {
            if (*np >= cap) { *overflow = 1; return next; }
            memcpy(out + (*np) * 32, layer + sib * 32, 32);
            (*np)++;
        }
        if (step == 0) break;
        i += (size_t)step;
    }
    return next;
}

long omerkle_prove(const omerkle *t, const size_t *queries, size_t nq, uint8_t *out, size_t cap) {
    size_t *q = (size_t *)malloc((nq ? nq : 1) * sizeof(size_t));
    memcpy(q, queries, nq * sizeof(size_t));
    size_t cnt = nq, np = 0;
    int overflow = 0;
    for (int d = 0; d < t->n_layers - 1; d++) {
        cnt = ith_layer(t->layers[d], cnt, q, out, cap, &np, &overflow);
        if (overflow) break;
    }
    free(q);
    return overflow ? -1 : (long)np;
}

static void hash_siblings(int kind, const uint8_t *l, const uint8_t *r, uint8_t out[32]) {
    /* merkle_tree.rs:113-118: Hasher::update(left); update(right) */
    if (kind == OMERKLE_SHA256) {
        uint8_t b[64];
        memcpy(b, l, 32);
        memcpy(b + 32, r, 32);
        hash_bytes(kind, b, 64, out);
        return;
    }
    ob3_hasher h;
    ob3_hasher_init(&h);
    ob3_hasher_update(&h, l, 32);
    ob3_hasher_update(&h, r, 32);
    ob3_hasher_finalize(&h, out);
}

int omerkle_verify(const uint8_t root[32], const uint8_t *proof, size_t n_proof, int depth,
                   const uint8_t *bytes, const size_t *offs, const size_t *indices, size_t nq,
                   int kind) {
    if (nq == 0) return 0;
    uint8_t(*layer)[32] = malloc(nq * 32);
    size_t *q = (size_t *)malloc(nq * sizeof(size_t));
    for (size_t i = 0; i < nq; i++) hash_bytes(kind, bytes + offs[i], offs[i + 1] - offs[i], layer[i]);
    memcpy(q, indices, nq * sizeof(size_t));
    size_t cnt = nq, pc = 0;
    int ok = 1;
    for (int d = 0; d < depth && ok; d++) { /* merkle_tree.rs:120-162 */
        size_t next = 0, i = 0;
        while (i < cnt) {
            size_t query = q[i], sib = query ^ 1;
            next++;
            q[next - 1] = query >> 1;
            uint8_t h[32];
            if (i == cnt - 1) {
                pc++;
                if (pc > n_proof) { ok = 0; break; }
                if (query % 2 != 0) hash_siblings(kind, proof + (pc - 1) * 32, layer[i], h);
                else hash_siblings(kind, layer[i], proof + (pc - 1) * 32, h);
                memcpy(layer[next - 1], h, 32);
                break;
            }
            if (query % 2 != 0) {
                pc++;
                if (pc > n_proof) { ok = 0; break; }
                hash_siblings(kind, proof + (pc - 1) * 32, layer[i], h);
                i += 1;
            } else if (q[i + 1] != sib) {
                pc++;
                if (pc > n_proof) { ok = 0; break; }
                hash_siblings(kind, layer[i], proof + (pc - 1) * 32, h);
                i += 1;
            } else {
                hash_siblings(kind, layer[i], layer[i + 1], h);
                i += 2;
            }
            memcpy(layer[next - 1], h, 32);
        }
        cnt = next;
    }
    int res = ok && cnt == 1 && memcmp(layer[0], root, 32) == 0; /* merkle_tree.rs:187 */
    free(layer);
    free(q);
    return res;
}
