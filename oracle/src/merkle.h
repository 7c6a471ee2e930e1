/* ORACLE — test infrastructure only.
 * BLAKE3 Merkle tree, restating src/merkle_tree.rs:4-188. */
#ifndef ORACLE_MERKLE_H
#define ORACLE_MERKLE_H
#include <stddef.h>
#include <stdint.h>

/* layers[0] = leaf hashes, layers[depth] = root; layer d holds n>>d nodes of
 * 32 bytes (Vec<Vec<u8>> of merkle_tree.rs:22-60). */
/* Hash used for leaves and parents.  The reference uses BLAKE3; its Merkle
 * known-answer tests (merkle_tree.rs:195-255) were produced by the Julia
 * implementation with SHA-256, so the SHA-256 variant exists only to check
 * the tree structure and proof order against those vectors. */
enum { OMERKLE_BLAKE3 = 0, OMERKLE_SHA256 = 1 };

typedef struct {
    int hash_kind;
    int n_layers;
    size_t n_leaves;
    uint8_t **layers;
} omerkle;

/* merkle_tree.rs:22-60.  Leaf i is bytes[offs[i] .. offs[i+1]).
 * Returns -1 unless n is a power of two. */
int omerkle_build(omerkle *t, const uint8_t *bytes, const size_t *offs, size_t n, int hash_kind);
/* Fixed-length leaves: leaf i = bytes[i*leaf_len ..]. */
int omerkle_build_fixed(omerkle *t, const uint8_t *bytes, size_t leaf_len, size_t n);
void omerkle_free(omerkle *t);
/* merkle_tree.rs:98-111: queries sorted ascending, distinct.  Writes the
 * batched proof (32-byte nodes) into out (capacity cap nodes); returns the
 * node count, or -1 on overflow. */
long omerkle_prove(const omerkle *t, const size_t *queries, size_t nq, uint8_t *out, size_t cap);
/* merkle_tree.rs:164-188. Leaf i = bytes[offs[i] .. offs[i+1]). */
int omerkle_verify(const uint8_t root[32], const uint8_t *proof, size_t n_proof, int depth,
                   const uint8_t *bytes, const size_t *offs, const size_t *indices, size_t nq,
                   int hash_kind);

#endif
