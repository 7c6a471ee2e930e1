/* ORACLE — test infrastructure only (see oracle/README.md).
 *
 * BLAKE3 (unkeyed hash mode, 32-byte output), written from the published
 * BLAKE3 specification (O'Connor, Aumasson, Neves, Wilcox-O'Hearn, 2020).
 * The reference crate calls the third-party `blake3 = 1.8.2` crate
 * (Cargo.lock) at src/merkle_tree.rs:15,50,114-117,176; that crate is not
 * vendored, so this restates the published algorithm.
 */
#ifndef ORACLE_BLAKE3_H
#define ORACLE_BLAKE3_H
#include <stddef.h>
#include <stdint.h>

typedef struct {
    uint32_t cv[8];
    uint64_t chunk_counter;
    uint8_t block[64];
    uint8_t block_len;
    uint8_t blocks_compressed;
    uint32_t flags;
} ob3_chunk_state;

typedef struct {
    ob3_chunk_state chunk;
    uint32_t cv_stack[54][8];
    uint8_t cv_stack_len;
} ob3_hasher;

void ob3_hasher_init(ob3_hasher *h);
void ob3_hasher_update(ob3_hasher *h, const void *input, size_t len);
void ob3_hasher_finalize(const ob3_hasher *h, uint8_t out[32]);
/* One-shot hash: blake3::hash(input). */
void ob3_hash(const void *input, size_t len, uint8_t out[32]);

#endif
