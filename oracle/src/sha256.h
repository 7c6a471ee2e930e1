/* ORACLE — test infrastructure only.
 * SHA-256 (FIPS 180-4), incremental with cheap copies, standing in for the
 * third-party `sha2 = 0.10.9` crate used by src/emulated_fs.rs:2,13-26. */
#ifndef ORACLE_SHA256_H
#define ORACLE_SHA256_H
#include <stddef.h>
#include <stdint.h>

typedef struct {
    uint32_t h[8];
    uint8_t buf[64];
    uint64_t total;   /* bytes absorbed */
    uint32_t buf_len;
} osha256_ctx;

void osha256_init(osha256_ctx *c);
void osha256_update(osha256_ctx *c, const void *data, size_t len);
/* Finalizes a copy; `c` itself stays usable (Sha256::clone().finalize()). */
void osha256_final_copy(const osha256_ctx *c, uint8_t out[32]);

#endif
