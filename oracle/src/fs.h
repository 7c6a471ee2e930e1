/* ORACLE — test infrastructure only.
 * Fiat–Shamir transcript, restating src/emulated_fs.rs:6-75. */
#ifndef ORACLE_FS_H
#define ORACLE_FS_H
#include <stddef.h>
#include <stdint.h>
#include "sha256.h"

typedef unsigned __int128 ou128;

typedef struct {
    osha256_ctx hasher;
    uint32_t counter;
} ofs;

void ofs_new(ofs *fs, uint32_t seed);                          /* emulated_fs.rs:12-16 */
void ofs_absorb(ofs *fs, const void *bytes, size_t len);       /* emulated_fs.rs:18-20 */
ou128 ofs_squeeze(ofs *fs);                                    /* emulated_fs.rs:22-32 */
size_t ofs_get_query(ofs *fs, size_t n);                       /* emulated_fs.rs:56-59 */
/* emulated_fs.rs:61-75; writes s sorted distinct queries into out.
 * Returns 0, or -1 when s > n (the reference loops forever there). */
int ofs_get_distinct_queries(ofs *fs, size_t n, size_t s, size_t *out);

#endif
