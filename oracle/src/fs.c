/* ORACLE — test infrastructure only.  Restates src/emulated_fs.rs. */
#include "fs.h"
#include <stdlib.h>
#include <string.h>

void ofs_new(ofs *fs, uint32_t seed) {
    uint8_t b[4] = {(uint8_t)seed, (uint8_t)(seed >> 8), (uint8_t)(seed >> 16), (uint8_t)(seed >> 24)};
    osha256_init(&fs->hasher);
    osha256_update(&fs->hasher, b, 4);
    fs->counter = 0;
}

void ofs_absorb(ofs *fs, const void *bytes, size_t len) { osha256_update(&fs->hasher, bytes, len); }

ou128 ofs_squeeze(ofs *fs) {
    uint32_t c = fs->counter;
    uint8_t b[4] = {(uint8_t)c, (uint8_t)(c >> 8), (uint8_t)(c >> 16), (uint8_t)(c >> 24)};
    osha256_update(&fs->hasher, b, 4);
    fs->counter++;
    uint8_t d[32];
    osha256_final_copy(&fs->hasher, d);
    ou128 v = 0;
    for (int i = 15; i >= 0; i--) v = (v << 8) | d[i];
    return v;
}

size_t ofs_get_query(ofs *fs, size_t n) {
    /* `squeeze() as usize` keeps the low 64 bits on a 64-bit target. */
    uint64_t v = (uint64_t)ofs_squeeze(fs);
    return (size_t)(v % (uint64_t)n);
}

static int cmp_size(const void *a, const void *b) {
    size_t x = *(const size_t *)a, y = *(const size_t *)b;
    return x < y ? -1 : (x > y);
}

int ofs_get_distinct_queries(ofs *fs, size_t n, size_t s, size_t *out) {
    if (s > n) return -1;
    uint8_t *seen = (uint8_t *)calloc((n + 7) / 8, 1);
    size_t got = 0;
    while (got < s) {
        size_t q = ofs_get_query(fs, n);
        if (!(seen[q >> 3] & (1u << (q & 7)))) {
            seen[q >> 3] |= (uint8_t)(1u << (q & 7));
            out[got++] = q;
        }
    }
    free(seen);
    qsort(out, s, sizeof(size_t), cmp_size);
    return 0;
}
