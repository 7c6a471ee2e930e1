/* ORACLE — test infrastructure only.  BLAKE3 hash mode from the published
 * specification; used where the reference calls blake3::hash and
 * blake3::Hasher (src/merkle_tree.rs:15,50,114-117,176). */
#include "blake3.h"
#include <string.h>

enum { B3_CHUNK_START = 1, B3_CHUNK_END = 2, B3_PARENT = 4, B3_ROOT = 8 };
enum { B3_BLOCK_LEN = 64, B3_CHUNK_LEN = 1024 };

static const uint32_t B3_IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                  0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static const uint8_t B3_PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};

static uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static void g(uint32_t *s, int a, int b, int c, int d, uint32_t mx, uint32_t my) {
    s[a] = s[a] + s[b] + mx;
    s[d] = rotr32(s[d] ^ s[a], 16);
    s[c] = s[c] + s[d];
    s[b] = rotr32(s[b] ^ s[c], 12);
    s[a] = s[a] + s[b] + my;
    s[d] = rotr32(s[d] ^ s[a], 8);
    s[c] = s[c] + s[d];
    s[b] = rotr32(s[b] ^ s[c], 7);
}

static void compress(const uint32_t cv[8], const uint32_t block[16], uint64_t counter,
                     uint32_t block_len, uint32_t flags, uint32_t out[16]) {
    uint32_t s[16], m[16], t[16];
    for (int i = 0; i < 8; i++) s[i] = cv[i];
    for (int i = 0; i < 4; i++) s[8 + i] = B3_IV[i];
    s[12] = (uint32_t)counter;
    s[13] = (uint32_t)(counter >> 32);
    s[14] = block_len;
    s[15] = flags;
    memcpy(m, block, 64);
    for (int r = 0; r < 7; r++) {
        g(s, 0, 4, 8, 12, m[0], m[1]);
        g(s, 1, 5, 9, 13, m[2], m[3]);
        g(s, 2, 6, 10, 14, m[4], m[5]);
        g(s, 3, 7, 11, 15, m[6], m[7]);
        g(s, 0, 5, 10, 15, m[8], m[9]);
        g(s, 1, 6, 11, 12, m[10], m[11]);
        g(s, 2, 7, 8, 13, m[12], m[13]);
        g(s, 3, 4, 9, 14, m[14], m[15]);
        if (r < 6) {
            for (int i = 0; i < 16; i++) t[i] = m[B3_PERM[i]];
            memcpy(m, t, 64);
        }
    }
    for (int i = 0; i < 8; i++) {
        out[i] = s[i] ^ s[i + 8];
        out[i + 8] = s[i + 8] ^ cv[i];
    }
}

static void words_from_le(const uint8_t *b, uint32_t *w, int n) {
    for (int i = 0; i < n; i++)
        w[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
               ((uint32_t)b[4 * i + 3] << 24);
}

typedef struct {
    uint32_t cv[8];
    uint32_t block[16];
    uint64_t counter;
    uint32_t block_len;
    uint32_t flags;
} b3_output;

static void output_cv(const b3_output *o, uint32_t cv[8]) {
    uint32_t out[16];
    compress(o->cv, o->block, o->counter, o->block_len, o->flags, out);
    memcpy(cv, out, 32);
}

static void chunk_init(ob3_chunk_state *c, uint64_t counter) {
    memcpy(c->cv, B3_IV, 32);
    c->chunk_counter = counter;
    memset(c->block, 0, 64);
    c->block_len = 0;
    c->blocks_compressed = 0;
    c->flags = 0;
}

static size_t chunk_len(const ob3_chunk_state *c) {
    return (size_t)B3_BLOCK_LEN * c->blocks_compressed + c->block_len;
}

static uint32_t chunk_start_flag(const ob3_chunk_state *c) {
    return c->blocks_compressed == 0 ? B3_CHUNK_START : 0;
}

static void chunk_update(ob3_chunk_state *c, const uint8_t *in, size_t len) {
    while (len > 0) {
        if (c->block_len == B3_BLOCK_LEN) {
            uint32_t w[16], out[16];
            words_from_le(c->block, w, 16);
            compress(c->cv, w, c->chunk_counter, B3_BLOCK_LEN, c->flags | chunk_start_flag(c), out);
            memcpy(c->cv, out, 32);
            c->blocks_compressed++;
            memset(c->block, 0, 64);
            c->block_len = 0;
        }
        size_t want = B3_BLOCK_LEN - c->block_len;
        size_t take = want < len ? want : len;
        memcpy(c->block + c->block_len, in, take);
        c->block_len += (uint8_t)take;
        in += take;
        len -= take;
    }
}

static b3_output chunk_output(const ob3_chunk_state *c) {
    b3_output o;
    memcpy(o.cv, c->cv, 32);
    words_from_le(c->block, o.block, 16);
    o.counter = c->chunk_counter;
    o.block_len = c->block_len;
    o.flags = c->flags | chunk_start_flag(c) | B3_CHUNK_END;
    return o;
}

static b3_output parent_output(const uint32_t l[8], const uint32_t r[8]) {
    b3_output o;
    memcpy(o.cv, B3_IV, 32);
    memcpy(o.block, l, 32);
    memcpy(o.block + 8, r, 32);
    o.counter = 0;
    o.block_len = B3_BLOCK_LEN;
    o.flags = B3_PARENT;
    return o;
}

void ob3_hasher_init(ob3_hasher *h) {
    chunk_init(&h->chunk, 0);
    h->cv_stack_len = 0;
}

static void add_chunk_cv(ob3_hasher *h, uint32_t cv[8], uint64_t total_chunks) {
    uint32_t cur[8];
    memcpy(cur, cv, 32);
    while ((total_chunks & 1) == 0) {
        h->cv_stack_len--;
        b3_output p = parent_output(h->cv_stack[h->cv_stack_len], cur);
        output_cv(&p, cur);
        total_chunks >>= 1;
    }
    memcpy(h->cv_stack[h->cv_stack_len], cur, 32);
    h->cv_stack_len++;
}

void ob3_hasher_update(ob3_hasher *h, const void *input, size_t len) {
    const uint8_t *in = (const uint8_t *)input;
    while (len > 0) {
        if (chunk_len(&h->chunk) == B3_CHUNK_LEN) {
            b3_output o = chunk_output(&h->chunk);
            uint32_t cv[8];
            output_cv(&o, cv);
            uint64_t total = h->chunk.chunk_counter + 1;
            add_chunk_cv(h, cv, total);
            chunk_init(&h->chunk, total);
        }
        size_t want = B3_CHUNK_LEN - chunk_len(&h->chunk);
        size_t take = want < len ? want : len;
        chunk_update(&h->chunk, in, take);
        in += take;
        len -= take;
    }
}

void ob3_hasher_finalize(const ob3_hasher *h, uint8_t out[32]) {
    b3_output o = chunk_output(&h->chunk);
    int remaining = h->cv_stack_len;
    while (remaining > 0) {
        remaining--;
        uint32_t cv[8];
        output_cv(&o, cv);
        o = parent_output(h->cv_stack[remaining], cv);
    }
    uint32_t words[16];
    compress(o.cv, o.block, 0, o.block_len, o.flags | B3_ROOT, words);
    for (int i = 0; i < 8; i++) {
        out[4 * i] = (uint8_t)words[i];
        out[4 * i + 1] = (uint8_t)(words[i] >> 8);
        out[4 * i + 2] = (uint8_t)(words[i] >> 16);
        out[4 * i + 3] = (uint8_t)(words[i] >> 24);
    }
}

void ob3_hash(const void *input, size_t len, uint8_t out[32]) {
    ob3_hasher h;
    ob3_hasher_init(&h);
    ob3_hasher_update(&h, input, len);
    ob3_hasher_finalize(&h, out);
}
