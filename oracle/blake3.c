/* Oracle: BLAKE3 restated from the public BLAKE3 specification (test infrastructure only).
 *
 * The reference links crate blake3 1.8.3 (Cargo.lock), called at src/utils/io.rs:33 (hash),
 * src/commitment/sis.rs:209,243 (keyed_hash) and src/protocol/transcript.rs:51-61 (Hasher).
 * The crate is not vendored; this file follows the spec: 7-round compression over a 16-word
 * state, 1 KiB chunks of 64-byte blocks, a left-balanced binary tree of parent nodes, and the
 * CHUNK_START/CHUNK_END/PARENT/ROOT/KEYED_HASH domain flags. */
#include "internal.h"
#include "oracle.h"

static const uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                               0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static const uint8_t PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
enum { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8, KEYED_HASH = 16 };

static inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static inline void g(uint32_t *s, int a, int b, int c, int d, uint32_t mx, uint32_t my) {
    s[a] = s[a] + s[b] + mx;
    s[d] = rotr(s[d] ^ s[a], 16);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 12);
    s[a] = s[a] + s[b] + my;
    s[d] = rotr(s[d] ^ s[a], 8);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 7);
}

static void compress(const uint32_t cv[8], const uint32_t block[16], uint64_t counter,
                     uint32_t block_len, uint32_t flags, uint32_t out[16]) {
    uint32_t s[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                      IV[0], IV[1], IV[2], IV[3], (uint32_t)counter, (uint32_t)(counter >> 32),
                      block_len, flags};
    uint32_t m[16], t[16];
    memcpy(m, block, sizeof(m));
    for (int r = 0; r < 7; ++r) {
        g(s, 0, 4, 8, 12, m[0], m[1]);
        g(s, 1, 5, 9, 13, m[2], m[3]);
        g(s, 2, 6, 10, 14, m[4], m[5]);
        g(s, 3, 7, 11, 15, m[6], m[7]);
        g(s, 0, 5, 10, 15, m[8], m[9]);
        g(s, 1, 6, 11, 12, m[10], m[11]);
        g(s, 2, 7, 8, 13, m[12], m[13]);
        g(s, 3, 4, 9, 14, m[14], m[15]);
        for (int i = 0; i < 16; ++i) t[i] = m[PERM[i]];
        memcpy(m, t, sizeof(m));
    }
    for (int i = 0; i < 8; ++i) {
        out[i] = s[i] ^ s[i + 8];
        out[i + 8] = s[i + 8] ^ cv[i];
    }
}

static void words_from_le(const uint8_t *bytes, uint32_t *words, int n) {
    for (int i = 0; i < n; ++i)
        words[i] = (uint32_t)bytes[4 * i] | ((uint32_t)bytes[4 * i + 1] << 8) |
                   ((uint32_t)bytes[4 * i + 2] << 16) | ((uint32_t)bytes[4 * i + 3] << 24);
}

static void parent_cv(const uint32_t l[8], const uint32_t r[8], const uint32_t key[8],
                      uint32_t flags, uint32_t out[8]) {
    uint32_t block[16], o[16];
    memcpy(block, l, 32);
    memcpy(block + 8, r, 32);
    compress(key, block, 0, 64, flags | PARENT, o);
    memcpy(out, o, 32);
}

void orc_b3_init(orc_b3 *h) {
    memset(h, 0, sizeof(*h));
    memcpy(h->key, IV, 32);
    memcpy(h->cv, IV, 32);
}

void orc_b3_init_keyed(orc_b3 *h, const uint8_t key[32]) {
    memset(h, 0, sizeof(*h));
    words_from_le(key, h->key, 8);
    memcpy(h->cv, h->key, 32);
    h->flags = KEYED_HASH;
}

static uint32_t chunk_len(const orc_b3 *h) { return 64 * h->blocks_compressed + h->buf_len; }

void orc_b3_update(orc_b3 *h, const void *in_v, size_t len) {
    const uint8_t *in = (const uint8_t *)in_v;
    while (len > 0) {
        if (chunk_len(h) == 1024) {
            /* finish the chunk: its buffered last block carries CHUNK_END */
            uint32_t block[16], o[16], cv[8];
            words_from_le(h->buf, block, 16);
            uint32_t fl = h->flags | CHUNK_END | (h->blocks_compressed == 0 ? CHUNK_START : 0);
            compress(h->cv, block, h->chunk_counter, h->buf_len, fl, o);
            memcpy(cv, o, 32);
            uint64_t total = h->chunk_counter + 1;
            while ((total & 1) == 0) {
                parent_cv(h->cv_stack[--h->cv_stack_len], cv, h->key, h->flags, cv);
                total >>= 1;
            }
            memcpy(h->cv_stack[h->cv_stack_len++], cv, 32);
            h->chunk_counter += 1;
            memcpy(h->cv, h->key, 32);
            h->blocks_compressed = 0;
            h->buf_len = 0;
            memset(h->buf, 0, 64);
        }
        if (h->buf_len == 64) {
            uint32_t block[16], o[16];
            words_from_le(h->buf, block, 16);
            uint32_t fl = h->flags | (h->blocks_compressed == 0 ? CHUNK_START : 0);
            compress(h->cv, block, h->chunk_counter, 64, fl, o);
            memcpy(h->cv, o, 32);
            h->blocks_compressed += 1;
            h->buf_len = 0;
            memset(h->buf, 0, 64);
        }
        size_t want = 64 - h->buf_len;
        size_t room = 1024 - chunk_len(h);
        if (want > room) want = room;
        if (want > len) want = len;
        memcpy(h->buf + h->buf_len, in, want);
        h->buf_len += (uint32_t)want;
        in += want;
        len -= want;
    }
}

void orc_b3_finalize(const orc_b3 *h, uint8_t *out, size_t out_len) {
    /* the output node: the current chunk's last block, then parents down the stack */
    uint32_t in_cv[8], block[16], o[16];
    uint64_t counter = h->chunk_counter;
    uint32_t block_len = h->buf_len;
    uint32_t flags = h->flags | CHUNK_END | (h->blocks_compressed == 0 ? CHUNK_START : 0);
    memcpy(in_cv, h->cv, 32);
    words_from_le(h->buf, block, 16);
    for (uint32_t i = h->cv_stack_len; i-- > 0;) {
        uint32_t cv[8];
        compress(in_cv, block, counter, block_len, flags, o);
        memcpy(cv, o, 32);
        memcpy(block, h->cv_stack[i], 32);
        memcpy(block + 8, cv, 32);
        memcpy(in_cv, h->key, 32);
        counter = 0;
        block_len = 64;
        flags = h->flags | PARENT;
    }
    /* root output: XOF blocks at increasing output counters */
    size_t pos = 0;
    uint64_t ctr = 0;
    while (pos < out_len) {
        compress(in_cv, block, ctr++, block_len, flags | ROOT, o);
        for (int i = 0; i < 16 && pos < out_len; ++i)
            for (int b = 0; b < 4 && pos < out_len; ++b) out[pos++] = (uint8_t)(o[i] >> (8 * b));
    }
}

void orc_blake3_hash(const uint8_t *in, size_t len, uint8_t out[32]) {
    orc_b3 h;
    orc_b3_init(&h);
    orc_b3_update(&h, in, len);
    orc_b3_finalize(&h, out, 32);
}

void orc_blake3_keyed_hash(const uint8_t key[32], const uint8_t *in, size_t len, uint8_t out[32]) {
    orc_b3 h;
    orc_b3_init_keyed(&h, key);
    orc_b3_update(&h, in, len);
    orc_b3_finalize(&h, out, 32);
}

void orc_blake3_hash_xof(const uint8_t *key_or_null, const uint8_t *in, size_t len, uint8_t *out,
                         size_t out_len) {
    orc_b3 h;
    if (key_or_null)
        orc_b3_init_keyed(&h, key_or_null);
    else
        orc_b3_init(&h);
    orc_b3_update(&h, in, len);
    orc_b3_finalize(&h, out, out_len);
}
