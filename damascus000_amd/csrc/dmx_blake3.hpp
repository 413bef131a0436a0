// BLAKE3 (the algorithm of crate blake3 1.8.3, used at src/utils/io.rs:33, src/commitment/sis.rs:209,243
// and src/protocol/transcript.rs:51-61), written from the public BLAKE3 specification.
// The compression function is shared by host and device code; the host incremental hasher
// serves small host-side hashes (the g0 sub-seed), the device tree hash lives in dmx_hash.hip.
#pragma once

#include <stdint.h>
#include <string.h>

#include "dmx_arith.hpp"

namespace dmx {
namespace b3 {

enum : uint32_t { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8, KEYED_HASH = 16 };
constexpr uint32_t CHUNK_LEN = 1024;
constexpr uint32_t BLOCK_LEN = 64;

DMX_HD uint32_t iv(int i) {
    const uint32_t v[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                           0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
    return v[i];
}

DMX_HD uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

#define DMX_B3_G(a, b, c, d, mx, my)  \
    s[a] = s[a] + s[b] + (mx);        \
    s[d] = rotr(s[d] ^ s[a], 16);     \
    s[c] = s[c] + s[d];               \
    s[b] = rotr(s[b] ^ s[c], 12);     \
    s[a] = s[a] + s[b] + (my);        \
    s[d] = rotr(s[d] ^ s[a], 8);      \
    s[c] = s[c] + s[d];               \
    s[b] = rotr(s[b] ^ s[c], 7);

// One round with the message schedule applied by index: round r reads m[SCHED[r][i]].
// The permutation [2,6,3,10,7,0,4,13,1,11,12,5,9,14,15,8] composed r times gives these rows.
#define DMX_B3_ROUND(m0, m1, m2, m3, m4, m5, m6, m7, m8, m9, m10, m11, m12, m13, m14, m15) \
    DMX_B3_G(0, 4, 8, 12, m[m0], m[m1])                                                    \
    DMX_B3_G(1, 5, 9, 13, m[m2], m[m3])                                                    \
    DMX_B3_G(2, 6, 10, 14, m[m4], m[m5])                                                   \
    DMX_B3_G(3, 7, 11, 15, m[m6], m[m7])                                                   \
    DMX_B3_G(0, 5, 10, 15, m[m8], m[m9])                                                   \
    DMX_B3_G(1, 6, 11, 12, m[m10], m[m11])                                                 \
    DMX_B3_G(2, 7, 8, 13, m[m12], m[m13])                                                  \
    DMX_B3_G(3, 4, 9, 14, m[m14], m[m15])

// full 16-word output of the compression function
DMX_HD void compress(const uint32_t cv[8], const uint32_t m[16], uint64_t counter,
                     uint32_t block_len, uint32_t flags, uint32_t out[16]) {
    uint32_t s[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                      iv(0), iv(1), iv(2), iv(3), (uint32_t)counter, (uint32_t)(counter >> 32),
                      block_len, flags};
    DMX_B3_ROUND(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15)
    DMX_B3_ROUND(2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8)
    DMX_B3_ROUND(3, 4, 10, 12, 13, 2, 7, 14, 6, 5, 9, 0, 11, 15, 8, 1)
    DMX_B3_ROUND(10, 7, 12, 9, 14, 3, 13, 15, 4, 0, 11, 2, 5, 8, 1, 6)
    DMX_B3_ROUND(12, 13, 9, 11, 15, 10, 14, 8, 7, 2, 5, 3, 0, 1, 6, 4)
    DMX_B3_ROUND(9, 14, 11, 5, 8, 12, 15, 1, 13, 3, 0, 10, 2, 6, 4, 7)
    DMX_B3_ROUND(11, 15, 5, 0, 1, 9, 8, 6, 14, 10, 2, 12, 3, 4, 7, 13)
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        out[i] = s[i] ^ s[i + 8];
        out[i + 8] = s[i + 8] ^ cv[i];
    }
}

// chaining value only (first 8 output words)
DMX_HD void compress_cv(const uint32_t cv[8], const uint32_t m[16], uint64_t counter,
                        uint32_t block_len, uint32_t flags, uint32_t out[8]) {
    uint32_t o[16];
    compress(cv, m, counter, block_len, flags, o);
#pragma unroll
    for (int i = 0; i < 8; ++i) out[i] = o[i];
}

DMX_HD void parent_cv(const uint32_t l[8], const uint32_t r[8], const uint32_t key[8],
                      uint32_t flags, uint32_t out[8]) {
    uint32_t m[16];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        m[i] = l[i];
        m[i + 8] = r[i];
    }
    compress_cv(key, m, 0, BLOCK_LEN, flags | PARENT, out);
}

inline void words_le(const uint8_t* b, uint32_t* w, int n) {
    for (int i = 0; i < n; ++i)
        w[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
               ((uint32_t)b[4 * i + 3] << 24);
}

// Host one-shot hash for short inputs (<= one chunk): keyed when key != nullptr.
inline void host_hash_short(const uint8_t* key, const uint8_t* in, size_t len, uint8_t out[32]) {
    uint32_t kw[8], cv[8], m[16];
    uint32_t flags = 0;
    if (key) {
        words_le(key, kw, 8);
        flags = KEYED_HASH;
    } else {
        for (int i = 0; i < 8; ++i) kw[i] = iv(i);
    }
    memcpy(cv, kw, 32);
    size_t nblocks = len == 0 ? 1 : (len + 63) / 64;
    for (size_t b = 0; b < nblocks; ++b) {
        uint8_t blk[64] = {0};
        size_t off = b * 64, n = len - off < 64 ? len - off : 64;
        if (len) memcpy(blk, in + off, n);
        words_le(blk, m, 16);
        uint32_t fl = flags | (b == 0 ? CHUNK_START : 0) | (b + 1 == nblocks ? CHUNK_END | ROOT : 0);
        uint32_t o[16];
        compress(cv, m, 0, (uint32_t)(len ? n : 0), fl, o);
        memcpy(cv, o, 32);
    }
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 4; ++j) out[4 * i + j] = (uint8_t)(cv[i] >> (8 * j));
}

}  // namespace b3
}  // namespace dmx
