// Transcript message bytes as aligned 8-byte words, computed from their sources.
//
// The challenge message (src/protocol/transcript.rs:50-61, 77-90) is
//   epoch_seed[32] | u64 len(label) | label | u64 32 | file_id[32] | u64 epoch (0) | u64 round |
//   3 x ( u64 K | K x ( u64 len | len x 16-byte LE coefficients ) ) | u32 counter
// The header is H = 96 + len(label) bytes. Every field after it has a length that is a multiple
// of 8, so from byte H on the message is a sequence of 8-byte "body words": a module's u64 K, a
// coordinate's u64 len, or one half of a coefficient. A message word at an aligned offset 8w >= H
// is two neighbouring body words shifted by (8w - H) mod 8 bytes, so a hash kernel can write the
// message into LDS with aligned stores instead of reading a packed copy from memory. Past the
// last module the body reads as zero: the counter (0 until a zero challenge is resampled) and
// the block padding.
//
// Pointers in the TranscriptMsg are used as given (already offset to the job); epoch_seed and
// file_id must be 8-byte aligned and the label 1 to 4 bytes long. Shared by host and device code
// so the word layout can be checked on the host against the reference encoding.
#pragma once

#include "dmx_kernels.hpp"

namespace dmx {

DMX_HD uint32_t tx_header_len(const TranscriptMsg& t) { return 96 + t.label_len; }

// whole message, counter included
DMX_HD size_t tx_message_len(uint32_t label_len, const uint32_t lens[3]) {
    size_t len = 96 + label_len + 4;
    for (int i = 0; i < 3; ++i) len += 8 + (size_t)K * (8 + 16 * (size_t)lens[i]);
    return len;
}

// Body word v (message bytes [H + 8v, H + 8v + 8)) as a source: the address of a coefficient
// half when `load`, else the immediate value (a length, K, or 0 past the last module). The
// address is always a valid one (module 0's first coefficient for an immediate), so callers can
// issue every load of a group unconditionally and select afterwards.
struct TxRef {
    const uint64_t* p;
    uint64_t imm;
    bool load;
};

DMX_HD TxRef tx_body_ref(const TranscriptMsg& t, uint32_t v) {
    TxRef ref{reinterpret_cast<const uint64_t*>(t.mods[0]), 0, false};
    for (int i = 0; i < 3; ++i) {
        const uint32_t n = t.lens[i], per = 1 + 2 * n, words = 1 + K * per;
        if (v < words) {
            if (v == 0) {
                ref.imm = K;
                return ref;
            }
            uint32_t r = v - 1, k = 0;
#pragma unroll
            for (uint32_t j = 1; j < (uint32_t)K; ++j) k += r >= j * per ? 1u : 0u;
            r -= k * per;
            if (r == 0) {
                ref.imm = n;
                return ref;
            }
            r -= 1;
            ref.p = reinterpret_cast<const uint64_t*>(t.mods[i] + (size_t)k * n + (r >> 1)) + (r & 1);
            ref.load = true;
            return ref;
        }
        v -= words;
    }
    return ref;
}

DMX_HD uint64_t tx_body_word(const TranscriptMsg& t, uint32_t v) {
    const TxRef r = tx_body_ref(t, v);
    return r.load ? *r.p : r.imm;
}

// message word w with 8w >= H
DMX_HD uint64_t tx_body_msg_word(const TranscriptMsg& t, uint32_t w) {
    const uint32_t b = 8 * w - tx_header_len(t), v = b >> 3, sh = 8 * (b & 7);
    const uint64_t lo = tx_body_word(t, v);
    return sh ? (lo >> sh) | (tx_body_word(t, v + 1) << (64 - sh)) : lo;
}

// Message word j < 13 (bytes [8j, 8j + 8)) of the header, computed alone: epoch_seed, u64
// len(label), then from byte 40 the label followed by u64 32, file_id, u64 epoch 0 and u64 round,
// each shifted by len(label) bytes, then from byte H the first body word. Needs 1 <= len(label)
// <= 4 and 8-byte aligned epoch_seed and file_id.
DMX_HD uint64_t tx_head_word(const TranscriptMsg& t, uint32_t j) {
    const uint64_t* e = reinterpret_cast<const uint64_t*>(t.epoch_seed);
    const uint64_t* f = reinterpret_cast<const uint64_t*>(t.file_id);
    if (j < 4) return e[j];
    if (j == 4) return t.label_len;
    const uint32_t sl = 8 * t.label_len;
    // S = {32, f[0..3], 0, round, first body word}: the fields after the label, as words
    auto S = [&](uint32_t i) -> uint64_t {
        return i == 0 ? 32u : i <= 4 ? f[i - 1] : i == 5 ? 0u : i == 6 ? t.round : tx_body_word(t, 0);
    };
    if (j == 5) {  // the label's bytes with constant indices, so the message stays in registers
        const uint64_t lw = (uint64_t)t.label[0] | ((uint64_t)t.label[1] << 8) | ((uint64_t)t.label[2] << 16) |
                            ((uint64_t)t.label[3] << 24);
        const uint64_t lab = lw & ((1ull << sl) - 1);
        return lab | (S(0) << sl);
    }
    return (S(j - 6) >> (64 - sl)) | (S(j - 5) << sl);
}

// message words w0 .. w0+3; past the header the five body words are loaded together
DMX_HD void tx_words4(const TranscriptMsg& t, uint32_t w0, uint64_t out[4]) {
    const uint32_t H = tx_header_len(t);
    if (8 * w0 < H) {  // w0 < 13
        for (uint32_t i = 0; i < 4; ++i) out[i] = w0 + i < 13 ? tx_head_word(t, w0 + i) : tx_body_msg_word(t, w0 + i);
        return;
    }
    const uint32_t b = 8 * w0 - H, v = b >> 3, sh = 8 * (b & 7);
    TxRef r[5];
    uint64_t bw[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) r[i] = tx_body_ref(t, v + i);
#pragma unroll
    for (int i = 0; i < 5; ++i) bw[i] = *r[i].p;
#pragma unroll
    for (int i = 0; i < 5; ++i) bw[i] = r[i].load ? bw[i] : r[i].imm;
#pragma unroll
    for (int i = 0; i < 4; ++i) out[i] = sh ? (bw[i] >> sh) | (bw[i + 1] << (64 - sh)) : bw[i];
}

}  // namespace dmx
