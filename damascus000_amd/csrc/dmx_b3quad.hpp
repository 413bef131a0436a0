// Quad-cooperative BLAKE3 on the device and the workgroup-level transcript challenge, shared by
// the hash kernels and the persistent tail kernel.
#pragma once

#include "dmx_blake3.hpp"
#include "dmx_kernels.hpp"
#include "dmx_transcript.hpp"

namespace dmx {

struct KeyW {
    uint32_t w[8];
};

// ---------------------------------------------------------------------------------------
// Quad-cooperative compression: four lanes (q = lane & 3) hold the four columns of the 4x4
// BLAKE3 state, lane q keeping (s[q], s[4+q], s[8+q], s[12+q]). The column step is one G per
// lane; the diagonal step rotates b, c, d within the quad by 1, 2, 3 lanes (DPP quad_perm),
// applies G and rotates back. Every lane holds the 16 message words and picks its pair per
// round with two lane-mask selects. A compression costs ~300 VALU instructions per lane
// instead of ~680 on one lane, which is what bounds a 16-block chunk's latency.
// All four lanes of a quad must be active together.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t qrot1(uint32_t v) { return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x39, 0xF, 0xF, false); }
__device__ __forceinline__ uint32_t qrot2(uint32_t v) { return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false); }
__device__ __forceinline__ uint32_t qrot3(uint32_t v) { return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x93, 0xF, 0xF, false); }

struct QuadSel {
    bool b0, b1;  // bits of the lane's column index
    __device__ __forceinline__ uint32_t operator()(uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3) const {
        const uint32_t t0 = b0 ? v1 : v0, t1 = b0 ? v3 : v2;
        return b1 ? t1 : t0;
    }
};

#define DMX_QG(mx, my)                 \
    a = a + b + (mx);                  \
    d = b3::rotr(d ^ a, 16);           \
    c = c + d;                         \
    b = b3::rotr(b ^ c, 12);           \
    a = a + b + (my);                  \
    d = b3::rotr(d ^ a, 8);            \
    c = c + d;                         \
    b = b3::rotr(b ^ c, 7);

#define DMX_QROUND(s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, s10, s11, s12, s13, s14, s15) \
    {                                                                                     \
        const uint32_t mx = sel(m[s0], m[s2], m[s4], m[s6]);                              \
        const uint32_t my = sel(m[s1], m[s3], m[s5], m[s7]);                              \
        DMX_QG(mx, my)                                                                    \
    }                                                                                     \
    b = qrot1(b);                                                                         \
    c = qrot2(c);                                                                         \
    d = qrot3(d);                                                                         \
    {                                                                                     \
        const uint32_t mx = sel(m[s8], m[s10], m[s12], m[s14]);                           \
        const uint32_t my = sel(m[s9], m[s11], m[s13], m[s15]);                           \
        DMX_QG(mx, my)                                                                    \
    }                                                                                     \
    b = qrot3(b);                                                                         \
    c = qrot2(c);                                                                         \
    d = qrot1(d);

// (ca, cb) = lane q's chaining-value words (cv[q], cv[4+q]); returns them updated
__device__ __forceinline__ void compress_quad(uint32_t& ca, uint32_t& cb, const uint32_t m[16], uint64_t counter,
                                              uint32_t block_len, uint32_t flags, const QuadSel& sel) {
    uint32_t a = ca, b = cb;
    uint32_t c = sel(b3::iv(0), b3::iv(1), b3::iv(2), b3::iv(3));
    uint32_t d = sel((uint32_t)counter, (uint32_t)(counter >> 32), block_len, flags);
    DMX_QROUND(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15)
    DMX_QROUND(2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8)
    DMX_QROUND(3, 4, 10, 12, 13, 2, 7, 14, 6, 5, 9, 0, 11, 15, 8, 1)
    DMX_QROUND(10, 7, 12, 9, 14, 3, 13, 15, 4, 0, 11, 2, 5, 8, 1, 6)
    DMX_QROUND(12, 13, 9, 11, 15, 10, 14, 8, 7, 2, 5, 3, 0, 1, 6, 4)
    DMX_QROUND(9, 14, 11, 5, 8, 12, 15, 1, 13, 3, 0, 10, 2, 6, 4, 7)
    DMX_QROUND(11, 15, 5, 0, 1, 9, 8, 6, 14, 10, 2, 12, 3, 4, 7, 13)
    ca = a ^ c;
    cb = b ^ d;
}

__host__ __device__ inline size_t nchunks_of(size_t len) { return len == 0 ? 1 : (len + 1023) / 1024; }

// ---------------------------------------------------------------------------------------
// Transcript messages hashed straight from their sources: a workgroup writes its window of the
// message (16 chunks, 16 KiB) into LDS as aligned 8-byte words (tx_words4) and its quads
// compress the chunks from there, so no packed copy goes through memory first.
// ---------------------------------------------------------------------------------------
constexpr uint32_t TX_WINDOW = 16 * 1024;  // bytes: 16 chunks, one per quad of the first wave

__device__ __forceinline__ void tx_to_job(TranscriptMsg& t, size_t jo) {
    t.epoch_seed = jp(t.epoch_seed, jo);
    t.file_id = jp(t.file_id, jo);
    for (int i = 0; i < 3; ++i) t.mods[i] = jp(t.mods[i], jo);
    t.reject = jp(t.reject, jo);
}

// the job's rejection count (TranscriptMsg::reject), 0 without one
__device__ __forceinline__ uint32_t reject_below(const uint32_t* r) { return r ? *r : 0u; }

// Candidate `counter` of sample_nonzero_challenge (transcript.rs:66-75) is taken when non-zero
// and not below the job's rejection count. Resampling is bounded on the device: past
// RESAMPLE_LIMIT the job reports ERR_RESAMPLE (about 2^-7000 for real digests).
constexpr uint32_t RESAMPLE_LIMIT = 64;
constexpr uint32_t ERR_RESAMPLE = 0x80000000u;
__device__ __forceinline__ bool candidate_taken(const zq& x, uint32_t counter, uint32_t rj) {
    return !zq_is_zero(x) && counter >= rj;
}
// A failed job still runs to its end: its challenge slot gets 1 (and inverse 1), so the kernels
// after it compute on defined values; the error bit fails the job.
__device__ __forceinline__ void exhausted_challenge(zq* chal_out, int want_inv) {
    chal_out[0] = zq_from_u64(1);
    if (want_inv) chal_out[1] = zq_from_u64(1);
}

// message words [w0, w0 + nw) into win[0 .. nw), nw a multiple of 4, by the whole workgroup
__device__ void tx_pack_lds(const TranscriptMsg& t, uint32_t w0, uint32_t nw, uint64_t* win) {
    for (uint32_t g = 4 * threadIdx.x; g < nw; g += 4 * blockDim.x) {
        uint64_t o[4];
        tx_words4(t, w0 + g, o);
        *reinterpret_cast<ulonglong2*>(win + g) = make_ulonglong2(o[0], o[1]);
        *reinterpret_cast<ulonglong2*>(win + g + 2) = make_ulonglong2(o[2], o[3]);
    }
}

// chunk_cv_quad over a message window in LDS that starts at message byte `base` and holds every
// block of the chunk (bytes past `len` read as zero)
__device__ void chunk_cv_quad_lds(const uint64_t* win, size_t base, size_t len, size_t ci, const uint32_t key[8],
                                  uint32_t flags, bool root, const QuadSel& sel, uint32_t& ca, uint32_t& cb) {
    ca = sel(key[0], key[1], key[2], key[3]);
    cb = sel(key[4], key[5], key[6], key[7]);
    const size_t start = ci * b3::CHUNK_LEN;
    const size_t end = (start + b3::CHUNK_LEN < len) ? start + b3::CHUNK_LEN : len;
    const size_t nb = end > start ? (end - start + 63) / 64 : 1;
    const uint4* p = reinterpret_cast<const uint4*>(win) + (start - base) / 16;
    // block b + 1 is read from LDS while block b is compressed
    uint4 nx[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) nx[i] = p[i];
    for (size_t b = 0; b < nb; ++b) {
        uint32_t m[16];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            m[4 * i] = nx[i].x;
            m[4 * i + 1] = nx[i].y;
            m[4 * i + 2] = nx[i].z;
            m[4 * i + 3] = nx[i].w;
        }
        if (b + 1 < nb) {
#pragma unroll
            for (int i = 0; i < 4; ++i) nx[i] = p[4 * (b + 1) + i];
        }
        const size_t off = start + 64 * b;
        const uint32_t blen = (uint32_t)(end - off < 64 ? end - off : 64);
        uint32_t fl = flags | (b == 0 ? b3::CHUNK_START : 0) | (b + 1 == nb ? b3::CHUNK_END : 0);
        if (root && b + 1 == nb) fl |= b3::ROOT;
        compress_quad(ca, cb, m, ci, end > start ? blen : 0, fl, sel);
    }
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// merge_lds by wave 0 alone, for n <= 32 chaining values (at most 16 parents per level, one quad
// each): the levels synchronise within the wave instead of across the workgroup. Every lane of
// wave 0 calls it.
__device__ uint32_t* merge_lds_wave(uint32_t* a, uint32_t* b, uint32_t n, const uint32_t key[8], uint32_t flags,
                                    bool final_root) {
    const uint32_t q = threadIdx.x & 3, quad = threadIdx.x >> 2;
    const QuadSel sel{(q & 1) != 0, (q & 2) != 0};
    while (n > 1) {
        const uint32_t pairs = n >> 1;
        const uint32_t fl = flags | b3::PARENT | ((n == 2 && final_root) ? b3::ROOT : 0);
        if (quad < pairs) {  // uniform within a quad
            uint32_t m[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) m[i] = a[2 * quad * 8 + i];
            uint32_t ca = sel(key[0], key[1], key[2], key[3]), cb = sel(key[4], key[5], key[6], key[7]);
            compress_quad(ca, cb, m, 0, b3::BLOCK_LEN, fl, sel);
            b[quad * 8 + q] = ca;
            b[quad * 8 + 4 + q] = cb;
        }
        if ((n & 1) && threadIdx.x < 8) b[pairs * 8 + threadIdx.x] = a[(n - 1) * 8 + threadIdx.x];
        wave_sync_lds();
        uint32_t* t = a;
        a = b;
        b = t;
        n = pairs + (n & 1);
    }
    return a;
}

// Pairwise bottom-up merge of the n CVs in `a` (LDS), one quad per parent node, ping-ponging
// between `a` and `b`; returns the buffer holding the single result. With `final_root` the last
// compression carries ROOT (n >= 2). Levels with more than 16 parents use the whole workgroup
// and its barriers; the last levels (at most 32 values) run in wave 0 alone with wave-level
// synchronisation (merge_lds_wave), and a final barrier publishes the result. Loop bounds are
// uniform, so every thread reaches every barrier; the caller has synchronised after filling `a`.
__device__ uint32_t* merge_lds(uint32_t* a, uint32_t* b, uint32_t n, const uint32_t key[8], uint32_t flags,
                               bool final_root) {
    const uint32_t q = threadIdx.x & 3, quad = threadIdx.x >> 2, nquads = blockDim.x >> 2;
    const QuadSel sel{(q & 1) != 0, (q & 2) != 0};
    while (n > 32) {
        const uint32_t pairs = n >> 1;
        const uint32_t fl = flags | b3::PARENT;  // n > 2: never the root here
        for (uint32_t base = 0; base < pairs; base += nquads) {
            const uint32_t pr = base + quad;
            if (pr < pairs) {
                uint32_t m[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) m[i] = a[2 * pr * 8 + i];
                uint32_t ca = sel(key[0], key[1], key[2], key[3]), cb = sel(key[4], key[5], key[6], key[7]);
                compress_quad(ca, cb, m, 0, b3::BLOCK_LEN, fl, sel);
                b[pr * 8 + q] = ca;
                b[pr * 8 + 4 + q] = cb;
            }
        }
        if ((n & 1) && threadIdx.x < 8) b[pairs * 8 + threadIdx.x] = a[(n - 1) * 8 + threadIdx.x];
        __syncthreads();
        uint32_t* t = a;
        a = b;
        b = t;
        n = pairs + (n & 1);
    }
    if (n <= 1) return a;
    // the buffer the wave's levels end in: one swap per level
    uint32_t levels = 0;
    for (uint32_t v = n; v > 1; v = (v >> 1) + (v & 1)) ++levels;
    uint32_t* const res = (levels & 1) ? b : a;
    if (threadIdx.x < 64) (void)merge_lds_wave(a, b, n, key, flags, final_root);
    __syncthreads();
    return res;
}

// A whole challenge (transcript.rs:42-75) in one workgroup, for messages of at most
// CHALLENGE_MAX_CHUNKS chunks and workgroups of at least 64 threads: the workgroup writes the
// message into LDS, the 16 quads of wave 0 hash the chunks and merge them to the root (wave-level
// synchronisation only), and wave 0 reduces the digest mod q (and inverts it with want_inv). A zero challenge bumps
// the counter in LDS and hashes the message again. `t` is already offset to the job. Every thread
// returns the challenge; with chal_out, thread 0 also stores x (and x^-1). A sampler that runs
// out sets ERR_RESAMPLE in *err (when given) and yields 1.
constexpr size_t CHALLENGE_MAX_CHUNKS = 16;
struct __align__(16) WgChal {
    uint64_t win[TX_WINDOW / 8];
    uint32_t sa[16 * 8], sb[8 * 8 + 8];
    uint32_t dig[8];
    zq x;
    int again;
};

__device__ __forceinline__ zq wg_tx_challenge(const TranscriptMsg& t, size_t len, const uint32_t key[8], uint32_t flags, WgChal& sm,
                              zq* chal_out, uint32_t* err, int want_inv) {
    const uint32_t rj = reject_below(t.reject);
    tx_pack_lds(t, 0, (uint32_t)(((len + 63) & ~(size_t)63) / 8), sm.win);
    __syncthreads();
    const uint32_t q = threadIdx.x & 3, quad = threadIdx.x >> 2;
    const QuadSel sel{(q & 1) != 0, (q & 2) != 0};
    const size_t n = nchunks_of(len);
    for (uint32_t counter = 0;; ++counter) {  // every thread counts the same passes
        if (threadIdx.x < 64) {  // wave 0 hashes, merges and reduces
            if (n == 1) {
                if (threadIdx.x < 4) {
                    uint32_t ca, cb;
                    chunk_cv_quad_lds(sm.win, 0, len, 0, key, flags, true, sel, ca, cb);
                    sm.dig[q] = ca;
                    sm.dig[4 + q] = cb;
                }
            } else {
                if (quad < n) {
                    uint32_t ca, cb;
                    chunk_cv_quad_lds(sm.win, 0, len, quad, key, flags, false, sel, ca, cb);
                    sm.sa[quad * 8 + q] = ca;
                    sm.sa[quad * 8 + 4 + q] = cb;
                }
                wave_sync_lds();
                const uint32_t* r = merge_lds_wave(sm.sa, sm.sb, (uint32_t)n, key, flags, true);
                if (threadIdx.x < 8) sm.dig[threadIdx.x] = r[threadIdx.x];
            }
            wave_sync_lds();
            uint32_t h[8];
            // identity quad permutation: same value, but a per-lane register to the compiler
            for (int j = 0; j < 8; ++j) h[j] = (uint32_t)__builtin_amdgcn_mov_dpp((int)sm.dig[j], 0xE4, 0xF, 0xF, false);
            const zq x = zq_from_digest(h);
            const bool taken = candidate_taken(x, counter, rj);
            if (taken) {
                const zq xi = want_inv ? zq_inv_divsteps(x) : x;
                if (threadIdx.x == 0) {
                    sm.x = x;
                    if (chal_out) {
                        chal_out[0] = x;
                        if (want_inv) chal_out[1] = xi;
                    }
                }
            }
            if (threadIdx.x == 0) {
                sm.again = 0;
                if (!taken) {
                    if (counter + 1 >= RESAMPLE_LIMIT) {
                        if (err) atomicOr(err, ERR_RESAMPLE);
                        if (chal_out) exhausted_challenge(chal_out, want_inv);
                        sm.x = zq_from_u64(1);
                    } else {
                        const uint32_t c1 = counter + 1;
                        uint8_t* wb = reinterpret_cast<uint8_t*>(sm.win);
                        for (uint32_t b = 0; b < 4; ++b) wb[len - 4 + b] = (uint8_t)(c1 >> (8 * b));
                        sm.again = 1;
                    }
                }
            }
        }
        __syncthreads();
        if (!sm.again) break;  // uniform: every thread reads the same shared flag
    }
    return sm.x;
}

}  // namespace dmx
