// BLAKE3 tree hashing on the device: file_id (src/utils/io.rs:33) and the Fiat-Shamir
// challenges (src/protocol/transcript.rs:42-90), so a whole proof runs without host round trips.
//
// Chunks (1 KiB) are compressed one per thread; aligned groups of 1024 chunk chaining values
// are merged by one workgroup each (an aligned power-of-two run of chunks is a complete subtree
// of BLAKE3's left-balanced tree); a last workgroup merges the group values pairwise bottom-up,
// carrying an odd node upward, which reproduces the left-balanced tree, and sets ROOT on the
// final compression.
#include "dmx_blake3.hpp"
#include "dmx_kernels.hpp"

namespace dmx {

struct KeyW {
    uint32_t w[8];
};

// ---------------------------------------------------------------------------------------
// transcript message layout (transcript.rs:50-61, 77-90):
//   epoch_seed[32] | u64 len(label) | label | u64 32 | file_id[32] | u64 0 | u64 round |
//   3 x ( u64 8 | 8 x ( u64 len | len x 16-byte LE coefficients ) ) | u32 counter
// ---------------------------------------------------------------------------------------
static inline size_t module_len(uint32_t n) { return 8 + (size_t)K * (8 + 16 * (size_t)n); }

size_t transcript_len(uint32_t label_len, const uint32_t lens[3]) {
    return 96 + label_len + module_len(lens[0]) + module_len(lens[1]) + module_len(lens[2]) + 4;
}

__device__ __forceinline__ uint8_t le_byte(uint64_t v, uint32_t i) { return (uint8_t)(v >> (8 * i)); }

__device__ uint8_t module_byte(const zq* m, uint32_t n, size_t o) {
    if (o < 8) return le_byte(K, (uint32_t)o);
    o -= 8;
    const size_t blk = 8 + 16 * (size_t)n;
    const size_t k = o / blk, r = o - k * blk;
    if (r < 8) return le_byte(n, (uint32_t)r);
    const size_t e = r - 8, c = e >> 4;
    const uint32_t* w = m[k * n + c].w;
    return (uint8_t)(w[(e >> 2) & 3] >> (8 * (e & 3)));
}

__device__ uint8_t transcript_byte(const TranscriptMsg& t, size_t o, uint32_t counter) {
    if (o < 32) return t.epoch_seed[o];
    o -= 32;
    if (o < 8) return le_byte(t.label_len, (uint32_t)o);
    o -= 8;
    if (o < t.label_len) return t.label[o];
    o -= t.label_len;
    if (o < 8) return le_byte(32, (uint32_t)o);
    o -= 8;
    if (o < 32) return t.file_id[o];
    o -= 32;
    if (o < 8) return 0;  // epoch index
    o -= 8;
    if (o < 8) return le_byte(t.round, (uint32_t)o);
    o -= 8;
    for (int i = 0; i < 3; ++i) {
        const size_t ml = 8 + (size_t)K * (8 + 16 * (size_t)t.lens[i]);
        if (o < ml) return module_byte(t.mods[i], t.lens[i], o);
        o -= ml;
    }
    return le_byte(counter, (uint32_t)o);
}

__global__ void k_transcript_pack(TranscriptMsg t, uint8_t* __restrict__ msg, size_t len) {
    const size_t w = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t o = w * 4;
    if (o >= len) return;
    uint32_t v = 0;
    for (uint32_t b = 0; b < 4 && o + b < len; ++b) v |= (uint32_t)transcript_byte(t, o + b, 0) << (8 * b);
    if (o + 4 <= len) {
        *reinterpret_cast<uint32_t*>(msg + o) = v;
    } else {
        for (uint32_t b = 0; o + b < len; ++b) msg[o + b] = (uint8_t)(v >> (8 * b));
    }
}

void launch_transcript_pack(const TranscriptMsg& t, uint8_t* msg, size_t len, hipStream_t s) {
    const size_t words = (len + 3) / 4;
    hipLaunchKernelGGL(k_transcript_pack, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s, t, msg,
                       len);
}

// ---------------------------------------------------------------------------------------
// chunk chaining values
// ---------------------------------------------------------------------------------------
__device__ void load_block(const uint8_t* msg, size_t len, size_t off, uint32_t m[16], uint32_t& blen) {
    const size_t n = len - off < 64 ? len - off : 64;
    blen = (uint32_t)n;
    if (n == 64 && ((reinterpret_cast<uintptr_t>(msg) & 15) == 0)) {
        const uint4* p = reinterpret_cast<const uint4*>(msg + off);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint4 v = p[i];
            m[4 * i] = v.x;
            m[4 * i + 1] = v.y;
            m[4 * i + 2] = v.z;
            m[4 * i + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) m[i] = 0;
        for (size_t b = 0; b < n; ++b) m[b >> 2] |= (uint32_t)msg[off + b] << (8 * (b & 3));
    }
}

// CV of chunk `ci`; with root != 0 the last block carries ROOT and the full digest is returned
__device__ void chunk_cv(const uint8_t* msg, size_t len, size_t ci, const uint32_t key[8],
                         uint32_t flags, bool root, uint32_t out[8]) {
    uint32_t cv[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) cv[i] = key[i];
    const size_t start = ci * b3::CHUNK_LEN;
    const size_t end = (start + b3::CHUNK_LEN < len) ? start + b3::CHUNK_LEN : len;
    const size_t nb = end > start ? (end - start + 63) / 64 : 1;
    for (size_t b = 0; b < nb; ++b) {
        uint32_t m[16], blen;
        if (end > start)
            load_block(msg, len, start + 64 * b, m, blen);
        else {
#pragma unroll
            for (int i = 0; i < 16; ++i) m[i] = 0;
            blen = 0;
        }
        uint32_t fl = flags | (b == 0 ? b3::CHUNK_START : 0) | (b + 1 == nb ? b3::CHUNK_END : 0);
        if (root && b + 1 == nb) fl |= b3::ROOT;
        b3::compress_cv(cv, m, ci, blen, fl, cv);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) out[i] = cv[i];
}

__global__ void __launch_bounds__(64) k_b3_chunks(const uint8_t* __restrict__ msg, size_t len, KeyW key,
                                                  uint32_t flags, uint32_t* __restrict__ cvs, size_t n) {
    const size_t ci = (size_t)blockIdx.x * 64 + threadIdx.x;
    if (ci >= n) return;
    uint32_t cv[8];
    chunk_cv(msg, len, ci, key.w, flags, false, cv);
    uint4* o = reinterpret_cast<uint4*>(cvs + ci * 8);
    o[0] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
    o[1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
}

static inline size_t nchunks_of(size_t len) { return len == 0 ? 1 : (len + 1023) / 1024; }

void launch_b3_chunks(const uint8_t* msg, size_t len, const uint32_t key[8], uint32_t flags,
                      uint32_t* cvs, hipStream_t s) {
    KeyW k;
    for (int i = 0; i < 8; ++i) k.w[i] = key[i];
    const size_t n = nchunks_of(len);
    hipLaunchKernelGGL(k_b3_chunks, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, msg, len, k, flags,
                       cvs, n);
}

// pairwise bottom-up merge of `n` <= 2 * blockDim.x CVs held in LDS; returns with the single
// CV in sh[0..8); with `final_root` the last compression carries ROOT (n >= 2).
// Every thread reaches every barrier: a thread owns at most one pair per level.
__device__ void merge_lds(uint32_t* sh, uint32_t n, const uint32_t key[8], uint32_t flags, bool final_root) {
    while (n > 1) {
        const uint32_t pairs = n >> 1;
        const bool last = (n == 2) && final_root;
        const uint32_t i = threadIdx.x;
        uint32_t o[8];
        if (i < pairs) {
            uint32_t l[8], r[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                l[j] = sh[(2 * i) * 8 + j];
                r[j] = sh[(2 * i + 1) * 8 + j];
            }
            b3::parent_cv(l, r, key, flags | (last ? b3::ROOT : 0), o);
        }
        __syncthreads();  // all reads of this level happen before any write
        if (i < pairs) {
#pragma unroll
            for (int j = 0; j < 8; ++j) sh[i * 8 + j] = o[j];
        }
        __syncthreads();
        if ((n & 1) && threadIdx.x == 0) {
#pragma unroll
            for (int j = 0; j < 8; ++j) sh[pairs * 8 + j] = sh[(n - 1) * 8 + j];
        }
        __syncthreads();
        n = pairs + (n & 1);
    }
}

__global__ void __launch_bounds__(1024) k_b3_group_merge(const uint32_t* __restrict__ cvs, size_t n, KeyW key,
                                                         uint32_t flags, uint32_t* __restrict__ out) {
    __shared__ uint32_t sh[1024 * 8];
    const size_t base = (size_t)blockIdx.x * 1024;
    const uint32_t cnt = (uint32_t)((n - base) < 1024 ? (n - base) : 1024);
    for (uint32_t i = threadIdx.x; i < cnt * 8; i += blockDim.x) sh[i] = cvs[base * 8 + i];
    __syncthreads();
    merge_lds(sh, cnt, key.w, flags, false);
    if (threadIdx.x < 8) out[blockIdx.x * 8 + threadIdx.x] = sh[threadIdx.x];
}

void launch_b3_group_merge(const uint32_t* cvs, size_t n, const uint32_t key[8], uint32_t flags,
                           uint32_t* out, hipStream_t s) {
    KeyW k;
    for (int i = 0; i < 8; ++i) k.w[i] = key[i];
    hipLaunchKernelGGL(k_b3_group_merge, dim3((unsigned)((n + 1023) / 1024)), dim3(1024), 0, s, cvs, n, k,
                       flags, out);
}

// modular inverse in Z_q by the binary extended Euclid algorithm (same value as x^(q-2))
__device__ zq zq_inv_binary(zq a) {
    const zq qv = zq{{Q0, Q1, Q2, Q3}};
    zq u = a, v = qv, x1 = zq_from_u64(1), x2 = zq_zero();
    auto is_one = [](const zq& z) { return z.w[0] == 1 && (z.w[1] | z.w[2] | z.w[3]) == 0; };
    auto shr1 = [](zq& z) {
        z.w[0] = (z.w[0] >> 1) | (z.w[1] << 31);
        z.w[1] = (z.w[1] >> 1) | (z.w[2] << 31);
        z.w[2] = (z.w[2] >> 1) | (z.w[3] << 31);
        z.w[3] >>= 1;
    };
    auto half = [&](zq& x) {  // x / 2 mod q
        if (x.w[0] & 1) {
            // (x + q) / 2 with the 113-bit intermediate
            uint64_t t = (uint64_t)x.w[0] + Q0;
            uint32_t r0 = (uint32_t)t;
            t = (uint64_t)x.w[1] + Q1 + (t >> 32);
            uint32_t r1 = (uint32_t)t;
            t = (uint64_t)x.w[2] + Q2 + (t >> 32);
            uint32_t r2 = (uint32_t)t;
            uint32_t r3 = x.w[3] + Q3 + (uint32_t)(t >> 32);
            x = zq{{r0, r1, r2, r3}};
        }
        shr1(x);
    };
    auto ge = [](const zq& a, const zq& b) {
        for (int i = 3; i >= 0; --i)
            if (a.w[i] != b.w[i]) return a.w[i] > b.w[i];
        return true;
    };
    auto sub_raw = [](zq& a, const zq& b) {
        uint64_t t = (uint64_t)a.w[0] - b.w[0];
        a.w[0] = (uint32_t)t;
        uint64_t br = t >> 63;
        t = (uint64_t)a.w[1] - b.w[1] - br;
        a.w[1] = (uint32_t)t;
        br = t >> 63;
        t = (uint64_t)a.w[2] - b.w[2] - br;
        a.w[2] = (uint32_t)t;
        br = t >> 63;
        a.w[3] = a.w[3] - b.w[3] - (uint32_t)br;
    };
    while (!is_one(u) && !is_one(v)) {
        while (!(u.w[0] & 1)) {
            shr1(u);
            half(x1);
        }
        while (!(v.w[0] & 1)) {
            shr1(v);
            half(x2);
        }
        if (ge(u, v)) {
            sub_raw(u, v);
            x1 = zq_sub(x1, x2);
        } else {
            sub_raw(v, u);
            x2 = zq_sub(x2, x1);
        }
    }
    return is_one(u) ? x1 : x2;
}

__device__ void write_challenge_pair(zq* chal_out, const zq& x) {
    chal_out[0] = x;
    chal_out[1] = zq_inv_binary(x);
}

// Root merge; one workgroup of 1024 threads over n <= 1024 CVs. `chunk_level` says the CVs
// are chunk CVs (needed to redo the counter chunks of a zero challenge).
__global__ void __launch_bounds__(1024) k_b3_root(const uint32_t* __restrict__ cvs, size_t n,
                                                  uint8_t* msg, size_t len, KeyW key, uint32_t flags,
                                                  uint32_t* digest_out, zq* chal_out, uint32_t* err,
                                                  int chunk_level) {
    __shared__ uint32_t sh[1024 * 8];
    __shared__ uint32_t dig[8];
    __shared__ int again;
    uint32_t counter = 0;
    for (;;) {
        if (n == 1) {
            if (threadIdx.x == 0) {
                uint32_t d[8];
                chunk_cv(msg, len, 0, key.w, flags, true, d);
                for (int j = 0; j < 8; ++j) dig[j] = d[j];
            }
            __syncthreads();
        } else {
            for (uint32_t i = threadIdx.x; i < n * 8; i += blockDim.x) sh[i] = cvs[i];
            __syncthreads();
            merge_lds(sh, (uint32_t)n, key.w, flags, true);
            if (threadIdx.x < 8) dig[threadIdx.x] = sh[threadIdx.x];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            again = 0;
            if (digest_out)
                for (int j = 0; j < 8; ++j) digest_out[j] = dig[j];
            if (chal_out) {
                uint32_t h[8];
                for (int j = 0; j < 8; ++j) h[j] = dig[j];
                const zq x = zq_from_digest(h);
                if (!zq_is_zero(x)) {
                    write_challenge_pair(chal_out, x);
                } else {
                    // sample_nonzero_challenge (transcript.rs:66-75): counter + 1, hash again
                    counter += 1;
                    for (uint32_t b = 0; b < 4; ++b) msg[len - 4 + b] = (uint8_t)(counter >> (8 * b));
                    again = 1;
                }
            }
        }
        __syncthreads();
        if (!again) break;
        if (!chunk_level) {  // group CVs cannot be patched here; reported to the host
            if (threadIdx.x == 0) atomicOr(err, 0x40000000u);
            break;
        }
        // recompute the chunk(s) holding the counter bytes
        if (threadIdx.x == 0 && n > 1) {
            const size_t c0 = (len - 4) / 1024, c1 = (len - 1) / 1024;
            uint32_t* cw = const_cast<uint32_t*>(cvs);
            for (size_t c = c0; c <= c1; ++c) {
                uint32_t d[8];
                chunk_cv(msg, len, c, key.w, flags, false, d);
                for (int j = 0; j < 8; ++j) cw[c * 8 + j] = d[j];
            }
            __threadfence_block();
        }
        __syncthreads();
        if (counter > 64 && threadIdx.x == 0) {  // unreachable in practice (2^-112 per try)
            atomicOr(err, 0x80000000u);
            break;
        }
    }
}

void launch_b3_root(const uint32_t* cvs, size_t n, const uint8_t* msg, size_t len,
                    const uint32_t key[8], uint32_t flags, uint32_t* digest_out, zq* chal_out,
                    uint32_t* err_flags, int chunk_level, hipStream_t s) {
    KeyW k;
    for (int i = 0; i < 8; ++i) k.w[i] = key[i];
    hipLaunchKernelGGL(k_b3_root, dim3(1), dim3(1024), 0, s, cvs, n, const_cast<uint8_t*>(msg), len, k,
                       flags, digest_out, chal_out, err_flags, chunk_level);
}

// Hash a device buffer: chunk CVs, group merges while more than 1024 values remain, root.
void device_hash(const uint8_t* msg, size_t len, const uint32_t key[8], uint32_t flags,
                 uint32_t* cv_scratch, size_t cv_scratch_words, uint32_t* digest_out, zq* chal_out,
                 uint32_t* err_flags, hipStream_t s) {
    const size_t n = nchunks_of(len);
    launch_b3_chunks(msg, len, key, flags, cv_scratch, s);
    uint32_t* cur = cv_scratch;
    size_t cnt = n;
    uint32_t* next = cv_scratch + n * 8;
    while (cnt > 1024) {
        launch_b3_group_merge(cur, cnt, key, flags, next, s);
        cnt = (cnt + 1023) / 1024;
        cur = next;
        next = next + cnt * 8;
    }
    (void)cv_scratch_words;
    launch_b3_root(cur, cnt, msg, len, key, flags, digest_out, chal_out, err_flags, cur == cv_scratch, s);
}

size_t device_hash_scratch_words(size_t len) {
    size_t n = nchunks_of(len), words = n * 8;
    while (n > 1024) {
        n = (n + 1023) / 1024;
        words += n * 8;
    }
    return words;
}

}  // namespace dmx
