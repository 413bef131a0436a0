// BLAKE3 tree hashing on the device: file_id (src/utils/io.rs:33) and the Fiat-Shamir
// challenges (src/protocol/transcript.rs:42-90), so a whole proof runs without host round trips.
//
// Chunks (1 KiB) are compressed one per thread; aligned groups of 1024 chunk chaining values
// are merged by one workgroup each (an aligned power-of-two run of chunks is a complete subtree
// of BLAKE3's left-balanced tree); a last workgroup merges the group values pairwise bottom-up,
// carrying an odd node upward, which reproduces the left-balanced tree, and sets ROOT on the
// final compression.
#include "dmx_b3quad.hpp"

namespace dmx {

size_t transcript_len(uint32_t label_len, const uint32_t lens[3]) { return tx_message_len(label_len, lens); }

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

// CV of chunk `ci` computed by one quad; with `root` the last block carries ROOT, giving the
// digest. Lane q returns (cv[q], cv[4+q]).
__device__ void chunk_cv_quad(const uint8_t* msg, size_t len, size_t ci, const uint32_t key[8], uint32_t flags,
                              bool root, const QuadSel& sel, uint32_t& ca, uint32_t& cb) {
    ca = sel(key[0], key[1], key[2], key[3]);
    cb = sel(key[4], key[5], key[6], key[7]);
    const size_t start = ci * b3::CHUNK_LEN;
    const size_t end = (start + b3::CHUNK_LEN < len) ? start + b3::CHUNK_LEN : len;
    const size_t nb = end > start ? (end - start + 63) / 64 : 1;
    // block b + 1 is loaded while block b is compressed, so only the first load's latency shows
    uint32_t m[16], nxt[16], blen = 0, nlen = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) nxt[i] = 0;
    if (end > start) load_block(msg, len, start, nxt, nlen);
    for (size_t b = 0; b < nb; ++b) {
#pragma unroll
        for (int i = 0; i < 16; ++i) m[i] = nxt[i];
        blen = nlen;
        if (b + 1 < nb) load_block(msg, len, start + 64 * (b + 1), nxt, nlen);
        uint32_t fl = flags | (b == 0 ? b3::CHUNK_START : 0) | (b + 1 == nb ? b3::CHUNK_END : 0);
        if (root && b + 1 == nb) fl |= b3::ROOT;
        compress_quad(ca, cb, m, ci, blen, fl, sel);
    }
}



// Chunk CVs of a transcript message: 512 threads write the window, its first 16 quads hash the
// chunks. The chunks holding the counter are also written to `msg`, where a root that meets a
// zero challenge patches the counter and hashes them again.
__global__ void __launch_bounds__(512) k_b3_tx_chunks(TranscriptMsg t, size_t len, KeyW key, uint32_t flags,
                                                     uint32_t* __restrict__ cvs, uint8_t* __restrict__ msg, size_t js) {
    critical_path_priority();
    __shared__ __align__(16) uint64_t win[TX_WINDOW / 8];
    const size_t jo = job_off(js);
    tx_to_job(t, jo);
    cvs = jp(cvs, jo);
    msg = jp(msg, jo);
    const size_t n = nchunks_of(len);
    const size_t base = (size_t)blockIdx.x * TX_WINDOW;
    const size_t wend = base + TX_WINDOW < len ? base + TX_WINDOW : len;
    const size_t stop = (wend + 63) & ~(size_t)63;  // through the end of the last block
    tx_pack_lds(t, (uint32_t)(base / 8), (uint32_t)((stop - base) / 8), win);
    __syncthreads();
    const size_t tail = ((len - 4) / b3::CHUNK_LEN) * b3::CHUNK_LEN;
    const uint8_t* wb = reinterpret_cast<const uint8_t*>(win);
    for (size_t o = (tail > base ? tail : base) + threadIdx.x; o < wend; o += blockDim.x) msg[o] = wb[o - base];
    const uint32_t q = threadIdx.x & 3;
    const size_t ci = (size_t)blockIdx.x * 16 + (threadIdx.x >> 2);
    if (threadIdx.x >= 64 || ci >= n) return;  // uniform within a quad; no barrier follows
    const QuadSel sel{(q & 1) != 0, (q & 2) != 0};
    uint32_t ca, cb;
    chunk_cv_quad_lds(win, base, len, ci, key.w, flags, false, sel, ca, cb);
    cvs[ci * 8 + q] = ca;
    cvs[ci * 8 + 4 + q] = cb;
}

// a job's own length when lengths live on the device (bounded by the launch capacity)
__device__ __forceinline__ size_t job_len(size_t cap, const uint64_t* dlen, size_t jo) {
    return dlen ? (size_t)min((uint64_t)cap, *jp(dlen, jo)) : cap;
}

// one quad per chunk, 16 chunks per 64-thread block; the guard is uniform within a quad
__global__ void __launch_bounds__(64) k_b3_chunks(const uint8_t* __restrict__ msg, size_t len, KeyW key,
                                                  uint32_t flags, uint32_t* __restrict__ cvs, size_t n, size_t js,
                                                  const uint64_t* dlen) {
    const uint32_t q = threadIdx.x & 3;
    const size_t ci = (size_t)blockIdx.x * 16 + (threadIdx.x >> 2);
    const size_t jo = job_off(js);
    if (dlen) {
        len = job_len(len, dlen, jo);
        n = nchunks_of(len);
    }
    if (ci >= n) return;
    msg = jp(msg, jo);
    cvs = jp(cvs, jo);
    const QuadSel sel{(q & 1) != 0, (q & 2) != 0};
    uint32_t ca, cb;
    chunk_cv_quad(msg, len, ci, key.w, flags, false, sel, ca, cb);
    cvs[ci * 8 + q] = ca;
    cvs[ci * 8 + 4 + q] = cb;
}


static void launch_b3_chunks(const uint8_t* msg, size_t len, const uint32_t key[8], uint32_t flags,
                             uint32_t* cvs, hipStream_t s, const Jobs& jb, const uint64_t* dlen) {
    KeyW k;
    for (int i = 0; i < 8; ++i) k.w[i] = key[i];
    const size_t n = nchunks_of(len);
    hipLaunchKernelGGL(k_b3_chunks, dim3((unsigned)((n + 15) / 16), jb.n), dim3(64), 0, s, msg, len, k, flags,
                       cvs, n, jb.stride, dlen);
}


// n chaining values (8 words each, 32-byte aligned) from global memory into LDS with 16-byte
// loads, a thread's loads issued together
__device__ __forceinline__ void load_cvs_lds(uint32_t* dst, const uint32_t* src, uint32_t n) {
    const uint4* s4 = reinterpret_cast<const uint4*>(src);
    uint4* d4 = reinterpret_cast<uint4*>(dst);
#pragma unroll 4
    for (uint32_t i = threadIdx.x; i < 2 * n; i += blockDim.x) d4[i] = s4[i];
}

// With `dlen`, `n` is the capacity's chunk count and each job uses its own; a job whose chunk
// values fit one group has no group level and skips.
__global__ void __launch_bounds__(1024) k_b3_group_merge(const uint32_t* __restrict__ cvs, size_t n, KeyW key,
                                                         uint32_t flags, uint32_t* __restrict__ out, size_t js,
                                                         const uint64_t* dlen, size_t cap_len) {
    __shared__ __align__(16) uint32_t sa[1024 * 8];
    __shared__ uint32_t sb[512 * 8 + 8];
    const size_t base = (size_t)blockIdx.x * 1024;
    const size_t jo = job_off(js);
    if (dlen) {
        n = nchunks_of(job_len(cap_len, dlen, jo));
        if (n <= 1024 || base >= n) return;  // uniform over the workgroup
    }
    cvs = jp(cvs, jo);
    out = jp(out, jo);
    const uint32_t cnt = (uint32_t)((n - base) < 1024 ? (n - base) : 1024);
    load_cvs_lds(sa, cvs + base * 8, cnt);
    __syncthreads();
    const uint32_t* r = merge_lds(sa, sb, cnt, key.w, flags, false);
    if (threadIdx.x < 8) out[blockIdx.x * 8 + threadIdx.x] = r[threadIdx.x];
}

static void launch_b3_group_merge(const uint32_t* cvs, size_t n, const uint32_t key[8], uint32_t flags,
                                  uint32_t* out, hipStream_t s, const Jobs& jb, const uint64_t* dlen, size_t cap_len) {
    KeyW k;
    for (int i = 0; i < 8; ++i) k.w[i] = key[i];
    hipLaunchKernelGGL(k_b3_group_merge, dim3((unsigned)((n + 1023) / 1024), jb.n), dim3(1024), 0, s, cvs, n, k,
                       flags, out, jb.stride, dlen, cap_len);
}

// Root merge; one workgroup of 512 threads per job over n <= 1024 CVs. (Its waves need about 120
// VGPRs each: at 1024 threads the workgroup would take a whole CU's register file and wait for
// any concurrent kernel to leave one; at 512 it shares a CU.) `level` says what the CVs at `cvs`
// are: 0 the chunk values (the same as `chunk_cvs`), 1 the group values of the nchunks chunk
// values at `chunk_cvs`, 2 anything deeper. With `dlen` the job's own length picks the level:
// chunk values at `cvs` (<= 1024 chunks) or group values at cvs + l1_off.
//
// A rejected challenge candidate (sample_nonzero_challenge, transcript.rs:66-75) bumps the
// counter in the message tail kept at `msg`, recomputes the chunk(s) holding it and, at level 1,
// the group value(s) over those chunks, then merges again.
__global__ void __launch_bounds__(512) k_b3_root(uint32_t* __restrict__ cvs, size_t n, uint32_t* chunk_cvs,
                                                  size_t nchunks, uint8_t* msg, size_t len, KeyW key, uint32_t flags,
                                                  uint32_t* digest_out, zq* chal_out, uint32_t* err,
                                                  const uint32_t* reject, int level, size_t js, const uint64_t* dlen,
                                                  size_t l1_off, int want_inv) {
    critical_path_priority();
    __shared__ __align__(16) uint32_t sa[1024 * 8];
    __shared__ uint32_t sb[512 * 8 + 8];
    __shared__ uint32_t dig[8];
    __shared__ int again;
    const size_t jo = job_off(js);
    cvs = jp(cvs, jo);
    chunk_cvs = jp(chunk_cvs, jo);
    msg = jp(msg, jo);
    digest_out = jp(digest_out, jo);
    chal_out = jp(chal_out, jo);
    err = jp(err, jo);
    const uint32_t rj = reject_below(jp(reject, jo));
    if (dlen) {
        len = job_len(len, dlen, jo);
        nchunks = nchunks_of(len);
        chunk_cvs = cvs;
        if (nchunks > 1024) {
            cvs += l1_off;
            n = (nchunks + 1023) / 1024;
            level = 1;
        } else {
            n = nchunks;
            level = 0;
        }
    }
    const uint32_t q = threadIdx.x & 3;
    const QuadSel sel{(q & 1) != 0, (q & 2) != 0};
    for (uint32_t counter = 0;; ++counter) {  // every thread counts the same passes
        if (n == 1) {
            if (threadIdx.x < 4) {
                uint32_t ca, cb;
                chunk_cv_quad(msg, len, 0, key.w, flags, true, sel, ca, cb);
                dig[q] = ca;
                dig[4 + q] = cb;
            }
            __syncthreads();
        } else {
            load_cvs_lds(sa, cvs, (uint32_t)n);
            __syncthreads();
            const uint32_t* r = merge_lds(sa, sb, (uint32_t)n, key.w, flags, true);
            if (threadIdx.x < 8) dig[threadIdx.x] = r[threadIdx.x];
            __syncthreads();
        }
        // The challenge and its inverse are computed by all of wave 0 (identical values in every
        // lane); lane 0 stores them. The multi-word products of the inverse stay in vector
        // registers, where a 64-bit multiply-add is one instruction; its divstep loop reads the
        // low words back into scalar registers. Without want_inv only x is written (the caller
        // inverts it off its critical path).
        if (threadIdx.x < 64 && chal_out) {
            uint32_t h[8];
            // identity quad permutation: same value, but a per-lane register to the compiler
            for (int j = 0; j < 8; ++j) h[j] = (uint32_t)__builtin_amdgcn_mov_dpp((int)dig[j], 0xE4, 0xF, 0xF, false);
            const zq x = zq_from_digest(h);
            if (candidate_taken(x, counter, rj)) {
                const zq xi = want_inv ? zq_inv_divsteps(x) : x;
                if (threadIdx.x == 0) {
                    chal_out[0] = x;
                    if (want_inv) chal_out[1] = xi;
                }
            }
        }
        if (threadIdx.x == 0) {
            again = 0;
            if (digest_out)
                for (int j = 0; j < 8; ++j) digest_out[j] = dig[j];
            if (chal_out) {
                uint32_t h[8];
                for (int j = 0; j < 8; ++j) h[j] = dig[j];
                if (candidate_taken(zq_from_digest(h), counter, rj)) {
                    // written above
                } else if (level > 1) {
                    atomicOr(err, 0x40000000u);  // deeper trees are not patched (no message has them)
                } else if (counter + 1 >= RESAMPLE_LIMIT) {
                    atomicOr(err, ERR_RESAMPLE);
                    exhausted_challenge(chal_out, want_inv);
                } else {
                    // sample_nonzero_challenge (transcript.rs:66-75): counter + 1, hash again
                    const uint32_t c1 = counter + 1;
                    for (uint32_t b = 0; b < 4; ++b) msg[len - 4 + b] = (uint8_t)(c1 >> (8 * b));
                    again = 1;
                }
            }
            __threadfence();
        }
        __syncthreads();
        if (!again) break;  // uniform: every thread reads the same shared flag
        const size_t c0 = (len - 4) / 1024, c1 = (len - 1) / 1024;
        if (threadIdx.x < 4 && n > 1) {  // quad 0 recomputes the chunk(s) holding the counter
            for (size_t c = c0; c <= c1; ++c) {
                uint32_t ca, cb;
                chunk_cv_quad(msg, len, c, key.w, flags, false, sel, ca, cb);
                chunk_cvs[c * 8 + q] = ca;
                chunk_cvs[c * 8 + 4 + q] = cb;
            }
            __threadfence();
        }
        __syncthreads();
        if (level == 1) {  // the group value(s) over those chunks, from their chunk values
            for (size_t g = c0 / 1024; g <= c1 / 1024; ++g) {  // uniform bounds
                const uint32_t cnt = (uint32_t)min((size_t)1024, nchunks - g * 1024);
                load_cvs_lds(sa, chunk_cvs + g * 1024 * 8, cnt);
                __syncthreads();
                const uint32_t* r = merge_lds(sa, sb, cnt, key.w, flags, false);
                if (threadIdx.x < 8) cvs[g * 8 + threadIdx.x] = r[threadIdx.x];
                __threadfence();
                __syncthreads();
            }
        }
    }
}

static void launch_b3_root(const uint32_t* cvs, size_t n, const uint32_t* chunk_cvs, size_t nchunks, int level,
                           const uint8_t* msg, size_t len, const uint32_t key[8], uint32_t flags, uint32_t* digest_out,
                           zq* chal_out, uint32_t* err_flags, const uint32_t* reject, hipStream_t s, const Jobs& jb,
                           const uint64_t* dlen, size_t l1_off, int want_inv = 1) {
    KeyW k;
    for (int i = 0; i < 8; ++i) k.w[i] = key[i];
    hipLaunchKernelGGL(k_b3_root, dim3(1, jb.n), dim3(512), 0, s, const_cast<uint32_t*>(cvs), n,
                       const_cast<uint32_t*>(chunk_cvs), nchunks, const_cast<uint8_t*>(msg), len, k, flags,
                       digest_out, chal_out, err_flags, reject, level, jb.stride, dlen, l1_off, want_inv);
}

// Hash a device buffer: chunk CVs, group merges while more than 1024 values remain, root.
// A whole challenge in one workgroup for messages of at most 16 chunks (wg_tx_challenge).
// (Larger messages need more than one CU to hash their chunks quickly and take the chunk and
// root kernels.)
__global__ void __launch_bounds__(512) k_b3_tx_challenge(TranscriptMsg t, size_t len, KeyW key, uint32_t flags,
                                                        zq* chal_out, uint32_t* err, size_t js, int want_inv) {
    critical_path_priority();
    __shared__ WgChal sm;
    const size_t jo = job_off(js);
    tx_to_job(t, jo);
    (void)wg_tx_challenge(t, len, key.w, flags, sm, jp(chal_out, jo), jp(err, jo), want_inv);
}

static void launch_b3_tx_chunks(const TranscriptMsg& t, size_t len, const uint32_t key[8], uint32_t flags,
                                uint32_t* cvs, uint8_t* msg, hipStream_t s, const Jobs& jb) {
    KeyW k;
    for (int i = 0; i < 8; ++i) k.w[i] = key[i];
    hipLaunchKernelGGL(k_b3_tx_chunks, dim3((unsigned)((len + TX_WINDOW - 1) / TX_WINDOW), jb.n), dim3(512), 0, s, t,
                       len, k, flags, cvs, msg, jb.stride);
}

static void hash_tree(uint32_t* cv_scratch, size_t n, const uint8_t* msg, size_t len, const uint32_t key[8],
                      uint32_t flags, uint32_t* digest_out, zq* chal_out, uint32_t* err_flags,
                      const uint32_t* reject, hipStream_t s, const Jobs& jb, int want_inv = 1);

void transcript_challenge(const TranscriptMsg& t, const uint32_t key[8], uint32_t flags, uint8_t* msg,
                          uint32_t* cv_scratch, zq* chal_out, uint32_t* err_flags, hipStream_t s, const Jobs& jb,
                          bool inverse) {
    if (jb.n == 0) return;
    const size_t len = tx_message_len(t.label_len, t.lens);
    if (nchunks_of(len) <= CHALLENGE_MAX_CHUNKS) {
        KeyW k;
        for (int i = 0; i < 8; ++i) k.w[i] = key[i];
        hipLaunchKernelGGL(k_b3_tx_challenge, dim3(1, jb.n), dim3(512), 0, s, t, len, k, flags, chal_out, err_flags,
                           jb.stride, (int)inverse);
        return;
    }
    launch_b3_tx_chunks(t, len, key, flags, cv_scratch, msg, s, jb);
    hash_tree(cv_scratch, nchunks_of(len), msg, len, key, flags, nullptr, chal_out, err_flags, t.reject, s, jb,
              (int)inverse);
}

void device_hash(const uint8_t* msg, size_t len, const uint32_t key[8], uint32_t flags,
                 uint32_t* cv_scratch, size_t cv_scratch_words, uint32_t* digest_out, zq* chal_out,
                 uint32_t* err_flags, hipStream_t s, const Jobs& jb, const uint64_t* dlen) {
    if (jb.n == 0) return;
    const size_t n = nchunks_of(len);
    (void)cv_scratch_words;
    launch_b3_chunks(msg, len, key, flags, cv_scratch, s, jb, dlen);
    if (dlen) {  // per-job lengths: at most one group level
        if (n > 1024) launch_b3_group_merge(cv_scratch, n, key, flags, cv_scratch + n * 8, s, jb, dlen, len);
        launch_b3_root(cv_scratch, n, cv_scratch, n, 0, msg, len, key, flags, digest_out, chal_out, err_flags, nullptr,
                       s, jb, dlen, n * 8);
        return;
    }
    hash_tree(cv_scratch, n, msg, len, key, flags, digest_out, chal_out, err_flags, nullptr, s, jb);
}

// group merges while more than 1024 values remain, then the root
static void hash_tree(uint32_t* cv_scratch, size_t n, const uint8_t* msg, size_t len, const uint32_t key[8],
                      uint32_t flags, uint32_t* digest_out, zq* chal_out, uint32_t* err_flags,
                      const uint32_t* reject, hipStream_t s, const Jobs& jb, int want_inv) {
    uint32_t* cur = cv_scratch;
    size_t cnt = n;
    uint32_t* next = cv_scratch + n * 8;
    int level = 0;
    while (cnt > 1024) {
        launch_b3_group_merge(cur, cnt, key, flags, next, s, jb, nullptr, 0);
        cnt = (cnt + 1023) / 1024;
        cur = next;
        next = next + cnt * 8;
        ++level;
    }
    launch_b3_root(cur, cnt, cv_scratch, n, level, msg, len, key, flags, digest_out, chal_out, err_flags, reject, s, jb,
                   nullptr, 0, want_inv);
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
