// Negacyclic CRT-NTT kernels for gfx950: forward transforms with fused row producers
// (fold / generator hash), NTT-domain multiply-accumulate reductions for the commitment
// mat-vec, inverse transforms and the exact centred CRT lift to Z_q.
//
// Reference behaviour: src/algebra/ntt.rs:112-224 (ring product via CRT over eight primes),
// ntt.rs:400-427 (centred lift), src/commitment/sis.rs:146-167 (commitment sum).
// The reference transforms each product separately with a 2n-point cyclic NTT; here each ring
// element is transformed once with an n-point negacyclic NTT and the commitment sum is formed
// in the NTT domain, so a commitment costs (rows*(1+K) + K) transforms instead of 3*rows*K.
#include <cstdlib>
#include <cstring>

#include "dmx_b3quad.hpp"
#include "dmx_blake3.hpp"
#include "dmx_kernels.hpp"

namespace dmx {

__constant__ PrimeConsts c_pc[NPRIMES];
__constant__ CrtConsts c_crt;

void set_device_constants(const PrimeConsts* pc, const CrtConsts* crt, hipStream_t s) {
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(c_pc), pc, sizeof(PrimeConsts) * NPRIMES, 0,
                                 hipMemcpyHostToDevice, s);
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(c_crt), crt, sizeof(CrtConsts), 0, hipMemcpyHostToDevice, s);
}

constexpr int NTT_THREADS = 256;
constexpr int LDS_WORDS = 16384;  // 64 KiB: all 8 primes up to degree 2048

__device__ __forceinline__ zq load_zq(const zq* p) {
    uint4 v = *reinterpret_cast<const uint4*>(p);
    return zq{{v.x, v.y, v.z, v.w}};
}
__device__ __forceinline__ void store_zq(zq* p, const zq& x) {
    *reinterpret_cast<uint4*>(p) = make_uint4(x.w[0], x.w[1], x.w[2], x.w[3]);
}

// keyed BLAKE3 of "g" || u64le(i) || u64le(k) || u64le(c) (25 bytes, one block), reduced mod q
// (src/commitment/sis.rs:231-245)
__device__ __forceinline__ zq gen_coeff(const uint32_t* key, uint64_t i, uint32_t k, uint64_t c) {
    uint32_t m[16];
    m[0] = 0x67u | ((uint32_t)i << 8);
    m[1] = (uint32_t)(i >> 24);
    m[2] = (uint32_t)(i >> 56) | (k << 8);
    m[3] = 0;  // k < 2^24
    // the only word that changes along a row; kept opaque so the compiler does not split it into
    // lane and loop parts and re-add those parts at each of its seven uses (about 85 extra adds)
    uint32_t m4 = ((uint32_t)c) << 8;
    asm("" : "+v"(m4));
    m[4] = m4;
    m[5] = (uint32_t)(c >> 24);
    m[6] = (uint32_t)(c >> 56);
#pragma unroll
    for (int t = 7; t < 16; ++t) m[t] = 0;
    // byte 17 starts u64le(c): bytes 17..24 -> word 4 bits 8.., word 5, word 6 bits 0..7
    uint32_t h[8];
    b3::compress_cv(key, m, 0, 25,
                    b3::CHUNK_START | b3::CHUNK_END | b3::ROOT | b3::KEYED_HASH, h);
    return zq_from_digest(h);
}

// produce coefficient c of row `row` per the row source
// `key` holds the ROW_GEN key words (registers or device memory)
__device__ __forceinline__ zq row_value(const RowSrc& src, const uint32_t* key, size_t row, uint32_t c, int log_m) {
    const size_t m = (size_t)1 << log_m;
    switch (src.mode) {
        case ROW_PLAIN:
            return load_zq(src.a + row * m + c);
        case ROW_VFOLD: {
            zq a = load_zq(src.a + row * m + c);
            zq b = load_zq(src.a + (row + src.b_off) * m + c);
            return zq_add(a, zq_mul(b, *src.scal));
        }
        case ROW_PFOLD_W: {
            const zq* base = src.a + row * (2 * m);
            zq e = load_zq(base + 2 * c), o = load_zq(base + 2 * c + 1);
            return zq_add(e, zq_mul(o, *src.scal));
        }
        case ROW_PFOLD_G: {
            const zq* base = src.a + row * (2 * m);
            zq e = load_zq(base + 2 * c);
            // (X * odd)[c] in Z_q[X]/(X^m + 1): odd[c-1] for c >= 1, -odd[m-1] at c = 0
            zq xo = c ? load_zq(base + 2 * c - 1) : zq_neg(load_zq(base + 2 * m - 1));
            return zq_add(e, zq_mul(xo, *src.scal));
        }
        case ROW_VFOLD_L: {
            zq a = load_zq(src.a + row * m + c);
            zq b = load_zq(src.a + (row + src.b_off) * m + c);
            return zq_add(zq_mul(a, *src.scal), b);
        }
        case ROW_PFOLD_GL: {
            const zq* base = src.a + row * (2 * m);
            zq e = load_zq(base + 2 * c);
            zq xo = c ? load_zq(base + 2 * c - 1) : zq_neg(load_zq(base + 2 * m - 1));
            return zq_add(zq_mul(e, *src.scal), xo);
        }
        default: {  // ROW_GEN
            uint64_t i = row / K;
            uint32_t k = (uint32_t)(row % K);
            return gen_coeff(key, i, k, c);
        }
    }
}

// ---------------------------------------------------------------------------------------
// forward negacyclic NTT (Cooley-Tukey, natural order in, bit-reversed order out)
// butterfly: (a, b) -> (a + z*b, a - z*b), z = zeta[m/(2*len) + block]
// ---------------------------------------------------------------------------------------
// InvTask: wave 0 of the extra workgroup 0 (every lane computes the same value; lane 0 stores)
__device__ void inverse_task(const InvTask& it, size_t jo) {
    if (threadIdx.x >= 64) return;
    zq* chal = jp(it.chal, jo);
    const zq x = chal[0];
    const zq xi = zq_inv_divsteps(x);
    const zq u = it.u_prev ? zq_mul(*jp(it.u_prev, jo), xi) : xi;
    if (threadIdx.x == 0) {
        chal[1] = xi;
        *jp(it.u_out, jo) = u;
    }
}

__global__ void __launch_bounds__(NTT_THREADS) k_ntt_fwd_rows(FwdSeg s1, FwdSeg s2, uint32_t blocks1, int log_m,
                                                               NttTables tw, size_t js, InvTask it) {
    extern __shared__ uint32_t sh[];
    uint32_t bx = blockIdx.x;
    if (it.chal) {  // uniform over the workgroup; the task workgroup meets no barrier
        if (bx == 0) return inverse_task(it, job_off(js));
        --bx;
    }
    const bool second = bx >= blocks1;
    RowSrc src = second ? s2.src : s1.src;
    uint32_t* out = second ? s2.out : s1.out;
    const size_t jo = job_off(js);
    if (!src.a_shared) src.a = jp(src.a, jo);
    src.dst = jp(src.dst, jo);
    src.dst2 = jp(src.dst2, jo);
    src.scal = jp(src.scal, jo);
    src.gseed = jp(src.gseed, jo);
    out = jp(out, jo);
    uint32_t key[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (src.mode == ROW_GEN) {
#pragma unroll
        for (int i = 0; i < 8; ++i) key[i] = src.gseed[i];
    }
    const int m = 1 << log_m;
    const size_t row = second ? bx - blocks1 : bx;
    const int pg = min(NPRIMES, LDS_WORDS >> log_m);  // primes per LDS pass
    for (int p0 = 0; p0 < NPRIMES; p0 += pg) {
        // 1) produce the row and reduce it into the primes of this group
        for (int c = threadIdx.x; c < m; c += NTT_THREADS) {
            zq x = row_value(src, key, row, c, log_m);
            if (src.dst && p0 == 0) store_zq(src.dst + row * m + c, x);
            if (src.dst2 && p0 == 0) store_zq(src.dst2 + row * m + c, x);
#pragma unroll 1
            for (int q = 0; q < pg; ++q) sh[q * m + c] = zq_mod_p_centred(x, c_pc[p0 + q]);
        }
        __syncthreads();
        // 2) log_m butterfly layers in LDS
        const int half = m >> 1;
        for (int len = half; len >= 1; len >>= 1) {
            const int lg = __builtin_ctz(len);
            for (int t = threadIdx.x; t < pg * half; t += NTT_THREADS) {
                const int q = t >> (log_m - 1);
                const int b = t & (half - 1);
                const int blk = b >> lg;
                const int j = (blk << (lg + 1)) | (b & (len - 1));
                const int kz = (half >> lg) + blk;  // m/(2 len) + blk
                const int pi = p0 + q;
                const uint32_t p = c_pc[pi].p;
                const uint32_t zm = tw.zeta_m[pi * TW_SIZE + kz];
                const uint32_t zmp = tw.zeta_mp[pi * TW_SIZE + kz];
                uint32_t* s = sh + q * m;
                const uint32_t u = s[j];
                const uint32_t v = mulc(s[j + len], zm, zmp, p);
                s[j] = addp(u, v, p);
                s[j + len] = subp(u, v, p);
            }
            __syncthreads();
        }
        // 3) write out
        for (int t = threadIdx.x; t < pg * m; t += NTT_THREADS) {
            const int q = t >> log_m, c = t & (m - 1);
            out[(row * NPRIMES + p0 + q) * m + c] = sh[t];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// Register-blocked forward transform for degrees 2^6..2^11. A row's 8 prime slices are each
// handled by TP threads holding E = m/TP elements (TP <= E). Pass 1: thread l of a slice holds
// elements l + TP*i and runs every layer with len >= TP in registers (its twiddles depend only
// on i, so they are shared by the slice's threads). One LDS transpose later, pass 2: thread l
// holds the contiguous elements E*l .. E*l+E-1 and runs the layers len < TP. Two LDS round
// trips per row instead of one per layer. A slice's LDS words are padded by one word after
// every E elements, and slices are TP * (E + 1) words apart: pass 1 reads slot l + TP*i + i*TP/E
// and pass 2 slot (E + 1)*l + e, both a per-thread base plus a constant offset, and in both
// passes (and for the slices sharing a 32-lane LDS pass) the lanes hit distinct banks.
// ---------------------------------------------------------------------------------------
template <int E, int TP>
struct RegNtt {
    static constexpr int M = E * TP;
    static constexpr int LOG_E = __builtin_ctz(E);
    static constexpr int LOG_TP = __builtin_ctz(TP);
    static constexpr int LOG_M = LOG_E + LOG_TP;
    static constexpr int TPR = NPRIMES * TP;  // threads per row
    static constexpr int RPB = NTT_THREADS / TPR;
    static constexpr int SL = TP * (E + 1);  // LDS words per prime slice
    static constexpr size_t LDS_BYTES = (size_t)RPB * NPRIMES * SL * 4;
    __device__ static __forceinline__ int slot(int q, int x) { return q * SL + x + (x >> LOG_E); }
};

__device__ __forceinline__ void ct_butterfly(uint32_t& x, uint32_t& y, uint32_t w, uint32_t wp, uint32_t p) {
    const uint32_t u = x;
    const uint32_t v = mulc(y, w, wp, p);
    x = addp(u, v, p);
    y = subp(u, v, p);
}

template <int E, int TP>
__global__ void __launch_bounds__(NTT_THREADS) k_ntt_fwd_reg(FwdSeg s1, FwdSeg s2, uint32_t blocks1, NttTables tw,
                                                             size_t js, InvTask it) {
    using G = RegNtt<E, TP>;
    extern __shared__ uint32_t sh[];
    uint32_t bx = blockIdx.x;
    if (it.chal) {  // uniform over the workgroup; the task workgroup meets no barrier
        if (bx == 0) return inverse_task(it, job_off(js));
        --bx;
    }
    const bool second = bx >= blocks1;
    RowSrc src = second ? s2.src : s1.src;
    uint32_t* out = second ? s2.out : s1.out;
    const uint32_t rows = second ? s2.rows : s1.rows;
    const size_t jo = job_off(js);
    if (!src.a_shared) src.a = jp(src.a, jo);
    src.dst = jp(src.dst, jo);
    src.dst2 = jp(src.dst2, jo);
    src.scal = jp(src.scal, jo);
    src.gseed = jp(src.gseed, jo);
    out = jp(out, jo);
    uint32_t key[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (src.mode == ROW_GEN) {
#pragma unroll
        for (int i = 0; i < 8; ++i) key[i] = src.gseed[i];
    }
    const int tr = threadIdx.x % G::TPR;
    const int rsub = threadIdx.x / G::TPR;
    const size_t row = (size_t)(second ? bx - blocks1 : bx) * G::RPB + rsub;
    const bool valid = row < rows;
    uint32_t* S = sh + rsub * NPRIMES * G::SL;
    // produce the row once, reduce into the 8 primes
    if (valid) {
        for (int c = tr; c < G::M; c += G::TPR) {
            const zq x = row_value(src, key, row, c, G::LOG_M);
            if (src.dst) store_zq(src.dst + row * G::M + c, x);
            if (src.dst2) store_zq(src.dst2 + row * G::M + c, x);
#pragma unroll
            for (int q = 0; q < NPRIMES; ++q) S[G::slot(q, c)] = zq_mod_p_centred(x, c_pc[q]);
        }
    }
    __syncthreads();
    const int q = tr >> G::LOG_TP, l = tr & (TP - 1);
    const uint32_t p = c_pc[q].p;
    const uint32_t* zt = tw.zeta_m + q * TW_SIZE;
    const uint32_t* ztp = tw.zeta_mp + q * TW_SIZE;
    uint32_t a[E];
#pragma unroll
    for (int i = 0; i < E; ++i) a[i] = S[G::slot(q, l + TP * i)];
    // pass 1: len = s*TP for s = E/2 .. 1; block b of 2s register slots uses zeta[E/(2s) + b]
#pragma unroll
    for (int s = E / 2; s >= 1; s >>= 1) {
#pragma unroll
        for (int b = 0; b < E / (2 * s); ++b) {
            const uint32_t w = zt[E / (2 * s) + b], wp = ztp[E / (2 * s) + b];
#pragma unroll
            for (int k = 0; k < s; ++k) ct_butterfly(a[2 * s * b + k], a[2 * s * b + k + s], w, wp, p);
        }
    }
#pragma unroll
    for (int i = 0; i < E; ++i) S[G::slot(q, l + TP * i)] = a[i];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) a[e] = S[G::slot(q, E * l + e)];
    // pass 2: len = TP/2 .. 1 inside the thread's E contiguous elements;
    // block b uses zeta[M/(2 len) + E*l/(2 len) + b]
#pragma unroll
    for (int len = TP / 2; len >= 1; len >>= 1) {
        const int base = G::M / (2 * len) + (E * l) / (2 * len);
#pragma unroll
        for (int b = 0; b < E / (2 * len); ++b) {
            const uint32_t w = zt[base + b], wp = ztp[base + b];
#pragma unroll
            for (int k = 0; k < len; ++k) ct_butterfly(a[2 * len * b + k], a[2 * len * b + k + len], w, wp, p);
        }
    }
    if (valid) {
        uint4* o = reinterpret_cast<uint4*>(out + (row * NPRIMES + q) * G::M + E * l);
#pragma unroll
        for (int e = 0; e < E; e += 4) o[e / 4] = make_uint4(a[e], a[e + 1], a[e + 2], a[e + 3]);
    }
}

template <int E, int TP>
static void launch_reg(const FwdSeg& s1, const FwdSeg& s2, const NttTables& tw, hipStream_t s, const Jobs& jb,
                       const InvTask& it) {
    using G = RegNtt<E, TP>;
    const size_t b1 = (s1.rows + G::RPB - 1) / G::RPB, b2 = (s2.rows + G::RPB - 1) / G::RPB;
    if (G::LDS_BYTES > 65536)  // past the default dynamic LDS limit
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_ntt_fwd_reg<E, TP>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)G::LDS_BYTES);
    hipLaunchKernelGGL((k_ntt_fwd_reg<E, TP>), dim3((unsigned)(b1 + b2 + (it.chal ? 1 : 0)), jb.n), dim3(NTT_THREADS),
                       G::LDS_BYTES, s, s1, s2, (uint32_t)b1, tw, jb.stride, it);
}

// ---------------------------------------------------------------------------------------
// Three-pass register transform for degrees 2^8..2^11: TP threads per prime slice (32 at 2^8,
// 64 at 2^9..2^11, 128 as the other shape at 2^11) holding E = m / TP elements each (8, 16 or
// 32), 8 TP threads per row, one row per workgroup. Against the two-pass blockings this takes a fraction of the LDS words per
// resident thread, so a CU holds more waves. With G = TP / E:
//   pass 1: thread l holds x = l + TP i (layers len = m/2 .. TP);
//   pass 2: x = b + G e + TP h (layers len = TP/2 .. G), with l = 2 h + b for G = 2,
//           l = h + E b for G = 4 and l = 8 h + b for G = 8;
//   pass 3: x = E l + e (layers len = G/2 .. 1).
// One LDS layout per shape serves the producer and all three passes without bank conflicts: the
// padded word x + (x >> 5) for TP = 64 at 2^10 and 2^11, and where no simple pad does, a swizzle
// x ^ s(x >> 5) inside each 32-word row: s(r) = (r & 7) | ((r >> 1) & 3) << 3 at 2^9,
// s(r) = (r & 15) | ((r >> 2) & 1) << 4 for TP = 128 at 2^11 and s(r) = r ^ (r << 2) at 2^8.
// Each pass writes back the words it read, so three barriers order the row.
// ---------------------------------------------------------------------------------------
template <int LOG_M, int TP_ = 64>
struct Reg3 {
    static constexpr int M = 1 << LOG_M;
    static constexpr int TP = TP_;
    static constexpr int LOG_TP = __builtin_ctz(TP);
    static constexpr int E = M / TP;
    static constexpr int G = TP / E;
    static constexpr int THREADS = NPRIMES * TP;
    static constexpr int LOG_E = LOG_M - LOG_TP;
    static constexpr int SWIZZLE = LOG_M == 9 ? 1 : TP == 128 ? 2 : TP == 32 ? 3 : 0;  // 0: padded rows
    static constexpr int SL = SWIZZLE ? M : M + M / 32;                 // LDS words per prime slice
    static constexpr size_t LDS_BYTES = (size_t)NPRIMES * SL * 4;
    static constexpr int MIN_WAVES = E == 32 ? 4 : 8;  // waves per SIMD asked of the register allocator
    static_assert((G == 2 || G == 4 || G == 8) &&
                      (TP == 64 || (TP == 128 && LOG_M == 11) || (TP == 32 && LOG_M == 8)), "shape");
    __device__ static __forceinline__ int slot(int x) {
        const int r = x >> 5;
        if (SWIZZLE == 1) return x ^ ((r & 7) | (((r >> 1) & 3) << 3));
        if (SWIZZLE == 2) return x ^ ((r & 15) | (((r >> 2) & 1) << 4));
        if (SWIZZLE == 3) return x ^ (r ^ (r << 2));
        return x + r;
    }
    // pass-2 lane split, chosen so the 32 lanes of an LDS access hit 32 banks
    __device__ static __forceinline__ int p2_h(int l) { return G == 2 ? l >> 1 : G == 4 ? l & (E - 1) : l >> 3; }
    __device__ static __forceinline__ int p2_b(int l) { return G == 2 ? l & 1 : G == 4 ? l >> LOG_E : l & 7; }
};

template <int LOG_M, int TP>
__global__ void __launch_bounds__((Reg3<LOG_M, TP>::THREADS), (Reg3<LOG_M, TP>::MIN_WAVES))
    k_ntt_fwd_reg3(FwdSeg s1, FwdSeg s2, uint32_t blocks1, NttTables tw, size_t js, InvTask it) {
    using C = Reg3<LOG_M, TP>;
    constexpr int M = C::M, E = C::E, G = C::G;
    extern __shared__ uint32_t sh[];
    uint32_t bx = blockIdx.x;
    if (it.chal) {  // uniform over the workgroup; the task workgroup meets no barrier
        if (bx == 0) return inverse_task(it, job_off(js));
        --bx;
    }
    const bool second = bx >= blocks1;
    RowSrc src = second ? s2.src : s1.src;
    uint32_t* out = second ? s2.out : s1.out;
    const size_t jo = job_off(js);
    if (!src.a_shared) src.a = jp(src.a, jo);
    src.dst = jp(src.dst, jo);
    src.dst2 = jp(src.dst2, jo);
    src.scal = jp(src.scal, jo);
    src.gseed = jp(src.gseed, jo);
    out = jp(out, jo);
    uint32_t key[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (src.mode == ROW_GEN) {
#pragma unroll
        for (int i = 0; i < 8; ++i) key[i] = src.gseed[i];
    }
    const size_t row = second ? bx - blocks1 : bx;  // one row per workgroup, every row valid
    const int tid = threadIdx.x;
    for (int c = tid; c < M; c += C::THREADS) {
        const zq x = row_value(src, key, row, c, LOG_M);
        if (src.dst) store_zq(src.dst + row * M + c, x);
        if (src.dst2) store_zq(src.dst2 + row * M + c, x);
#pragma unroll
        for (int q = 0; q < NPRIMES; ++q) sh[q * C::SL + C::slot(c)] = zq_mod_p_centred(x, c_pc[q]);
    }
    __syncthreads();
    const int q = tid >> C::LOG_TP, l = tid & (TP - 1);
    uint32_t* S = sh + q * C::SL;
    const uint32_t p = c_pc[q].p;
    const uint32_t* zt = tw.zeta_m + q * TW_SIZE;
    const uint32_t* ztp = tw.zeta_mp + q * TW_SIZE;
    uint32_t a[E];
    // pass 1: registers i hold x = l + TP i; layer len = TP s pairs i, i + s and uses
    // zeta[E / (2 s) + i / (2 s)]
#pragma unroll
    for (int i = 0; i < E; ++i) a[i] = S[C::slot(l + TP * i)];
#pragma unroll
    for (int s = E / 2; s >= 1; s >>= 1) {
#pragma unroll
        for (int b = 0; b < E / (2 * s); ++b) {
            const uint32_t w = zt[E / (2 * s) + b], wp = ztp[E / (2 * s) + b];
#pragma unroll
            for (int k = 0; k < s; ++k) ct_butterfly(a[2 * s * b + k], a[2 * s * b + k + s], w, wp, p);
        }
    }
#pragma unroll
    for (int i = 0; i < E; ++i) S[C::slot(l + TP * i)] = a[i];
    __syncthreads();
    // pass 2: registers e hold x = b + G e + TP h; layer len = G s pairs e, e + s and uses
    // zeta[m / (2 G s) + (TP / 2) h / (G s) + e / (2 s)]
    const int h = C::p2_h(l), b0 = C::p2_b(l);
#pragma unroll
    for (int e = 0; e < E; ++e) a[e] = S[C::slot(b0 + G * e + TP * h)];
#pragma unroll
    for (int s = E / 2; s >= 1; s >>= 1) {
        const int base = M / (2 * G * s) + (TP / 2 / (G * s)) * h;
#pragma unroll
        for (int b = 0; b < E / (2 * s); ++b) {
            const uint32_t w = zt[base + b], wp = ztp[base + b];
#pragma unroll
            for (int k = 0; k < s; ++k) ct_butterfly(a[2 * s * b + k], a[2 * s * b + k + s], w, wp, p);
        }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) S[C::slot(b0 + G * e + TP * h)] = a[e];
    __syncthreads();
    // pass 3: registers e hold x = E l + e; layer len = t pairs e, e + t and uses
    // zeta[m / (2 t) + E l / (2 t) + e / (2 t)]
#pragma unroll
    for (int e = 0; e < E; ++e) a[e] = S[C::slot(E * l + e)];
#pragma unroll
    for (int t = G / 2; t >= 1; t >>= 1) {
        const int base = M / (2 * t) + (E / (2 * t)) * l;
#pragma unroll
        for (int b = 0; b < E / (2 * t); ++b) {
            const uint32_t w = zt[base + b], wp = ztp[base + b];
#pragma unroll
            for (int k = 0; k < t; ++k) ct_butterfly(a[2 * t * b + k], a[2 * t * b + k + t], w, wp, p);
        }
    }
    uint4* o = reinterpret_cast<uint4*>(out + (row * NPRIMES + q) * M + E * l);
#pragma unroll
    for (int e = 0; e < E; e += 4) o[e / 4] = make_uint4(a[e], a[e + 1], a[e + 2], a[e + 3]);
}

template <int LOG_M, int TP = 64>
static void launch_reg3(const FwdSeg& s1, const FwdSeg& s2, const NttTables& tw, hipStream_t s, const Jobs& jb,
                        const InvTask& it) {
    using C = Reg3<LOG_M, TP>;
    if (C::LDS_BYTES > 65536)  // past the default dynamic LDS limit
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_ntt_fwd_reg3<LOG_M, TP>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::LDS_BYTES);
    hipLaunchKernelGGL((k_ntt_fwd_reg3<LOG_M, TP>), dim3((unsigned)(s1.rows + s2.rows + (it.chal ? 1 : 0)), jb.n),
                       dim3(C::THREADS), C::LDS_BYTES, s, s1, s2, (uint32_t)s1.rows, tw, jb.stride, it);
}

// DAMASCUS_NTT_KERNEL=layer keeps every degree on the per-layer kernel (for comparisons)
static bool reg_ntt_enabled() {
    static const bool on = [] {
        const char* e = getenv("DAMASCUS_NTT_KERNEL");
        return !(e && strcmp(e, "layer") == 0);
    }();
    return on;
}

// Variants per degree: variant 0 and 1 are register blockings (E elements x TP threads per prime
// slice; TP <= E, TP <= 32, E * TP = m; from 2^8 on one of them is the three-pass kernel), the
// last is the per-layer kernel. Every variant computes the same residues; the context picks the fastest
// per degree by timing.
int fwd_variant_count(int log_m) {
    if (!reg_ntt_enabled()) return 1;
    return log_m >= 6 && log_m <= 11 ? 3 : 1;
}

void launch_ntt_fwd_rows2(const RowSrc& src1, size_t rows1, uint32_t* out1, const RowSrc& src2, size_t rows2,
                          uint32_t* out2, int log_m, const NttTables& tw, hipStream_t s, const Jobs& jb,
                          const InvTask& it, int variant) {
    if (rows1 + rows2 == 0 || jb.n == 0) return;
    const FwdSeg s1{src1, out1, (uint32_t)rows1}, s2{src2, out2, (uint32_t)rows2};
    const int nv = fwd_variant_count(log_m);
    const int v = variant >= 0 && variant < nv ? variant : 0;
    if (v < nv - 1) {  // a register blocking (nv > 1 only for degrees 2^6..2^11 with them enabled)
        switch (log_m * 2 + v) {
            case 12: return launch_reg<8, 8>(s1, s2, tw, s, jb, it);
            case 13: return launch_reg<16, 4>(s1, s2, tw, s, jb, it);
            case 14: return launch_reg<16, 8>(s1, s2, tw, s, jb, it);
            case 15: return launch_reg<32, 4>(s1, s2, tw, s, jb, it);
            case 16: return launch_reg<16, 16>(s1, s2, tw, s, jb, it);
            case 17: return launch_reg3<8, 32>(s1, s2, tw, s, jb, it);
            case 18: return launch_reg<32, 16>(s1, s2, tw, s, jb, it);
            case 19: return launch_reg3<9>(s1, s2, tw, s, jb, it);
            case 20: return launch_reg<32, 32>(s1, s2, tw, s, jb, it);
            case 21: return launch_reg3<10>(s1, s2, tw, s, jb, it);
            case 22: return launch_reg3<11, 128>(s1, s2, tw, s, jb, it);
            case 23: return launch_reg3<11>(s1, s2, tw, s, jb, it);
            default: break;
        }
    }
    const int m = 1 << log_m;
    const int pg = std::min(NPRIMES, LDS_WORDS >> log_m);
    const size_t lds = (size_t)pg * m * sizeof(uint32_t);
    hipLaunchKernelGGL(k_ntt_fwd_rows, dim3((unsigned)(rows1 + rows2 + (it.chal ? 1 : 0)), jb.n), dim3(NTT_THREADS),
                       lds, s, s1, s2, (uint32_t)rows1, log_m, tw, jb.stride, it);
}

void launch_ntt_fwd_rows(const RowSrc& src, size_t rows, int log_m, uint32_t* out, const NttTables& tw,
                         hipStream_t s, const Jobs& jb, int variant) {
    launch_ntt_fwd_rows2(src, rows, out, RowSrc{}, 0, nullptr, log_m, tw, s, jb, InvTask{}, variant);
}

// ---------------------------------------------------------------------------------------
// inverse (Gentleman-Sande): (x, y) -> (x + y, (x - y) * zinv), then * m^-1
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(NTT_THREADS) k_ntt_inv(uint32_t* __restrict__ data, int log_m,
                                                         NttTables tw, size_t js) {
    extern __shared__ uint32_t sh[];
    data = jp(data, job_off(js));
    const int m = 1 << log_m;
    const size_t slice = blockIdx.x;  // (slice, prime) pairs: blockIdx.x = slice*8 + p
    const int pi = (int)(slice % NPRIMES);
    const uint32_t p = c_pc[pi].p;
    uint32_t* d = data + slice * m;
    for (int c = threadIdx.x; c < m; c += NTT_THREADS) sh[c] = d[c];
    __syncthreads();
    const int half = m >> 1;
    for (int len = 1; len < m; len <<= 1) {
        const int lg = __builtin_ctz(len);
        for (int b = threadIdx.x; b < half; b += NTT_THREADS) {
            const int blk = b >> lg;
            const int j = (blk << (lg + 1)) | (b & (len - 1));
            const int kz = (half >> lg) + blk;
            const uint32_t u = sh[j], v = sh[j + len];
            sh[j] = addp(u, v, p);
            sh[j + len] = mulc(subp(u, v, p), tw.izeta_m[pi * TW_SIZE + kz],
                               tw.izeta_mp[pi * TW_SIZE + kz], p);
        }
        __syncthreads();
    }
    const uint32_t im = tw.invn_m[pi * (MAX_LOG_N + 1) + log_m];
    const uint32_t imp = tw.invn_mp[pi * (MAX_LOG_N + 1) + log_m];
    for (int c = threadIdx.x; c < m; c += NTT_THREADS) d[c] = mulc(sh[c], im, imp, p);
}

void launch_ntt_inv(uint32_t* data, size_t slices, int log_m, const NttTables& tw, hipStream_t s,
                    const Jobs& jb) {
    if (slices == 0 || jb.n == 0) return;
    const size_t lds = ((size_t)1 << log_m) * sizeof(uint32_t);
    hipLaunchKernelGGL(k_ntt_inv, dim3((unsigned)(slices * NPRIMES), jb.n), dim3(NTT_THREADS), lds, s, data,
                       log_m, tw, jb.stride);
}

// ---------------------------------------------------------------------------------------
// NTT-domain multiply-accumulate: sums of residue products kept exact in 96 bits
// ---------------------------------------------------------------------------------------
struct Acc96 {
    uint64_t lo;
    uint32_t hi;
    __device__ __forceinline__ void zero() {
        lo = 0;
        hi = 0;
    }
    __device__ __forceinline__ void mac(uint32_t a, uint32_t b) {
        const uint64_t pr = (uint64_t)a * b;
        lo += pr;
        hi += (lo < pr) ? 1u : 0u;
    }
};

constexpr int MAC_THREADS = 256;

// thread = (k, p, c) with c fastest
__global__ void __launch_bounds__(MAC_THREADS) k_mac_pairs(const uint32_t* __restrict__ What,
                                                           const uint32_t* __restrict__ Ghat,
                                                           uint32_t rows, int log_m,
                                                           uint32_t* __restrict__ out_inv,
                                                           uint32_t* __restrict__ out_L,
                                                           uint32_t* __restrict__ out_R, size_t js, size_t gjs) {
    const size_t m = (size_t)1 << log_m;
    const size_t gid = (size_t)blockIdx.x * MAC_THREADS + threadIdx.x;
    if (gid >= (size_t)K * NPRIMES * m) return;
    const size_t jo = job_off(js);
    What = jp(What, jo);
    Ghat = jp(Ghat, (size_t)blockIdx.y * gjs);
    out_inv = jp(out_inv, jo);
    out_L = jp(out_L, jo);
    out_R = jp(out_R, jo);
    const uint32_t c = (uint32_t)(gid & (m - 1));
    const uint32_t p = (uint32_t)((gid >> log_m) % NPRIMES);
    const uint32_t k = (uint32_t)((gid >> log_m) / NPRIMES);
    Acc96 inv, L, R;
    inv.zero();
    L.zero();
    R.zero();
    const size_t wstride = NPRIMES * m;             // per witness row
    const size_t gstride = (size_t)K * NPRIMES * m;  // per generator row i
    const uint32_t* wp = What + p * m + c;
    const uint32_t* gp = Ghat + ((size_t)k * NPRIMES + p) * m + c;
    const bool cross = out_L != nullptr && rows > 1;  // rows is even when cross terms are asked
    if (!cross) {
        for (uint32_t i = 0; i < rows; ++i) inv.mac(wp[i * wstride], gp[i * gstride]);
    } else {
        const uint32_t h = rows >> 1;
        for (uint32_t i = 0; i < h; ++i) {
            const uint32_t w0 = wp[i * wstride], w1 = wp[(i + h) * wstride];
            const uint32_t g0 = gp[i * gstride], g1 = gp[(i + h) * gstride];
            inv.mac(w0, g0);
            inv.mac(w1, g1);
            L.mac(g1, w0);
            R.mac(g0, w1);
        }
    }
    const PrimeConsts pc = c_pc[p];
    const size_t o = ((size_t)k * NPRIMES + p) * m + c;
    out_inv[o] = reduce_u96(inv.lo, inv.hi, pc);
    if (cross) {
        out_L[o] = reduce_u96(L.lo, L.hi, pc);
        out_R[o] = reduce_u96(R.lo, R.hi, pc);
    }
}


// thread = (k, p, t), t < m/2, over the position pair (2t, 2t+1).
// With F = NTT_m(f): F[2t] = f(r_t), F[2t+1] = f(-r_t), r_t = zeta[m/2 + t]; the degree-m/2
// transforms of the even/odd parts are (F[2t]+F[2t+1])/2 and (F[2t]-F[2t+1])/(2 r_t), and
// multiplying by Y = X^2 is multiplying by r_t^2. Hence
//   L_poly[t] = r_t/4     * sum_i (W0+W1)(G0-G1)     (even(m~) against X*odd(g~))
//   R_poly[t] = 1/(4 r_t) * sum_i (W0-W1)(G0+G1)     (odd(m~)  against even(g~))
// which four exact sums S00, S11, S01, S10 provide together with the invariant (S00, S11).
__global__ void __launch_bounds__(MAC_THREADS) k_mac_split(const uint32_t* __restrict__ What,
                                                           const uint32_t* __restrict__ Ghat,
                                                           uint32_t rows, int log_m,
                                                           uint32_t* __restrict__ out_inv,
                                                           uint32_t* __restrict__ out_L,
                                                           uint32_t* __restrict__ out_R,
                                                           NttTables tw, size_t js, size_t gjs) {
    const size_t m = (size_t)1 << log_m;
    const size_t h = m >> 1;
    const size_t gid = (size_t)blockIdx.x * MAC_THREADS + threadIdx.x;
    if (gid >= (size_t)K * NPRIMES * h) return;
    const size_t jo = job_off(js);
    What = jp(What, jo);
    Ghat = jp(Ghat, (size_t)blockIdx.y * gjs);
    out_inv = jp(out_inv, jo);
    out_L = jp(out_L, jo);
    out_R = jp(out_R, jo);
    const uint32_t t = (uint32_t)(gid & (h - 1));
    const uint32_t p = (uint32_t)((gid >> (log_m - 1)) % NPRIMES);
    const uint32_t k = (uint32_t)((gid >> (log_m - 1)) / NPRIMES);
    Acc96 s00, s11, s01, s10;
    s00.zero();
    s11.zero();
    s01.zero();
    s10.zero();
    const size_t wstride = NPRIMES * m;
    const size_t gstride = (size_t)K * NPRIMES * m;
    const uint2* wp = reinterpret_cast<const uint2*>(What + p * m) + t;
    const uint2* gp = reinterpret_cast<const uint2*>(Ghat + ((size_t)k * NPRIMES + p) * m) + t;
    for (uint32_t i = 0; i < rows; ++i) {
        const uint2 w = wp[i * (wstride / 2)];
        const uint2 g = gp[i * (gstride / 2)];
        s00.mac(w.x, g.x);
        s11.mac(w.y, g.y);
        s01.mac(w.x, g.y);
        s10.mac(w.y, g.x);
    }
    const PrimeConsts pc = c_pc[p];
    const uint32_t P = pc.p;
    const uint32_t a00 = reduce_u96(s00.lo, s00.hi, pc), a11 = reduce_u96(s11.lo, s11.hi, pc);
    const uint32_t a01 = reduce_u96(s01.lo, s01.hi, pc), a10 = reduce_u96(s10.lo, s10.hi, pc);
    const size_t o = ((size_t)k * NPRIMES + p) * m + 2 * t;
    out_inv[o] = a00;
    out_inv[o + 1] = a11;
    const uint32_t d = subp(a00, a11, P);
    const uint32_t Lp = addp(subp(d, a01, P), a10, P);
    const uint32_t Rp = subp(addp(d, a01, P), a10, P);
    // * r_t and * r_t^-1, then * 1/4 (inv4 = invn for log 2)
    const uint32_t kz = (uint32_t)h + t;
    const uint32_t i4m = tw.invn_m[p * (MAX_LOG_N + 1) + 2], i4mp = tw.invn_mp[p * (MAX_LOG_N + 1) + 2];
    const uint32_t lv = mulc(mulc(Lp, tw.zeta_m[p * TW_SIZE + kz], tw.zeta_mp[p * TW_SIZE + kz], P), i4m, i4mp, P);
    const uint32_t rv = mulc(mulc(Rp, tw.izeta_m[p * TW_SIZE + kz], tw.izeta_mp[p * TW_SIZE + kz], P), i4m, i4mp, P);
    const size_t oh = ((size_t)k * NPRIMES + p) * h + t;
    out_L[oh] = lv;
    out_R[oh] = rv;
}


// Row segments per MAC pass: doubled while each segment keeps at least 2^log_min_rows rows and
// the pass stays under 2^log_max_threads threads (at most 32 segments). The default shape (16
// rows, 2^18 threads) keeps HBM busy at every degree; the context can tune it per degree.
static inline uint32_t mac_segments(size_t rows, int log_m, MacShape sh) {
    const size_t quads = ((size_t)K * NPRIMES << log_m) / 4;  // threads per segment
    size_t seg = 1;
    while (seg < 32 && rows / (2 * seg) >= ((size_t)1 << sh.log_min_rows) &&
           quads * seg < (size_t)1 << sh.log_max_threads)
        seg *= 2;
    return (uint32_t)seg;
}

__device__ __forceinline__ uint4 ld4(const uint32_t* p) { return *reinterpret_cast<const uint4*>(p); }

// Pass over generator rows for the post-poly invariant and the next round's vector cross
// terms: thread = (segment, k, p, 4 consecutive positions). Partial sums, reduced mod p, go
// to part[seg][slot][k][p][c] (slot 0 = inv, 1 = L, 2 = R).
__global__ void __launch_bounds__(MAC_THREADS) k_mac_pairs4(const uint32_t* __restrict__ What,
                                                            const uint32_t* __restrict__ Ghat, uint32_t rows,
                                                            int log_m, uint32_t nseg, uint32_t seg_rows, int cross,
                                                            uint32_t* __restrict__ part, size_t js, size_t gjs) {
    const size_t m = (size_t)1 << log_m;
    const size_t per_seg = (size_t)K * NPRIMES * m / 4;
    const size_t gid = (size_t)blockIdx.x * MAC_THREADS + threadIdx.x;
    if (gid >= (size_t)nseg * per_seg) return;
    const size_t jo = job_off(js);
    What = jp(What, jo);
    Ghat = jp(Ghat, (size_t)blockIdx.y * gjs);
    part = jp(part, jo);
    const uint32_t seg = (uint32_t)(gid / per_seg);
    const size_t r = gid - (size_t)seg * per_seg;
    const uint32_t c = (uint32_t)((r * 4) & (m - 1));
    const uint32_t kp = (uint32_t)((r * 4) >> log_m);  // k * 8 + p
    const uint32_t p = kp % NPRIMES;
    const size_t wstride = NPRIMES * m, gstride = (size_t)K * NPRIMES * m;
    const uint32_t* wp = What + p * m + c;
    const uint32_t* gp = Ghat + (size_t)kp * m + c;
    Acc96 inv[4], L[4], R[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        inv[e].zero();
        L[e].zero();
        R[e].zero();
    }
    if (!cross) {
        const uint32_t i0 = seg * seg_rows, i1 = min(rows, i0 + seg_rows);
#pragma unroll 4
        for (uint32_t i = i0; i < i1; ++i) {
            const uint4 w = ld4(wp + i * wstride), g = ld4(gp + i * gstride);
            inv[0].mac(w.x, g.x);
            inv[1].mac(w.y, g.y);
            inv[2].mac(w.z, g.z);
            inv[3].mac(w.w, g.w);
        }
    } else {
        const uint32_t h = rows >> 1;
        const uint32_t i0 = seg * seg_rows, i1 = min(h, i0 + seg_rows);
        auto pair_mac = [&](const uint4& w0, const uint4& w1, const uint4& g0, const uint4& g1) {
            const uint32_t W0[4] = {w0.x, w0.y, w0.z, w0.w}, W1[4] = {w1.x, w1.y, w1.z, w1.w};
            const uint32_t G0[4] = {g0.x, g0.y, g0.z, g0.w}, G1[4] = {g1.x, g1.y, g1.z, g1.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                inv[e].mac(W0[e], G0[e]);
                inv[e].mac(W1[e], G1[e]);
                L[e].mac(G1[e], W0[e]);
                R[e].mac(G0[e], W1[e]);
            }
        };
        // two rows' eight loads are issued before either row's multiply-adds (the scheduler
        // barrier keeps them there), so a wave has twice the bytes in flight
        uint32_t i = i0;
        for (; i + 1 < i1; i += 2) {
            const uint4 a0 = ld4(wp + i * wstride), a1 = ld4(wp + (i + h) * wstride);
            const uint4 b0 = ld4(gp + i * gstride), b1 = ld4(gp + (i + h) * gstride);
            const uint4 c0 = ld4(wp + (i + 1) * wstride), c1 = ld4(wp + (i + 1 + h) * wstride);
            const uint4 d0 = ld4(gp + (i + 1) * gstride), d1 = ld4(gp + (i + 1 + h) * gstride);
            __builtin_amdgcn_sched_barrier(0);
            pair_mac(a0, a1, b0, b1);
            pair_mac(c0, c1, d0, d1);
        }
        if (i < i1)
            pair_mac(ld4(wp + i * wstride), ld4(wp + (i + h) * wstride), ld4(gp + i * gstride),
                     ld4(gp + (i + h) * gstride));
    }
    const PrimeConsts pc = c_pc[p];
    const size_t slot = (size_t)K * NPRIMES * m;
    uint32_t* o = part + (size_t)seg * 3 * slot + (size_t)kp * m + c;
    uint4 v = make_uint4(reduce_u96(inv[0].lo, inv[0].hi, pc), reduce_u96(inv[1].lo, inv[1].hi, pc),
                         reduce_u96(inv[2].lo, inv[2].hi, pc), reduce_u96(inv[3].lo, inv[3].hi, pc));
    *reinterpret_cast<uint4*>(o) = v;
    if (cross) {
        v = make_uint4(reduce_u96(L[0].lo, L[0].hi, pc), reduce_u96(L[1].lo, L[1].hi, pc),
                       reduce_u96(L[2].lo, L[2].hi, pc), reduce_u96(L[3].lo, L[3].hi, pc));
        *reinterpret_cast<uint4*>(o + slot) = v;
        v = make_uint4(reduce_u96(R[0].lo, R[0].hi, pc), reduce_u96(R[1].lo, R[1].hi, pc),
                       reduce_u96(R[2].lo, R[2].hi, pc), reduce_u96(R[3].lo, R[3].hi, pc));
        *reinterpret_cast<uint4*>(o + 2 * slot) = v;
    }
}

// sum the segment partials: out[slot][k][p][c] for nslots slots
__global__ void k_mac_pairs_reduce(const uint32_t* __restrict__ part, uint32_t nseg, int log_m, int nslots,
                                   uint32_t* __restrict__ out_inv, uint32_t* __restrict__ out_L,
                                   uint32_t* __restrict__ out_R, size_t js) {
    const size_t m = (size_t)1 << log_m;
    const size_t slot = (size_t)K * NPRIMES * m;
    const size_t gid = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= slot * nslots) return;
    const size_t jo = job_off(js);
    part = jp(part, jo);
    out_inv = jp(out_inv, jo);
    out_L = jp(out_L, jo);
    out_R = jp(out_R, jo);
    const size_t sl = gid / slot, e = gid - sl * slot;
    const PrimeConsts pc = c_pc[(e >> log_m) % NPRIMES];
    uint64_t acc = 0;
    for (uint32_t sg = 0; sg < nseg; ++sg) acc += part[(size_t)sg * 3 * slot + sl * slot + e];
    uint32_t* o = sl == 0 ? out_inv : (sl == 1 ? out_L : out_R);
    o[e] = reduce_u96(acc, 0, pc);
}

void launch_mac_pairs(const uint32_t* What, const uint32_t* Ghat, size_t rows, int log_m, uint32_t* out_inv,
                      uint32_t* out_L, uint32_t* out_R, uint32_t* part, hipStream_t s, const Jobs& jb, bool g_shared,
                      MacShape sh) {
    if (jb.n == 0) return;
    const size_t gjs = g_shared ? 0 : jb.stride;
    const bool cross = out_L != nullptr && rows > 1;
    const size_t half = cross ? rows / 2 : rows;
    const uint32_t nseg = log_m >= 2 ? mac_segments(half, log_m, sh) : 1;
    const uint32_t seg_rows = (uint32_t)((half + nseg - 1) / nseg);
    const size_t m = (size_t)1 << log_m;
    if (log_m < 2) {  // degrees 1 and 2: the scalar kernel
        const size_t n = (size_t)K * NPRIMES << log_m;
        hipLaunchKernelGGL(k_mac_pairs, dim3((unsigned)((n + MAC_THREADS - 1) / MAC_THREADS), jb.n),
                           dim3(MAC_THREADS), 0, s, What, Ghat, (uint32_t)rows, log_m, out_inv, out_L, out_R,
                           jb.stride, gjs);
        return;
    }
    const size_t threads = (size_t)nseg * K * NPRIMES * m / 4;
    hipLaunchKernelGGL(k_mac_pairs4, dim3((unsigned)((threads + MAC_THREADS - 1) / MAC_THREADS), jb.n),
                       dim3(MAC_THREADS), 0, s, What, Ghat, (uint32_t)rows, log_m, nseg, seg_rows, (int)cross, part,
                       jb.stride, gjs);
    const int nslots = cross ? 3 : 1;
    const size_t n = (size_t)K * NPRIMES * m * nslots;
    hipLaunchKernelGGL(k_mac_pairs_reduce, dim3((unsigned)((n + 255) / 256), jb.n), dim3(256), 0, s, part, nseg,
                       log_m, nslots, out_inv, out_L, out_R, jb.stride);
}

// Vector-stage invariant (degree m) and poly-stage cross terms (degree m/2) in one pass;
// thread = (segment, k, p, two position pairs). Partial sums S00, S11, S01, S10 per pair go to
// part[seg][4][k][p][t].
__global__ void __launch_bounds__(MAC_THREADS) k_mac_split4(const uint32_t* __restrict__ What,
                                                            const uint32_t* __restrict__ Ghat, uint32_t rows,
                                                            int log_m, uint32_t nseg, uint32_t seg_rows,
                                                            uint32_t* __restrict__ part, size_t js, size_t gjs) {
    const size_t m = (size_t)1 << log_m, h = m >> 1;
    const size_t per_seg = (size_t)K * NPRIMES * m / 4;
    const size_t gid = (size_t)blockIdx.x * MAC_THREADS + threadIdx.x;
    if (gid >= (size_t)nseg * per_seg) return;
    const size_t jo = job_off(js);
    What = jp(What, jo);
    Ghat = jp(Ghat, (size_t)blockIdx.y * gjs);
    part = jp(part, jo);
    const uint32_t seg = (uint32_t)(gid / per_seg);
    const size_t r = gid - (size_t)seg * per_seg;
    const uint32_t c = (uint32_t)((r * 4) & (m - 1));  // positions c..c+3 = pairs t, t+1
    const uint32_t kp = (uint32_t)((r * 4) >> log_m);
    const uint32_t p = kp % NPRIMES;
    const size_t wstride = NPRIMES * m, gstride = (size_t)K * NPRIMES * m;
    const uint32_t* wp = What + p * m + c;
    const uint32_t* gp = Ghat + (size_t)kp * m + c;
    Acc96 s00[2], s11[2], s01[2], s10[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        s00[e].zero();
        s11[e].zero();
        s01[e].zero();
        s10[e].zero();
    }
    const uint32_t i0 = seg * seg_rows, i1 = min(rows, i0 + seg_rows);
    auto row_mac = [&](const uint4& w, const uint4& g) {
        s00[0].mac(w.x, g.x);
        s11[0].mac(w.y, g.y);
        s01[0].mac(w.x, g.y);
        s10[0].mac(w.y, g.x);
        s00[1].mac(w.z, g.z);
        s11[1].mac(w.w, g.w);
        s01[1].mac(w.z, g.w);
        s10[1].mac(w.w, g.z);
    };
    // four rows' loads are issued before their multiply-adds (the scheduler barrier keeps them
    // there), so a wave has four rows in flight
    uint32_t i = i0;
    for (; i + 3 < i1; i += 4) {
        uint4 w[4], g[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            w[u] = ld4(wp + (i + u) * wstride);
            g[u] = ld4(gp + (i + u) * gstride);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 4; ++u) row_mac(w[u], g[u]);
    }
    for (; i < i1; ++i) row_mac(ld4(wp + i * wstride), ld4(gp + i * gstride));
    const PrimeConsts pc = c_pc[p];
    const size_t slot = (size_t)K * NPRIMES * h;
    uint32_t* o = part + (size_t)seg * 4 * slot + (size_t)kp * h + c / 2;
    *reinterpret_cast<uint2*>(o) = make_uint2(reduce_u96(s00[0].lo, s00[0].hi, pc), reduce_u96(s00[1].lo, s00[1].hi, pc));
    *reinterpret_cast<uint2*>(o + slot) = make_uint2(reduce_u96(s11[0].lo, s11[0].hi, pc), reduce_u96(s11[1].lo, s11[1].hi, pc));
    *reinterpret_cast<uint2*>(o + 2 * slot) = make_uint2(reduce_u96(s01[0].lo, s01[0].hi, pc), reduce_u96(s01[1].lo, s01[1].hi, pc));
    *reinterpret_cast<uint2*>(o + 3 * slot) = make_uint2(reduce_u96(s10[0].lo, s10[0].hi, pc), reduce_u96(s10[1].lo, s10[1].hi, pc));
}

__global__ void k_mac_split_reduce(const uint32_t* __restrict__ part, uint32_t nseg, int log_m,
                                   uint32_t* __restrict__ out_inv, uint32_t* __restrict__ out_L,
                                   uint32_t* __restrict__ out_R, NttTables tw, size_t js) {
    const size_t m = (size_t)1 << log_m, h = m >> 1;
    const size_t slot = (size_t)K * NPRIMES * h;
    const size_t gid = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= slot) return;
    const size_t jo = job_off(js);
    part = jp(part, jo);
    out_inv = jp(out_inv, jo);
    out_L = jp(out_L, jo);
    out_R = jp(out_R, jo);
    const uint32_t t = (uint32_t)(gid & (h - 1));
    const uint32_t kp = (uint32_t)(gid >> (log_m - 1));
    const uint32_t p = kp % NPRIMES;
    const PrimeConsts pc = c_pc[p];
    const uint32_t P = pc.p;
    uint64_t a[4] = {0, 0, 0, 0};
    for (uint32_t sg = 0; sg < nseg; ++sg)
        for (int j = 0; j < 4; ++j) a[j] += part[(size_t)sg * 4 * slot + j * slot + gid];
    const uint32_t a00 = reduce_u96(a[0], 0, pc), a11 = reduce_u96(a[1], 0, pc), a01 = reduce_u96(a[2], 0, pc),
                   a10 = reduce_u96(a[3], 0, pc);
    const size_t o = (size_t)kp * m + 2 * t;
    out_inv[o] = a00;
    out_inv[o + 1] = a11;
    const uint32_t d = subp(a00, a11, P);
    const uint32_t Lp = addp(subp(d, a01, P), a10, P);
    const uint32_t Rp = subp(addp(d, a01, P), a10, P);
    const uint32_t kz = (uint32_t)h + t;
    const uint32_t i4m = tw.invn_m[p * (MAX_LOG_N + 1) + 2], i4mp = tw.invn_mp[p * (MAX_LOG_N + 1) + 2];
    out_L[(size_t)kp * h + t] =
        mulc(mulc(Lp, tw.zeta_m[p * TW_SIZE + kz], tw.zeta_mp[p * TW_SIZE + kz], P), i4m, i4mp, P);
    out_R[(size_t)kp * h + t] =
        mulc(mulc(Rp, tw.izeta_m[p * TW_SIZE + kz], tw.izeta_mp[p * TW_SIZE + kz], P), i4m, i4mp, P);
}

void launch_mac_split(const uint32_t* What, const uint32_t* Ghat, size_t rows, int log_m, uint32_t* out_inv,
                      uint32_t* out_L, uint32_t* out_R, uint32_t* part, const NttTables& tw, hipStream_t s,
                      const Jobs& jb, MacShape sh) {
    if (jb.n == 0) return;
    const size_t gjs = jb.stride;
    if (log_m < 2) {  // degree 2: the scalar kernel
        const size_t n = (size_t)K * NPRIMES << (log_m - 1);
        hipLaunchKernelGGL(k_mac_split, dim3((unsigned)((n + MAC_THREADS - 1) / MAC_THREADS), jb.n),
                           dim3(MAC_THREADS), 0, s, What, Ghat, (uint32_t)rows, log_m, out_inv, out_L, out_R, tw,
                           jb.stride, gjs);
        return;
    }
    const uint32_t nseg = mac_segments(rows * 2, log_m, sh);
    const uint32_t seg_rows = (uint32_t)((rows + nseg - 1) / nseg);
    const size_t m = (size_t)1 << log_m;
    const size_t threads = (size_t)nseg * K * NPRIMES * m / 4;
    hipLaunchKernelGGL(k_mac_split4, dim3((unsigned)((threads + MAC_THREADS - 1) / MAC_THREADS), jb.n),
                       dim3(MAC_THREADS), 0, s, What, Ghat, (uint32_t)rows, log_m, nseg, seg_rows, part, jb.stride,
                       gjs);
    const size_t n = (size_t)K * NPRIMES * (m / 2);
    hipLaunchKernelGGL(k_mac_split_reduce, dim3((unsigned)((n + 255) / 256), jb.n), dim3(256), 0, s, part, nseg,
                       log_m, out_inv, out_L, out_R, tw, jb.stride);
}

// ---------------------------------------------------------------------------------------
// MAC passes with their reductions and inverse transforms fused: the segment partials of one
// (k, p) slice are summed into LDS, transformed back there and scaled by 1/m. One workgroup
// per (k, p); degrees 1 and 2 keep the scalar kernels and separate inverse launches. With only
// 64 workgroups per job, each one has 16 waves with their segment loads in flight together, so
// the reads of the partials are not bound by one wave's memory latency.
// ---------------------------------------------------------------------------------------
constexpr int FIN_THREADS = 1024;

// Gentleman-Sande inverse layers on `cnt` adjacent slices of 2^log_n words in LDS (no scaling)
__device__ __forceinline__ void lds_intt_slices(uint32_t* s, int cnt, int log_n, int pi, const NttTables& tw) {
    const int n = 1 << log_n, half = n >> 1;
    const uint32_t p = c_pc[pi].p;
    for (int len = 1; len < n; len <<= 1) {
        const int lg = __builtin_ctz(len);
        for (int t = threadIdx.x; t < cnt * half; t += FIN_THREADS) {
            const int sl = t >> (log_n - 1), b = t & (half - 1);
            const int blk = b >> lg;
            const int j = (blk << (lg + 1)) | (b & (len - 1));
            const int kz = (half >> lg) + blk;
            uint32_t* sq = s + sl * n;
            const uint32_t u = sq[j], v = sq[j + len];
            sq[j] = addp(u, v, p);
            sq[j + len] = mulc(subp(u, v, p), tw.izeta_m[pi * TW_SIZE + kz], tw.izeta_mp[pi * TW_SIZE + kz], p);
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(FIN_THREADS) k_mac_pairs_finish(const uint32_t* __restrict__ part, uint32_t nseg,
                                                                  int log_m, int nslots, uint32_t* __restrict__ out_inv,
                                                                  uint32_t* __restrict__ out_L,
                                                                  uint32_t* __restrict__ out_R, NttTables tw,
                                                                  size_t js) {
    extern __shared__ uint32_t sh[];
    const size_t jo = job_off(js);
    part = jp(part, jo);
    out_inv = jp(out_inv, jo);
    out_L = jp(out_L, jo);
    out_R = jp(out_R, jo);
    const int m = 1 << log_m;
    const uint32_t kp = blockIdx.x, p = kp % NPRIMES;
    const PrimeConsts pc = c_pc[p];
    const size_t slot = (size_t)K * NPRIMES * m;
    if ((m & 3) == 0) {  // four positions per thread, one 16-byte load per segment
        const int qm = m >> 2;
        for (int e = threadIdx.x; e < nslots * qm; e += FIN_THREADS) {
            const int sl = e >> (log_m - 2), c = (e & (qm - 1)) << 2;
            const uint32_t* b = part + sl * slot + (size_t)kp * m + c;
            uint64_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
#pragma unroll 8
            for (uint32_t sg = 0; sg < nseg; ++sg) {
                const uint4 v = ld4(b + (size_t)sg * 3 * slot);
                a0 += v.x;
                a1 += v.y;
                a2 += v.z;
                a3 += v.w;
            }
            uint32_t* o = sh + sl * m + c;
            o[0] = reduce_u96(a0, 0, pc);
            o[1] = reduce_u96(a1, 0, pc);
            o[2] = reduce_u96(a2, 0, pc);
            o[3] = reduce_u96(a3, 0, pc);
        }
    } else {
        for (int e = threadIdx.x; e < nslots * m; e += FIN_THREADS) {
            const int sl = e >> log_m, c = e & (m - 1);
            uint64_t acc = 0;
#pragma unroll 8
            for (uint32_t sg = 0; sg < nseg; ++sg) acc += part[(size_t)sg * 3 * slot + sl * slot + (size_t)kp * m + c];
            sh[e] = reduce_u96(acc, 0, pc);
        }
    }
    __syncthreads();
    lds_intt_slices(sh, nslots, log_m, (int)p, tw);
    const uint32_t im = tw.invn_m[p * (MAX_LOG_N + 1) + log_m], imp = tw.invn_mp[p * (MAX_LOG_N + 1) + log_m];
    for (int e = threadIdx.x; e < nslots * m; e += FIN_THREADS) {
        const int sl = e >> log_m, c = e & (m - 1);
        uint32_t* o = sl == 0 ? out_inv : (sl == 1 ? out_L : out_R);
        o[(size_t)kp * m + c] = mulc(sh[e], im, imp, pc.p);
    }
}

// split pass: inv (degree m) from (S00, S11); L, R (degree m/2) from the four sums, as in
// k_mac_split_reduce; then the three inverse transforms. out_lr holds L then R ([k][p][m/2] each).
__global__ void __launch_bounds__(FIN_THREADS) k_mac_split_finish(const uint32_t* __restrict__ part, uint32_t nseg,
                                                                  int log_m, uint32_t* __restrict__ out_inv,
                                                                  uint32_t* __restrict__ out_lr, NttTables tw,
                                                                  size_t js) {
    extern __shared__ uint32_t sh[];
    const size_t jo = job_off(js);
    part = jp(part, jo);
    out_inv = jp(out_inv, jo);
    out_lr = jp(out_lr, jo);
    const int m = 1 << log_m, h = m >> 1;
    const uint32_t kp = blockIdx.x, p = kp % NPRIMES;
    const PrimeConsts pc = c_pc[p];
    const uint32_t P = pc.p;
    const size_t slot = (size_t)K * NPRIMES * h;
    uint32_t* si = sh;
    uint32_t* sl = sh + m;  // L then R, adjacent
    uint32_t* sr = sl + h;
    uint32_t* sums = sh + 2 * m;  // S00, S11, S01, S10, h words each
    const uint32_t i4m = tw.invn_m[p * (MAX_LOG_N + 1) + 2], i4mp = tw.invn_mp[p * (MAX_LOG_N + 1) + 2];
    if ((h & 3) == 0) {  // four pairs of one sum per thread, one 16-byte load per segment
        const int qh = h >> 2;
        for (int e = threadIdx.x; e < 4 * qh; e += FIN_THREADS) {
            const int j = e >> (log_m - 3), t = (e & (qh - 1)) << 2;
            const uint32_t* b = part + j * slot + (size_t)kp * h + t;
            uint64_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
#pragma unroll 8
            for (uint32_t sg = 0; sg < nseg; ++sg) {
                const uint4 v = ld4(b + (size_t)sg * 4 * slot);
                a0 += v.x;
                a1 += v.y;
                a2 += v.z;
                a3 += v.w;
            }
            uint32_t* o = sums + j * h + t;
            o[0] = reduce_u96(a0, 0, pc);
            o[1] = reduce_u96(a1, 0, pc);
            o[2] = reduce_u96(a2, 0, pc);
            o[3] = reduce_u96(a3, 0, pc);
        }
    } else {
        for (int e = threadIdx.x; e < 4 * h; e += FIN_THREADS) {  // one (sum, pair) per thread
            const int j = e >> (log_m - 1), t = e & (h - 1);
            uint64_t a = 0;
#pragma unroll 8
            for (uint32_t sg = 0; sg < nseg; ++sg) a += part[(size_t)sg * 4 * slot + j * slot + (size_t)kp * h + t];
            sums[e] = reduce_u96(a, 0, pc);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < h; t += FIN_THREADS) {
        const uint32_t a00 = sums[t], a11 = sums[h + t], a01 = sums[2 * h + t], a10 = sums[3 * h + t];
        si[2 * t] = a00;
        si[2 * t + 1] = a11;
        const uint32_t d = subp(a00, a11, P);
        const uint32_t Lp = addp(subp(d, a01, P), a10, P);
        const uint32_t Rp = subp(addp(d, a01, P), a10, P);
        const uint32_t kz = (uint32_t)h + t;
        sl[t] = mulc(mulc(Lp, tw.zeta_m[p * TW_SIZE + kz], tw.zeta_mp[p * TW_SIZE + kz], P), i4m, i4mp, P);
        sr[t] = mulc(mulc(Rp, tw.izeta_m[p * TW_SIZE + kz], tw.izeta_mp[p * TW_SIZE + kz], P), i4m, i4mp, P);
    }
    __syncthreads();
    lds_intt_slices(si, 1, log_m, (int)p, tw);
    lds_intt_slices(sl, 2, log_m - 1, (int)p, tw);
    const uint32_t im = tw.invn_m[p * (MAX_LOG_N + 1) + log_m], imp = tw.invn_mp[p * (MAX_LOG_N + 1) + log_m];
    const uint32_t ih = tw.invn_m[p * (MAX_LOG_N + 1) + log_m - 1], ihp = tw.invn_mp[p * (MAX_LOG_N + 1) + log_m - 1];
    for (int c = threadIdx.x; c < m; c += FIN_THREADS) out_inv[(size_t)kp * m + c] = mulc(si[c], im, imp, P);
    for (int t = threadIdx.x; t < h; t += FIN_THREADS) {
        out_lr[(size_t)kp * h + t] = mulc(sl[t], ih, ihp, P);
        out_lr[slot + (size_t)kp * h + t] = mulc(sr[t], ih, ihp, P);
    }
}

void launch_mac_pairs_inv(const uint32_t* What, const uint32_t* Ghat, size_t rows, int log_m, uint32_t* out_inv,
                          uint32_t* out_L, uint32_t* out_R, uint32_t* part, const NttTables& tw, hipStream_t s,
                          const Jobs& jb, bool g_shared, MacShape sh) {
    if (jb.n == 0) return;
    const bool cross = out_L != nullptr && rows > 1;
    if (log_m < 2) {
        launch_mac_pairs(What, Ghat, rows, log_m, out_inv, out_L, out_R, part, s, jb, g_shared, sh);
        launch_ntt_inv(out_inv, K, log_m, tw, s, jb);
        if (cross) {
            launch_ntt_inv(out_L, K, log_m, tw, s, jb);
            launch_ntt_inv(out_R, K, log_m, tw, s, jb);
        }
        return;
    }
    const size_t gjs = g_shared ? 0 : jb.stride;
    const size_t half = cross ? rows / 2 : rows;
    const uint32_t nseg = mac_segments(half, log_m, sh);
    const uint32_t seg_rows = (uint32_t)((half + nseg - 1) / nseg);
    const size_t m = (size_t)1 << log_m;
    const size_t threads = (size_t)nseg * K * NPRIMES * m / 4;
    hipLaunchKernelGGL(k_mac_pairs4, dim3((unsigned)((threads + MAC_THREADS - 1) / MAC_THREADS), jb.n),
                       dim3(MAC_THREADS), 0, s, What, Ghat, (uint32_t)rows, log_m, nseg, seg_rows, (int)cross, part,
                       jb.stride, gjs);
    const int nslots = cross ? 3 : 1;
    hipLaunchKernelGGL(k_mac_pairs_finish, dim3(K * NPRIMES, jb.n), dim3(FIN_THREADS), (size_t)nslots * m * 4, s,
                       part, nseg, log_m, nslots, out_inv, out_L, out_R, tw, jb.stride);
}

void launch_mac_split_inv(const uint32_t* What, const uint32_t* Ghat, size_t rows, int log_m, uint32_t* out_inv,
                          uint32_t* out_lr, uint32_t* part, const NttTables& tw, hipStream_t s, const Jobs& jb,
                          MacShape sh) {
    if (jb.n == 0) return;
    const size_t m = (size_t)1 << log_m;
    if (log_m < 2) {
        launch_mac_split(What, Ghat, rows, log_m, out_inv, out_lr, out_lr + (size_t)K * NPRIMES * (m / 2), part, tw, s,
                         jb, sh);
        launch_ntt_inv(out_inv, K, log_m, tw, s, jb);
        launch_ntt_inv(out_lr, 2 * K, log_m - 1, tw, s, jb);
        return;
    }
    const uint32_t nseg = mac_segments(rows * 2, log_m, sh);
    const uint32_t seg_rows = (uint32_t)((rows + nseg - 1) / nseg);
    const size_t threads = (size_t)nseg * K * NPRIMES * m / 4;
    hipLaunchKernelGGL(k_mac_split4, dim3((unsigned)((threads + MAC_THREADS - 1) / MAC_THREADS), jb.n),
                       dim3(MAC_THREADS), 0, s, What, Ghat, (uint32_t)rows, log_m, nseg, seg_rows, part, jb.stride,
                       jb.stride);
    hipLaunchKernelGGL(k_mac_split_finish, dim3(K * NPRIMES, jb.n), dim3(FIN_THREADS), (size_t)4 * m * 4, s, part,
                       nseg, log_m, out_inv, out_lr, tw, jb.stride);
}

size_t mac_part_words(int max_log_m) {
    // a launch holds at most max(2^18, quads) threads of 4 positions, in at most 3 slots
    size_t w = 0;
    for (int lg = 2; lg <= max_log_m; ++lg)
        w = std::max(w, (size_t)mac_segments(SIZE_MAX / 2, lg, mac_shape_of(MAC_SHAPES - 1)) * 3 * K * NPRIMES << lg);
    return std::max(w, (size_t)1);
}

// ---------------------------------------------------------------------------------------
// centred CRT lift (ntt.rs:400-427): Garner mixed-radix digits v_i in CRT_PRIMES order,
// x = sum v_i * (p_0..p_{i-1}); if x > floor(P/2) the result is (x - P) mod q.
// ---------------------------------------------------------------------------------------
// residues r[i] mod CRT_PRIMES[i] -> the centred lift in Z_q
__device__ __forceinline__ zq crt_lift_residues(const uint32_t r[NPRIMES]) {
    uint32_t v[NPRIMES];
    v[0] = r[0];
#pragma unroll
    for (int i = 1; i < NPRIMES; ++i) {
        const uint32_t pi = c_pc[i].p;
        // acc = (v_0 + v_1 p_0 + ... + v_{i-1} p_0..p_{i-2}) mod p_i; mulc reduces any 32-bit v_j
        uint32_t acc = 0;
#pragma unroll
        for (int j = 0; j < i; ++j) acc = addp(acc, mulc(v[j], c_crt.pre_m[i][j], c_crt.pre_mp[i][j], pi), pi);
        v[i] = mulc(subp(r[i], acc, pi), c_crt.inv_m[i], c_crt.inv_mp[i], pi);
    }
    // centred comparison on digits, most significant first
    int gt = 0;
#pragma unroll
    for (int i = NPRIMES - 1; i >= 0; --i) {
        if (gt == 0 && v[i] != c_crt.half_digits[i]) gt = v[i] > c_crt.half_digits[i] ? 1 : -1;
    }
    // x mod q = sum v_i * prefix_q[i]  (< 2^148), reduced as a 256-bit value
    uint64_t acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < NPRIMES; ++i) {
        uint64_t carry = 0;
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const uint64_t t = (uint64_t)v[i] * c_crt.prefix_q[i].w[l] + (uint32_t)acc[l] + carry;
            acc[l] = (uint32_t)t;
            carry = t >> 32;
        }
        uint64_t t = (uint32_t)acc[4] + carry;
        acc[4] = (uint32_t)t;
        acc[5] += t >> 32;
    }
    uint32_t x8[8] = {(uint32_t)acc[0], (uint32_t)acc[1], (uint32_t)acc[2], (uint32_t)acc[3],
                      (uint32_t)acc[4], (uint32_t)acc[5], 0, 0};
    zq xq = zq_reduce256(x8);
    if (gt > 0) xq = zq_sub(xq, c_crt.P_q);
    return xq;
}

__global__ void __launch_bounds__(256) k_crt_lift(const uint32_t* __restrict__ res,
                                                  zq* __restrict__ out, size_t total, int log_m, size_t js) {
    const size_t gid = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= total) return;
    res = jp(res, job_off(js));
    out = jp(out, job_off(js));
    const size_t m = (size_t)1 << log_m;
    const size_t slice = gid >> log_m, c = gid & (m - 1);
    uint32_t r[NPRIMES];
#pragma unroll
    for (int i = 0; i < NPRIMES; ++i) r[i] = res[(slice * NPRIMES + i) * m + c];
    store_zq(out + gid, crt_lift_residues(r));
}

void launch_crt_lift(const uint32_t* res, zq* out, size_t slices, int log_m, hipStream_t s, const Jobs& jb) {
    const size_t total = slices << log_m;
    if (!total || jb.n == 0) return;
    hipLaunchKernelGGL(k_crt_lift, dim3((unsigned)((total + 255) / 256), jb.n), dim3(256), 0, s, res, out,
                       total, log_m, jb.stride);
}

// Lifts a stage's commitment residues and its cross terms in one launch. Part A (K slices of
// degree 2^log_ma): the lifted commitment is compared with the commitment update
// c' = c + s^-1 l + s r (poly: even(c) + ...), which is written to com_out, and a mismatch sets
// `bit` in flags (the folding invariant); with cu.c == nullptr the lift itself is written to
// com_out (and com_out2). Part B: lr_slices slices of degree 2^log_mb lifted to out_lr.
// a lifted commitment coefficient v (unscaled) at gid = k * 2^log_ma + i: stored when cu.c is
// null, else checked against the update of `cu` with the scalars (s, sinv), which is stored
__device__ __forceinline__ void commit_check(size_t gid, const zq& v, int log_ma, const ComUpd& cu, const zq& s,
                                             const zq& sinv, zq* com_out, zq* com_out2, uint32_t* flags,
                                             uint32_t bit, size_t jo) {
    if (!cu.c) {
        store_zq(jp(com_out, jo) + gid, v);
        if (com_out2) store_zq(jp(com_out2, jo) + gid, v);
        return;
    }
    const size_t m = (size_t)1 << log_ma;
    const size_t k = gid >> log_ma, i = gid & (m - 1);
    const zq* c = jp(cu.c, jo);
    const zq base = cu.poly ? load_zq(c + k * (2 * m) + 2 * i) : load_zq(c + gid);
    zq e = zq_add(base, zq_mul(load_zq(jp(cu.l, jo) + gid), sinv));
    e = zq_add(e, zq_mul(load_zq(jp(cu.r, jo) + gid), s));
    store_zq(jp(com_out, jo) + gid, e);
    if (!zq_eq(v, e)) atomicOr(jp(flags, jo), bit);
}

// one output of a lift launch: gid < K * 2^log_ma lifts a commitment coefficient (and checks it
// against the update of `cu`, or stores it when cu.c is null), the rest lift cross-term slices
__device__ __forceinline__ void lift_check_elem(size_t gid, const uint32_t* res_a, int log_ma, const ComUpd& cu,
                                                zq* com_out, zq* com_out2, uint32_t* flags, uint32_t bit,
                                                const uint32_t* res_b, zq* out_lr, int log_mb, size_t jo) {
    const size_t ta = (size_t)K << log_ma;
    uint32_t r[NPRIMES];
    if (gid < ta) {
        const size_t m = (size_t)1 << log_ma;
        const size_t k = gid >> log_ma, i = gid & (m - 1);
        res_a = jp(res_a, jo);
#pragma unroll
        for (int p = 0; p < NPRIMES; ++p) r[p] = res_a[(k * NPRIMES + p) * m + i];
        zq v = crt_lift_residues(r);
        if (cu.unscale) v = zq_mul(v, *jp(cu.unscale, jo));
        zq s = zq_zero(), sinv = zq_zero();
        if (cu.c) {
            const zq* scal = jp(cu.scal, jo);
            s = scal[0];
            sinv = scal[1];
        }
        commit_check(gid, v, log_ma, cu, s, sinv, com_out, com_out2, flags, bit, jo);
    } else {
        const size_t g = gid - ta;
        const size_t m = (size_t)1 << log_mb;
        const size_t sl = g >> log_mb, c = g & (m - 1);
        res_b = jp(res_b, jo);
#pragma unroll
        for (int p = 0; p < NPRIMES; ++p) r[p] = res_b[(sl * NPRIMES + p) * m + c];
        zq v = crt_lift_residues(r);
        if (cu.unscale) v = zq_mul(v, *jp(cu.unscale, jo));
        store_zq(jp(out_lr, jo) + g, v);
    }
}

__global__ void __launch_bounds__(256) k_lift_check(const uint32_t* __restrict__ res_a, int log_ma, ComUpd cu,
                                                    zq* com_out, zq* com_out2, uint32_t* flags, uint32_t bit,
                                                    const uint32_t* __restrict__ res_b, zq* __restrict__ out_lr,
                                                    uint32_t lr_slices, int log_mb, size_t js) {
    const size_t gid = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t ta = (size_t)K << log_ma, tb = (size_t)lr_slices << log_mb;
    if (gid >= ta + tb) return;
    lift_check_elem(gid, res_a, log_ma, cu, com_out, com_out2, flags, bit, res_b, out_lr, log_mb, job_off(js));
}

void launch_lift_check(const uint32_t* res_a, int log_ma, const ComUpd& cu, zq* com_out, zq* com_out2,
                       uint32_t* flags, uint32_t bit, const uint32_t* res_b, zq* out_lr, size_t lr_slices, int log_mb,
                       hipStream_t s, const Jobs& jb) {
    const size_t total = ((size_t)K << log_ma) + (lr_slices << log_mb);
    if (!total || jb.n == 0) return;
    hipLaunchKernelGGL(k_lift_check, dim3((unsigned)((total + 255) / 256), jb.n), dim3(256), 0, s, res_a, log_ma, cu,
                       com_out, com_out2, flags, bit, res_b, out_lr, (uint32_t)lr_slices, log_mb, jb.stride);
}

// ---------------------------------------------------------------------------------------
// One prover stage at small degree in one launch. At m <= 32 the stage's work is a few hundred
// thousand instructions, so the five-launch chain (fold + transform, MAC pass, finish, lift)
// costs its launch latencies rather than its work. Workgroup k (of K) folds the stage's witness
// rows and the generator rows of coordinate k, reduces and transforms them in LDS, runs the
// MAC pass of coordinate k (split or pairs), transforms the sums back and lifts coordinate k of
// the invariant commitment (checked as k_lift_check does) and of the cross terms: every
// residue of coordinate k is in its own LDS. Workgroup 0 computes the challenge's inverse
// meanwhile and publishes it with a device-scope release increment of the job's stage counter;
// the coordinate workgroups need it only for the lifts, so by then it is normally there (their
// wait is bounded and flags the job on expiry). Same values as the multi-launch chain.
// ---------------------------------------------------------------------------------------
constexpr int SMALL_THREADS = 1024;

// forward layers (as k_ntt_fwd_rows) on cnt slices of 2^log_n words; slice s uses prime s % 8
template <int NT>
__device__ void lds_ntt_fwd_slices_pp(uint32_t* sh, int cnt, int log_n, const NttTables& tw) {
    const int n = 1 << log_n, half = n >> 1;
    for (int len = half; len >= 1; len >>= 1) {
        const int lg = __builtin_ctz(len);
        for (int t = threadIdx.x; t < cnt * half; t += NT) {
            const int sl = t >> (log_n - 1), b = t & (half - 1);
            const int pi = sl & (NPRIMES - 1);
            const int blk = b >> lg;
            const int j = (blk << (lg + 1)) | (b & (len - 1));
            const int kz = (half >> lg) + blk;
            const uint32_t p = c_pc[pi].p;
            uint32_t* q = sh + sl * n;
            const uint32_t u = q[j];
            const uint32_t v = mulc(q[j + len], tw.zeta_m[pi * TW_SIZE + kz], tw.zeta_mp[pi * TW_SIZE + kz], p);
            q[j] = addp(u, v, p);
            q[j + len] = subp(u, v, p);
        }
        __syncthreads();
    }
}

// inverse layers (as lds_intt_slices, no scaling) on cnt slices; slice s uses prime s % 8
template <int NT>
__device__ void lds_intt_slices_pp(uint32_t* sh, int cnt, int log_n, const NttTables& tw) {
    const int n = 1 << log_n, half = n >> 1;
    for (int len = 1; len < n; len <<= 1) {
        const int lg = __builtin_ctz(len);
        for (int t = threadIdx.x; t < cnt * half; t += NT) {
            const int sl = t >> (log_n - 1), b = t & (half - 1);
            const int pi = sl & (NPRIMES - 1);
            const int blk = b >> lg;
            const int j = (blk << (lg + 1)) | (b & (len - 1));
            const int kz = (half >> lg) + blk;
            const uint32_t p = c_pc[pi].p;
            uint32_t* q = sh + sl * n;
            const uint32_t u = q[j], v = q[j + len];
            q[j] = addp(u, v, p);
            q[j + len] = mulc(subp(u, v, p), tw.izeta_m[pi * TW_SIZE + kz], tw.izeta_mp[pi * TW_SIZE + kz], p);
        }
        __syncthreads();
    }
}

// a fold row's coefficient c with the challenge s given by value (the fold modes of row_value)
__device__ __forceinline__ zq fold_value(const RowSrc& src, size_t row, uint32_t c, int log_m, const zq& s) {
    const size_t m = (size_t)1 << log_m;
    switch (src.mode) {
        case ROW_VFOLD:
            return zq_add(load_zq(src.a + row * m + c), zq_mul(load_zq(src.a + (row + src.b_off) * m + c), s));
        case ROW_VFOLD_L:
            return zq_add(zq_mul(load_zq(src.a + row * m + c), s), load_zq(src.a + (row + src.b_off) * m + c));
        case ROW_PFOLD_W: {
            const zq* base = src.a + row * (2 * m);
            return zq_add(load_zq(base + 2 * c), zq_mul(load_zq(base + 2 * c + 1), s));
        }
        case ROW_PFOLD_G: {
            const zq* base = src.a + row * (2 * m);
            const zq xo = c ? load_zq(base + 2 * c - 1) : zq_neg(load_zq(base + 2 * m - 1));
            return zq_add(load_zq(base + 2 * c), zq_mul(xo, s));
        }
        case ROW_PFOLD_GL: {
            const zq* base = src.a + row * (2 * m);
            const zq xo = c ? load_zq(base + 2 * c - 1) : zq_neg(load_zq(base + 2 * m - 1));
            return zq_add(zq_mul(load_zq(base + 2 * c), s), xo);
        }
        default:
            return load_zq(src.a + row * m + c);
    }
}

// Steps 1-4 of a small stage for coordinate k, with the stage's challenge s: fold the witness
// rows and coordinate k's generator rows, reduce and transform them in LDS (sh), run the MAC
// pass, transform the sums back and scale them. Leaves the commitment slots and cross terms in
// sh for small_stage_lift; the caller synchronises between the two.
template <int NT>
__device__ void small_stage_compute(const SmallStage& st, const NttTables& tw, size_t jo, uint32_t k, uint32_t* sh,
                                    const zq& s) {
    const int poly = st.poly;
    const int lnd = poly ? st.log_m - 1 : st.log_m;  // degree of the transforms
    const int nd = 1 << lnd;
    const uint32_t H = st.rows;
    const int log_ma = lnd, log_mb = poly ? lnd : lnd - 1;  // commitment / cross-term degrees
    const bool lr = !poly || st.cross;
    RowSrc w = st.w, g = st.g;
    if (!w.a_shared) w.a = jp(w.a, jo);
    if (!g.a_shared) g.a = jp(g.a, jo);
    w.dst = jp(w.dst, jo);
    g.dst = jp(g.dst, jo);
    // 1) rows 0..H-1 witness, H..2H-1 generators of coordinate k, as residues T[row][p][c]
    uint32_t* T = sh;
    for (uint32_t e = threadIdx.x; e < (2 * H) << lnd; e += NT) {
        const uint32_t rl = e >> lnd, c = e & (nd - 1);
        zq x;
        if (rl < H) {
            x = fold_value(w, rl, c, lnd, s);
            if (k == 0 && w.dst) store_zq(w.dst + ((size_t)rl << lnd) + c, x);
        } else {
            const size_t grow = (size_t)(rl - H) * K + k;
            x = fold_value(g, grow, c, lnd, s);
            if (g.dst) store_zq(g.dst + (grow << lnd) + c, x);
        }
#pragma unroll
        for (int q = 0; q < NPRIMES; ++q) T[((rl * NPRIMES + q) << lnd) + c] = zq_mod_p_centred(x, c_pc[q]);
    }
    __syncthreads();
    // 2) forward transforms
    lds_ntt_fwd_slices_pp<NT>(T, (int)(2 * H) * NPRIMES, lnd, tw);
    // 3) the MAC pass of coordinate k: O = commitment slots [p][nd], then L [p][nb], R [p][nb]
    uint32_t* O = T + ((size_t)(2 * H) * NPRIMES << lnd);
    uint32_t* OL = O + (NPRIMES << lnd);
    uint32_t* OR = OL + (NPRIMES << log_mb);
    const uint32_t* Wt = T;
    const uint32_t* Gt = T + ((size_t)H * NPRIMES << lnd);
    if (!poly) {  // split: (S00, S11, S01, S10) per pair, as k_mac_split4 and its finish
        const int hm = nd >> 1;
        for (int e = threadIdx.x; e < NPRIMES * hm; e += NT) {
            const int p = e / hm, t = e - p * hm;
            Acc96 s00, s11, s01, s10;
            s00.zero();
            s11.zero();
            s01.zero();
            s10.zero();
            for (uint32_t i = 0; i < H; ++i) {
                const uint32_t* wr = Wt + ((i * NPRIMES + p) << lnd) + 2 * t;
                const uint32_t* gr = Gt + ((i * NPRIMES + p) << lnd) + 2 * t;
                s00.mac(wr[0], gr[0]);
                s11.mac(wr[1], gr[1]);
                s01.mac(wr[0], gr[1]);
                s10.mac(wr[1], gr[0]);
            }
            const PrimeConsts pc = c_pc[p];
            const uint32_t P = pc.p;
            const uint32_t a00 = reduce_u96(s00.lo, s00.hi, pc), a11 = reduce_u96(s11.lo, s11.hi, pc);
            const uint32_t a01 = reduce_u96(s01.lo, s01.hi, pc), a10 = reduce_u96(s10.lo, s10.hi, pc);
            O[(p << lnd) + 2 * t] = a00;
            O[(p << lnd) + 2 * t + 1] = a11;
            const uint32_t d = subp(a00, a11, P);
            const uint32_t Lp = addp(subp(d, a01, P), a10, P);
            const uint32_t Rp = subp(addp(d, a01, P), a10, P);
            const uint32_t kz = (uint32_t)hm + t;
            const uint32_t i4m = tw.invn_m[p * (MAX_LOG_N + 1) + 2], i4mp = tw.invn_mp[p * (MAX_LOG_N + 1) + 2];
            OL[(p << log_mb) + t] =
                mulc(mulc(Lp, tw.zeta_m[p * TW_SIZE + kz], tw.zeta_mp[p * TW_SIZE + kz], P), i4m, i4mp, P);
            OR[(p << log_mb) + t] =
                mulc(mulc(Rp, tw.izeta_m[p * TW_SIZE + kz], tw.izeta_mp[p * TW_SIZE + kz], P), i4m, i4mp, P);
        }
    } else {  // pairs: the invariant over every row, or with cross terms over the two halves
        const uint32_t hh = st.cross ? H / 2 : H;
        for (int e = threadIdx.x; e < (NPRIMES << lnd); e += NT) {
            const int p = e >> lnd, c = e & (nd - 1);
            Acc96 inv, L, R;
            inv.zero();
            L.zero();
            R.zero();
            for (uint32_t i = 0; i < hh; ++i) {
                const uint32_t w0 = Wt[((i * NPRIMES + p) << lnd) + c], g0 = Gt[((i * NPRIMES + p) << lnd) + c];
                inv.mac(w0, g0);
                if (st.cross) {
                    const uint32_t w1 = Wt[(((i + hh) * NPRIMES + p) << lnd) + c];
                    const uint32_t g1 = Gt[(((i + hh) * NPRIMES + p) << lnd) + c];
                    inv.mac(w1, g1);
                    L.mac(g1, w0);
                    R.mac(g0, w1);
                }
            }
            const PrimeConsts pc = c_pc[p];
            O[e] = reduce_u96(inv.lo, inv.hi, pc);
            if (st.cross) {
                OL[e] = reduce_u96(L.lo, L.hi, pc);
                OR[e] = reduce_u96(R.lo, R.hi, pc);
            }
        }
    }
    __syncthreads();
    // 4) inverse transforms, scaled by the inverse degrees
    lds_intt_slices_pp<NT>(O, NPRIMES, log_ma, tw);
    if (lr) lds_intt_slices_pp<NT>(OL, 2 * NPRIMES, log_mb, tw);
    for (int e = threadIdx.x; e < (NPRIMES << log_ma); e += NT) {
        const int p = e >> log_ma;
        O[e] = mulc(O[e], tw.invn_m[p * (MAX_LOG_N + 1) + log_ma], tw.invn_mp[p * (MAX_LOG_N + 1) + log_ma], c_pc[p].p);
    }
    if (lr) {
        for (int e = threadIdx.x; e < (2 * NPRIMES << log_mb); e += NT) {
            const int p = (e >> log_mb) & (NPRIMES - 1);
            OL[e] = mulc(OL[e], tw.invn_m[p * (MAX_LOG_N + 1) + log_mb], tw.invn_mp[p * (MAX_LOG_N + 1) + log_mb],
                         c_pc[p].p);
        }
    }
}

// Step 5: lift coordinate k of the commitment (checked against the update of st.cu with the
// scalars (cx, cxi), as k_lift_check does) and of the cross terms, each unscaled by u.
template <int NT>
__device__ void small_stage_lift(const SmallStage& st, size_t jo, uint32_t k, uint32_t* sh, const zq& u, const zq& cx,
                                 const zq& cxi) {
    const int poly = st.poly;
    const int lnd = poly ? st.log_m - 1 : st.log_m;
    const uint32_t H = st.rows;
    const int log_ma = lnd, log_mb = poly ? lnd : lnd - 1;
    const int nb = 1 << log_mb;
    const bool lr = !poly || st.cross;
    uint32_t* O = sh + ((size_t)(2 * H) * NPRIMES << lnd);
    uint32_t* OL = O + (NPRIMES << lnd);
    const int na = 1 << log_ma;
    const int cnt = na + (lr ? 2 * nb : 0);
    for (int e = threadIdx.x; e < cnt; e += NT) {
        uint32_t r[NPRIMES];
        if (e < na) {
#pragma unroll
            for (int p = 0; p < NPRIMES; ++p) r[p] = O[(p << log_ma) + e];
            const zq v = zq_mul(crt_lift_residues(r), u);
            commit_check(((size_t)k << log_ma) + e, v, log_ma, st.cu, cx, cxi, st.com_out, nullptr, st.flags, st.bit,
                         jo);
        } else {
            const int f = e - na, side = f >> log_mb, c = f & (nb - 1);  // side 0: L, 1: R
#pragma unroll
            for (int p = 0; p < NPRIMES; ++p) r[p] = OL[(((side * NPRIMES) + p) << log_mb) + c];
            const size_t slot = side ? K + k : k;
            store_zq(jp(st.out_lr, jo) + (slot << log_mb) + c, zq_mul(crt_lift_residues(r), u));
        }
    }
}

__global__ void __launch_bounds__(SMALL_THREADS) k_stage_small(SmallStage st, NttTables tw, size_t js) {
    critical_path_priority();
    extern __shared__ uint32_t sh[];
    __shared__ zq s_x, s_xi, s_u;
    const size_t jo = job_off(js);
    if (blockIdx.x == 0) {  // the challenge inverse, published by a release increment
        if (threadIdx.x < 64) {
            zq* chal = jp(st.it.chal, jo);
            const zq x = chal[0];  // every lane loads it: a per-lane register, as the inverse wants
            const zq xi = zq_inv_divsteps(x);
            const zq u = st.it.u_prev ? zq_mul(*jp(st.it.u_prev, jo), xi) : xi;
            if (threadIdx.x == 0) {
                chal[1] = xi;
                *jp(st.it.u_out, jo) = u;
                __hip_atomic_fetch_add(jp(st.ctr, jo), 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        return;
    }
    const uint32_t k = blockIdx.x - 1;
    small_stage_compute<SMALL_THREADS>(st, tw, jo, k, sh, *jp(st.w.scal, jo));
    // wait (bounded) for the inverse, then lift coordinate k, unscaled by u
    if (threadIdx.x == 0) {
        uint32_t spins = 0;
        while (__hip_atomic_load(jp(st.ctr, jo), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < st.target) {
            if (++spins > (1u << 22)) {
                atomicOr(jp(st.flags, jo), STAGE_WAIT_EXPIRED);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        s_x = jp(st.it.chal, jo)[0];
        s_xi = jp(st.it.chal, jo)[1];
        s_u = *jp(st.it.u_out, jo);
    }
    __syncthreads();
    small_stage_lift<SMALL_THREADS>(st, jo, k, sh, s_u, s_x, s_xi);
}

void launch_stage_small(const SmallStage& st, const NttTables& tw, hipStream_t s, const Jobs& jb) {
    if (jb.n == 0) return;
    const int lnd = st.poly ? st.log_m - 1 : st.log_m;
    const int log_mb = st.poly ? lnd : lnd - 1;
    const size_t lds = 4 * ((2 * (size_t)st.rows * NPRIMES << lnd) + ((size_t)NPRIMES << lnd) +
                            ((size_t)2 * NPRIMES << (log_mb > 0 ? log_mb : 0)));
    hipLaunchKernelGGL(k_stage_small, dim3(K + 1, jb.n), dim3(SMALL_THREADS), lds, s, st, tw, jb.stride);
}

// ---------------------------------------------------------------------------------------
// The persistent tail. In a proof's last rounds a stage's work is small, and as a chain of
// launches (a challenge kernel, then the stage kernel, each a few microseconds of launch and
// drain, with waits across hardware queues for the verifier) it costs more than its work; spread
// over several workgroups, each stage would also pay for hand-offs that write back and invalidate
// caches. So one workgroup per job runs the last stages, every coordinate: per stage it hashes
// the transcript message (the previous stage's lift outputs, kept in LDS), then wave 0 inverts the
// challenge x while the other waves fold the witness rows and, in groups of coordinates that fit
// LDS, the generator rows with x, transform them, run the MAC pass and transform the sums back;
// then the whole workgroup lifts every coordinate, unscaled by the running inverse product u. Its
// re-reads of the rows it stored hit a warm cache. After each stage it publishes with a release
// increment of the job's stage counter. With `verifier`, a second workgroup replays the same
// stages (verifier.rs:57-120) behind it: it hashes each message from its own commitment and the
// prover's cross terms and, once the prover's stage is published, updates its commitment with the
// checked inverse candidate. The prover never waits for it. Values are those of the multi-launch
// chain.
// ---------------------------------------------------------------------------------------
// Waits until the job's stage counter reaches `target`, bounded by the device's constant-rate
// clock (s_memrealtime, 100 MHz): after TAIL_WAIT_TICKS the job is flagged and the workgroup goes
// on, so a launch always ends.
constexpr uint64_t TAIL_WAIT_TICKS = 50000000;  // 0.5 s
__device__ __forceinline__ void tail_wait(const uint32_t* ctr, uint32_t target, uint32_t* flags) {
    if (threadIdx.x == 0) {
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        while (__hip_atomic_load(ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__builtin_amdgcn_s_memrealtime() - t0 > TAIL_WAIT_TICKS) {
                atomicOr(flags, STAGE_WAIT_EXPIRED);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
}

// the verifier's commitment update of one replay step (as k_com_update), by the whole workgroup
__device__ __forceinline__ void tail_com_update(const TailVerStep& vs, const zq& s, const zq* hint, size_t jo) {
    zq sinv = load_zq(hint);  // uniform: every thread checks the same candidate
    if (!zq_eq(zq_mul(s, sinv), zq_from_u64(1))) sinv = zq_inv_divsteps(s);
    zq* scal = jp(vs.scal, jo);
    if (threadIdx.x == 0) scal[1] = sinv;
    const zq* c = jp(vs.c, jo);
    const zq* l = jp(vs.l, jo);
    const zq* r = jp(vs.r, jo);
    zq* out = jp(vs.out, jo);
    const uint32_t m = 1u << vs.log_m;
    for (uint32_t gid = threadIdx.x; gid < K * m; gid += blockDim.x) {
        const uint32_t k = gid >> vs.log_m, i = gid & (m - 1);
        const zq base = vs.poly ? load_zq(c + k * (2 * m) + 2 * i) : load_zq(c + gid);
        zq v = zq_add(base, zq_mul(load_zq(l + gid), sinv));
        v = zq_add(v, zq_mul(load_zq(r + gid), s));
        store_zq(out + gid, v);
    }
}

// 512 threads: the challenge hash needs more than the 128 VGPRs a 1024-thread workgroup allows
constexpr int TAIL_THREADS = 512;
constexpr int TAIL_COMPUTE_THREADS = TAIL_THREADS - 64;  // waves 1.. while wave 0 inverts
constexpr size_t TAIL_DESC_LDS = (sizeof(TailDesc) + 15) & ~(size_t)15;
constexpr size_t TAIL_SCRATCH_BUDGET = 128 * 1024;  // bytes of LDS for one stage's residues
constexpr int TAIL_MAX_N = 32;                      // the tail's largest transform degree

// The twiddles of degrees <= 32 (the first 32 entries of each prime's nested tables) and the
// inverse degrees, copied to LDS once per launch.
struct TailTw {
    uint32_t z[NPRIMES][4][TAIL_MAX_N];  // zeta_m, zeta_mp, izeta_m, izeta_mp
    uint32_t inv[NPRIMES][2][8];         // invn_m, invn_mp for log degrees 0..5 (and 2 = 1/4)
};

// A stage's residue scratch in words, with the generator rows transformed `grp` coordinates at a
// time: witness rows Wt [H][8][nd], commitment slots O [K][8][nd], cross terms OL [K][2][8][nb],
// generator rows Gt [grp][H][8][nd].
__host__ __device__ inline size_t tail_stage_words(const SmallStage& st, uint32_t grp) {
    const int lnd = st.poly ? st.log_m - 1 : st.log_m;
    const int log_mb = st.poly ? lnd : lnd - 1;
    const size_t H = st.rows;
    return (((H + K + grp * H) * NPRIMES) << lnd) + (((size_t)K * 2 * NPRIMES) << (log_mb > 0 ? log_mb : 0));
}
__host__ __device__ inline uint32_t tail_group(const SmallStage& st) {
    uint32_t g = K;
    while (g > 1 && 4 * tail_stage_words(st, g) > TAIL_SCRATCH_BUDGET) g >>= 1;
    return g;
}
// the lift outputs kept for the next message: K commitment coordinates and 2K cross-term slices
__host__ __device__ inline size_t tail_msg_zq(const SmallStage& st) {
    const int lnd = st.poly ? st.log_m - 1 : st.log_m;
    const int log_mb = st.poly ? lnd : lnd - 1;
    return ((size_t)K << lnd) + ((size_t)2 * K << (log_mb > 0 ? log_mb : 0));
}

// A barrier among the workgroup's waves 1.. (wave 0 computes the challenge inverse meanwhile):
// every wave adds its arrival to an LDS counter and waits for the count of this phase. All
// participating waves pass the same barriers in the same order.
struct SubBarrier {
    uint32_t* cnt;
    uint32_t target;
    __device__ __forceinline__ void sync() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        target += TAIL_COMPUTE_THREADS / 64;
        if ((threadIdx.x & 63) == 0) __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < target)
            __builtin_amdgcn_s_sleep(1);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
};

// forward / inverse layers (as lds_ntt_fwd_slices_pp / lds_intt_slices_pp) on cnt slices of the
// compute waves, twiddles from LDS
__device__ __forceinline__ void tail_fwd_layers(uint32_t* sh, int cnt, int log_n, const TailTw& tt, SubBarrier& bar) {
    const int n = 1 << log_n, half = n >> 1;
    for (int len = half; len >= 1; len >>= 1) {
        const int lg = __builtin_ctz(len);
        for (int t = (int)threadIdx.x - 64; t < cnt * half; t += TAIL_COMPUTE_THREADS) {
            const int sl = t >> (log_n - 1), b = t & (half - 1);
            const int pi = sl & (NPRIMES - 1);
            const int blk = b >> lg;
            const int j = (blk << (lg + 1)) | (b & (len - 1));
            const int kz = (half >> lg) + blk;
            const uint32_t p = c_pc[pi].p;
            uint32_t* q = sh + sl * n;
            const uint32_t u = q[j];
            const uint32_t v = mulc(q[j + len], tt.z[pi][0][kz], tt.z[pi][1][kz], p);
            q[j] = addp(u, v, p);
            q[j + len] = subp(u, v, p);
        }
        bar.sync();
    }
}
__device__ __forceinline__ void tail_inv_layers(uint32_t* sh, int cnt, int log_n, const TailTw& tt, SubBarrier& bar) {
    const int n = 1 << log_n, half = n >> 1;
    for (int len = 1; len < n; len <<= 1) {
        const int lg = __builtin_ctz(len);
        for (int t = (int)threadIdx.x - 64; t < cnt * half; t += TAIL_COMPUTE_THREADS) {
            const int sl = t >> (log_n - 1), b = t & (half - 1);
            const int pi = sl & (NPRIMES - 1);
            const int blk = b >> lg;
            const int j = (blk << (lg + 1)) | (b & (len - 1));
            const int kz = (half >> lg) + blk;
            const uint32_t p = c_pc[pi].p;
            uint32_t* q = sh + sl * n;
            const uint32_t u = q[j], v = q[j + len];
            q[j] = addp(u, v, p);
            q[j + len] = mulc(subp(u, v, p), tt.z[pi][2][kz], tt.z[pi][3][kz], p);
        }
        bar.sync();
    }
}

// A fold's operands, loaded apart from the combination so that a thread's loads for several
// coefficients are in flight together.
struct FoldIn {
    zq a, b;
};
__device__ __forceinline__ FoldIn fold_load(const RowSrc& src, size_t row, uint32_t c, int log_m) {
    const size_t m = (size_t)1 << log_m;
    FoldIn o;
    switch (src.mode) {
        case ROW_VFOLD:
        case ROW_VFOLD_L:
            o.a = load_zq(src.a + row * m + c);
            o.b = load_zq(src.a + (row + src.b_off) * m + c);
            break;
        case ROW_PFOLD_W: {
            const zq* base = src.a + row * (2 * m);
            o.a = load_zq(base + 2 * c);
            o.b = load_zq(base + 2 * c + 1);
            break;
        }
        case ROW_PFOLD_G:
        case ROW_PFOLD_GL: {  // b: odd[c - 1], or odd[m - 1] (negated in fold_combine) at c = 0
            const zq* base = src.a + row * (2 * m);
            o.a = load_zq(base + 2 * c);
            o.b = load_zq(base + (c ? 2 * c - 1 : 2 * m - 1));
            break;
        }
        default:
            o.a = load_zq(src.a + row * m + c);
            o.b = zq_zero();
    }
    return o;
}
// the same values as fold_value
__device__ __forceinline__ zq fold_combine(int mode, const FoldIn& o, uint32_t c, const zq& s) {
    switch (mode) {
        case ROW_VFOLD:
        case ROW_PFOLD_W:
            return zq_add(o.a, zq_mul(o.b, s));
        case ROW_VFOLD_L:
            return zq_add(zq_mul(o.a, s), o.b);
        case ROW_PFOLD_G:
            return zq_add(o.a, zq_mul(c ? o.b : zq_neg(o.b), s));
        case ROW_PFOLD_GL:
            return zq_add(zq_mul(o.a, s), c ? o.b : zq_neg(o.b));
        default:
            return o.a;
    }
}

// `total` fold outputs by the compute waves, FOLD_BATCH coefficients per thread at a time: output e
// is row row_of(e) coefficient e mod nd, stored at dst (when set) and reduced into the residue
// slots T[(slot_of(e) * 8 + q) * nd + c].
constexpr int FOLD_BATCH = 4;
template <class RowOf, class SlotOf>
__device__ __forceinline__ void tail_fold(const RowSrc& src, uint32_t total, int lnd, const zq& x, RowOf row_of, SlotOf slot_of,
                          uint32_t* T) {
    const uint32_t nd = 1u << lnd;
    for (uint32_t e0 = threadIdx.x - 64; e0 < total; e0 += TAIL_COMPUTE_THREADS * FOLD_BATCH) {
        FoldIn in[FOLD_BATCH];
#pragma unroll
        for (int i = 0; i < FOLD_BATCH; ++i) {
            const uint32_t e = e0 + i * TAIL_COMPUTE_THREADS;
            if (e < total) in[i] = fold_load(src, row_of(e), e & (nd - 1), lnd);
        }
#pragma unroll
        for (int i = 0; i < FOLD_BATCH; ++i) {
            const uint32_t e = e0 + i * TAIL_COMPUTE_THREADS;
            if (e < total) {
                const uint32_t c = e & (nd - 1);
                const size_t row = row_of(e);
                const zq v = fold_combine(src.mode, in[i], c, x);
                if (src.dst) store_zq(src.dst + (row << lnd) + c, v);
                const uint32_t sl = slot_of(e);
#pragma unroll
                for (int q = 0; q < NPRIMES; ++q) T[((sl * NPRIMES + q) << lnd) + c] = zq_mod_p_centred(v, c_pc[q]);
            }
        }
    }
}

// Fold, transforms, MAC pass and inverse transforms of one stage for every coordinate, by the
// compute waves (the multi-coordinate form of small_stage_compute); leaves O and OL in T for
// tail_stage_lift.
// `w` and `g` are the stage's row sources with every address resolved (LDS or the job's memory).
__device__ __forceinline__ void tail_stage_compute(const SmallStage& st, const RowSrc& w, const RowSrc& g, uint32_t* T, const zq& x,
                                   const TailTw& tt, SubBarrier& bar) {
    const int poly = st.poly;
    const int lnd = poly ? st.log_m - 1 : st.log_m;
    const int nd = 1 << lnd;
    const uint32_t H = st.rows;
    const int log_ma = lnd, log_mb = poly ? lnd : lnd - 1;
    const bool lr = !poly || st.cross;
    const uint32_t grp = tail_group(st);
    uint32_t* Wt = T;
    uint32_t* O = Wt + ((H * NPRIMES) << lnd);
    uint32_t* OL = O + ((K * NPRIMES) << lnd);
    uint32_t* Gt = OL + (((uint32_t)K * 2 * NPRIMES) << (log_mb > 0 ? log_mb : 0));
    // 1) the witness rows (stored: the next stage folds them)
    tail_fold(w, H << lnd, lnd, x, [=](uint32_t e) { return (size_t)(e >> lnd); }, [=](uint32_t e) { return e >> lnd; },
              Wt);
    bar.sync();
    tail_fwd_layers(Wt, (int)H * NPRIMES, lnd, tt, bar);
    // 2) generator rows, grp coordinates at a time, and their MAC passes
    for (uint32_t k0 = 0; k0 < K; k0 += grp) {
        tail_fold(
            g, (grp * H) << lnd, lnd, x,
            [=](uint32_t e) {  // rk = kk * H + row
                const uint32_t rk = e >> lnd, kk = rk / H;
                return (size_t)(rk - kk * H) * K + k0 + kk;
            },
            [=](uint32_t e) { return e >> lnd; }, Gt);
        bar.sync();
        tail_fwd_layers(Gt, (int)(grp * H) * NPRIMES, lnd, tt, bar);
        if (!poly) {  // split: (S00, S11, S01, S10) per pair, as k_mac_split4 and its finish
            const int hm = nd >> 1;
            for (int e = (int)threadIdx.x - 64; e < (int)grp * NPRIMES * hm; e += TAIL_COMPUTE_THREADS) {
                const int kk = e / (NPRIMES * hm), r0 = e - kk * NPRIMES * hm;
                const int p = r0 / hm, t = r0 - p * hm;
                const uint32_t k = k0 + kk;
                const uint32_t* Gk = Gt + ((kk * H * NPRIMES) << lnd);
                Acc96 s00, s11, s01, s10;
                s00.zero();
                s11.zero();
                s01.zero();
                s10.zero();
                for (uint32_t i = 0; i < H; ++i) {
                    const uint32_t* wr = Wt + ((i * NPRIMES + p) << lnd) + 2 * t;
                    const uint32_t* gr = Gk + ((i * NPRIMES + p) << lnd) + 2 * t;
                    s00.mac(wr[0], gr[0]);
                    s11.mac(wr[1], gr[1]);
                    s01.mac(wr[0], gr[1]);
                    s10.mac(wr[1], gr[0]);
                }
                const PrimeConsts pc = c_pc[p];
                const uint32_t P = pc.p;
                const uint32_t a00 = reduce_u96(s00.lo, s00.hi, pc), a11 = reduce_u96(s11.lo, s11.hi, pc);
                const uint32_t a01 = reduce_u96(s01.lo, s01.hi, pc), a10 = reduce_u96(s10.lo, s10.hi, pc);
                uint32_t* Ok = O + ((k * NPRIMES) << lnd);
                Ok[(p << lnd) + 2 * t] = a00;
                Ok[(p << lnd) + 2 * t + 1] = a11;
                const uint32_t d = subp(a00, a11, P);
                const uint32_t Lp = addp(subp(d, a01, P), a10, P);
                const uint32_t Rp = subp(addp(d, a01, P), a10, P);
                const uint32_t kz = (uint32_t)hm + t;
                const uint32_t i4m = tt.inv[p][0][2], i4mp = tt.inv[p][1][2];
                uint32_t* OLk = OL + ((k * 2 * NPRIMES) << log_mb);
                OLk[(p << log_mb) + t] = mulc(mulc(Lp, tt.z[p][0][kz], tt.z[p][1][kz], P), i4m, i4mp, P);
                OLk[((NPRIMES + p) << log_mb) + t] = mulc(mulc(Rp, tt.z[p][2][kz], tt.z[p][3][kz], P), i4m, i4mp, P);
            }
        } else {  // pairs: the invariant over every row, or with cross terms over the two halves
            const uint32_t hh = st.cross ? H / 2 : H;
            for (int e = (int)threadIdx.x - 64; e < (int)((grp * NPRIMES) << lnd); e += TAIL_COMPUTE_THREADS) {
                const int kk = e >> (lnd + 3), r0 = e & ((NPRIMES << lnd) - 1);
                const int p = r0 >> lnd, c = r0 & (nd - 1);
                const uint32_t k = k0 + kk;
                const uint32_t* Gk = Gt + ((kk * H * NPRIMES) << lnd);
                Acc96 inv, L, R;
                inv.zero();
                L.zero();
                R.zero();
                for (uint32_t i = 0; i < hh; ++i) {
                    const uint32_t w0 = Wt[((i * NPRIMES + p) << lnd) + c], g0 = Gk[((i * NPRIMES + p) << lnd) + c];
                    inv.mac(w0, g0);
                    if (st.cross) {
                        const uint32_t w1 = Wt[(((i + hh) * NPRIMES + p) << lnd) + c];
                        const uint32_t g1 = Gk[(((i + hh) * NPRIMES + p) << lnd) + c];
                        inv.mac(w1, g1);
                        L.mac(g1, w0);
                        R.mac(g0, w1);
                    }
                }
                const PrimeConsts pc = c_pc[p];
                O[((k * NPRIMES) << lnd) + r0] = reduce_u96(inv.lo, inv.hi, pc);
                if (st.cross) {
                    uint32_t* OLk = OL + ((k * 2 * NPRIMES) << log_mb);  // poly: log_mb == lnd
                    OLk[r0] = reduce_u96(L.lo, L.hi, pc);
                    OLk[(NPRIMES << lnd) + r0] = reduce_u96(R.lo, R.hi, pc);
                }
            }
        }
        bar.sync();
    }
    // 3) inverse transforms, scaled by the inverse degrees
    tail_inv_layers(O, K * NPRIMES, log_ma, tt, bar);
    if (lr) tail_inv_layers(OL, K * 2 * NPRIMES, log_mb, tt, bar);
    for (int e = (int)threadIdx.x - 64; e < ((K * NPRIMES) << log_ma); e += TAIL_COMPUTE_THREADS) {
        const int p = (e >> log_ma) & (NPRIMES - 1);
        O[e] = mulc(O[e], tt.inv[p][0][log_ma], tt.inv[p][1][log_ma], c_pc[p].p);
    }
    if (lr) {
        for (int e = (int)threadIdx.x - 64; e < ((K * 2 * NPRIMES) << log_mb); e += TAIL_COMPUTE_THREADS) {
            const int p = (e >> log_mb) & (NPRIMES - 1);
            OL[e] = mulc(OL[e], tt.inv[p][0][log_mb], tt.inv[p][1][log_mb], c_pc[p].p);
        }
    }
}

// Lifts of every coordinate (as small_stage_lift with commit_check), each unscaled by u, by the
// whole workgroup; a commitment coefficient's update operands (cc, cl, cr: the previous stage's
// commitment and cross terms) are loaded before its lift. The commitment goes to msg only (the
// next stage's transcript message and update operands; nothing reads it after the tail), the
// cross terms to the round record olr and after the commitment in msg.
__device__ __forceinline__ void tail_stage_lift(const SmallStage& st, const zq* cc, const zq* cl, const zq* cr, zq* olr,
                                uint32_t* flags, const uint32_t* T, zq* msg, const zq& u, const zq& cx, const zq& cxi) {
    const int poly = st.poly;
    const int lnd = poly ? st.log_m - 1 : st.log_m;
    const uint32_t H = st.rows;
    const int log_ma = lnd, log_mb = poly ? lnd : lnd - 1;
    const bool lr = !poly || st.cross;
    const uint32_t* O = T + ((H * NPRIMES) << lnd);
    const uint32_t* OL = O + ((K * NPRIMES) << lnd);
    const int na = 1 << log_ma, nb = lr ? 1 << log_mb : 0;
    const int ta = K * na;
    for (int e = threadIdx.x; e < ta + 2 * K * nb; e += TAIL_THREADS) {
        uint32_t r[NPRIMES];
        if (e < ta) {  // the invariant: C' = c + s^-1 l + s r (poly: even(c) + ...), as commit_check
            const int k = e >> log_ma, i = e & (na - 1);
            const zq base = poly ? load_zq(cc + (size_t)k * (2 * na) + 2 * i) : load_zq(cc + e);
            const zq lv = load_zq(cl + e), rv = load_zq(cr + e);
#pragma unroll
            for (int p = 0; p < NPRIMES; ++p) r[p] = O[((k * NPRIMES + p) << log_ma) + i];
            const zq v = zq_mul(crt_lift_residues(r), u);
            const zq ex = zq_add(zq_add(base, zq_mul(lv, cxi)), zq_mul(rv, cx));
            msg[e] = ex;
            if (!zq_eq(v, ex)) atomicOr(flags, st.bit);
        } else {
            const int f = e - ta, k = f / (2 * nb), side = (f / nb) & 1, c = f & (nb - 1);
#pragma unroll
            for (int p = 0; p < NPRIMES; ++p) r[p] = OL[(((k * 2 + side) * NPRIMES + p) << log_mb) + c];
            const size_t slot = side ? K + k : k;
            const zq v = zq_mul(crt_lift_residues(r), u);
            store_zq(olr + (slot << log_mb) + c, v);
            msg[ta + (slot << log_mb) + c] = v;
        }
    }
}

// LDS of a tail launch: the descriptor, the twiddles, the message kept between stages, then the
// stage scratch (shared with the challenge's window)
constexpr size_t TAIL_TW_LDS = (sizeof(TailTw) + 15) & ~(size_t)15;

// The folded rows between stages, in LDS: X holds a vector stage's input (m rows of degree m for
// the witness, then K m rows for the generators; m <= TAIL_MAX_M), Y a poly stage's input (half
// the rows); a vector stage folds X into Y, a poly stage Y into the next round's X.
constexpr size_t TAIL_STATE_X = TAIL_MAX_M * TAIL_MAX_M * (1 + K);
constexpr size_t TAIL_STATE_Y = TAIL_STATE_X / 2;

// The LDS regions of the prover's workgroup (see k_tail).
struct TailProverLds {
    zq* msgs;       // two lift outputs: stage s writes msgs[s & 1], stage s + 1 reads it
    uint32_t msg_zq;
    zq* stX;        // folded rows: a vector stage's input, a poly stage's output
    zq* stY;        // a vector stage's output, a poly stage's input
    uint32_t* work; // stage scratch, shared with the challenge window
    const uint8_t* seeds;
    const uint32_t* reject;
    zq* s_xi;
    zq* s_u;
};

// One prover stage of the tail. The first stage reads the job's rows, commitment and cross terms
// in memory; later ones only LDS (the rows folded by the stage before, the previous lift
// outputs), instantiated separately so that every access there is an LDS instruction, which
// waits for LDS alone and not for the global stores still in flight (the round records).
template <bool FIRST>
__device__ __forceinline__ void tail_prover_stage(uint32_t s, const TailDesc& D, size_t jo, const TailProverLds& L,
                                                  WgChal& wc, const TailTw& tt, SubBarrier& bar,
                                                  const uint32_t key[8], zq& u) {
    const SmallStage& st = D.st[s];
    uint32_t* flags = jp(st.flags, jo);
    zq* const prev = L.msgs + (size_t)((s + 1) & 1) * L.msg_zq;
    zq* const next = L.msgs + (size_t)(s & 1) * L.msg_zq;
    TranscriptMsg t = D.ptx[s];
    tx_to_job(t, jo);
    t.epoch_seed = L.seeds;
    t.file_id = L.seeds + 32;
    t.reject = L.reject;
    // the update operands: the previous stage's commitment and cross terms
    const zq *cc, *cl, *cr;
    if (FIRST) {
        cc = jp(st.cu.c, jo);
        cl = jp(st.cu.l, jo);
        cr = jp(st.cu.r, jo);
    } else {  // the message is the previous stage's lift outputs (tail_stage_lift)
        t.mods[0] = prev;
        t.mods[1] = prev + (size_t)K * t.lens[0];
        t.mods[2] = t.mods[1] + (size_t)K * t.lens[1];
        cc = t.mods[0];
        cl = t.mods[1];
        cr = t.mods[2];
    }
    // rows: folded from the job's rows (first stage) or the LDS state, into the LDS state
    const int lnd = st.poly ? st.log_m - 1 : st.log_m;
    const size_t w_in = (size_t)(st.poly ? st.rows : 2 * st.rows) << st.log_m, w_out = (size_t)st.rows << lnd;
    zq* const in = st.poly ? L.stY : L.stX;
    zq* const out = st.poly ? L.stX : L.stY;
    RowSrc w = st.w, g = st.g;
    if (FIRST) {
        if (!w.a_shared) w.a = jp(w.a, jo);
        if (!g.a_shared) g.a = jp(g.a, jo);
    } else {
        w.a = in;
        g.a = in + w_in;
    }
    w.dst = out;
    g.dst = out + w_out;
    const size_t len = tx_message_len(t.label_len, t.lens);
    zq* chal = jp(st.it.chal, jo);
    const zq x = wg_tx_challenge(t, len, key, 0, wc, chal, flags, 0);
    __syncthreads();  // every thread holds x; the window's LDS is reused below
    if (threadIdx.x < 64) {  // wave 0 inverts x (every lane: a per-lane register, as the inverse wants)
        const zq xi = zq_inv_divsteps(x);
        const zq un = zq_mul(u, xi);
        if (threadIdx.x == 0) {
            *L.s_xi = xi;
            *L.s_u = un;
            chal[1] = xi;
            *jp(st.it.u_out, jo) = un;
        }
    } else {  // the other waves fold, transform and run the MAC pass meanwhile
        tail_stage_compute(st, w, g, L.work, x, tt, bar);
    }
    __syncthreads();
    u = *L.s_u;
    tail_stage_lift(st, cc, cl, cr, jp(st.out_lr, jo), flags, L.work, next, u, x, *L.s_xi);
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        __hip_atomic_fetch_add(jp(D.ctr, jo), 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void __launch_bounds__(TAIL_THREADS) k_tail(const TailDesc* __restrict__ gdesc, NttTables tw, size_t js,
                                                        int hint_off, uint32_t msg_zq) {
    critical_path_priority();
    extern __shared__ __align__(16) uint32_t dsh[];
    __shared__ zq s_xi, s_u;
    __shared__ uint32_t s_bar;
    __shared__ __align__(16) uint8_t s_seeds[64];  // this workgroup's epoch_seed | file_id
    __shared__ uint32_t s_reject;                 // and its resampling test count
    const size_t jo = job_off(js);
    TailTw& tt = *reinterpret_cast<TailTw*>(dsh + TAIL_DESC_LDS / 4);
    {  // the descriptor, the twiddles and the transcript seeds into LDS
        const uint32_t* src = reinterpret_cast<const uint32_t*>(gdesc);
        for (uint32_t i = threadIdx.x; i < sizeof(TailDesc) / 4; i += blockDim.x) dsh[i] = src[i];
        for (uint32_t i = threadIdx.x; i < NPRIMES * 4 * TAIL_MAX_N; i += blockDim.x) {
            const uint32_t p = i / (4 * TAIL_MAX_N), a = (i / TAIL_MAX_N) & 3, k = i & (TAIL_MAX_N - 1);
            const uint32_t* t = a == 0 ? tw.zeta_m : a == 1 ? tw.zeta_mp : a == 2 ? tw.izeta_m : tw.izeta_mp;
            tt.z[p][a][k] = t[p * TW_SIZE + k];
        }
        if (threadIdx.x < NPRIMES * 2 * 8) {
            const uint32_t p = threadIdx.x >> 4, a = (threadIdx.x >> 3) & 1, l = threadIdx.x & 7;
            const uint32_t* t = a ? tw.invn_mp : tw.invn_m;
            tt.inv[p][a][l] = t[p * (MAX_LOG_N + 1) + l];
        }
        if (threadIdx.x == 0) s_bar = 0;
    }
    __syncthreads();
    const TailDesc& D = *reinterpret_cast<const TailDesc*>(dsh);
    {
        const TranscriptMsg& t0 = blockIdx.x == 0 ? D.ptx[0] : D.vs[0].tx;
        if (threadIdx.x < 8) {
            const uint32_t* e = reinterpret_cast<const uint32_t*>(jp(t0.epoch_seed, jo));
            const uint32_t* f = reinterpret_cast<const uint32_t*>(jp(t0.file_id, jo));
            reinterpret_cast<uint32_t*>(s_seeds)[threadIdx.x] = e[threadIdx.x];
            reinterpret_cast<uint32_t*>(s_seeds)[8 + threadIdx.x] = f[threadIdx.x];
        }
        if (threadIdx.x == 0) s_reject = t0.reject ? *jp(t0.reject, jo) : 0u;
    }
    __syncthreads();
    zq* const msgs = reinterpret_cast<zq*>(dsh + (TAIL_DESC_LDS + TAIL_TW_LDS) / 4);  // two messages
    zq* const stX = msgs + 2 * (size_t)msg_zq;
    zq* const stY = stX + TAIL_STATE_X;
    uint32_t* work = reinterpret_cast<uint32_t*>(stY + TAIL_STATE_Y);
    WgChal& wc = *reinterpret_cast<WgChal*>(work);
    const uint32_t S = D.nstages;
    uint32_t key[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) key[i] = b3::iv(i);  // transcript hashes are unkeyed
    if (blockIdx.x == 0) {  // the prover
        SubBarrier bar;
        bar.cnt = &s_bar;
        bar.target = 0;
        zq u = *jp(D.st[0].it.u_prev, jo);  // the inverses of every earlier stage
        TailProverLds L;
        L.msgs = msgs;
        L.msg_zq = msg_zq;
        L.stX = stX;
        L.stY = stY;
        L.work = work;
        L.seeds = s_seeds;
        L.reject = &s_reject;
        L.s_xi = &s_xi;
        L.s_u = &s_u;
        tail_prover_stage<true>(0, D, jo, L, wc, tt, bar, key, u);
        for (uint32_t s = 1; s < S; ++s) tail_prover_stage<false>(s, D, jo, L, wc, tt, bar, key, u);
        if (threadIdx.x == 0) store_zq(jp(D.st[S - 1].w.dst, jo), load_zq(S & 1 ? stY : stX));  // m*
        return;
    }
    // the verifier: stage s's challenge needs the prover's stage s - 1 (its cross terms), the
    // update needs stage s (the inverse candidate)
    const uint32_t* ctr = jp(D.ctr, jo);
    uint32_t* vflags = jp(D.vflags, jo);
    for (uint32_t s = 0; s < S; ++s) {
        const TailVerStep& vs = D.vs[s];
        if (s) tail_wait(ctr, s, vflags);
        TranscriptMsg t = vs.tx;
        tx_to_job(t, jo);
        t.epoch_seed = s_seeds;
        t.file_id = s_seeds + 32;
        t.reject = &s_reject;
        const size_t len = tx_message_len(t.label_len, t.lens);
        const zq x = wg_tx_challenge(t, len, key, 0, wc, jp(vs.scal, jo), vflags, 0);
        tail_wait(ctr, s + 1, vflags);
        tail_com_update(vs, x, jp(vs.hint, jo) + hint_off, jo);
        __syncthreads();  // the next message reads the updated commitment
    }
}

size_t tail_lds_bytes(const TailDesc& d, uint32_t* msg_zq) {
    size_t work = sizeof(WgChal), msg = 0;
    for (uint32_t s = 0; s < d.nstages; ++s) {
        const SmallStage& st = d.st[s];
        const size_t w = 4 * tail_stage_words(st, tail_group(st));
        if (w > work) work = w;
        if (tail_msg_zq(st) > msg) msg = tail_msg_zq(st);
    }
    *msg_zq = (uint32_t)msg;
    return TAIL_DESC_LDS + TAIL_TW_LDS + 2 * 16 * msg + 16 * (TAIL_STATE_X + TAIL_STATE_Y) + work;
}

hipError_t tail_set_lds_limit(size_t lds_bytes) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tail), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds_bytes);
}

void launch_tail(const TailDesc* d_desc, size_t lds_bytes, uint32_t msg_zq, int verifier, int hint_off,
                 const NttTables& tw, hipStream_t s, const Jobs& jb) {
    if (jb.n == 0) return;
    hipLaunchKernelGGL(k_tail, dim3(1 + (verifier ? 1 : 0), jb.n), dim3(TAIL_THREADS), lds_bytes, s, d_desc, tw,
                       jb.stride, hint_off, msg_zq);
}

// Z_q-only row production (the verifier's generator folds, verifier.rs:141-179)
__global__ void k_fold_rows_zq(RowSrc src, size_t total, int log_m, size_t js) {
    const size_t gid = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= total) return;
    const size_t jo = job_off(js);
    if (!src.a_shared) src.a = jp(src.a, jo);
    src.dst = jp(src.dst, jo);
    src.scal = jp(src.scal, jo);
    src.gseed = jp(src.gseed, jo);
    const size_t row = gid >> log_m;
    const uint32_t c = (uint32_t)(gid & (((size_t)1 << log_m) - 1));
    store_zq(src.dst + gid, row_value(src, src.gseed, row, c, log_m));
}

void launch_fold_rows_zq(const RowSrc& src, size_t rows, int log_m, hipStream_t s, const Jobs& jb) {
    const size_t total = rows << log_m;
    if (!total || jb.n == 0) return;
    hipLaunchKernelGGL(k_fold_rows_zq, dim3((unsigned)((total + 255) / 256), jb.n), dim3(256), 0, s, src, total,
                       log_m, jb.stride);
}

// ---------------------------------------------------------------------------------------
// The verifier's terminal generator (verifier.rs:57-134) without folding the matrix round by
// round. Both folds are linear. The vector fold (verifier.rs:141-151) sends row i to row
// i mod (rows/2), scaled by x^-1 when i is in the right half. The poly fold g_e + y^-1 X g_o in
// Z_q[X]/(X^L + 1) (verifier.rs:153-179) sends coefficient c of a degree-2L element to c/2 when
// c is even, to (c+1)/2 scaled by y^-1 when c is odd and below 2L-1, and to 0 scaled by -y^-1
// when c = 2L-1. So after the last round
//   g_d[0]_k = sum_i w_i sum_c u_c g[i][k][c],
// w_i the product of x_t^-1 over the vector rounds t that found row i in the right half (round
// t reads bit log2(rows) - 1 - t of i), u_c the product of the poly rounds' factors along c's
// path. Both weights factor at any round T: the rounds before T fold the matrix to
//   H[i_lo][k][c_T] = sum_{i_hi} w_early(i_hi) sum_{c -> c_T} u_early(c) g[i_hi * R + i_lo][k][c]
// (R = rows >> early vector folds; c_T = c's position after the early poly folds), and the rounds
// from T on finish on H with their own weights. The early part reads the matrix once and runs as
// soon as the early challenges are known, while the proof's last rounds are still in flight; the
// late part is one small kernel at the end. chal holds (x, x^-1, y, y^-1) per round; the masks
// say which rounds folded rows (vec) and coefficients (poly), bit r for round r.
// ---------------------------------------------------------------------------------------

// product of x_r^-1 over the vector rounds that see row i in the right half; the first of the
// masked rounds reads bit nbits - 1 of i
__device__ zq row_weight(uint32_t i, int nbits, const zq* chal, uint32_t vmask) {
    zq acc = zq_from_u64(1);
    int bit = nbits - 1;
    for (uint32_t r = 0; vmask >> r; ++r) {
        if (!((vmask >> r) & 1)) continue;
        if ((i >> bit) & 1) acc = zq_mul(acc, load_zq(chal + 4 * r + 1));
        --bit;
    }
    return acc;
}

// position of coefficient c (of a length-len element) after the masked poly rounds; `wrapped`
// when its path took the X^L = -1 wrap (every later position is then 0)
__device__ __forceinline__ uint32_t coeff_path(uint32_t c, uint32_t len, uint32_t pmask, bool& wrapped) {
    wrapped = false;
    for (uint32_t r = 0; pmask >> r; ++r) {
        if (!((pmask >> r) & 1)) continue;
        if (c & 1) {
            const bool w = c == len - 1;
            wrapped = wrapped || w;
            c = w ? 0 : (c + 1) >> 1;
        } else {
            c >>= 1;
        }
        len >>= 1;
    }
    return c;
}

// the poly rounds' factor along c's path
__device__ zq coeff_weight(uint32_t c, uint32_t len, const zq* chal, uint32_t pmask) {
    zq acc = zq_from_u64(1);
    for (uint32_t r = 0; pmask >> r; ++r) {
        if (!((pmask >> r) & 1)) continue;
        if (c & 1) {
            const zq yi = load_zq(chal + 4 * r + 3);
            const bool wrap = c == len - 1;
            acc = zq_mul(acc, wrap ? zq_neg(yi) : yi);
            c = wrap ? 0 : (c + 1) >> 1;
        } else {
            c >>= 1;
        }
        len >>= 1;
    }
    return acc;
}

constexpr uint32_t EARLY_TILE = 1024;  // coefficients staged in LDS per step

uint32_t early_chunks(int log_n) { return log_n > 8 ? 1u << (log_n - 8) : 1u; }

// One workgroup per (i_lo, k) and chunk z of 256 coefficients (blockIdx.z; one chunk below 512):
// the chunk's partial sums H[z][(i_lo*K + k)*n_t + j], which the late fold adds up. The positions
// c -> c_T are nondecreasing in c up to the first coefficient whose path wraps and 0 from there
// on, so position j collects the run [lo_j, lo_{j+1}) (and position 0 also the wrapped tail), found
// by bisection. Splitting the coefficients puts several workgroups on each CU, enough loads in
// flight to stream the matrix.
__global__ void __launch_bounds__(256) k_gen_contract_early(const zq* __restrict__ G, size_t gjs, const zq* chal,
                                                            uint32_t vmask, uint32_t pmask, int log_hi, int log_lo,
                                                            int log_n, int log_nt, zq* __restrict__ H, size_t js) {
    __shared__ zq w[EARLY_MAX_HI];
    __shared__ zq sp[EARLY_TILE];
    const size_t jo = job_off(js);
    chal = jp(chal, jo);
    H = jp(H, jo) + ((size_t)blockIdx.z * ((size_t)K << log_lo) << log_nt);
    G = jp(G, (size_t)blockIdx.y * gjs);
    const uint32_t i_lo = blockIdx.x / K, k = blockIdx.x % K;
    const uint32_t n_hi = 1u << log_hi, n = 1u << log_n, nt = 1u << log_nt, t = threadIdx.x;
    for (uint32_t h = t; h < n_hi; h += blockDim.x) w[h] = row_weight(h, log_hi, chal, vmask);
    // run bounds of position t (t < nt): first coefficient of the run and one past its end
    uint32_t lo = 0, hi = 0, wrap_start = n;
    if (t < nt) {
        uint32_t a = 0, b = n;  // first c whose path wraps
        while (a < b) {
            const uint32_t mid = (a + b) >> 1;
            bool wr;
            (void)coeff_path(mid, n, pmask, wr);
            if (wr) b = mid; else a = mid + 1;
        }
        wrap_start = a;
        auto first_at_least = [&](uint32_t j) {  // first c < wrap_start with c_T >= j
            uint32_t x = 0, y = wrap_start;
            while (x < y) {
                const uint32_t mid = (x + y) >> 1;
                bool wr;
                if (coeff_path(mid, n, pmask, wr) >= j) y = mid; else x = mid + 1;
            }
            return x;
        };
        lo = first_at_least(t);
        hi = t + 1 < nt ? first_at_least(t + 1) : wrap_start;
    }
    __syncthreads();
    zq acc = zq_zero();
    const size_t row_stride = ((size_t)K << log_n) << log_lo;  // from row block h to h + 1
    const zq* g0 = G + (((size_t)i_lo * K + k) << log_n);
    const uint32_t cw = n / gridDim.z, cbeg = blockIdx.z * cw, cend = cbeg + cw;  // this chunk
    for (uint32_t c0 = cbeg; c0 < cend; c0 += EARLY_TILE) {
        const uint32_t tl = cend - c0 < EARLY_TILE ? cend - c0 : EARLY_TILE;
        for (uint32_t cc = t; cc < tl; cc += blockDim.x) {
            const uint32_t c = c0 + cc;
            zq s = zq_zero();
            // four row blocks' loads are issued before their multiply-adds (the scheduler barrier
            // keeps them there), so a thread has four coefficients in flight
            uint32_t h = 0;
            for (; h + 3 < n_hi; h += 4) {
                zq gv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) gv[u] = load_zq(g0 + (h + u) * row_stride + c);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int u = 0; u < 4; ++u) s = zq_add(s, zq_mul(w[h + u], gv[u]));
            }
            for (; h < n_hi; ++h) s = zq_add(s, zq_mul(w[h], load_zq(g0 + h * row_stride + c)));
            sp[cc] = zq_mul(s, coeff_weight(c, n, chal, pmask));
        }
        __syncthreads();
        if (t < nt) {
            const uint32_t a = lo > c0 ? lo : c0, b = hi < c0 + tl ? hi : c0 + tl;
            for (uint32_t c = a; c < b; ++c) acc = zq_add(acc, sp[c - c0]);
            if (t == 0)
                for (uint32_t c = (wrap_start > c0 ? wrap_start : c0); c < c0 + tl; ++c) acc = zq_add(acc, sp[c - c0]);
        }
        __syncthreads();
    }
    if (t < nt) store_zq(H + (((size_t)i_lo * K + k) << log_nt) + t, acc);
}

// H[0][e] += H[1][e] + ... + H[nch - 1][e]: the early fold's chunk partials into its first block
__global__ void __launch_bounds__(256) k_gen_sum_chunks(zq* __restrict__ H, uint32_t count, uint32_t nch, size_t js) {
    const uint32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e >= count) return;
    H = jp(H, job_off(js));
    zq acc = load_zq(H + e);
    for (uint32_t z = 1; z < nch; ++z) acc = zq_add(acc, load_zq(H + (size_t)z * count + e));
    store_zq(H + e, acc);
}

void launch_gen_contract_early(const zq* G, bool g_shared, const zq* chal, uint32_t vmask, uint32_t pmask,
                               int log_hi, int log_lo, int log_n, int log_nt, zq* H, hipStream_t s,
                               const Jobs& jb) {
    if (jb.n == 0) return;
    const uint32_t nch = early_chunks(log_n);
    hipLaunchKernelGGL(k_gen_contract_early, dim3((unsigned)(K << log_lo), jb.n, nch), dim3(256), 0, s, G,
                       g_shared ? (size_t)0 : jb.stride, chal, vmask, pmask, log_hi, log_lo, log_n, log_nt, H,
                       jb.stride);
    if (nch > 1) {
        const uint32_t count = ((uint32_t)K << log_lo) << log_nt;
        hipLaunchKernelGGL(k_gen_sum_chunks, dim3((count + 255) / 256, jb.n), dim3(256), 0, s, H, count, nch,
                           jb.stride);
    }
}

// g_d[0]_k from H (rows x K x n, n and rows <= 32: what the late rounds fold) with the late
// rounds' weights, and the final opening check (verifier.rs:122-134): flags |= bit when
// g_d[0]_k * opening != C_k. One workgroup per k.
__global__ void __launch_bounds__(256) k_gen_contract_late(const zq* __restrict__ H, size_t hjs, const zq* chal,
                                                           uint32_t vmask, uint32_t pmask, int log_rows, int log_n,
                                                           const zq* com, const zq* opening, uint32_t* flags,
                                                           uint32_t bit, zq* gfin, size_t js) {
    __shared__ zq wr[32], wc[32], red[256];
    const size_t jo = job_off(js);
    chal = jp(chal, jo);
    H = jp(H, (size_t)blockIdx.y * hjs);
    const uint32_t k = blockIdx.x, t = threadIdx.x, rows = 1u << log_rows, n = 1u << log_n;
    if (t < rows) wr[t] = row_weight(t, log_rows, chal, vmask);
    if (t >= 64 && t - 64 < n) wc[t - 64] = coeff_weight(t - 64, n, chal, pmask);
    __syncthreads();
    zq acc = zq_zero();
    for (uint32_t e = t; e < rows * n; e += blockDim.x) {
        const uint32_t i = e >> log_n, c = e & (n - 1);
        acc = zq_add(acc, zq_mul(zq_mul(wr[i], wc[c]), load_zq(H + (((size_t)i * K + k) << log_n) + c)));
    }
    red[t] = acc;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (t < h) red[t] = zq_add(red[t], red[t + h]);
        __syncthreads();
    }
    if (t == 0) {
        store_zq(jp(gfin, jo) + k, red[0]);
        if (!zq_eq(zq_mul(red[0], load_zq(jp(opening, jo))), load_zq(jp(com, jo) + k))) atomicOr(jp(flags, jo), bit);
    }
}

void launch_gen_contract_late(const zq* H, bool h_shared, const zq* chal, uint32_t vmask, uint32_t pmask,
                              int log_rows, int log_n, const zq* com, const zq* opening, uint32_t* flags,
                              uint32_t bit, zq* gfin, hipStream_t s, const Jobs& jb) {
    if (jb.n == 0) return;
    hipLaunchKernelGGL(k_gen_contract_late, dim3(K, jb.n), dim3(256), 0, s, H, h_shared ? (size_t)0 : jb.stride,
                       chal, vmask, pmask, log_rows, log_n, com, opening, flags, bit, gfin, jb.stride);
}

// ---------------------------------------------------------------------------------------
// commitment bookkeeping
// ---------------------------------------------------------------------------------------
__global__ void k_com_update(const zq* c, const zq* l, const zq* r, zq* scal, int poly, int log_m, zq* out,
                             const zq* hint, size_t js) {
    const size_t m = (size_t)1 << log_m;
    const size_t gid = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= K * m) return;
    const size_t jo = job_off(js);
    c = jp(c, jo);
    l = jp(l, jo);
    r = jp(r, jo);
    scal = jp(scal, jo);
    out = jp(out, jo);
    const zq s = scal[0];
    zq sinv;
    if (hint) {  // uniform: every thread checks the same candidate
        sinv = load_zq(jp(hint, jo));
        if (!zq_eq(zq_mul(s, sinv), zq_from_u64(1))) sinv = zq_inv_divsteps(s);
        if (gid == 0) scal[1] = sinv;
    } else {
        sinv = scal[1];
    }
    const size_t k = gid >> log_m, i = gid & (m - 1);
    const zq base = poly ? load_zq(c + k * (2 * m) + 2 * i) : load_zq(c + gid);
    zq v = zq_add(base, zq_mul(load_zq(l + gid), sinv));
    v = zq_add(v, zq_mul(load_zq(r + gid), s));
    store_zq(out + gid, v);
}

void launch_com_update(const zq* c, const zq* l, const zq* r, zq* scal, int poly, int log_m, zq* out,
                       hipStream_t s, const Jobs& jb, const zq* hint) {
    const size_t n = (size_t)K << log_m;
    if (jb.n == 0) return;
    hipLaunchKernelGGL(k_com_update, dim3((unsigned)((n + 255) / 256), jb.n), dim3(256), 0, s, c, l, r, scal, poly,
                       log_m, out, hint, jb.stride);
}

__global__ void k_result_head(uint32_t* head, const uint32_t* pflags, const uint32_t* vflags,
                              const uint32_t* file_id, const uint32_t* mstar, size_t js) {
    const size_t jo = job_off(js);
    const uint32_t t = threadIdx.x;
    uint32_t w = 0;
    if (t == RESULT_HEAD_PFLAGS)
        w = jp(pflags, jo)[0];
    else if (t == RESULT_HEAD_VFLAGS || t == RESULT_HEAD_VFLAGS + 1)
        w = jp(vflags, jo)[t - RESULT_HEAD_VFLAGS];
    else if (t >= RESULT_HEAD_FID && t < RESULT_HEAD_FID + 8)
        w = jp(file_id, jo)[t - RESULT_HEAD_FID];
    else if (t >= RESULT_HEAD_MSTAR && t < RESULT_HEAD_MSTAR + 4)
        w = jp(mstar, jo)[t - RESULT_HEAD_MSTAR];
    if (t < RESULT_HEAD_BYTES / 4) jp(head, jo)[t] = w;
}

void launch_result_head(uint32_t* head, const uint32_t* pflags, const uint32_t* vflags, const uint8_t* file_id,
                        const zq* mstar, hipStream_t s, const Jobs& jb) {
    if (jb.n == 0) return;
    hipLaunchKernelGGL(k_result_head, dim3(1, jb.n), dim3(64), 0, s, head, pflags, vflags,
                       reinterpret_cast<const uint32_t*>(file_id), reinterpret_cast<const uint32_t*>(mstar),
                       jb.stride);
}

__global__ void k_fill32(uint32_t* __restrict__ p, size_t n, uint32_t v, size_t js) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) jp(p, job_off(js))[i] = v;
}

void launch_fill32(uint32_t* p, size_t n, uint32_t v, hipStream_t s, const Jobs& jb) {
    if (!n || jb.n == 0) return;
    hipLaunchKernelGGL(k_fill32, dim3((unsigned)((n + 255) / 256), jb.n), dim3(256), 0, s, p, n, v, jb.stride);
}

__global__ void k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n, size_t js) {
    dst = jp(dst, job_off(js));
    src = jp(src, job_off(js));
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

void launch_copy16(void* dst, const void* src, size_t n16, hipStream_t s, const Jobs& jb) {
    if (!n16 || jb.n == 0) return;
    const size_t blocks = std::min<size_t>((n16 + 255) / 256, 4096);
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks, jb.n), dim3(256), 0, s, static_cast<uint4*>(dst),
                       static_cast<const uint4*>(src), n16, jb.stride);
}

// Batched file copy-in: job b's header holds (length, source address); the bytes go to the
// job's file buffer. A zero source means the file is already in place.
__global__ void k_copy_in(const uint64_t* __restrict__ meta, uint8_t* __restrict__ dst, size_t cap, size_t js) {
    const size_t jo = job_off(js);
    meta = jp(meta, jo);
    dst = jp(dst, jo);
    const uint8_t* src = reinterpret_cast<const uint8_t*>(meta[1]);
    const size_t len = (size_t)min((uint64_t)cap, meta[0]);
    const size_t i = ((size_t)blockIdx.x * 256 + threadIdx.x) * 16;
    if (!src || i >= len) return;
    if (i + 16 <= len && ((reinterpret_cast<uintptr_t>(src) & 15) == 0)) {
        *reinterpret_cast<uint4*>(dst + i) = *reinterpret_cast<const uint4*>(src + i);
    } else {
        const size_t e = min(len, i + 16);
        for (size_t b = i; b < e; ++b) dst[b] = src[b];
    }
}

void launch_copy_in(const uint64_t* meta, uint8_t* dst, size_t cap, hipStream_t s, const Jobs& jb) {
    const size_t threads = (cap + 15) / 16;
    if (!threads || jb.n == 0) return;
    hipLaunchKernelGGL(k_copy_in, dim3((unsigned)((threads + 255) / 256), jb.n), dim3(256), 0, s, meta, dst, cap,
                       jb.stride);
}

__global__ void k_zq_add(const zq* a, const zq* b, zq* out, size_t n) {
    const size_t gid = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= n) return;
    store_zq(out + gid, zq_add(load_zq(a + gid), load_zq(b + gid)));
}

void launch_zq_add(const zq* a, const zq* b, zq* out, size_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_zq_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, out, n);
}

__global__ void k_com_check(const zq* a, const zq* b, size_t n, uint32_t* flags, uint32_t bit, size_t js) {
    const size_t gid = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= n) return;
    const size_t jo = job_off(js);
    a = jp(a, jo);
    b = jp(b, jo);
    flags = jp(flags, jo);
    if (!zq_eq(load_zq(a + gid), load_zq(b + gid))) atomicOr(flags, bit);
}

void launch_com_check(const zq* a, const zq* b, size_t n, uint32_t* flags, uint32_t bit,
                      hipStream_t s, const Jobs& jb) {
    if (!n || jb.n == 0) return;
    hipLaunchKernelGGL(k_com_check, dim3((unsigned)((n + 255) / 256), jb.n), dim3(256), 0, s, a, b, n,
                       flags, bit, jb.stride);
}

// ---------------------------------------------------------------------------------------
// witness expansion (io.rs:51-54) and generator expansion (sis.rs:212-245)
// ---------------------------------------------------------------------------------------
__global__ void k_expand(const uint8_t* __restrict__ file, size_t len, zq* __restrict__ out,
                         size_t capacity, size_t js, const uint64_t* dlen) {
    const size_t j = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= capacity) return;
    const size_t jo = job_off(js);
    file = jp(file, jo);
    out = jp(out, jo);
    if (dlen) len = (size_t)min((uint64_t)len, *jp(dlen, jo));
    uint32_t w[4] = {0, 0, 0, 0};
    const size_t off = j * BYTES_PER_COEFF;
    if (off < len) {
        const size_t n = len - off < BYTES_PER_COEFF ? len - off : BYTES_PER_COEFF;
        for (size_t s = 0; s < n; ++s) w[s >> 2] |= (uint32_t)file[off + s] << (8 * (s & 3));
    }
    store_zq(out + j, zq{{w[0], w[1], w[2], w[3]}});  // < 2^104 < q: already canonical
}

void launch_expand(const uint8_t* file, size_t len, zq* out, size_t capacity, hipStream_t s, const Jobs& jb,
                   const uint64_t* dlen) {
    if (jb.n == 0) return;
    hipLaunchKernelGGL(k_expand, dim3((unsigned)((capacity + 255) / 256), jb.n), dim3(256), 0, s, file, len,
                       out, capacity, jb.stride, dlen);
}

__global__ void k_gen_generators(const uint32_t* gseed, size_t total, int log_m, zq* __restrict__ out,
                                 size_t js) {
    const size_t jo = job_off(js);
    uint32_t key[8];
    const uint32_t* g = jp(gseed, jo);
#pragma unroll
    for (int i = 0; i < 8; ++i) key[i] = g[i];
    out = jp(out, jo);
    const size_t m = (size_t)1 << log_m;
    const size_t gid = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= total) return;
    const size_t row = gid >> log_m, c = gid & (m - 1);
    store_zq(out + gid, gen_coeff(key, row / K, (uint32_t)(row % K), c));
}

void launch_gen_generators(const uint32_t* gseed, size_t rows, int log_m, zq* out, hipStream_t s, const Jobs& jb) {
    const size_t total = rows << log_m;
    if (!total || jb.n == 0) return;
    hipLaunchKernelGGL(k_gen_generators, dim3((unsigned)((total + 255) / 256), jb.n), dim3(256), 0, s, gseed, total,
                       log_m, out, jb.stride);
}

// ---------------------------------------------------------------------------------------
// The reference-compatible drop-ins (damascus_cuda_fold_batch / damascus_cuda_ntt_batch,
// src/utils/gpu.rs:69-93), on buffers the host side keeps per device.
// ---------------------------------------------------------------------------------------
// out = left + c * right mod q over SoA 64-bit halves (fold_kernel.cu:189-206). Inputs must be
// canonical (< q, as every Fp is); any that is not sets *flag and the call is declined.
__global__ void __launch_bounds__(256) k_fold_soa(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                                  size_t len, uint64_t c_lo, uint64_t c_hi, uint32_t* flag) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= len) return;
    const uint64_t ll = in[i], lh = in[len + i], rl = in[2 * len + i], rh = in[3 * len + i];
    const zq l{{(uint32_t)ll, (uint32_t)(ll >> 32), (uint32_t)lh, (uint32_t)(lh >> 32)}};
    const zq r{{(uint32_t)rl, (uint32_t)(rl >> 32), (uint32_t)rh, (uint32_t)(rh >> 32)}};
    const zq c{{(uint32_t)c_lo, (uint32_t)(c_lo >> 32), (uint32_t)c_hi, (uint32_t)(c_hi >> 32)}};
    if (zq_ge_q(l) || zq_ge_q(r)) atomicOr(flag, 1u);
    const zq o = zq_add(l, zq_mul(r, c));
    out[i] = (uint64_t)o.w[0] | ((uint64_t)o.w[1] << 32);
    out[len + i] = (uint64_t)o.w[2] | ((uint64_t)o.w[3] << 32);
}

void launch_fold_soa(const uint64_t* in, uint64_t* out, size_t len, uint64_t c_lo, uint64_t c_hi, uint32_t* flag,
                     hipStream_t s) {
    if (!len) return;
    hipLaunchKernelGGL(k_fold_soa, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, in, out, len, c_lo, c_hi,
                       flag);
}

// The cyclic product of damascus_cuda_ntt_batch (ntt_kernel.cu:106-293) for each of 8 primes, one
// workgroup per prime with both operands in LDS: bit-reversed load, radix-2 DIT stages whose
// twiddles are the powers of the caller's stage roots (stage s: root_s^j for j < 2^s, in the
// table as Montgomery pairs at [2^s - 1 + j]), pointwise product, bit reversal, the inverse stages
// with the caller's inverse roots, times the caller's inv_size. The same modular operations in
// the same order as the reference, on residues in [0, p). An input >= p sets *flag (the call is
// declined). Table per prime: forward pairs then inverse pairs, 4 (n - 1) words.
constexpr int CYC_THREADS = 1024;
__device__ __forceinline__ uint32_t brv_bits(uint32_t x, int bits) { return bits ? __brev(x) >> (32 - bits) : 0u; }

__global__ void __launch_bounds__(CYC_THREADS) k_cyc_lds(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b,
                                                         uint64_t* __restrict__ out, int log_n,
                                                         const CycPrime* __restrict__ cp, const uint32_t* __restrict__ tw,
                                                         uint32_t* flag) {
    extern __shared__ uint32_t sh[];
    const int q = blockIdx.x, n = 1 << log_n, half_n = n >> 1;
    const CycPrime c = cp[q];
    uint32_t* sa = sh;
    uint32_t* sb = sh + n;
    const uint64_t* aq = a + (size_t)q * n;
    const uint64_t* bq = b + (size_t)q * n;
    uint32_t bad = 0;
    for (int i = threadIdx.x; i < n; i += CYC_THREADS) {
        const uint64_t x = aq[i], y = bq[i];
        bad |= (x >= c.p || y >= c.p) ? 1u : 0u;
        const uint32_t r = brv_bits((uint32_t)i, log_n);
        sa[r] = (uint32_t)x;
        sb[r] = (uint32_t)y;
    }
    if (bad) atomicOr(flag, 1u);
    __syncthreads();
    const uint32_t* fw = tw + (size_t)q * 4 * (n - 1);
    const uint32_t* iw = fw + 2 * (n - 1);
    for (int st = 0; st < log_n; ++st) {
        const int half = 1 << st;
        for (int t = threadIdx.x; t < half_n; t += CYC_THREADS) {
            const int off = t & (half - 1), left = ((t >> st) << (st + 1)) + off;
            const uint32_t w = fw[2 * (half - 1 + off)], wp = fw[2 * (half - 1 + off) + 1];
            ct_butterfly(sa[left], sa[left + half], w, wp, c.p);
            ct_butterfly(sb[left], sb[left + half], w, wp, c.p);
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += CYC_THREADS)  // a * b: (a * b / R) * R, in place in sa
        sa[i] = mulc(redc((uint64_t)sa[i] * sb[i], c.p, c.pinv), c.r2m, c.r2mp, c.p);
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += CYC_THREADS) sb[brv_bits((uint32_t)i, log_n)] = sa[i];  // bit reversal
    __syncthreads();
    for (int st = 0; st < log_n; ++st) {
        const int half = 1 << st;
        for (int t = threadIdx.x; t < half_n; t += CYC_THREADS) {
            const int off = t & (half - 1), left = ((t >> st) << (st + 1)) + off;
            ct_butterfly(sb[left], sb[left + half], iw[2 * (half - 1 + off)], iw[2 * (half - 1 + off) + 1], c.p);
        }
        __syncthreads();
    }
    uint64_t* oq = out + (size_t)q * n;
    for (int i = threadIdx.x; i < n; i += CYC_THREADS) oq[i] = mulc(sb[i], c.invs_m, c.invs_mp, c.p);
}

void launch_cyc_lds(const uint64_t* a, const uint64_t* b, uint64_t* out, int log_n, const CycPrime* cp,
                    const uint32_t* tw, uint32_t* flag, hipStream_t s) {
    hipLaunchKernelGGL(k_cyc_lds, dim3(NPRIMES), dim3(CYC_THREADS), (size_t)2 * ((size_t)1 << log_n) * 4, s, a, b, out,
                       log_n, cp, tw, flag);
}

// ---------------------------------------------------------------------------------------
// ring products (ragged batches grouped by degree)
// ---------------------------------------------------------------------------------------
__global__ void k_pointwise(uint32_t* a, const uint32_t* b, size_t total, int log_m) {
    const size_t gid = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= total) return;
    const uint32_t p = (uint32_t)((gid >> log_m) % NPRIMES);
    const uint64_t pr = (uint64_t)a[gid] * b[gid];
    a[gid] = (uint32_t)(pr % c_pc[p].p);
}

void launch_pointwise(uint32_t* a, const uint32_t* b, size_t slices, int log_m, hipStream_t s) {
    const size_t total = slices * NPRIMES << log_m;
    if (!total) return;
    hipLaunchKernelGGL(k_pointwise, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a, b, total,
                       log_m);
}

// ---------------------------------------------------------------------------------------
// Persistent ragged ring products: one launch for a batch of negacyclic products of mixed
// degrees. Workgroups pull products from a queue ordered largest first (a device counter);
// per product, G primes at a time (G * n <= RING_LDS_WORDS): reduce both operands into LDS,
// forward transforms, pointwise product, inverse transform, residues to scratch; then one
// centred CRT lift per output coefficient. Loop bounds and the exit test are uniform over the
// workgroup.
// ---------------------------------------------------------------------------------------
constexpr int RING_THREADS = 256;

__device__ __forceinline__ void lds_fwd_layers(uint32_t* s, int G, int log_n, int p0, const NttTables& tw) {
    const int n = 1 << log_n, half = n >> 1;
    for (int len = half; len >= 1; len >>= 1) {
        const int lg = __builtin_ctz(len);
        for (int t = threadIdx.x; t < G * half; t += RING_THREADS) {
            const int q = t >> (log_n - 1), b = t & (half - 1);
            const int blk = b >> lg;
            const int j = (blk << (lg + 1)) | (b & (len - 1));
            const int kz = (half >> lg) + blk;
            const int pi = p0 + q;
            const uint32_t p = c_pc[pi].p;
            uint32_t* sq = s + q * n;
            const uint32_t u = sq[j];
            const uint32_t v = mulc(sq[j + len], tw.zeta_m[pi * TW_SIZE + kz], tw.zeta_mp[pi * TW_SIZE + kz], p);
            sq[j] = addp(u, v, p);
            sq[j + len] = subp(u, v, p);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void lds_inv_layers(uint32_t* s, int G, int log_n, int p0, const NttTables& tw) {
    const int n = 1 << log_n, half = n >> 1;
    for (int len = 1; len < n; len <<= 1) {
        const int lg = __builtin_ctz(len);
        for (int t = threadIdx.x; t < G * half; t += RING_THREADS) {
            const int q = t >> (log_n - 1), b = t & (half - 1);
            const int blk = b >> lg;
            const int j = (blk << (lg + 1)) | (b & (len - 1));
            const int kz = (half >> lg) + blk;
            const int pi = p0 + q;
            const uint32_t p = c_pc[pi].p;
            uint32_t* sq = s + q * n;
            const uint32_t u = sq[j], v = sq[j + len];
            sq[j] = addp(u, v, p);
            sq[j + len] = mulc(subp(u, v, p), tw.izeta_m[pi * TW_SIZE + kz], tw.izeta_mp[pi * TW_SIZE + kz], p);
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(RING_THREADS) k_ring_mul_persistent(
    const zq* __restrict__ A, const zq* __restrict__ B, zq* __restrict__ O, const uint64_t* __restrict__ offs,
    const uint32_t* __restrict__ logn, const uint32_t* __restrict__ order, uint32_t count, uint32_t* counter,
    uint32_t* __restrict__ res, NttTables tw, int lds_words) {
    extern __shared__ uint32_t sh[];
    __shared__ uint32_t item;
    uint32_t* sa = sh;
    uint32_t* sb = sh + lds_words / 2;
    for (;;) {
        if (threadIdx.x == 0) item = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t it = item;
        __syncthreads();  // every thread has read `item` before it is replaced
        if (it >= count) break;
        const uint32_t idx = order[it];
        const int log_n = (int)logn[idx];
        const int n = 1 << log_n;
        const size_t off = offs[idx];
        const zq* a = A + off;
        const zq* b = B + off;
        uint32_t* r = res + off * NPRIMES;
        if (n == 1) {  // degree 1: the product of two constants
            if (threadIdx.x < NPRIMES) {
                const PrimeConsts pc = c_pc[threadIdx.x];
                const uint32_t x = zq_mod_p_centred(load_zq(a), pc), y = zq_mod_p_centred(load_zq(b), pc);
                const uint32_t t = redc((uint64_t)x * y, pc.p, pc.pinv);
                r[threadIdx.x] = redc((uint64_t)t * pc.two64, pc.p, pc.pinv);
            }
        } else {
            const int G = min(NPRIMES, (lds_words / 2) >> log_n);
            for (int p0 = 0; p0 < NPRIMES; p0 += G) {
                for (int t = threadIdx.x; t < G * n; t += RING_THREADS) {
                    const int q = t >> log_n, c = t & (n - 1);
                    const PrimeConsts pc = c_pc[p0 + q];
                    sa[t] = zq_mod_p_centred(load_zq(a + c), pc);
                    sb[t] = zq_mod_p_centred(load_zq(b + c), pc);
                }
                __syncthreads();
                lds_fwd_layers(sa, G, log_n, p0, tw);
                lds_fwd_layers(sb, G, log_n, p0, tw);
                for (int t = threadIdx.x; t < G * n; t += RING_THREADS) {
                    const PrimeConsts pc = c_pc[p0 + (t >> log_n)];
                    const uint32_t u = redc((uint64_t)sa[t] * sb[t], pc.p, pc.pinv);  // a*b*R^-1
                    sa[t] = redc((uint64_t)u * pc.two64, pc.p, pc.pinv);             // * R
                }
                __syncthreads();
                lds_inv_layers(sa, G, log_n, p0, tw);
                for (int t = threadIdx.x; t < G * n; t += RING_THREADS) {
                    const int q = t >> log_n, c = t & (n - 1);
                    const int pi = p0 + q;
                    r[pi * n + c] = mulc(sa[t], tw.invn_m[pi * (MAX_LOG_N + 1) + log_n],
                                         tw.invn_mp[pi * (MAX_LOG_N + 1) + log_n], c_pc[pi].p);
                }
                __syncthreads();  // LDS reused by the next prime group
            }
        }
        __threadfence_block();
        __syncthreads();  // residues of this product visible to the whole workgroup
        for (int c = threadIdx.x; c < n; c += RING_THREADS) {
            uint32_t rr[NPRIMES];
#pragma unroll
            for (int i = 0; i < NPRIMES; ++i) rr[i] = r[i * n + c];
            store_zq(O + off + c, crt_lift_residues(rr));
        }
    }
}

void launch_ring_mul_persistent(const zq* A, const zq* B, zq* O, const uint64_t* offs, const uint32_t* logn,
                                const uint32_t* order, uint32_t count, uint32_t* counter, uint32_t* res,
                                int max_log_n, const NttTables& tw, hipStream_t s) {
    if (!count) return;
    const int lds_words = 2 * std::max(4096, 1 << max_log_n);  // two operands, one prime group
    const unsigned wgs = std::min<unsigned>(count, 1024);
    launch_fill32(counter, 1, 0, s);
    hipLaunchKernelGGL(k_ring_mul_persistent, dim3(wgs), dim3(RING_THREADS), (size_t)lds_words * 4, s, A, B, O,
                       offs, logn, order, count, counter, res, tw, lds_words);
}

// ---------------------------------------------------------------------------------------
// LDS-resident ragged ring products (degrees 2^0 .. 2^12) in one persistent launch.
//
// A queue item is up to RPB products of one degree 2^LOG_M = E * TP. Each product's 8 prime slices
// are held by TP lanes each (TP <= 64, so a slice never leaves its wave). The workgroup stages an
// operand into LDS (each coefficient loaded once and reduced into all 8 primes); lane l of slice
// q then takes the coefficients l + TP*i (i < E) of its prime, runs the transform layers
// len >= TP in registers, swaps to the contiguous layout (E*l .. E*l+E-1) through its wave's part
// of LDS, and runs the layers len < TP. With both operands transformed, the pointwise product is
// scaled by 1/m; the inverse runs the same two passes backwards. The residues of all 8 primes
// then meet in LDS, where each output coefficient gets its centred CRT lift. The transposes are
// wave-local; operands are read once and nothing goes through global scratch.
// ---------------------------------------------------------------------------------------
constexpr int RING_LDS_THREADS = 512;
constexpr int RING_LDS_MAX_LOG = 12;
// 8 primes at degree 4096, one pad word per 64 coefficients: 130 KiB
constexpr uint32_t RING_LDS_WORDS = NPRIMES * ((1u << RING_LDS_MAX_LOG) + 64);

template <int E, int TP>
struct RingCfg {
    static constexpr int M = E * TP;
    static constexpr int LOG_E = __builtin_ctz(E);
    static constexpr int LOG_TP = __builtin_ctz(TP);
    static constexpr int LOG_M = LOG_E + LOG_TP;
    static constexpr int TPR = NPRIMES * TP;                // lanes per product
    static constexpr int RPB = RING_LDS_THREADS / TPR;     // products per item
    static constexpr int SLICE = M + TP;                    // words per prime slice, padded
    static_assert(TP <= E && TP <= 64 && TPR >= 64 && RPB >= 1, "ring blocking");
    static_assert(RPB * NPRIMES * SLICE <= (int)RING_LDS_WORDS, "ring LDS");
    static_assert(RPB <= 8, "an item's product offsets are staged in 8 LDS words");
    // LDS word of coefficient x of prime slice q: one pad word after every E coefficients, so the
    // contiguous pattern (lane l at E*l + e) has lane stride E + 1 (odd: distinct banks) and the
    // strided pattern (lane l at l + TP*i) is the lane's base plus a constant per i
    __device__ static __forceinline__ int slot(int q, int x) { return q * SLICE + x + (x >> LOG_E); }
    __device__ static __forceinline__ int strided(int q, int l, int i) {  // slot(q, l + TP * i)
        return q * SLICE + l + TP * i + (i >> (LOG_E - LOG_TP));
    }
    __device__ static __forceinline__ int contiguous(int q, int l, int e) {  // slot(q, E * l + e)
        return q * SLICE + (E + 1) * l + e;
    }
};

// orders this lane's LDS accesses against the rest of its wave's (the transposes stay in a wave)
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// keeps the compiler's scheduler from moving instructions across a phase boundary (loads hoisted
// from a later phase would hold registers through this one)
__device__ __forceinline__ void phase_boundary() { __builtin_amdgcn_sched_barrier(0); }

__device__ __forceinline__ void gs_butterfly(uint32_t& x, uint32_t& y, uint32_t w, uint32_t wp, uint32_t p) {
    const uint32_t u = x, v = y;
    x = addp(u, v, p);
    y = mulc(subp(u, v, p), w, wp, p);
}

// forward negacyclic transform of one prime slice: x[i] holds coefficient l + TP*i on entry and
// bit-reversed position E*l + i on exit
template <int E, int TP>
__device__ __forceinline__ void ring_fwd(uint32_t* S, int q, int l, uint32_t (&x)[E], const uint32_t* zt,
                                         uint32_t p, uint32_t pinv) {
    using C = RingCfg<E, TP>;
#pragma unroll
    for (int s = E / 2; s >= 1; s >>= 1) {
        phase_boundary();  // one layer's twiddles in registers at a time
#pragma unroll
        for (int b = 0; b < E / (2 * s); ++b) {
            const uint32_t w = zt[E / (2 * s) + b], wp = w * pinv;
#pragma unroll
            for (int k = 0; k < s; ++k) ct_butterfly(x[2 * s * b + k], x[2 * s * b + k + s], w, wp, p);
        }
    }
#pragma unroll
    for (int i = 0; i < E; ++i) S[C::strided(q, l, i)] = x[i];
    wave_lds_sync();
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = S[C::contiguous(q, l, e)];
    wave_lds_sync();  // every lane has read its elements before the slice's LDS is written again
#pragma unroll
    for (int len = TP / 2; len >= 1; len >>= 1) {
        phase_boundary();
        const int base = C::M / (2 * len) + (E * l) / (2 * len);
#pragma unroll
        for (int b = 0; b < E / (2 * len); ++b) {
            const uint32_t w = zt[base + b], wp = w * pinv;
#pragma unroll
            for (int k = 0; k < len; ++k) ct_butterfly(x[2 * len * b + k], x[2 * len * b + k + len], w, wp, p);
        }
    }
}

// the inverse (Gentleman-Sande, no 1/m): bit-reversed positions E*l + i in, coefficients
// l + TP*i out
template <int E, int TP>
__device__ __forceinline__ void ring_inv(uint32_t* S, int q, int l, uint32_t (&x)[E], const uint32_t* izt,
                                         uint32_t p, uint32_t pinv) {
    using C = RingCfg<E, TP>;
#pragma unroll
    for (int len = 1; len < TP; len <<= 1) {
        phase_boundary();
        const int base = C::M / (2 * len) + (E * l) / (2 * len);
#pragma unroll
        for (int b = 0; b < E / (2 * len); ++b) {
            const uint32_t w = izt[base + b], wp = w * pinv;
#pragma unroll
            for (int k = 0; k < len; ++k) gs_butterfly(x[2 * len * b + k], x[2 * len * b + k + len], w, wp, p);
        }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) S[C::contiguous(q, l, e)] = x[e];
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < E; ++i) x[i] = S[C::strided(q, l, i)];
    wave_lds_sync();
#pragma unroll
    for (int s = 1; s < E; s <<= 1) {
        phase_boundary();
#pragma unroll
        for (int b = 0; b < E / (2 * s); ++b) {
            const uint32_t w = izt[E / (2 * s) + b], wp = w * pinv;
#pragma unroll
            for (int k = 0; k < s; ++k) gs_butterfly(x[2 * s * b + k], x[2 * s * b + k + s], w, wp, p);
        }
    }
}

// An operand of every product of the item into LDS: each coefficient is loaded once (16-byte
// loads, consecutive threads on consecutive coefficients) and reduced into all 8 primes, stored at
// its slot in the product's region. The whole workgroup takes part. A thread issues all of its
// loads (at most 8) before reducing any, so the item waits for one load latency, not one per
// coefficient; `item_off` (LDS) holds each product's first coefficient.
template <int E, int TP>
__device__ __forceinline__ void ring_stage(uint32_t* sh, const zq* __restrict__ src, const uint64_t* item_off,
                                           uint32_t cnt, int tid) {
    using C = RingCfg<E, TP>;
    constexpr int ITERS = (C::RPB * C::M + RING_LDS_THREADS - 1) / RING_LDS_THREADS;
    const int total = (int)cnt * C::M;
    zq x[ITERS];
#pragma unroll
    for (int u = 0; u < ITERS; ++u) {
        const int idx = tid + u * RING_LDS_THREADS;
        if (idx < total) x[u] = load_zq(src + item_off[idx >> C::LOG_M] + (idx & (C::M - 1)));
    }
#pragma unroll
    for (int u = 0; u < ITERS; ++u) {
        const int idx = tid + u * RING_LDS_THREADS;
        if (idx < total) {
            const int prod = idx >> C::LOG_M, c = idx & (C::M - 1);
            uint32_t* S = sh + prod * NPRIMES * C::SLICE;
#pragma unroll
            for (int q = 0; q < NPRIMES; ++q) S[C::slot(q, c)] = zq_mod_p_centred(x[u], c_pc[q]);
        }
    }
}

// one queue item of `cnt` products of degree E*TP; ends with a workgroup barrier
template <int E, int TP>
__device__ void ring_item(uint32_t* sh, const zq* __restrict__ A, const zq* __restrict__ B, zq* __restrict__ O,
                          const uint64_t* item_off, uint32_t cnt, const NttTables& tw, int tid) {
    using C = RingCfg<E, TP>;
    const int rsub = tid / C::TPR, tr = tid % C::TPR;
    const bool valid = rsub < (int)cnt;  // uniform per wave (TPR >= 64)
    const int q = tr >> C::LOG_TP, l = tr & (TP - 1);
    uint32_t* S = sh + rsub * NPRIMES * C::SLICE;
    const PrimeConsts pc = c_pc[q];
    const uint32_t* zt = tw.zeta_m + q * TW_SIZE;
    uint32_t a[E], b[E];
    ring_stage<E, TP>(sh, A, item_off, cnt, tid);
    __syncthreads();
    if (valid) {
#pragma unroll
        for (int i = 0; i < E; ++i) a[i] = S[C::strided(q, l, i)];
        ring_fwd<E, TP>(S, q, l, a, zt, pc.p, pc.pinv);
    }
    phase_boundary();
    __syncthreads();  // every slice has read its elements of a: the regions take b
    ring_stage<E, TP>(sh, B, item_off, cnt, tid);
    __syncthreads();
    phase_boundary();
    if (valid) {
#pragma unroll
        for (int i = 0; i < E; ++i) b[i] = S[C::strided(q, l, i)];
        ring_fwd<E, TP>(S, q, l, b, zt, pc.p, pc.pinv);
        // a * b / m: redc gives a*b/R, then * (R/m) as a mulc constant (R^2/m mod p)
        const uint32_t im = tw.invn_m[q * (MAX_LOG_N + 1) + C::LOG_M];  // R/m
        const uint32_t sm = mulc(im, pc.two64, pc.two64 * pc.pinv, pc.p), smp = sm * pc.pinv;
#pragma unroll
        for (int e = 0; e < E; ++e) a[e] = mulc(redc((uint64_t)a[e] * b[e], pc.p, pc.pinv), sm, smp, pc.p);
        phase_boundary();
        ring_inv<E, TP>(S, q, l, a, tw.izeta_m + q * TW_SIZE, pc.p, pc.pinv);
#pragma unroll
        for (int i = 0; i < E; ++i) S[C::strided(q, l, i)] = a[i];
    }
    __syncthreads();  // every prime's residues of the item in LDS
    phase_boundary();
    if (valid) {
        const size_t off = item_off[rsub];
#pragma unroll 1
        for (int c = tr; c < C::M; c += C::TPR) {
            uint32_t r[NPRIMES];
#pragma unroll
            for (int i = 0; i < NPRIMES; ++i) r[i] = S[C::slot(i, c)];
            store_zq(O + off + c, crt_lift_residues(r));
        }
    }
    __syncthreads();  // LDS and the item word are free for the next item
}

// degrees 2^0 .. 2^5: one product, 8 * 2^log_n lanes, transforms layer by layer in LDS
__device__ void ring_item_small(uint32_t* sh, const zq* __restrict__ A, const zq* __restrict__ B,
                                zq* __restrict__ O, uint64_t off, int log_n, const NttTables& tw, int tid) {
    const int n = 1 << log_n, half = n >> 1;
    const int t = tid;
    uint32_t* sa = sh;
    uint32_t* sb = sh + NPRIMES * n;
    if (t < NPRIMES * n) {
        const int q = t >> log_n, c = t & (n - 1);
        sa[t] = zq_mod_p_centred(load_zq(A + off + c), c_pc[q]);
        sb[t] = zq_mod_p_centred(load_zq(B + off + c), c_pc[q]);
    }
    __syncthreads();
    for (int len = half; len >= 1; len >>= 1) {
        const int lg = __builtin_ctz(len);
        if (t < NPRIMES * half) {
            const int q = t >> (log_n - 1), b = t & (half - 1);
            const int blk = b >> lg, j = (blk << (lg + 1)) | (b & (len - 1)), kz = (half >> lg) + blk;
            const uint32_t p = c_pc[q].p, w = tw.zeta_m[q * TW_SIZE + kz], wp = tw.zeta_mp[q * TW_SIZE + kz];
            ct_butterfly(sa[q * n + j], sa[q * n + j + len], w, wp, p);
            ct_butterfly(sb[q * n + j], sb[q * n + j + len], w, wp, p);
        }
        __syncthreads();
    }
    if (t < NPRIMES * n) {
        const int q = t >> log_n;
        const PrimeConsts pc = c_pc[q];
        const uint32_t im = tw.invn_m[q * (MAX_LOG_N + 1) + log_n];
        const uint32_t sm = mulc(im, pc.two64, pc.two64 * pc.pinv, pc.p);
        sa[t] = mulc(redc((uint64_t)sa[t] * sb[t], pc.p, pc.pinv), sm, sm * pc.pinv, pc.p);
    }
    __syncthreads();
    for (int len = 1; len < n; len <<= 1) {
        const int lg = __builtin_ctz(len);
        if (t < NPRIMES * half) {
            const int q = t >> (log_n - 1), b = t & (half - 1);
            const int blk = b >> lg, j = (blk << (lg + 1)) | (b & (len - 1)), kz = (half >> lg) + blk;
            const uint32_t p = c_pc[q].p, w = tw.izeta_m[q * TW_SIZE + kz], wp = tw.izeta_mp[q * TW_SIZE + kz];
            gs_butterfly(sa[q * n + j], sa[q * n + j + len], w, wp, p);
        }
        __syncthreads();
    }
    if (t < n) {
        uint32_t r[NPRIMES];
#pragma unroll
        for (int i = 0; i < NPRIMES; ++i) r[i] = sa[i * n + t];
        store_zq(O + off + t, crt_lift_residues(r));
    }
    __syncthreads();
}

// items[k]: first index into `order` (bits 0..31), log2 degree (bits 32..39), product count
// (bits 40..47). hdr[0] = number of items, hdr[1] = the queue counter (both written by
// k_ring_prep); every workgroup leaves once the items run out.
__global__ void __launch_bounds__(RING_LDS_THREADS) k_ring_mul_lds(const zq* __restrict__ A, const zq* __restrict__ B,
                                                                   zq* __restrict__ O, const uint64_t* __restrict__ offs,
                                                                   const uint32_t* __restrict__ order,
                                                                   const uint64_t* __restrict__ items, uint32_t* hdr,
                                                                   NttTables tw) {
    extern __shared__ uint32_t sh[];
    __shared__ uint32_t item_sh;
    __shared__ uint64_t item_off[8];  // each product's first coefficient (at most 8 products per item)
    const uint32_t n_items = hdr[0];
    for (;;) {
        if (threadIdx.x == 0) item_sh = atomicAdd(hdr + 1, 1u);
        __syncthreads();
        const uint32_t it = item_sh;
        if (it >= n_items) break;  // uniform: every thread read the same word
        const uint64_t d = items[it];
        const uint32_t first = (uint32_t)d, cnt = (uint32_t)(d >> 40) & 0xFFu;
        const int log_n = (int)((d >> 32) & 0xFFu);
        if (threadIdx.x < cnt) item_off[threadIdx.x] = offs[order[first + threadIdx.x]];
        __syncthreads();
        // the thread index as an opaque per-item value: what the item code derives from it is
        // computed per item instead of being hoisted out of this loop for every degree at once
        int tid = (int)threadIdx.x;
        __asm__ volatile("" : "+v"(tid));
        switch (log_n) {
            case 6: ring_item<8, 8>(sh, A, B, O, item_off, cnt, tw, tid); break;
            case 7: ring_item<16, 8>(sh, A, B, O, item_off, cnt, tw, tid); break;
            case 8: ring_item<16, 16>(sh, A, B, O, item_off, cnt, tw, tid); break;
            case 9: ring_item<32, 16>(sh, A, B, O, item_off, cnt, tw, tid); break;
            case 10: ring_item<32, 32>(sh, A, B, O, item_off, cnt, tw, tid); break;
            case 11: ring_item<64, 32>(sh, A, B, O, item_off, cnt, tw, tid); break;
            case 12: ring_item<64, 64>(sh, A, B, O, item_off, cnt, tw, tid); break;
            default: ring_item_small(sh, A, B, O, item_off[0], log_n, tw, tid); break;
        }
    }
}

__host__ __device__ constexpr uint32_t ring_per_item(int log_n) {
    return log_n == 6 ? RingCfg<8, 8>::RPB
         : log_n == 7 ? RingCfg<16, 8>::RPB
         : log_n == 8 ? RingCfg<16, 16>::RPB
         : log_n == 9 ? RingCfg<32, 16>::RPB
         : log_n == 10 ? RingCfg<32, 32>::RPB
         : log_n == 11 ? RingCfg<64, 32>::RPB
         : log_n == 12 ? RingCfg<64, 64>::RPB : 1;
}
uint32_t ring_lds_per_item(int log_n) { return ring_per_item(log_n); }

// The queue of a ragged batch, built on the device by one workgroup from the products' degrees
// (logn[i] <= 12, product i's coefficients back to back in input order): offs = exclusive prefix
// sums of the degrees, order = the products grouped by degree, largest degree first (a counting
// sort), items = each degree group cut into runs of ring_per_item products, hdr = (item count, 0).
constexpr int RING_PREP_THREADS = 1024;
__global__ void __launch_bounds__(RING_PREP_THREADS) k_ring_prep(const uint32_t* __restrict__ logn, uint32_t count,
                                                                 uint64_t* __restrict__ offs, uint32_t* __restrict__ order,
                                                                 uint64_t* __restrict__ items, uint32_t* hdr) {
    __shared__ uint64_t part[RING_PREP_THREADS];
    __shared__ uint32_t hist[RING_LDS_MAX_LOG + 1], start[RING_LDS_MAX_LOG + 1], fill[RING_LDS_MAX_LOG + 1];
    __shared__ uint32_t ibase[RING_LDS_MAX_LOG + 2];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (count + RING_PREP_THREADS - 1) / RING_PREP_THREADS;
    const uint32_t lo = min(count, t * per), hi = min(count, lo + per);
    if (t <= RING_LDS_MAX_LOG) {
        hist[t] = 0;
        fill[t] = 0;
    }
    uint64_t sum = 0;
    for (uint32_t i = lo; i < hi; ++i) sum += (uint64_t)1 << logn[i];
    part[t] = sum;
    __syncthreads();
    // degrees above 2^12 are refused by the host before the launch; skipped here all the same
    for (uint32_t i = lo; i < hi; ++i)
        if (logn[i] <= RING_LDS_MAX_LOG) atomicAdd(&hist[logn[i]], 1u);
    // inclusive scan of the per-thread sums (Hillis-Steele in LDS)
    for (uint32_t d = 1; d < RING_PREP_THREADS; d <<= 1) {
        const uint64_t v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint64_t run = part[t] - sum;  // exclusive prefix of this thread's range
    for (uint32_t i = lo; i < hi; ++i) {
        offs[i] = run;
        run += (uint64_t)1 << logn[i];
    }
    if (t == 0) {  // group starts, largest degree first, and each group's first item
        uint32_t at = 0, it = 0;
        for (int d = RING_LDS_MAX_LOG; d >= 0; --d) {
            start[d] = at;
            ibase[d] = it;
            at += hist[d];
            it += (hist[d] + ring_per_item(d) - 1) / ring_per_item(d);
        }
        ibase[RING_LDS_MAX_LOG + 1] = it;
        hdr[0] = it;
        hdr[1] = 0;
    }
    __syncthreads();
    for (uint32_t i = lo; i < hi; ++i) {
        const uint32_t d = logn[i];
        if (d <= RING_LDS_MAX_LOG) order[start[d] + atomicAdd(&fill[d], 1u)] = i;
    }
    for (int d = 0; d <= RING_LDS_MAX_LOG; ++d) {
        const uint32_t pi = ring_per_item(d), n_it = (hist[d] + pi - 1) / pi;
        for (uint32_t k = t; k < n_it; k += RING_PREP_THREADS) {
            const uint32_t first = start[d] + k * pi, cnt = min(pi, hist[d] - k * pi);
            items[ibase[d] + k] = (uint64_t)first | ((uint64_t)d << 32) | ((uint64_t)cnt << 40);
        }
    }
}

void launch_ring_mul_lds(const zq* A, const zq* B, zq* O, const uint32_t* logn, uint32_t count, RingScratch sc,
                         uint32_t workgroups, const NttTables& tw, hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL(k_ring_prep, dim3(1), dim3(RING_PREP_THREADS), 0, s, logn, count, sc.offs, sc.order, sc.items,
                       sc.hdr);
    hipLaunchKernelGGL(k_ring_mul_lds, dim3(std::min(workgroups, count)), dim3(RING_LDS_THREADS),
                       (size_t)RING_LDS_WORDS * 4, s, A, B, O, sc.offs, sc.order, sc.items, sc.hdr, tw);
}

// Schoolbook negacyclic products, the reference's product with ntt_enabled = false
// (naive_negacyclic, src/algebra/ntt.rs:226-240, chosen in RingElement::mul, src/algebra/poly.rs:
// 134-140). One workgroup per product; thread k sums a_i * b_(k-i) with the wrapped terms
// negated, as exact 256-bit sums (at most 8192 products below 2^224), reduced once.
__device__ __forceinline__ void acc256_add(uint32_t acc[8], const uint32_t t[8]) {
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        c += (uint64_t)acc[i] + t[i];
        acc[i] = (uint32_t)c;
        c >>= 32;
    }
}

__global__ void __launch_bounds__(256) k_ring_mul_naive(const zq* __restrict__ A, const zq* __restrict__ B,
                                                        zq* __restrict__ O, const uint64_t* __restrict__ offs,
                                                        const uint32_t* __restrict__ logn) {
    const uint32_t idx = blockIdx.x;
    const int n = 1 << logn[idx];
    const size_t off = offs[idx];
    const zq* a = A + off;
    const zq* b = B + off;
    for (int k = threadIdx.x; k < n; k += 256) {
        uint32_t pos[8] = {0, 0, 0, 0, 0, 0, 0, 0}, neg[8] = {0, 0, 0, 0, 0, 0, 0, 0}, t[8];
        for (int i = 0; i < n; ++i) {
            const int j = k - i;
            zq_mul_wide(load_zq(a + i), load_zq(b + (j >= 0 ? j : j + n)), t);
            acc256_add(j >= 0 ? pos : neg, t);
        }
        store_zq(O + off + k, zq_sub(zq_reduce256(pos), zq_reduce256(neg)));
    }
}

void launch_ring_mul_naive(const zq* A, const zq* B, zq* O, const uint64_t* offs, const uint32_t* logn,
                           uint32_t count, hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL(k_ring_mul_naive, dim3(count), dim3(256), 0, s, A, B, O, offs, logn);
}

__global__ void k_gather_rows(const zq* src, const uint64_t* offs, const uint32_t* idx, size_t total,
                              int log_m, zq* dst, int scatter) {
    const size_t gid = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (gid >= total) return;
    const size_t m = (size_t)1 << log_m;
    const size_t r = gid >> log_m, c = gid & (m - 1);
    const size_t o = offs[idx[r]] + c;
    if (scatter)
        store_zq(dst + o, load_zq(src + gid));
    else
        store_zq(dst + gid, load_zq(src + o));
}

void launch_gather_rows(const zq* src, const uint64_t* offs, const uint32_t* idx, size_t count,
                        int log_m, zq* dst, hipStream_t s) {
    const size_t total = count << log_m;
    if (!total) return;
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, src, offs,
                       idx, total, log_m, dst, 0);
}
void launch_scatter_rows(const zq* src, const uint64_t* offs, const uint32_t* idx, size_t count,
                         int log_m, zq* dst, hipStream_t s) {
    const size_t total = count << log_m;
    if (!total) return;
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, src, offs,
                       idx, total, log_m, dst, 1);
}

}  // namespace dmx
