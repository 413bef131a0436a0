// Exact integer arithmetic shared by host and device code.
//
// Z_q, q = 2^112 - 75 (src/utils/config.rs:4; src/algebra/field.rs), held as four 32-bit
// little-endian limbs in a 16-byte word: the same bytes as Fp::to_le_bytes (field.rs:81-83).
// Every routine returns the canonical residue in [0, q), so any correct reduction order gives
// the reference's bits.
//
// Mod-p arithmetic for the eight CRT primes (config.rs:9-18), all below 2^32, uses Montgomery
// reduction with R = 2^32 built on 32-bit mul_lo / mul_hi:
//   redc64(h,l) = (h*R + l) * R^-1 mod p  via t = h - mulhi(l * p^-1, p), plus p on a borrow
//   mulc(b,w)   = b * w mod p             for any 32-bit b, with w stored as (w*R mod p) and
//                                          (w*R*p^-1 mod 2^32) so one product is never formed.
// Half of the primes exceed 2^31, so sums of two residues can carry out of 32 bits: the
// corrections test the carry or borrow of the 32-bit add or subtract itself.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DMX_HD __host__ __device__ __forceinline__

namespace dmx {

constexpr int K = 8;            // MODULE_RANK (config.rs:7)
constexpr int NPRIMES = 8;      // CRT_PRIMES.len() (config.rs:9-18)
constexpr int BYTES_PER_COEFF = 13;
constexpr int MIN_FOLD_DEPTH = 6;
constexpr int MAX_LOG_N = 13;   // ring degrees up to 2^13 (witness depth d <= 13)
constexpr int TW_SIZE = 1 << MAX_LOG_N;

constexpr uint32_t PRIMES[NPRIMES] = {3892314113u, 2281701377u, 2013265921u, 2885681153u,
                                      2483027969u, 1811939329u, 469762049u,  4194304001u};

// ------------------------------------------------------------------------------------------
// Z_q
// ------------------------------------------------------------------------------------------
struct alignas(16) zq {
    uint32_t w[4];
};

constexpr uint32_t Q0 = 0xFFFFFFB5u;  // 2^112 - 75, limb 0
constexpr uint32_t Q1 = 0xFFFFFFFFu;
constexpr uint32_t Q2 = 0xFFFFFFFFu;
constexpr uint32_t Q3 = 0x0000FFFFu;

DMX_HD zq zq_zero() { return zq{{0, 0, 0, 0}}; }
DMX_HD zq zq_from_u64(uint64_t v) { return zq{{(uint32_t)v, (uint32_t)(v >> 32), 0, 0}}; }
DMX_HD bool zq_is_zero(const zq& a) { return (a.w[0] | a.w[1] | a.w[2] | a.w[3]) == 0; }
DMX_HD bool zq_eq(const zq& a, const zq& b) {
    return ((a.w[0] ^ b.w[0]) | (a.w[1] ^ b.w[1]) | (a.w[2] ^ b.w[2]) | (a.w[3] ^ b.w[3])) == 0;
}

// a >= q for a < 2^113 held in 4 limbs (limb 3 may carry bit 16)
DMX_HD bool zq_ge_q(const zq& a) {
    if (a.w[3] != Q3) return a.w[3] > Q3;
    if (a.w[2] != Q2) return a.w[2] > Q2;
    if (a.w[1] != Q1) return a.w[1] > Q1;
    return a.w[0] >= Q0;
}

DMX_HD zq zq_sub_q(const zq& a) {
    zq r;
    uint64_t t = (uint64_t)a.w[0] - Q0;
    r.w[0] = (uint32_t)t;
    uint64_t br = (t >> 63);
    t = (uint64_t)a.w[1] - Q1 - br;
    r.w[1] = (uint32_t)t;
    br = (t >> 63);
    t = (uint64_t)a.w[2] - Q2 - br;
    r.w[2] = (uint32_t)t;
    br = (t >> 63);
    r.w[3] = a.w[3] - Q3 - (uint32_t)br;
    return r;
}

// canonical add (field.rs:134-145)
DMX_HD zq zq_add(const zq& a, const zq& b) {
    zq r;
    uint64_t t = (uint64_t)a.w[0] + b.w[0];
    r.w[0] = (uint32_t)t;
    t = (uint64_t)a.w[1] + b.w[1] + (t >> 32);
    r.w[1] = (uint32_t)t;
    t = (uint64_t)a.w[2] + b.w[2] + (t >> 32);
    r.w[2] = (uint32_t)t;
    r.w[3] = a.w[3] + b.w[3] + (uint32_t)(t >> 32);
    return zq_ge_q(r) ? zq_sub_q(r) : r;
}

// canonical sub (field.rs:153-161): a - b, plus q when it borrows
DMX_HD zq zq_sub(const zq& a, const zq& b) {
    zq r;
    uint64_t t = (uint64_t)a.w[0] - b.w[0];
    r.w[0] = (uint32_t)t;
    uint64_t br = t >> 63;
    t = (uint64_t)a.w[1] - b.w[1] - br;
    r.w[1] = (uint32_t)t;
    br = t >> 63;
    t = (uint64_t)a.w[2] - b.w[2] - br;
    r.w[2] = (uint32_t)t;
    br = t >> 63;
    uint64_t t3 = (uint64_t)a.w[3] - b.w[3] - br;
    r.w[3] = (uint32_t)t3;
    if (t3 >> 63) {  // negative: add q back
        uint64_t s = (uint64_t)r.w[0] + Q0;
        r.w[0] = (uint32_t)s;
        s = (uint64_t)r.w[1] + Q1 + (s >> 32);
        r.w[1] = (uint32_t)s;
        s = (uint64_t)r.w[2] + Q2 + (s >> 32);
        r.w[2] = (uint32_t)s;
        r.w[3] = r.w[3] + Q3 + (uint32_t)(s >> 32);
    }
    return r;
}

// canonical negation (field.rs:183-193)
DMX_HD zq zq_neg(const zq& a) { return zq_is_zero(a) ? a : zq_sub(zq{{Q0, Q1, Q2, Q3}}, a); }

// Reduce a value given as 8 x 32-bit limbs (< 2^256) mod q, using 2^112 = 75 (mod q):
// x = lo112 + mid112 * 2^112 + top32 * 2^224  ==  lo + 75*mid + 5625*top.
DMX_HD zq zq_reduce256(const uint32_t x[8]) {
    // split into 112-bit pieces
    uint32_t lo[4] = {x[0], x[1], x[2], x[3] & 0xFFFFu};
    uint32_t mid[4] = {(x[3] >> 16) | (x[4] << 16), (x[4] >> 16) | (x[5] << 16),
                       (x[5] >> 16) | (x[6] << 16), (x[6] >> 16)};
    uint32_t top = x[7];
    // s = lo + 75*mid + 5625*top  (< 2^119)
    uint64_t c;
    uint32_t s[4];
    c = (uint64_t)lo[0] + (uint64_t)mid[0] * 75u + (uint64_t)top * 5625u;
    s[0] = (uint32_t)c;
    c = (c >> 32) + lo[1] + (uint64_t)mid[1] * 75u;
    s[1] = (uint32_t)c;
    c = (c >> 32) + lo[2] + (uint64_t)mid[2] * 75u;
    s[2] = (uint32_t)c;
    c = (c >> 32) + lo[3] + (uint64_t)mid[3] * 75u;
    s[3] = (uint32_t)c;  // < 2^23 since total < 2^119
    // fold bits >= 112 once more: s = (s mod 2^112) + 75 * (s >> 112)
    uint32_t hi = s[3] >> 16;
    zq r;
    c = (uint64_t)s[0] + (uint64_t)hi * 75u;
    r.w[0] = (uint32_t)c;
    c = (c >> 32) + s[1];
    r.w[1] = (uint32_t)c;
    c = (c >> 32) + s[2];
    r.w[2] = (uint32_t)c;
    r.w[3] = (s[3] & 0xFFFFu) + (uint32_t)(c >> 32);
    if (zq_ge_q(r)) r = zq_sub_q(r);
    if (zq_ge_q(r)) r = zq_sub_q(r);
    return r;
}

// 112 x 112 -> 224-bit schoolbook product into 8 limbs
DMX_HD void zq_mul_wide(const zq& a, const zq& b, uint32_t out[8]) {
    uint64_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        uint64_t carry = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uint64_t t = (uint64_t)a.w[i] * b.w[j] + (uint32_t)acc[i + j] + carry;
            acc[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        acc[i + 4] = carry;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) out[i] = (uint32_t)acc[i];
}

// canonical product (field.rs:229-258 computes the same residue)
DMX_HD zq zq_mul(const zq& a, const zq& b) {
    uint32_t t[8];
    zq_mul_wide(a, b, t);
    return zq_reduce256(t);
}

// a + c*b
DMX_HD zq zq_fma(const zq& a, const zq& b, const zq& c) { return zq_add(a, zq_mul(b, c)); }

// Fermat inverse x^(q-2) (field.rs:106-109)
DMX_HD zq zq_pow(zq base, const zq& e) {
    zq r = zq_from_u64(1);
    for (int limb = 0; limb < 4; ++limb) {
        uint32_t bits = e.w[limb];
        int nb = (limb == 3) ? 16 : 32;
        for (int i = 0; i < nb; ++i) {
            if (bits & 1u) r = zq_mul(r, base);
            base = zq_mul(base, base);
            bits >>= 1;
        }
    }
    return r;
}
DMX_HD zq zq_inv(const zq& a) { return zq_pow(a, zq{{Q0 - 2, Q1, Q2, Q3}}); }

// ------------------------------------------------------------------------------------------
// Z_q inverse by batched divsteps (Bernstein and Yang, "Fast constant-time gcd computation
// and modular inversion", TCHES 2019), in its variable-time form. The inverse is unique, so
// this returns the same value as x^(q-2) (field.rs:106-109) with about a third of the serial
// instructions of the square-and-multiply chain.
//
// divstep(delta, f, g) with f odd:
//   delta > 0 and g odd:  (1 - delta, g, (g - f) / 2)
//   g odd otherwise:      (1 + delta, f, (g + f) / 2)
//   g even:               (1 + delta, f, g / 2)
// Started from (1, q, x), g reaches 0 within 326 divsteps for 112-bit inputs (the paper's
// bound (49 d + 57) / 17 at d = 112) and f ends at +-1. Thirty divsteps depend only on the low
// 32 bits of f and g and combine into a 2x2 integer matrix T with 2^30 (f', g') = T (f, g);
// each row of T has L1 norm <= 2^30. Alongside, d and e with f = d x and g = e x (mod q) move by
// (d', e') = T (d, e) / 2^30 mod q; at the end x^-1 = +-d.
//
// Inside a batch, runs of even g are skipped with one count of trailing zeros, and after an odd
// step the next min(1 - delta, remaining, 6) divsteps, which cannot swap, are taken at once: they
// add w f to g for the unique w < 2^k with g + w f = 0 (mod 2^k), w = -g f^-1 (mod 2^k).
//
// Multi-word values use four signed 30-bit limbs: v = l0 + l1 2^30 + l2 2^60 + l3 2^90 with
// l0..l2 in [0, 2^30) and l3 signed; |f|, |g| <= q and |d|, |e| < 13 q over at most 12 batches.
// ------------------------------------------------------------------------------------------
namespace divstep_detail {
constexpr int32_t M30 = (1 << 30) - 1;
constexpr int32_t QL0 = (1 << 30) - 75, QL1 = M30, QL2 = M30, QL3 = (1 << 22) - 1;  // q in 30-bit limbs
constexpr uint32_t neg_q_inv30() {  // -q^-1 mod 2^30 by Newton's iteration on 2-adic inverses
    uint32_t x = Q0;                 // q * q = 1 (mod 8)
    for (int i = 0; i < 5; ++i) x *= 2u - Q0 * x;
    return (0u - x) & (uint32_t)M30;
}
constexpr uint32_t NQI = neg_q_inv30();
static_assert(((NQI * Q0) & (uint32_t)M30) == (uint32_t)M30, "-q^-1 mod 2^30");

struct Mat {
    int32_t u, v, q, r;
};

DMX_HD uint32_t uniform(uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);  // every lane holds the same value
#else
    return v;
#endif
}

// Thirty divsteps on the low words f, g (f odd); delta is carried across batches.
DMX_HD Mat divsteps30(int32_t& delta, uint32_t f, uint32_t g) {
    int32_t u = 1, v = 0, q = 0, r = 1;
    int32_t i = 30;
    for (;;) {
        const int32_t z = __builtin_ctz(g | (1u << i));  // even steps, at most i of them
        g >>= z;
        u *= 1 << z;
        v *= 1 << z;
        delta += z;
        i -= z;
        if (i == 0) break;
        if (delta > 0) {  // (f, g) <- (g, -f): the first of these steps is the swapping case
            delta = -delta;
            const uint32_t tf = f;
            f = g;
            g = 0u - tf;
            const int32_t tu = u, tv = v;
            u = q;
            v = r;
            q = -tu;
            r = -tv;
        }
        const int32_t lim = i < 1 - delta ? (i < 6 ? i : 6) : (1 - delta < 6 ? 1 - delta : 6);
        const uint32_t finv = f * (2u - f * f);  // f^-1 mod 2^6
        const uint32_t w = (0u - g * finv) & ((1u << lim) - 1u);
        g += w * f;
        q += (int32_t)w * u;
        r += (int32_t)w * v;
    }
    return Mat{u, v, q, r};
}

// (a, b) <- T (a, b) / 2^30, exact
DMX_HD void apply_fg(int32_t a[4], int32_t b[4], const Mat& t) {
    int64_t ca = (int64_t)t.u * a[0] + (int64_t)t.v * b[0];
    int64_t cb = (int64_t)t.q * a[0] + (int64_t)t.r * b[0];
    ca >>= 30;
    cb >>= 30;
    for (int i = 1; i < 4; ++i) {
        ca += (int64_t)t.u * a[i] + (int64_t)t.v * b[i];
        cb += (int64_t)t.q * a[i] + (int64_t)t.r * b[i];
        a[i - 1] = (int32_t)ca & M30;
        b[i - 1] = (int32_t)cb & M30;
        ca >>= 30;
        cb >>= 30;
    }
    a[3] = (int32_t)ca;
    b[3] = (int32_t)cb;
}

// (a, b) <- T (a, b) / 2^30 mod q: adds m q with m = -c q^-1 (mod 2^30) so each division is exact
DMX_HD void apply_de(int32_t a[4], int32_t b[4], const Mat& t) {
    int64_t ca = (int64_t)t.u * a[0] + (int64_t)t.v * b[0];
    int64_t cb = (int64_t)t.q * a[0] + (int64_t)t.r * b[0];
    const int64_t ma = (int64_t)(((uint32_t)ca * NQI) & (uint32_t)M30);
    const int64_t mb = (int64_t)(((uint32_t)cb * NQI) & (uint32_t)M30);
    ca = (ca + ma * QL0) >> 30;
    cb = (cb + mb * QL0) >> 30;
    const int32_t ql[4] = {QL0, QL1, QL2, QL3};
    for (int i = 1; i < 4; ++i) {
        ca += (int64_t)t.u * a[i] + (int64_t)t.v * b[i] + ma * ql[i];
        cb += (int64_t)t.q * a[i] + (int64_t)t.r * b[i] + mb * ql[i];
        a[i - 1] = (int32_t)ca & M30;
        b[i - 1] = (int32_t)cb & M30;
        ca >>= 30;
        cb >>= 30;
    }
    a[3] = (int32_t)ca;
    b[3] = (int32_t)cb;
}

// canonical residue of a signed-30 value with |v| < 16 q
DMX_HD zq from_s30(const int32_t a[4]) {
    const int32_t ql[4] = {QL0, QL1, QL2, QL3};
    int64_t c = 0;
    uint64_t l[4];
    for (int i = 0; i < 4; ++i) {  // v + 16 q, now in (0, 2^117)
        c += (int64_t)a[i] + 16 * (int64_t)ql[i];
        l[i] = (uint64_t)(c & M30);
        c >>= 30;
    }
    l[3] |= (uint64_t)c << 30;
    const uint64_t lo = l[0] | (l[1] << 30) | (l[2] << 60);
    const uint64_t hi = (l[2] >> 4) | (l[3] << 26);  // bits 64 and up
    // 2^112 = 75 (mod q): fold bits 112 and up once, then one conditional subtraction
    uint64_t r0 = lo + 75 * (hi >> 48);
    uint64_t r1 = (hi & 0xFFFFFFFFFFFFull) + (r0 < lo ? 1 : 0);
    zq r{{(uint32_t)r0, (uint32_t)(r0 >> 32), (uint32_t)r1, (uint32_t)(r1 >> 32)}};
    return zq_ge_q(r) ? zq_sub_q(r) : r;
}
}  // namespace divstep_detail

// x^-1 mod q for canonical x; 0 maps to 0 as in x^(q-2).
DMX_HD zq zq_inv_divsteps(const zq& x) {
    using namespace divstep_detail;
    if (zq_is_zero(x)) return x;
    int32_t f[4] = {QL0, QL1, QL2, QL3};
    int32_t g[4] = {(int32_t)(x.w[0] & M30), (int32_t)(((x.w[0] >> 30) | (x.w[1] << 2)) & M30),
                    (int32_t)(((x.w[1] >> 28) | (x.w[2] << 4)) & M30),
                    (int32_t)((x.w[2] >> 26) | (x.w[3] << 6))};
    int32_t d[4] = {0, 0, 0, 0}, e[4] = {1, 0, 0, 0};
    int32_t delta = 1;
    for (int batch = 0; batch < 12; ++batch) {
        const uint32_t fl = uniform((uint32_t)f[0] | ((uint32_t)f[1] << 30));
        const uint32_t gl = uniform((uint32_t)g[0] | ((uint32_t)g[1] << 30));
        const Mat t = divsteps30(delta, fl, gl);
        apply_fg(f, g, t);
        apply_de(d, e, t);
        if (uniform((uint32_t)(g[0] | g[1] | g[2] | g[3])) == 0) break;
    }
    const zq r = from_s30(d);
    return f[3] < 0 ? zq_neg(r) : r;  // f = -1: x^-1 = -d
}

// LE256 bytes (as 8 words) mod q: Fp::from_le_bytes_mod_order on a 32-byte digest (field.rs:45-63)
DMX_HD zq zq_from_digest(const uint32_t h[8]) { return zq_reduce256(h); }

// 13-byte little-endian chunk (< 2^104 < q): Fp::from_le_chunk (field.rs:65-71)

// ------------------------------------------------------------------------------------------
// mod-p (Montgomery, R = 2^32)
// ------------------------------------------------------------------------------------------
struct PrimeConsts {
    uint32_t p;
    uint32_t pinv;       // p^-1 mod 2^32
    uint32_t r4_m, r4_mp;  // R^3 as a mulc constant: (R^4 mod p, (R^4 mod p) * pinv)
    uint32_t two64;      // 2^64 mod p
    uint32_t q_mod_p;    // q mod p
    uint32_t r3_m, r3_mp;  // R^2 as a mulc constant: (R^3 mod p, (R^3 mod p) * pinv)
};

#if defined(__HIP_DEVICE_COMPILE__)
DMX_HD uint32_t mulhi32(uint32_t a, uint32_t b) { return __umulhi(a, b); }
#else
DMX_HD uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }
#endif

// a - b, plus p when the subtraction borrows. Canonical for a, b in [0, p); for any 32-bit a
// and b < p the result is the representative of a - b in [0, 2^32) (below p when a < p).
DMX_HD uint32_t sub_fix(uint32_t a, uint32_t b, uint32_t p) {
    uint32_t t;
    const bool borrow = __builtin_usub_overflow(a, b, &t);
    return t + (borrow ? p : 0u);
}

// (hi * R + lo) * R^-1 mod p for any 32-bit hi and lo: the result lies in [0, 2^32), and in
// [0, p) when hi < p. With m = lo * p^-1 mod R, hi*R + lo - m*p = (hi - mulhi(m, p)) * R exactly.
DMX_HD uint32_t redc64(uint32_t hi, uint32_t lo, uint32_t p, uint32_t pinv) {
    const uint32_t m = lo * pinv;
    return sub_fix(hi, mulhi32(m, p), p);
}

// T * R^-1 mod p for T < p*R (result in [0, p))
DMX_HD uint32_t redc(uint64_t T, uint32_t p, uint32_t pinv) { return redc64((uint32_t)(T >> 32), (uint32_t)T, p, pinv); }

// b * w mod p for any 32-bit b; (wm, wmp) = (w*R mod p, wm*pinv mod 2^32)
DMX_HD uint32_t mulc(uint32_t b, uint32_t wm, uint32_t wmp, uint32_t p) {
    return sub_fix(mulhi32(b, wm), mulhi32(b * wmp, p), p);
}

// a + b mod p for a, b in [0, p): the 32-bit sum may carry (p > 2^31), then sum - p is right
DMX_HD uint32_t addp(uint32_t a, uint32_t b, uint32_t p) {
    uint32_t s, t;
    const bool carry = __builtin_uadd_overflow(a, b, &s);
    const bool below = __builtin_usub_overflow(s, p, &t);
    return (carry || !below) ? t : s;
}
DMX_HD uint32_t subp(uint32_t a, uint32_t b, uint32_t p) { return sub_fix(a, b, p); }

// a - b - bin and the borrow out of that subtraction
DMX_HD uint32_t sub_borrow(uint32_t a, uint32_t b, bool bin, bool& bout) {
    uint32_t t, r;
    const bool b1 = __builtin_usub_overflow(a, b, &t);
    const bool b2 = __builtin_usub_overflow(t, (uint32_t)bin, &r);
    bout = b1 || b2;
    return r;
}

// A chain of REDC steps over the limbs w0, w1, ... of a value: step k forms
// d_k = w_k - mulhi(t_{k-1} * p^-1, p) - b_{k-1}, where t_{k-1} is the previous d mod 2^32
// and b_{k-1} its sign. Leaving out each step's +p correction and subtracting the borrow from
// the next high word instead keeps d_k = (w_k R + d_{k-1}) R^-1 (mod p) with d_k in [-p, 2^32).
// Starts with t_0 = w0, b_0 = 0.
struct RedcChain {
    uint32_t t;
    bool b;
    DMX_HD void step(uint32_t hi, uint32_t p, uint32_t pinv) { t = sub_borrow(hi, mulhi32(t * pinv, p), b, b); }
    DMX_HD uint32_t value(uint32_t p) const { return t + (b ? p : 0u); }  // in [0, 2^32), below p if d < 0
};

// x mod p for x < 2^128 given as four limbs w0..w3 (w3 < 2^16 here): the chain gives
// d_3 = x R^-3 (mod p), d_3 in [-p, 2^16); then * R^3.
DMX_HD uint32_t zq_mod_p(const zq& x, const PrimeConsts& c) {
    RedcChain r{x.w[0], false};
    r.step(x.w[1], c.p, c.pinv);
    r.step(x.w[2], c.p, c.pinv);
    r.step(x.w[3], c.p, c.pinv);
    return mulc(r.value(c.p), c.r4_m, c.r4_mp, c.p);
}

// Residue of the centred representative of x (x - q when x > (q-1)/2). Products of centred
// representatives are bounded by m*(q/2)^2, which keeps every commitment sum of up to 2^24
// coefficient products inside the CRT range |V| < P/2 (P ~ 2^248).
DMX_HD bool zq_gt_half(const zq& x) {  // x > (q-1)/2 = 2^111 - 38
    if (x.w[3] != 0x7FFFu) return x.w[3] > 0x7FFFu;
    if (x.w[2] != 0xFFFFFFFFu || x.w[1] != 0xFFFFFFFFu) return false;
    return x.w[0] > 0xFFFFFFDAu;
}
DMX_HD uint32_t zq_mod_p_centred(const zq& x, const PrimeConsts& c) {
    const uint32_t r = zq_mod_p(x, c);
    return zq_gt_half(x) ? subp(r, c.q_mod_p, c.p) : r;
}

// (hi * 2^64 + lo) mod p for the u96 accumulators of the MAC kernels: two REDC steps give
// V * R^-2 (mod p), then * R^2
DMX_HD uint32_t reduce_u96(uint64_t lo, uint32_t hi, const PrimeConsts& c) {
    RedcChain r{(uint32_t)lo, false};
    r.step((uint32_t)(lo >> 32), c.p, c.pinv);
    r.step(hi, c.p, c.pinv);
    return mulc(r.value(c.p), c.r3_m, c.r3_mp, c.p);
}

// ------------------------------------------------------------------------------------------
// host-side helpers for constants
// ------------------------------------------------------------------------------------------
inline uint64_t host_powmod(uint64_t b, uint64_t e, uint64_t p) {
    unsigned __int128 r = 1, x = b % p;
    while (e) {
        if (e & 1) r = (r * x) % p;
        x = (x * x) % p;
        e >>= 1;
    }
    return (uint64_t)r;
}
inline uint32_t host_inv32(uint32_t p) {  // p^-1 mod 2^32 (Newton)
    uint32_t x = p;
    for (int i = 0; i < 5; ++i) x *= 2u - p * x;
    return x;
}

}  // namespace dmx
