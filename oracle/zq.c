/* Oracle: Z_q arithmetic, q = 2^112 - 75 (test infrastructure only).
 * Restates src/algebra/field.rs of the reference. */
#include "internal.h"
#include "oracle.h"

/* field.rs:229-248 — schoolbook over 56-bit limbs */
u128 zq_mul(u128 lhs, u128 rhs) {
    const u128 LIMB_MASK = (((u128)1) << 56) - 1;
    u128 a0 = lhs & LIMB_MASK, a1 = lhs >> 56;
    u128 b0 = rhs & LIMB_MASK, b1 = rhs >> 56;
    u128 c0 = a0 * b0;
    u128 c1 = a0 * b1 + a1 * b0;
    u128 c2 = a1 * b1;
    u128 c1_lo = c1 & LIMB_MASK, c1_hi = c1 >> 56;
    u128 low_sum = c0 + (c1_lo << 56);
    u128 low = low_sum & ZQ_MASK112;
    u128 carry = low_sum >> 112;
    u128 high = c2 + c1_hi + carry;
    /* field.rs:250-258 reduce_pseudo_mersenne */
    u128 acc = low + high * 75;
    u128 hi = acc >> 112;
    acc = (acc & ZQ_MASK112) + hi * 75;
    while (acc >= ZQ_Q) acc -= ZQ_Q;
    return acc;
}

/* field.rs:89-104 */
u128 zq_pow(u128 base, u128 exp) {
    u128 result = 1;
    while (exp > 0) {
        if (exp & 1) result = zq_mul(result, base);
        base = zq_mul(base, base);
        exp >>= 1;
    }
    return result;
}

/* field.rs:106-109 — Fermat inverse x^(q-2) */
u128 zq_inv(u128 a) { return zq_pow(a, ZQ_Q - 2); }

/* field.rs:45-63 — little-endian bytes reduced mod q (BigUint semantics).
 * Horner from the most significant byte keeps every partial value < 2^120. */
u128 zq_from_le_bytes_mod_order(const uint8_t *bytes, size_t len) {
    u128 acc = 0;
    for (size_t i = len; i-- > 0;) acc = ((acc << 8) | bytes[i]) % ZQ_Q;
    return acc;
}

void orc_zq_modulus(uint64_t out[2]) { zq_store(out, ZQ_Q); }

void orc_zq_add(const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n) {
    for (size_t i = 0; i < n; ++i) zq_store(out + 2 * i, zq_add(zq_load(a + 2 * i), zq_load(b + 2 * i)));
}
void orc_zq_sub(const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n) {
    for (size_t i = 0; i < n; ++i) zq_store(out + 2 * i, zq_sub(zq_load(a + 2 * i), zq_load(b + 2 * i)));
}
void orc_zq_neg(const uint64_t *a, uint64_t *out, size_t n) {
    for (size_t i = 0; i < n; ++i) zq_store(out + 2 * i, zq_neg(zq_load(a + 2 * i)));
}
void orc_zq_mul(const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n) {
    for (size_t i = 0; i < n; ++i) zq_store(out + 2 * i, zq_mul(zq_load(a + 2 * i), zq_load(b + 2 * i)));
}
void orc_zq_inv(const uint64_t *a, uint64_t *out, size_t n) {
    for (size_t i = 0; i < n; ++i) zq_store(out + 2 * i, zq_inv(zq_load(a + 2 * i)));
}
/* field.rs:45-47 from_u128 = value % q */
void orc_zq_from_u128(const uint64_t *a, uint64_t *out, size_t n) {
    for (size_t i = 0; i < n; ++i) zq_store(out + 2 * i, zq_load(a + 2 * i) % ZQ_Q);
}
void orc_zq_from_le_bytes_mod_order(const uint8_t *bytes, size_t len, uint64_t out[2]) {
    zq_store(out, len == 0 ? 0 : zq_from_le_bytes_mod_order(bytes, len));
}
/* field.rs:65-71 — up to 16 bytes packed little-endian, then % q */
void orc_zq_from_le_chunk(const uint8_t *bytes, size_t len, uint64_t out[2]) {
    u128 acc = 0;
    for (size_t s = 0; s < len && s < 16; ++s) acc |= ((u128)bytes[s]) << (8 * s);
    zq_store(out, acc % ZQ_Q);
}
