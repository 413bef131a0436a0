/* Oracle internals shared between the oracle's C files (test infrastructure only). */
#ifndef DAMASCUS_ORACLE_INTERNAL_H
#define DAMASCUS_ORACLE_INTERNAL_H

#include <stddef.h>
#include <stdint.h>
#include <string.h>

typedef unsigned __int128 u128;

/* q = 2^112 - 75 (src/utils/config.rs:4) */
#define ZQ_Q ((((u128)1) << 112) - 75)
#define ZQ_MASK112 ((((u128)1) << 112) - 1)

static inline u128 zq_load(const uint64_t *w) { return ((u128)w[1] << 64) | w[0]; }
static inline void zq_store(uint64_t *w, u128 v) {
    w[0] = (uint64_t)v;
    w[1] = (uint64_t)(v >> 64);
}

/* field.rs:134-145 */
static inline u128 zq_add(u128 a, u128 b) {
    u128 s = a + b;
    return s >= ZQ_Q ? s - ZQ_Q : s;
}
/* field.rs:153-161 */
static inline u128 zq_sub(u128 a, u128 b) { return a >= b ? a - b : ZQ_Q - (b - a); }
/* field.rs:183-193 */
static inline u128 zq_neg(u128 a) { return a == 0 ? 0 : ZQ_Q - a; }

u128 zq_mul(u128 a, u128 b);
u128 zq_pow(u128 base, u128 exp);
u128 zq_inv(u128 a);
u128 zq_from_le_bytes_mod_order(const uint8_t *bytes, size_t len);

/* BLAKE3 incremental hasher */
typedef struct {
    uint32_t key[8];
    uint32_t cv[8];
    uint64_t chunk_counter;
    uint8_t buf[64];
    uint32_t buf_len;
    uint32_t blocks_compressed;
    uint32_t flags;
    uint32_t cv_stack[54][8];
    uint32_t cv_stack_len;
} orc_b3;
void orc_b3_init(orc_b3 *h);
void orc_b3_init_keyed(orc_b3 *h, const uint8_t key[32]);
void orc_b3_update(orc_b3 *h, const void *in, size_t len);
void orc_b3_finalize(const orc_b3 *h, uint8_t *out, size_t out_len);

/* ring product entry points on u128 arrays */
int ring_mul(const u128 *a, const u128 *b, size_t n, int ntt_enabled, u128 *out);
int negacyclic_multiply_u128(const u128 *a, const u128 *b, size_t n, u128 *out);
void naive_negacyclic_u128(const u128 *a, const u128 *b, size_t n, u128 *out);
u128 crt_reconstruct_u128(const uint64_t residues[8]);

extern const uint64_t ORC_PRIMES[8];
extern int orc_threads;

#endif
