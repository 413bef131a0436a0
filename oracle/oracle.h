/*
 * Oracle: a plain-C CPU restatement of the Damascus Fold commit-and-fold path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, the smoke check and the CPU-baseline
 * leg of bench.py may load this library, and only as the checker or as the
 * timed CPU baseline. The product (damascus000_amd/) never links or calls it.
 *
 * Every function restates the reference algorithm (fengzelinnn/damascus000)
 * and cites the file:line it follows. Z_q elements cross this API as 16-byte
 * little-endian words (two uint64_t: lo, hi), the same encoding as
 * Fp::to_le_bytes (src/algebra/field.rs:81-83).
 *
 * Commitments / module elements are K=8 coordinates of a shared ring length,
 * stored coordinate-major: elem[k * ring_len + c].
 * Generators are stored [vec_idx][k][coeff]: g[(i * 8 + k) * ring_len + c].
 */
#ifndef DAMASCUS_ORACLE_H
#define DAMASCUS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_K 8
#define ORC_NUM_PRIMES 8
#define ORC_BYTES_PER_COEFF 13
#define ORC_MIN_FOLD_DEPTH 6

/* status codes */
#define ORC_OK 0
#define ORC_ERR_ARG -1
#define ORC_ERR_MEMORY_LIMIT -2   /* "split the file into smaller statements" */
#define ORC_ERR_ROUND -3          /* round index mismatch / all rounds done */
#define ORC_ERR_INVARIANT -4      /* folding invariant failed */
#define ORC_ERR_SHAPE -5          /* ring length mismatch */
#define ORC_ERR_OPENING -6        /* final opening mismatch */
#define ORC_ERR_DECODE -7         /* bincode decode failure */
#define ORC_ERR_NOT_FINAL -8      /* opening before all rounds */
#define ORC_ERR_ZERO_TERMS -9     /* scalar stage emitted non-zero cross terms */
#define ORC_ERR_OOM -10

/* ---- Z_q, q = 2^112 - 75 (src/algebra/field.rs) ---- */
void orc_zq_modulus(uint64_t out[2]);
void orc_zq_add(const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n);
void orc_zq_sub(const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n);
void orc_zq_neg(const uint64_t *a, uint64_t *out, size_t n);
void orc_zq_mul(const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n);
void orc_zq_inv(const uint64_t *a, uint64_t *out, size_t n);
void orc_zq_from_u128(const uint64_t *a, uint64_t *out, size_t n);
void orc_zq_from_le_bytes_mod_order(const uint8_t *bytes, size_t len, uint64_t out[2]);
void orc_zq_from_le_chunk(const uint8_t *bytes, size_t len, uint64_t out[2]);

/* ---- BLAKE3 (crate blake3 1.8.3, restated from the public BLAKE3 spec) ---- */
void orc_blake3_hash(const uint8_t *in, size_t len, uint8_t out[32]);
void orc_blake3_keyed_hash(const uint8_t key[32], const uint8_t *in, size_t len, uint8_t out[32]);
/* extended output (XOF) from the root node, for spec test vectors */
void orc_blake3_hash_xof(const uint8_t *key_or_null, const uint8_t *in, size_t len,
                         uint8_t *out, size_t out_len);

/* ---- NTT over one CRT prime (src/algebra/ntt.rs:10-96, 243-370) ---- */
int orc_ntt_forward(uint64_t *values, size_t size, uint64_t modulus);
int orc_ntt_inverse(uint64_t *values, size_t size, uint64_t modulus);
uint64_t orc_multiplicative_generator(uint64_t modulus);
void orc_crt_primes(uint64_t out[8]);

/* ---- ring products (src/algebra/ntt.rs:98-240) ---- */
int orc_convolution(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out);
int orc_negacyclic_multiply(const uint64_t *lhs, const uint64_t *rhs, size_t n, uint64_t *out);
int orc_naive_negacyclic(const uint64_t *lhs, const uint64_t *rhs, size_t n, uint64_t *out);
/* reconstruct_mod_q on one coefficient (ntt.rs:400-427) */
void orc_crt_reconstruct(const uint64_t residues[8], uint64_t out[2]);

/* ---- poly helpers (src/algebra/poly.rs) ---- */
void orc_poly_mul_by_x(const uint64_t *in, size_t n, uint64_t *out);

/* ---- commitment (src/commitment/sis.rs) ---- */
void orc_derive_subseed(const uint8_t seed[32], const uint8_t *label, size_t label_len,
                        uint8_t out[32]);
void orc_derive_generators(const uint8_t g_seed[32], size_t vector_len, size_t ring_len,
                           uint64_t *out);
/* commit_with_generators_ntt (sis.rs:146-167). rows x ring_len witness,
 * rows x 8 x ring_len generators, out 8 x ring_len. rows == 0 gives zero(1). */
int orc_commit_with_generators(const uint64_t *witness, const uint64_t *gens, size_t rows,
                               size_t ring_len, int ntt_enabled, uint64_t *out,
                               size_t *out_ring_len);
/* commit with generators derived from (seed -> g_0_seed) (sis.rs:111-132). */
int orc_commit(const uint8_t seed[32], const uint64_t *witness, size_t rows, size_t ring_len,
               int ntt_enabled, uint64_t *out, size_t *out_ring_len);

/* ---- transcript (src/protocol/transcript.rs) ---- */
/* label: "vec" or "poly". Each commitment is 8 x len_x coefficients. */
void orc_challenge(const uint8_t epoch_seed[32], const char *label, const uint8_t file_id[32],
                   uint64_t round, const uint64_t *cur, size_t len_cur, const uint64_t *lhs,
                   size_t len_l, const uint64_t *rhs, size_t len_r, uint64_t out[2]);

/* ---- io (src/utils/io.rs) ---- */
size_t orc_coeff_count_for_byte_len(uint64_t byte_len);
int orc_layout_for_byte_len(uint64_t byte_len, size_t *depth, size_t *dimension,
                            size_t *coeff_count, size_t *capacity);
/* expand into a dimension x dimension witness (out must hold capacity coeffs) */
int orc_expand(const uint8_t *file, size_t len, size_t max_preprocess_bytes, uint8_t file_id[32],
               uint64_t *out, size_t out_capacity_coeffs);

/* ---- prover (src/protocol/prover.rs) ---- */
typedef struct orc_prover orc_prover;
int orc_prover_new(const uint8_t *file, size_t len, const uint8_t seed_generators[32],
                   const uint8_t epoch_seed[32], size_t max_preprocess_bytes, int ntt_enabled,
                   orc_prover **out);
void orc_prover_free(orc_prover *p);
size_t orc_prover_rounds(const orc_prover *p);
size_t orc_prover_round(const orc_prover *p);
void orc_prover_dims(const orc_prover *p, size_t *rows, size_t *ring_len);
/* statement as bincode bytes; returns the needed length (writes if cap suffices) */
size_t orc_prover_statement(const orc_prover *p, uint8_t *out, size_t cap);
/* current commitment (8 x ring_len) and witness/generator state, for white-box tests */
void orc_prover_current_commitment(const orc_prover *p, uint64_t *out);
void orc_prover_message(const orc_prover *p, uint64_t *out);
void orc_prover_generators(const orc_prover *p, uint64_t *out);
/* runs fold_round(round); writes the bincode RoundRecord into out (cap bytes);
 * *out_len receives the record length. */
int orc_prover_fold_round(orc_prover *p, size_t round, uint8_t *out, size_t cap,
                          size_t *out_len);
int orc_prover_final_opening(const orc_prover *p, uint64_t out[2]);

/* ---- verifier (src/protocol/verifier.rs) ---- */
typedef struct orc_verifier orc_verifier;
int orc_verifier_new(const uint8_t epoch_seed[32], const uint8_t *statement, size_t len,
                     orc_verifier **out);
orc_verifier *orc_verifier_clone(const orc_verifier *v);
void orc_verifier_free(orc_verifier *v);
void orc_verifier_dims(const orc_verifier *v, size_t *rows, size_t *ring_len);
size_t orc_verifier_round(const orc_verifier *v);
void orc_verifier_current_commitment(const orc_verifier *v, uint64_t *out);
/* record given as bincode bytes (decode errors -> ORC_ERR_DECODE) */
int orc_verifier_update(orc_verifier *v, const uint8_t *record, size_t len);
int orc_verifier_verify_final_opening(const orc_verifier *v, const uint64_t opening[2]);

/* ---- bincode 1.3 fixint-LE helpers (serde derive layouts) ---- */
size_t orc_record_size(size_t ring_len_vec, size_t ring_len_poly);

/* ---- full flow on one file: prove, replay, verify (examples/full_flow.rs) ----
 * Returns 0 when the verifier accepts. statement/records/opening are written
 * when the out pointers are non-NULL (records concatenated bincode). */
int orc_full_flow(const uint8_t *file, size_t len, const uint8_t seed_generators[32],
                  const uint8_t epoch_seed[32], uint8_t *statement_out, size_t statement_cap,
                  size_t *statement_len, uint8_t *records_out, size_t records_cap,
                  size_t *records_len, uint64_t opening_out[2]);

/* transcript.rs:66-75 driven by the mock candidate of the reference's test (transcript.rs:130-143):
 * zero below `zero_below`, then `value`; records the counters it was called with */
uint64_t orc_sample_nonzero_mock(uint32_t zero_below, uint64_t value, uint32_t *seen, size_t cap,
                                 size_t *n_seen);
/* test knob: every challenge rejects the candidates of counters below n (0: the reference) */
void orc_set_challenge_reject(uint32_t n);

/* number of worker threads used by the commit loop (1 = the reference's sequential loop) */
void orc_set_threads(int threads);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
