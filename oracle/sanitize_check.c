/*
 * Sanitizer driver for the oracle (test infrastructure, like the oracle itself).
 *
 * Built and run by `make -C oracle run-asan` (AddressSanitizer + UndefinedBehaviorSanitizer) and
 * `make -C oracle run-tsan` (ThreadSanitizer). It runs every public oracle entry point on small
 * inputs, the threaded commitment loop among them, and checks results that must agree (threaded
 * against sequential proofs, NTT round trips, the NTT product against the schoolbook product,
 * prover output against the verifier), so a sanitizer run also fails on a wrong answer.
 * Exit status 0: clean and consistent.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int failures;

#define CHECK(cond)                                                      \
    do {                                                                 \
        if (!(cond)) {                                                   \
            fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                  \
        }                                                                \
    } while (0)

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64(void) { /* SplitMix64 */
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* n canonical Z_q values (lo, hi words) */
static void random_zq(uint64_t *out, size_t n) {
    uint64_t q[2];
    orc_zq_modulus(q);
    for (size_t i = 0; i < n; ++i) {
        uint64_t raw[2] = {next_u64(), next_u64() & ((1ull << 48) - 1)};
        orc_zq_from_u128(raw, out + 2 * i, 1);
    }
}

static void check_field(void) {
    enum { N = 257 };
    uint64_t a[2 * N], b[2 * N], s[2 * N], d[2 * N], p[2 * N], inv[2 * N], t[2 * N];
    random_zq(a, N);
    random_zq(b, N);
    orc_zq_add(a, b, s, N);
    orc_zq_sub(s, b, d, N);
    CHECK(memcmp(d, a, sizeof a) == 0); /* (a + b) - b = a */
    orc_zq_mul(a, b, p, N);
    orc_zq_inv(b, inv, N);
    orc_zq_mul(p, inv, t, N);
    for (size_t i = 0; i < N; ++i) /* b = 0 has no inverse; a random b is non-zero */
        CHECK(t[2 * i] == a[2 * i] && t[2 * i + 1] == a[2 * i + 1]);
    orc_zq_neg(a, t, N);
    orc_zq_add(a, t, s, N);
    for (size_t i = 0; i < 2 * N; ++i) CHECK(s[i] == 0);
    uint8_t bytes[77];
    for (size_t i = 0; i < sizeof bytes; ++i) bytes[i] = (uint8_t)next_u64();
    uint64_t v[2];
    for (size_t len = 0; len <= sizeof bytes; ++len) orc_zq_from_le_bytes_mod_order(bytes, len, v);
    for (size_t len = 0; len <= 13; ++len) orc_zq_from_le_chunk(bytes, len, v);
}

static void check_blake3(void) {
    uint8_t key[32], out1[32], out2[32], xof[200];
    static uint8_t msg[5000];
    for (size_t i = 0; i < sizeof msg; ++i) msg[i] = (uint8_t)(i % 251);
    for (size_t i = 0; i < 32; ++i) key[i] = (uint8_t)(3 * i + 1);
    const size_t lens[] = {0, 1, 63, 64, 65, 1023, 1024, 1025, 2048, 3073, 5000};
    for (size_t i = 0; i < sizeof lens / sizeof lens[0]; ++i) {
        orc_blake3_hash(msg, lens[i], out1);
        orc_blake3_hash_xof(NULL, msg, lens[i], xof, sizeof xof);
        CHECK(memcmp(out1, xof, 32) == 0); /* the XOF's first 32 bytes are the hash */
        orc_blake3_keyed_hash(key, msg, lens[i], out2);
        orc_blake3_hash_xof(key, msg, lens[i], xof, 64);
        CHECK(memcmp(out2, xof, 32) == 0);
    }
}

static void check_ntt(void) {
    uint64_t primes[8];
    orc_crt_primes(primes);
    for (size_t size = 1; size <= 512; size *= 2) {
        uint64_t *v = malloc(size * sizeof *v), *w = malloc(size * sizeof *w);
        for (int pi = 0; pi < 8; ++pi) {
            for (size_t i = 0; i < size; ++i) v[i] = w[i] = next_u64() % primes[pi];
            CHECK(orc_ntt_forward(v, size, primes[pi]) == 0);
            CHECK(orc_ntt_inverse(v, size, primes[pi]) == 0);
            CHECK(memcmp(v, w, size * sizeof *v) == 0);
        }
        free(v);
        free(w);
    }
    for (size_t n = 1; n <= 64; n *= 2) { /* CRT-NTT product against the schoolbook product */
        uint64_t *a = malloc(2 * n * sizeof *a), *b = malloc(2 * n * sizeof *b);
        uint64_t *x = malloc(2 * n * sizeof *x), *y = malloc(2 * n * sizeof *y), *z = malloc(2 * n * sizeof *z);
        random_zq(a, n);
        random_zq(b, n);
        CHECK(orc_negacyclic_multiply(a, b, n, x) == 0);
        CHECK(orc_naive_negacyclic(a, b, n, y) == 0);
        CHECK(memcmp(x, y, 2 * n * sizeof *x) == 0);
        orc_poly_mul_by_x(a, n, z);
        uint64_t *c = malloc((4 * n + 2) * sizeof *c);
        CHECK(orc_convolution(a, n, b, n, c) == 0);
        free(c);
        free(a);
        free(b);
        free(x);
        free(y);
        free(z);
    }
    uint64_t res[8] = {1, 2, 3, 4, 5, 6, 7, 8}, out[2];
    orc_crt_reconstruct(res, out);
}

/* prove + verify one file with 1 and 4 commitment threads; the proofs must be identical */
static void check_flow(size_t len) {
    uint8_t *file = malloc(len ? len : 1);
    for (size_t i = 0; i < len; ++i) file[i] = (uint8_t)next_u64();
    uint8_t seed[32], epoch[32];
    memset(seed, 7, 32);
    memset(epoch, 11, 32);
    size_t st_cap = 1 << 20, rec_cap = 8 << 20;
    uint8_t *st[2], *rec[2];
    size_t st_len[2], rec_len[2];
    uint64_t op[2][2];
    for (int t = 0; t < 2; ++t) {
        st[t] = malloc(st_cap);
        rec[t] = malloc(rec_cap);
        orc_set_threads(t ? 4 : 1);
        int rc = orc_full_flow(file, len, seed, epoch, st[t], st_cap, &st_len[t], rec[t], rec_cap, &rec_len[t],
                               op[t]);
        CHECK(rc == 0);
    }
    CHECK(st_len[0] == st_len[1] && memcmp(st[0], st[1], st_len[0]) == 0);
    CHECK(rec_len[0] == rec_len[1] && memcmp(rec[0], rec[1], rec_len[0]) == 0);
    CHECK(op[0][0] == op[1][0] && op[0][1] == op[1][1]);

    /* the step-by-step prover and verifier API on the same file */
    orc_prover *p = NULL;
    CHECK(orc_prover_new(file, len, seed, epoch, 1 << 30, 1, &p) == 0);
    if (p) {
        size_t slen = orc_prover_statement(p, NULL, 0);
        uint8_t *sb = malloc(slen);
        CHECK(orc_prover_statement(p, sb, slen) == slen);
        orc_verifier *v = NULL;
        CHECK(orc_verifier_new(epoch, sb, slen, &v) == 0);
        size_t rows, ring;
        orc_prover_dims(p, &rows, &ring);
        uint64_t *cur = malloc(2 * ORC_K * ring * sizeof *cur);
        orc_prover_current_commitment(p, cur);
        size_t rounds = orc_prover_rounds(p);
        uint8_t *r = malloc(rec_cap);
        for (size_t j = 0; j < rounds && v; ++j) {
            size_t rl = 0;
            CHECK(orc_prover_fold_round(p, j, r, rec_cap, &rl) == 0);
            orc_verifier *vc = orc_verifier_clone(v);
            CHECK(orc_verifier_update(v, r, rl) == 0);
            if (rl > 1) { /* a record with a flipped byte: decoding and the update stay in bounds */
                r[rl / 2] ^= 0x5A;
                (void)orc_verifier_update(vc, r, rl);
            }
            orc_verifier_free(vc);
        }
        uint64_t opening[2];
        CHECK(orc_prover_final_opening(p, opening) == 0);
        if (v) CHECK(orc_verifier_verify_final_opening(v, opening) == 0);
        orc_verifier_free(v);
        free(r);
        free(cur);
        free(sb);
        orc_prover_free(p);
    }
    for (int t = 0; t < 2; ++t) {
        free(st[t]);
        free(rec[t]);
    }
    free(file);
}

static void check_commit_and_transcript(void) {
    enum { ROWS = 4, RING = 16 };
    uint8_t seed[32], epoch[32], fid[32];
    memset(seed, 5, 32);
    memset(epoch, 9, 32);
    memset(fid, 1, 32);
    uint64_t w[2 * ROWS * RING], g[2 * ROWS * ORC_K * RING], c1[2 * ORC_K * RING], c2[2 * ORC_K * RING];
    random_zq(w, ROWS * RING);
    size_t out_ring = 0;
    orc_derive_generators(seed, ROWS, RING, g);
    for (int threads = 1; threads <= 3; threads += 2) {
        orc_set_threads(threads);
        CHECK(orc_commit_with_generators(w, g, ROWS, RING, 1, c1, &out_ring) == 0);
        CHECK(orc_commit_with_generators(w, g, ROWS, RING, 0, c2, &out_ring) == 0);
        CHECK(memcmp(c1, c2, sizeof c1) == 0); /* NTT on and off agree */
    }
    uint64_t x0[2], x1[2];
    orc_challenge(epoch, "vec", fid, 3, c1, RING, c1, RING, c2, RING, x0);
    orc_set_challenge_reject(2);
    orc_challenge(epoch, "vec", fid, 3, c1, RING, c1, RING, c2, RING, x1);
    orc_set_challenge_reject(0);
    CHECK(x0[0] != x1[0] || x0[1] != x1[1]);
    uint32_t seen[8];
    size_t n_seen = 0;
    CHECK(orc_sample_nonzero_mock(3, 42, seen, 8, &n_seen) == 42 && n_seen == 4);
    size_t depth, dim, coeffs, cap;
    for (uint64_t n = 0; n < 70000; n += 4099) CHECK(orc_layout_for_byte_len(n, &depth, &dim, &coeffs, &cap) == 0);
}

int main(void) {
    check_field();
    check_blake3();
    check_ntt();
    check_commit_and_transcript();
    const size_t lens[] = {0, 1, 1025, 53248, 53249};
    for (size_t i = 0; i < sizeof lens / sizeof lens[0]; ++i) check_flow(lens[i]);
    orc_set_threads(1);
    if (failures) {
        fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    printf("sanitize_check: all checks passed\n");
    return 0;
}
