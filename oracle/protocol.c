/* Oracle: commitment, transcript, witness expansion, prover and verifier (test
 * infrastructure only). Restates src/commitment/sis.rs, src/protocol/{transcript,prover,
 * verifier}.rs and src/utils/io.rs of the reference, step for step. */
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>

#include "internal.h"
#include "oracle.h"

#define K ORC_K

int orc_threads = 1;
void orc_set_threads(int t) { orc_threads = t < 1 ? 1 : (t > 256 ? 256 : t); }
int orc_get_threads(void) { return orc_threads; }

static void *xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

/* A parallel loop over [0, n) on orc_threads POSIX threads: each worker takes the next index from
 * a shared counter and calls body(ctx, worker, index). Workers are joined before it returns, so
 * everything they wrote is visible to the caller afterwards. */
typedef void (*par_body)(void *ctx, int worker, size_t index);
typedef struct {
    par_body body;
    void *ctx;
    size_t n;
    atomic_size_t next;
    int worker;
} par_job;
typedef struct {
    par_job *job;
    int worker;
} par_arg;

static void *par_worker(void *p) {
    const par_arg *a = (const par_arg *)p;
    for (;;) {
        const size_t i = atomic_fetch_add(&a->job->next, 1);
        if (i >= a->job->n) break;
        a->job->body(a->job->ctx, a->worker, i);
    }
    return NULL;
}

static void par_for(size_t n, int workers, par_body body, void *ctx) {
    par_job job = {body, ctx, n, 0, 0};
    pthread_t tid[256];
    par_arg args[256];
    int started = 0;
    for (int w = 1; w < workers; ++w) { /* worker 0 is the calling thread */
        args[w] = (par_arg){&job, w};
        if (pthread_create(&tid[w], NULL, par_worker, &args[w]) != 0) break;
        started = w;
    }
    args[0] = (par_arg){&job, 0};
    par_worker(&args[0]);
    for (int w = 1; w <= started; ++w) pthread_join(tid[w], NULL);
}

/* ---------------- generators (sis.rs:198-245) ---------------- */

/* sis.rs:208-210 */
void orc_derive_subseed(const uint8_t seed[32], const uint8_t *label, size_t label_len,
                        uint8_t out[32]) {
    orc_blake3_keyed_hash(seed, label, label_len, out);
}

static void put_u64le(uint8_t *dst, uint64_t v) {
    for (int i = 0; i < 8; ++i) dst[i] = (uint8_t)(v >> (8 * i));
}

/* sis.rs:231-245 derive_field: keyed_blake3(seed, "g" || i || k || c) mod q */
static u128 derive_field(const uint8_t seed[32], size_t i, size_t k, size_t c) {
    uint8_t msg[25], h[32];
    msg[0] = 'g';
    put_u64le(msg + 1, i);
    put_u64le(msg + 9, k);
    put_u64le(msg + 17, c);
    orc_blake3_keyed_hash(seed, msg, sizeof msg, h);
    return zq_from_le_bytes_mod_order(h, 32);
}

typedef struct {
    const uint8_t *seed;
    size_t ring_len;
    u128 *g;
} gen_ctx;

static void gen_row(void *p, int worker, size_t i) {
    const gen_ctx *c = (const gen_ctx *)p;
    for (size_t k = 0; k < K; ++k)
        for (size_t j = 0; j < c->ring_len; ++j) c->g[(i * K + k) * c->ring_len + j] = derive_field(c->seed, i, k, j);
}

static void derive_generators_u128(const uint8_t seed[32], size_t rows, size_t ring_len, u128 *g) {
    gen_ctx c = {seed, ring_len, g};
    par_for(rows, orc_threads, gen_row, &c);
}

void orc_derive_generators(const uint8_t g_seed[32], size_t rows, size_t ring_len, uint64_t *out) {
    u128 *g = (u128 *)xcalloc(rows * K * ring_len, sizeof(u128));
    if (!g) return;
    derive_generators_u128(g_seed, rows, ring_len, g);
    for (size_t i = 0; i < rows * K * ring_len; ++i) zq_store(out + 2 * i, g[i]);
    free(g);
}

/* ---------------- commitment (sis.rs:146-167) ---------------- */

/* one term g[i]_k * w[i] (t = i * K + k) added into the worker's own accumulator */
typedef struct {
    const u128 *w, *g;
    size_t m;
    int ntt;
    u128 *prod, *acc; /* per worker: m and K * m values */
    int *rc;          /* per worker */
} commit_ctx;

static void commit_term(void *p, int worker, size_t t) {
    const commit_ctx *c = (const commit_ctx *)p;
    if (c->rc[worker] != ORC_OK) return;
    const size_t i = t / K, k = t % K, m = c->m;
    u128 *prod = c->prod + (size_t)worker * m, *acc = c->acc + (size_t)worker * K * m;
    c->rc[worker] = ring_mul(c->g + (i * K + k) * m, c->w + i * m, m, c->ntt, prod);
    for (size_t j = 0; j < m; ++j) acc[k * m + j] = zq_add(acc[k * m + j], prod[j]);
}

/* acc_k = sum_i g[i]_k * w[i]; rows == 0 -> zero(1). out holds 8*m (or 8 when rows == 0). */
static int commit_u128(const u128 *w, const u128 *g, size_t rows, size_t m, int ntt, u128 *out,
                       size_t *out_len) {
    if (rows == 0) {
        for (int k = 0; k < K; ++k) out[k] = 0;
        *out_len = 1;
        return ORC_OK;
    }
    *out_len = m;
    memset(out, 0, K * m * sizeof(u128));
    int rc = ORC_OK;
    if (orc_threads <= 1) {
        /* the reference's sequential accumulate loop (sis.rs:163-165) */
        u128 *prod = (u128 *)xcalloc(m, sizeof(u128));
        if (!prod) return ORC_ERR_OOM;
        for (size_t i = 0; i < rows && rc == ORC_OK; ++i)
            for (int k = 0; k < K && rc == ORC_OK; ++k) {
                rc = ring_mul(g + (i * K + k) * m, w + i * m, m, ntt, prod);
                for (size_t c = 0; c < m; ++c) out[k * m + c] = zq_add(out[k * m + c], prod[c]);
            }
        free(prod);
        return rc;
    }
    /* same products, summed mod q in another order: Z_q addition is exact, so the
     * result is identical to the sequential loop */
    const int workers = orc_threads;
    commit_ctx c = {w, g, m, ntt, (u128 *)xcalloc((size_t)workers * m, sizeof(u128)),
                    (u128 *)xcalloc((size_t)workers * K * m, sizeof(u128)),
                    (int *)xcalloc((size_t)workers, sizeof(int))};
    if (c.prod && c.acc && c.rc) {
        par_for(rows * K, workers, commit_term, &c);
        for (int t = 0; t < workers; ++t) {
            if (c.rc[t] != ORC_OK) rc = c.rc[t];
            for (size_t j = 0; j < K * m; ++j) out[j] = zq_add(out[j], c.acc[(size_t)t * K * m + j]);
        }
    } else {
        rc = ORC_ERR_OOM;
    }
    free(c.prod);
    free(c.acc);
    free(c.rc);
    return rc;
}

int orc_commit_with_generators(const uint64_t *witness, const uint64_t *gens, size_t rows,
                               size_t ring_len, int ntt_enabled, uint64_t *out,
                               size_t *out_ring_len) {
    u128 *w = (u128 *)xcalloc(rows * ring_len, sizeof(u128));
    u128 *g = (u128 *)xcalloc(rows * K * ring_len, sizeof(u128));
    u128 *o = (u128 *)xcalloc(K * (ring_len ? ring_len : 1), sizeof(u128));
    int rc = ORC_ERR_OOM;
    if (w && g && o) {
        for (size_t i = 0; i < rows * ring_len; ++i) w[i] = zq_load(witness + 2 * i);
        for (size_t i = 0; i < rows * K * ring_len; ++i) g[i] = zq_load(gens + 2 * i);
        size_t ol = 0;
        rc = commit_u128(w, g, rows, ring_len, ntt_enabled, o, &ol);
        if (rc == ORC_OK) {
            for (size_t i = 0; i < K * ol; ++i) zq_store(out + 2 * i, o[i]);
            *out_ring_len = ol;
        }
    }
    free(w);
    free(g);
    free(o);
    return rc;
}

/* sis.rs:111-132: generators derived from g_0_seed = keyed(seed, "g0") */
int orc_commit(const uint8_t seed[32], const uint64_t *witness, size_t rows, size_t ring_len,
               int ntt_enabled, uint64_t *out, size_t *out_ring_len) {
    uint8_t gs[32];
    orc_derive_subseed(seed, (const uint8_t *)"g0", 2, gs);
    uint64_t *g = (uint64_t *)xcalloc(rows * K * ring_len * 2, sizeof(uint64_t));
    if (!g) return ORC_ERR_OOM;
    orc_derive_generators(gs, rows, ring_len, g);
    int rc = orc_commit_with_generators(witness, g, rows, ring_len, ntt_enabled, out, out_ring_len);
    free(g);
    return rc;
}

/* ---------------- transcript (transcript.rs:42-90) ---------------- */

typedef struct {
    const u128 *v; /* K x len */
    size_t len;
} modref;

static void absorb_u64(orc_b3 *h, uint64_t v) {
    uint8_t b[8];
    put_u64le(b, v);
    orc_b3_update(h, b, 8);
}

/* transcript.rs:82-90 */
static void absorb_commitment(orc_b3 *h, modref m) {
    absorb_u64(h, K);
    for (int k = 0; k < K; ++k) {
        absorb_u64(h, m.len);
        uint8_t buf[16 * 64];
        size_t pos = 0;
        for (size_t c = 0; c < m.len; ++c) {
            u128 v = m.v[k * m.len + c];
            for (int b = 0; b < 16; ++b) buf[pos++] = (uint8_t)(v >> (8 * b));
            if (pos == sizeof buf) {
                orc_b3_update(h, buf, pos);
                pos = 0;
            }
        }
        if (pos) orc_b3_update(h, buf, pos);
    }
}

/* transcript.rs:66-75: sample_nonzero_challenge. Candidates are derived for counter = 0, 1, ...
 * until one is non-zero. */
typedef u128 (*candidate_fn)(uint32_t counter, void *ctx);
static u128 sample_nonzero_challenge(candidate_fn derive, void *ctx) {
    uint32_t counter = 0;
    for (;;) {
        u128 c = derive(counter, ctx);
        if (c != 0) return c;
        counter += 1; /* wrapping_add */
    }
}

/* Test knob: challenges treat the candidates of counters below orc_challenge_reject as zero, the
 * branch a real digest reaches with probability about 2^-112 per candidate. 0 (the default) is
 * the reference. The device honours the same knob, so both sides can be checked on it. */
static uint32_t orc_challenge_reject = 0;
void orc_set_challenge_reject(uint32_t n) { orc_challenge_reject = n; }

typedef struct {
    const uint8_t *epoch_seed, *file_id;
    const char *label;
    uint64_t round;
    modref cur, l, r;
} tx_ctx;

/* transcript.rs:50-61: the candidate for one counter */
static u128 tx_candidate(uint32_t counter, void *vctx) {
    const tx_ctx *t = (const tx_ctx *)vctx;
    size_t llen = strlen(t->label);
    orc_b3 h;
    orc_b3_init(&h);
    orc_b3_update(&h, t->epoch_seed, 32);
    absorb_u64(&h, llen);
    orc_b3_update(&h, t->label, llen);
    absorb_u64(&h, 32);
    orc_b3_update(&h, t->file_id, 32);
    absorb_u64(&h, 0); /* epoch_index */
    absorb_u64(&h, t->round);
    absorb_commitment(&h, t->cur);
    absorb_commitment(&h, t->l);
    absorb_commitment(&h, t->r);
    uint8_t c4[4] = {(uint8_t)counter, (uint8_t)(counter >> 8), (uint8_t)(counter >> 16),
                     (uint8_t)(counter >> 24)};
    orc_b3_update(&h, c4, 4);
    uint8_t dig[32];
    orc_b3_finalize(&h, dig, 32);
    if (counter < orc_challenge_reject) return 0;
    return zq_from_le_bytes_mod_order(dig, 32);
}

/* transcript.rs:42-63: stateless challenge */
static u128 challenge_u128(const uint8_t epoch_seed[32], const char *label,
                           const uint8_t file_id[32], uint64_t round, modref cur, modref l,
                           modref r) {
    tx_ctx t = {epoch_seed, file_id, label, round, cur, l, r};
    return sample_nonzero_challenge(tx_candidate, &t);
}

/* The reference's sampler test (transcript.rs:130-143) drives sample_nonzero_challenge with a
 * mock candidate: zero for counters below `zero_below`, `value` after. Records the counters
 * seen (at most `cap`) and returns the sampled value. */
typedef struct {
    uint32_t zero_below;
    uint64_t value;
    uint32_t *seen;
    size_t cap, n;
} mock_ctx;
static u128 mock_candidate(uint32_t counter, void *vctx) {
    mock_ctx *m = (mock_ctx *)vctx;
    if (m->n < m->cap) m->seen[m->n] = counter;
    m->n += 1;
    return counter < m->zero_below ? 0 : (u128)m->value;
}
uint64_t orc_sample_nonzero_mock(uint32_t zero_below, uint64_t value, uint32_t *seen, size_t cap,
                                 size_t *n_seen) {
    mock_ctx m = {zero_below, value, seen, cap, 0};
    u128 v = sample_nonzero_challenge(mock_candidate, &m);
    if (n_seen) *n_seen = m.n;
    return (uint64_t)v;
}

static u128 *load_mod(const uint64_t *w, size_t len) {
    u128 *v = (u128 *)xcalloc(K * len, sizeof(u128));
    if (v)
        for (size_t i = 0; i < K * len; ++i) v[i] = zq_load(w + 2 * i);
    return v;
}

void orc_challenge(const uint8_t epoch_seed[32], const char *label, const uint8_t file_id[32],
                   uint64_t round, const uint64_t *cur, size_t len_cur, const uint64_t *lhs,
                   size_t len_l, const uint64_t *rhs, size_t len_r, uint64_t out[2]) {
    u128 *c = load_mod(cur, len_cur), *l = load_mod(lhs, len_l), *r = load_mod(rhs, len_r);
    modref mc = {c, len_cur}, ml = {l, len_l}, mr = {r, len_r};
    zq_store(out, challenge_u128(epoch_seed, label, file_id, round, mc, ml, mr));
    free(c);
    free(l);
    free(r);
}

/* ---------------- io (io.rs:29-120) ---------------- */

size_t orc_coeff_count_for_byte_len(uint64_t byte_len) {
    uint64_t c = (byte_len + ORC_BYTES_PER_COEFF - 1) / ORC_BYTES_PER_COEFF;
    return c < 1 ? 1 : (size_t)c;
}

int orc_layout_for_byte_len(uint64_t byte_len, size_t *depth, size_t *dimension,
                            size_t *coeff_count, size_t *capacity) {
    size_t cc = orc_coeff_count_for_byte_len(byte_len);
    size_t d = ORC_MIN_FOLD_DEPTH;
    for (;; ++d) {
        if (d >= 32) return ORC_ERR_ARG;
        size_t dim = (size_t)1 << d;
        if (dim * dim >= cc) {
            if (depth) *depth = d;
            if (dimension) *dimension = dim;
            if (coeff_count) *coeff_count = cc;
            if (capacity) *capacity = dim * dim;
            return ORC_OK;
        }
    }
}

static int expand_u128(const uint8_t *file, size_t len, size_t max_bytes, uint8_t file_id[32],
                       u128 **msg, size_t *dim_out, size_t *depth_out) {
    orc_blake3_hash(file, len, file_id);
    size_t depth, dim, cc, cap;
    if (orc_layout_for_byte_len(len, &depth, &dim, &cc, &cap) != ORC_OK) return ORC_ERR_ARG;
    /* io.rs:42-49 memory guard: capacity * sizeof(Fp) * 2 */
    if (cap > SIZE_MAX / 32 || cap * 32 > max_bytes) return ORC_ERR_MEMORY_LIMIT;
    u128 *m = (u128 *)xcalloc(cap, sizeof(u128));
    if (!m) return ORC_ERR_OOM;
    for (size_t idx = 0; idx * ORC_BYTES_PER_COEFF < len; ++idx) {
        size_t off = idx * ORC_BYTES_PER_COEFF;
        size_t n = len - off < ORC_BYTES_PER_COEFF ? len - off : ORC_BYTES_PER_COEFF;
        u128 acc = 0;
        for (size_t s = 0; s < n; ++s) acc |= ((u128)file[off + s]) << (8 * s);
        m[idx] = acc % ZQ_Q;
    }
    *msg = m;
    *dim_out = dim;
    *depth_out = depth;
    return ORC_OK;
}

int orc_expand(const uint8_t *file, size_t len, size_t max_preprocess_bytes, uint8_t file_id[32],
               uint64_t *out, size_t out_capacity_coeffs) {
    u128 *m = NULL;
    size_t dim, depth;
    int rc = expand_u128(file, len, max_preprocess_bytes, file_id, &m, &dim, &depth);
    if (rc != ORC_OK) return rc;
    if (out_capacity_coeffs < dim * dim) {
        free(m);
        return ORC_ERR_ARG;
    }
    for (size_t i = 0; i < dim * dim; ++i) zq_store(out + 2 * i, m[i]);
    free(m);
    return ORC_OK;
}

/* ---------------- module helpers (module.rs, poly.rs) ---------------- */

/* dst = a + s*b over n elements (Poly::add(&r.scale(s))) */
static void axpy(const u128 *a, const u128 *b, u128 s, size_t n, u128 *dst) {
    for (size_t i = 0; i < n; ++i) dst[i] = zq_add(a[i], zq_mul(b[i], s));
}

/* poly.rs:144-154 mul_by_x in Z_q[X]/(X^n+1) */
static void mul_by_x(const u128 *in, size_t n, u128 *out) {
    u128 last = in[n - 1];
    for (size_t i = n - 1; i >= 1; --i) out[i] = in[i - 1];
    out[0] = zq_neg(last);
}

void orc_poly_mul_by_x(const uint64_t *in, size_t n, uint64_t *out) {
    if (n == 0) return;
    u128 *a = (u128 *)xcalloc(n, sizeof(u128)), *o = (u128 *)xcalloc(n, sizeof(u128));
    for (size_t i = 0; i < n; ++i) a[i] = zq_load(in + 2 * i);
    mul_by_x(a, n, o);
    for (size_t i = 0; i < n; ++i) zq_store(out + 2 * i, o[i]);
    free(a);
    free(o);
}

/* ---------------- bincode 1.3 (fixint, little endian) ---------------- */

static size_t module_bytes(size_t len) { return 8 + K * (8 + 16 * len); }

size_t orc_record_size(size_t lv, size_t lp) { return 8 + 2 * module_bytes(lv) + 2 * module_bytes(lp); }

static uint8_t *put_module(uint8_t *p, const u128 *v, size_t len) {
    put_u64le(p, K);
    p += 8;
    for (int k = 0; k < K; ++k) {
        put_u64le(p, len);
        p += 8;
        for (size_t c = 0; c < len; ++c) {
            u128 x = v[k * len + c];
            for (int b = 0; b < 16; ++b) *p++ = (uint8_t)(x >> (8 * b));
        }
    }
    return p;
}

typedef struct {
    const uint8_t *p;
    size_t left;
} reader;

static int rd_u64(reader *r, uint64_t *v) {
    if (r->left < 8) return ORC_ERR_DECODE;
    uint64_t x = 0;
    for (int i = 0; i < 8; ++i) x |= (uint64_t)r->p[i] << (8 * i);
    r->p += 8;
    r->left -= 8;
    *v = x;
    return ORC_OK;
}

/* ModuleElement<8> deserialize (module.rs:104-115): Vec<Poly> of exactly 8 coords
 * with a shared ring length; each Fp is from_u128 (reduced mod q, field.rs:121-129). */
static int rd_module(reader *r, u128 **out, size_t *len_out) {
    uint64_t n;
    if (rd_u64(r, &n)) return ORC_ERR_DECODE;
    /* serde reads every element before the rank check fires */
    u128 *vals = NULL;
    size_t len = 0;
    int rc = ORC_OK;
    for (uint64_t k = 0; k < n; ++k) {
        uint64_t l;
        if (rd_u64(r, &l) || l > r->left / 16) {
            rc = ORC_ERR_DECODE;
            break;
        }
        if (k == 0) {
            len = (size_t)l;
            vals = (u128 *)xcalloc(K * (len ? len : 1), sizeof(u128));
            if (!vals) return ORC_ERR_OOM;
        } else if (l != len) {
            rc = ORC_ERR_SHAPE; /* keep reading to mirror serde, but the result is an error */
        }
        for (uint64_t c = 0; c < l; ++c) {
            u128 x = 0;
            for (int b = 0; b < 16; ++b) x |= ((u128)r->p[b]) << (8 * b);
            r->p += 16;
            r->left -= 16;
            if (k < K && l == len) vals[k * len + c] = x % ZQ_Q;
        }
    }
    if (rc == ORC_OK && n != K) rc = ORC_ERR_DECODE; /* "module rank mismatch" */
    if (rc != ORC_OK) {
        free(vals);
        return rc == ORC_ERR_SHAPE ? ORC_ERR_DECODE : rc;
    }
    *out = vals;
    *len_out = len;
    return ORC_OK;
}

typedef struct {
    uint64_t round;
    u128 *m[4]; /* l_vec, r_vec, l_poly, r_poly */
    size_t len[4];
} record_t;

static void record_free(record_t *r) {
    for (int i = 0; i < 4; ++i) free(r->m[i]);
}

static int decode_record(const uint8_t *buf, size_t n, record_t *rec) {
    reader r = {buf, n};
    memset(rec, 0, sizeof *rec);
    if (rd_u64(&r, &rec->round)) return ORC_ERR_DECODE;
    for (int i = 0; i < 4; ++i) {
        int rc = rd_module(&r, &rec->m[i], &rec->len[i]);
        if (rc) {
            record_free(rec);
            return rc;
        }
    }
    return ORC_OK; /* bincode::deserialize allows trailing bytes */
}

/* ---------------- prover (prover.rs) ---------------- */

struct orc_prover {
    uint8_t epoch_seed[32], g_seed[32], file_id[32];
    uint64_t original_len;
    size_t d, rounds, round;
    size_t rows, ring_len; /* message rows x ring_len; generators rows x K x ring_len */
    u128 *msg, *g;
    u128 *com; /* K x com_len */
    size_t com_len;
    u128 *com0;
    size_t com0_len;
    int ntt;
};

static size_t statement_bytes(size_t com_len) { return 32 + 8 + 8 + module_bytes(com_len) + 32; }

size_t orc_prover_statement(const orc_prover *p, uint8_t *out, size_t cap) {
    size_t need = statement_bytes(p->com0_len);
    if (out && cap >= need) {
        uint8_t *q = out;
        memcpy(q, p->file_id, 32);
        q += 32;
        put_u64le(q, p->original_len);
        q += 8;
        put_u64le(q, p->d);
        q += 8;
        q = put_module(q, p->com0, p->com0_len);
        memcpy(q, p->g_seed, 32);
    }
    return need;
}

/* prover.rs:61-111 initialize_with_config */
int orc_prover_new(const uint8_t *file, size_t len, const uint8_t seed_generators[32],
                   const uint8_t epoch_seed[32], size_t max_preprocess_bytes, int ntt_enabled,
                   orc_prover **out) {
    orc_prover *p = (orc_prover *)calloc(1, sizeof *p);
    if (!p) return ORC_ERR_OOM;
    size_t dim, depth;
    int rc = expand_u128(file, len, max_preprocess_bytes, p->file_id, &p->msg, &dim, &depth);
    if (rc != ORC_OK) {
        free(p);
        return rc;
    }
    memcpy(p->epoch_seed, epoch_seed, 32);
    orc_derive_subseed(seed_generators, (const uint8_t *)"g0", 2, p->g_seed);
    p->original_len = len;
    p->d = depth;
    p->rounds = depth;
    p->rows = dim;
    p->ring_len = dim;
    p->ntt = ntt_enabled;
    p->g = (u128 *)xcalloc(dim * K * dim, sizeof(u128));
    p->com = (u128 *)xcalloc(K * dim, sizeof(u128));
    p->com0 = (u128 *)xcalloc(K * dim, sizeof(u128));
    if (!p->g || !p->com || !p->com0) {
        orc_prover_free(p);
        return ORC_ERR_OOM;
    }
    derive_generators_u128(p->g_seed, dim, dim, p->g);
    rc = commit_u128(p->msg, p->g, dim, dim, ntt_enabled, p->com0, &p->com0_len);
    if (rc != ORC_OK) {
        orc_prover_free(p);
        return rc;
    }
    memcpy(p->com, p->com0, K * dim * sizeof(u128));
    p->com_len = p->com0_len;
    *out = p;
    return ORC_OK;
}

void orc_prover_free(orc_prover *p) {
    if (!p) return;
    free(p->msg);
    free(p->g);
    free(p->com);
    free(p->com0);
    free(p);
}

size_t orc_prover_rounds(const orc_prover *p) { return p->rounds; }
size_t orc_prover_round(const orc_prover *p) { return p->round; }
void orc_prover_dims(const orc_prover *p, size_t *rows, size_t *ring_len) {
    *rows = p->rows;
    *ring_len = p->ring_len;
}
void orc_prover_current_commitment(const orc_prover *p, uint64_t *out) {
    for (size_t i = 0; i < K * p->com_len; ++i) zq_store(out + 2 * i, p->com[i]);
}
void orc_prover_message(const orc_prover *p, uint64_t *out) {
    for (size_t i = 0; i < p->rows * p->ring_len; ++i) zq_store(out + 2 * i, p->msg[i]);
}
void orc_prover_generators(const orc_prover *p, uint64_t *out) {
    for (size_t i = 0; i < p->rows * K * p->ring_len; ++i) zq_store(out + 2 * i, p->g[i]);
}

static int mod_eq(const u128 *a, size_t la, const u128 *b, size_t lb) {
    return la == lb && memcmp(a, b, K * la * sizeof(u128)) == 0;
}

/* prover.rs:140-313 fold_round */
int orc_prover_fold_round(orc_prover *p, size_t round, uint8_t *out, size_t cap, size_t *out_len) {
    if (round != p->round || p->round >= p->rounds) return ORC_ERR_ROUND;
    const size_t rows = p->rows, n = p->ring_len;
    int rc = ORC_OK;
    /* ---- vector stage (prover.rs:148-227) ---- */
    size_t vrows = rows, vlen_cross = n;
    u128 *l_vec = (u128 *)xcalloc(K * n, sizeof(u128)), *r_vec = (u128 *)xcalloc(K * n, sizeof(u128));
    u128 *vfc = (u128 *)xcalloc(K * n, sizeof(u128));
    u128 *fm = NULL, *fg = NULL;
    size_t ll = 0, rl = 0;
    if (!l_vec || !r_vec || !vfc) return ORC_ERR_OOM;
    if (rows > 1) {
        size_t mid = rows / 2;
        const u128 *mL = p->msg, *mR = p->msg + mid * n;
        const u128 *gL = p->g, *gR = p->g + mid * K * n;
        rc = commit_u128(mL, gR, mid, n, p->ntt, l_vec, &ll);
        if (!rc) rc = commit_u128(mR, gL, mid, n, p->ntt, r_vec, &rl);
        if (rc) goto fail;
        modref mc = {p->com, p->com_len}, ml = {l_vec, ll}, mr = {r_vec, rl};
        u128 x = challenge_u128(p->epoch_seed, "vec", p->file_id, round, mc, ml, mr);
        u128 xi = zq_inv(x);
        for (size_t i = 0; i < K * n; ++i)
            vfc[i] = zq_add(zq_add(p->com[i], zq_mul(l_vec[i], xi)), zq_mul(r_vec[i], x));
        fm = (u128 *)xcalloc(mid * n, sizeof(u128));
        fg = (u128 *)xcalloc(mid * K * n, sizeof(u128));
        if (!fm || !fg) {
            rc = ORC_ERR_OOM;
            goto fail;
        }
        axpy(mL, mR, x, mid * n, fm);      /* fold_vec_poly: l + x*r */
        axpy(gL, gR, xi, mid * K * n, fg); /* fold_vec_module: l + x^-1*r */
        u128 *chk = (u128 *)xcalloc(K * n, sizeof(u128));
        size_t cl = 0;
        rc = chk ? commit_u128(fm, fg, mid, n, p->ntt, chk, &cl) : ORC_ERR_OOM;
        if (!rc && !mod_eq(vfc, n, chk, cl)) rc = ORC_ERR_INVARIANT;
        free(chk);
        if (rc) goto fail;
        vrows = mid;
    } else {
        /* unreachable for square layouts; kept for parity (prover.rs:215-227) */
        ll = rl = n;
        modref mc = {p->com, p->com_len}, mz = {l_vec, n};
        (void)challenge_u128(p->epoch_seed, "vec", p->file_id, round, mc, mz, mz);
        memcpy(vfc, p->com, K * n * sizeof(u128));
        fm = (u128 *)xcalloc(rows * n, sizeof(u128));
        fg = (u128 *)xcalloc(rows * K * n, sizeof(u128));
        memcpy(fm, p->msg, rows * n * sizeof(u128));
        memcpy(fg, p->g, rows * K * n * sizeof(u128));
    }
    (void)vlen_cross;
    /* ---- poly stage (prover.rs:230-277) ---- */
    u128 *l_poly = NULL, *r_poly = NULL, *nm = NULL, *ng = NULL, *ncom = NULL;
    size_t lpl = 0, rpl = 0, nlen;
    if (n > 1) {
        size_t h = n / 2;
        u128 *e = (u128 *)xcalloc(vrows * h, sizeof(u128)), *o = (u128 *)xcalloc(vrows * h, sizeof(u128));
        u128 *ge = (u128 *)xcalloc(vrows * K * h, sizeof(u128));
        u128 *gox = (u128 *)xcalloc(vrows * K * h, sizeof(u128));
        u128 *tmp = (u128 *)xcalloc(h, sizeof(u128));
        l_poly = (u128 *)xcalloc(K * h, sizeof(u128));
        r_poly = (u128 *)xcalloc(K * h, sizeof(u128));
        ncom = (u128 *)xcalloc(K * h, sizeof(u128));
        nm = (u128 *)xcalloc(vrows * h, sizeof(u128));
        ng = (u128 *)xcalloc(vrows * K * h, sizeof(u128));
        if (!e || !o || !ge || !gox || !tmp || !l_poly || !r_poly || !ncom || !nm || !ng) rc = ORC_ERR_OOM;
        if (!rc) {
            for (size_t i = 0; i < vrows; ++i)
                for (size_t c = 0; c < h; ++c) { /* poly.rs:87-96 */
                    e[i * h + c] = fm[i * n + 2 * c];
                    o[i * h + c] = fm[i * n + 2 * c + 1];
                }
            for (size_t i = 0; i < vrows * K; ++i) { /* prover.rs:408-419 */
                for (size_t c = 0; c < h; ++c) {
                    ge[i * h + c] = fg[i * n + 2 * c];
                    tmp[c] = fg[i * n + 2 * c + 1];
                }
                mul_by_x(tmp, h, gox + i * h);
            }
            rc = commit_u128(e, gox, vrows, h, p->ntt, l_poly, &lpl);
            if (!rc) rc = commit_u128(o, ge, vrows, h, p->ntt, r_poly, &rpl);
        }
        if (!rc) {
            modref mc = {vfc, n}, ml = {l_poly, lpl}, mr = {r_poly, rpl};
            u128 y = challenge_u128(p->epoch_seed, "poly", p->file_id, round, mc, ml, mr);
            u128 yi = zq_inv(y);
            for (int k = 0; k < K; ++k)
                for (size_t c = 0; c < h; ++c) {
                    size_t i = k * h + c;
                    ncom[i] = zq_add(zq_add(vfc[k * n + 2 * c], zq_mul(l_poly[i], yi)), zq_mul(r_poly[i], y));
                }
            axpy(e, o, y, vrows * h, nm);        /* fold_poly_poly: e + y*o */
            axpy(ge, gox, yi, vrows * K * h, ng); /* fold_poly_module: g_e + y^-1*X*g_o */
        }
        free(e);
        free(o);
        free(ge);
        free(gox);
        free(tmp);
        nlen = h;
        if (rc) goto fail2;
    } else {
        /* unreachable for square layouts (prover.rs:265-277) */
        lpl = rpl = 1;
        l_poly = (u128 *)xcalloc(K, sizeof(u128));
        r_poly = (u128 *)xcalloc(K, sizeof(u128));
        modref mc = {vfc, n}, mz = {l_poly, 1};
        (void)challenge_u128(p->epoch_seed, "poly", p->file_id, round, mc, mz, mz);
        nm = fm;
        ng = fg;
        fm = fg = NULL;
        ncom = vfc;
        vfc = NULL;
        nlen = n;
    }
    /* ---- post-poly invariant (prover.rs:279-291) ---- */
    {
        u128 *chk = (u128 *)xcalloc(K * nlen, sizeof(u128));
        size_t cl = 0;
        rc = chk ? commit_u128(nm, ng, vrows, nlen, p->ntt, chk, &cl) : ORC_ERR_OOM;
        if (!rc && !mod_eq(ncom, nlen, chk, cl)) rc = ORC_ERR_INVARIANT;
        free(chk);
        if (rc) goto fail2;
    }
    /* ---- emit the record (bincode) and advance state ---- */
    {
        size_t need = orc_record_size(ll, lpl);
        if (out_len) *out_len = need;
        if (out && cap >= need) {
            uint8_t *q = out;
            put_u64le(q, round);
            q += 8;
            q = put_module(q, l_vec, ll);
            q = put_module(q, r_vec, rl);
            q = put_module(q, l_poly, lpl);
            q = put_module(q, r_poly, rpl);
        }
    }
    free(p->msg);
    free(p->g);
    free(p->com);
    p->msg = nm;
    p->g = ng;
    p->com = ncom;
    p->com_len = nlen;
    p->rows = vrows;
    p->ring_len = nlen;
    p->round += 1;
    free(l_vec);
    free(r_vec);
    free(vfc);
    free(fm);
    free(fg);
    free(l_poly);
    free(r_poly);
    return ORC_OK;
fail2:
    free(l_poly);
    free(r_poly);
    free(nm);
    free(ng);
    free(ncom);
fail:
    free(l_vec);
    free(r_vec);
    free(vfc);
    free(fm);
    free(fg);
    return rc;
}

/* prover.rs:133-138 */
int orc_prover_final_opening(const orc_prover *p, uint64_t out[2]) {
    if (!(p->round == p->rounds && p->rows == 1 && p->ring_len == 1)) return ORC_ERR_NOT_FINAL;
    zq_store(out, p->msg[0]);
    return ORC_OK;
}

/* ---------------- verifier (verifier.rs) ---------------- */

struct orc_verifier {
    uint8_t epoch_seed[32], file_id[32];
    size_t rounds, round;
    size_t g_rows, g_len;
    u128 *g; /* g_rows x K x g_len */
    u128 *com;
    size_t com_len;
};

/* verifier.rs:22-43 */
int orc_verifier_new(const uint8_t epoch_seed[32], const uint8_t *st, size_t st_len,
                     orc_verifier **out) {
    reader r = {st, st_len};
    if (r.left < 32) return ORC_ERR_DECODE;
    orc_verifier *v = (orc_verifier *)calloc(1, sizeof *v);
    if (!v) return ORC_ERR_OOM;
    memcpy(v->epoch_seed, epoch_seed, 32);
    memcpy(v->file_id, r.p, 32);
    r.p += 32;
    r.left -= 32;
    uint64_t orig_len, d;
    uint8_t g_seed[32];
    int rc = rd_u64(&r, &orig_len);
    if (!rc) rc = rd_u64(&r, &d);
    if (!rc) rc = rd_module(&r, &v->com, &v->com_len);
    if (!rc && r.left < 32) rc = ORC_ERR_DECODE;
    if (rc) {
        free(v->com);
        free(v);
        return rc;
    }
    memcpy(g_seed, r.p, 32);
    size_t depth, dim;
    orc_layout_for_byte_len(orig_len, &depth, &dim, NULL, NULL);
    v->rounds = (size_t)d;
    v->g_rows = dim;
    v->g_len = v->com_len;
    if (v->g_len == 0 || (v->g_len & (v->g_len - 1))) { /* Poly::new asserts a power of two */
        free(v->com);
        free(v);
        return ORC_ERR_SHAPE;
    }
    v->g = (u128 *)xcalloc(dim * K * v->g_len, sizeof(u128));
    if (!v->g) {
        free(v->com);
        free(v);
        return ORC_ERR_OOM;
    }
    derive_generators_u128(g_seed, dim, v->g_len, v->g);
    *out = v;
    return ORC_OK;
}

orc_verifier *orc_verifier_clone(const orc_verifier *v) {
    orc_verifier *c = (orc_verifier *)malloc(sizeof *c);
    if (!c) return NULL;
    *c = *v;
    c->g = (u128 *)xcalloc(v->g_rows * K * v->g_len, sizeof(u128));
    c->com = (u128 *)xcalloc(K * v->com_len, sizeof(u128));
    memcpy(c->g, v->g, v->g_rows * K * v->g_len * sizeof(u128));
    memcpy(c->com, v->com, K * v->com_len * sizeof(u128));
    return c;
}

void orc_verifier_free(orc_verifier *v) {
    if (!v) return;
    free(v->g);
    free(v->com);
    free(v);
}

void orc_verifier_dims(const orc_verifier *v, size_t *rows, size_t *ring_len) {
    *rows = v->g_rows;
    *ring_len = v->com_len;
}
size_t orc_verifier_round(const orc_verifier *v) { return v->round; }
void orc_verifier_current_commitment(const orc_verifier *v, uint64_t *out) {
    for (size_t i = 0; i < K * v->com_len; ++i) zq_store(out + 2 * i, v->com[i]);
}

static int is_zero_mod(const u128 *m, size_t len, size_t want_len) {
    if (len != want_len) return 0;
    for (size_t i = 0; i < K * len; ++i)
        if (m[i]) return 0;
    return 1;
}

/* verifier.rs:57-120 update_commitment; state changes only on success */
int orc_verifier_update(orc_verifier *v, const uint8_t *buf, size_t len) {
    record_t rec;
    int rc = decode_record(buf, len, &rec);
    if (rc) return rc;
    if (rec.round != v->round) {
        record_free(&rec);
        return ORC_ERR_ROUND;
    }
    const size_t n = v->com_len;
    u128 *vfc = NULL, *fg = NULL, *ncom = NULL, *ng = NULL;
    size_t frows = v->g_rows;
    if (v->g_rows > 1) {
        size_t mid = v->g_rows / 2;
        modref mc = {v->com, n}, ml = {rec.m[0], rec.len[0]}, mr = {rec.m[1], rec.len[1]};
        u128 x = challenge_u128(v->epoch_seed, "vec", v->file_id, v->round, mc, ml, mr);
        if (rec.len[0] != n || rec.len[1] != n) {
            rc = ORC_ERR_SHAPE;
            goto done;
        }
        u128 xi = zq_inv(x);
        vfc = (u128 *)xcalloc(K * n, sizeof(u128));
        fg = (u128 *)xcalloc(mid * K * v->g_len, sizeof(u128));
        for (size_t i = 0; i < K * n; ++i)
            vfc[i] = zq_add(zq_add(v->com[i], zq_mul(rec.m[0][i], xi)), zq_mul(rec.m[1][i], x));
        axpy(v->g, v->g + mid * K * v->g_len, xi, mid * K * v->g_len, fg);
        frows = mid;
    } else {
        u128 *z = (u128 *)xcalloc(K * n, sizeof(u128));
        modref mc = {v->com, n}, mz = {z, n};
        (void)challenge_u128(v->epoch_seed, "vec", v->file_id, v->round, mc, mz, mz);
        free(z);
        if (!is_zero_mod(rec.m[0], rec.len[0], n) || !is_zero_mod(rec.m[1], rec.len[1], n)) {
            rc = ORC_ERR_ZERO_TERMS;
            goto done;
        }
        vfc = (u128 *)xcalloc(K * n, sizeof(u128));
        fg = (u128 *)xcalloc(v->g_rows * K * v->g_len, sizeof(u128));
        memcpy(vfc, v->com, K * n * sizeof(u128));
        memcpy(fg, v->g, v->g_rows * K * v->g_len * sizeof(u128));
    }
    size_t glen = v->g_len, nlen;
    if (glen > 1) {
        size_t h = glen / 2;
        modref mc = {vfc, n}, ml = {rec.m[2], rec.len[2]}, mr = {rec.m[3], rec.len[3]};
        u128 y = challenge_u128(v->epoch_seed, "poly", v->file_id, v->round, mc, ml, mr);
        if (rec.len[2] != n / 2 || rec.len[3] != n / 2 || n != glen) {
            rc = ORC_ERR_SHAPE;
            goto done;
        }
        u128 yi = zq_inv(y);
        ncom = (u128 *)xcalloc(K * h, sizeof(u128));
        ng = (u128 *)xcalloc(frows * K * h, sizeof(u128));
        u128 *tmp = (u128 *)xcalloc(h, sizeof(u128)), *sh = (u128 *)xcalloc(h, sizeof(u128));
        for (int k = 0; k < K; ++k)
            for (size_t c = 0; c < h; ++c) {
                size_t i = k * h + c;
                ncom[i] = zq_add(zq_add(vfc[k * n + 2 * c], zq_mul(rec.m[2][i], yi)), zq_mul(rec.m[3][i], y));
            }
        for (size_t i = 0; i < frows * K; ++i) {
            for (size_t c = 0; c < h; ++c) tmp[c] = fg[i * glen + 2 * c + 1];
            mul_by_x(tmp, h, sh);
            for (size_t c = 0; c < h; ++c) ng[i * h + c] = zq_add(fg[i * glen + 2 * c], zq_mul(sh[c], yi));
        }
        free(tmp);
        free(sh);
        nlen = h;
    } else {
        u128 *z = (u128 *)xcalloc(K, sizeof(u128));
        modref mc = {vfc, n}, mz = {z, 1};
        (void)challenge_u128(v->epoch_seed, "poly", v->file_id, v->round, mc, mz, mz);
        free(z);
        if (!is_zero_mod(rec.m[2], rec.len[2], 1) || !is_zero_mod(rec.m[3], rec.len[3], 1)) {
            rc = ORC_ERR_ZERO_TERMS;
            goto done;
        }
        ncom = vfc;
        ng = fg;
        vfc = fg = NULL;
        nlen = n;
    }
    free(v->com);
    free(v->g);
    v->com = ncom;
    v->g = ng;
    v->com_len = nlen;
    v->g_len = glen > 1 ? glen / 2 : glen;
    v->g_rows = frows;
    v->round += 1;
    ncom = ng = NULL;
done:
    free(vfc);
    free(fg);
    free(ncom);
    free(ng);
    record_free(&rec);
    return rc;
}

/* verifier.rs:122-134 */
int orc_verifier_verify_final_opening(const orc_verifier *v, const uint64_t opening[2]) {
    if (v->round != v->rounds) return ORC_ERR_NOT_FINAL;
    if (v->g_rows != 1) return ORC_ERR_NOT_FINAL;
    u128 s = zq_load(opening);
    if (v->g_len != v->com_len) return ORC_ERR_OPENING;
    for (size_t i = 0; i < K * v->g_len; ++i)
        if (zq_mul(v->g[i], s) != v->com[i]) return ORC_ERR_OPENING;
    return ORC_OK;
}

/* ---------------- full flow (examples/full_flow.rs:47-104) ---------------- */
int orc_full_flow(const uint8_t *file, size_t len, const uint8_t seed_generators[32],
                  const uint8_t epoch_seed[32], uint8_t *statement_out, size_t statement_cap,
                  size_t *statement_len, uint8_t *records_out, size_t records_cap,
                  size_t *records_len, uint64_t opening_out[2]) {
    orc_prover *p = NULL;
    int rc = orc_prover_new(file, len, seed_generators, epoch_seed, SIZE_MAX, 1, &p);
    if (rc) return rc;
    size_t sl = orc_prover_statement(p, NULL, 0);
    uint8_t *st = (uint8_t *)malloc(sl);
    orc_prover_statement(p, st, sl);
    if (statement_len) *statement_len = sl;
    if (statement_out && statement_cap >= sl) memcpy(statement_out, st, sl);
    orc_verifier *v = NULL;
    rc = orc_verifier_new(epoch_seed, st, sl, &v);
    free(st);
    if (rc) {
        orc_prover_free(p);
        return rc;
    }
    size_t pos = 0;
    for (size_t j = 0; j < p->rounds && !rc; ++j) {
        size_t n = p->ring_len;
        size_t rl = orc_record_size(n, n / 2 ? n / 2 : 1);
        uint8_t *rec = (uint8_t *)malloc(rl);
        size_t got = 0;
        rc = orc_prover_fold_round(p, j, rec, rl, &got);
        if (!rc) rc = orc_verifier_update(v, rec, got);
        if (!rc && records_out && pos + got <= records_cap) memcpy(records_out + pos, rec, got);
        pos += got;
        free(rec);
    }
    if (records_len) *records_len = pos;
    uint64_t open[2] = {0, 0};
    if (!rc) rc = orc_prover_final_opening(p, open);
    if (!rc) rc = orc_verifier_verify_final_opening(v, open);
    if (opening_out) {
        opening_out[0] = open[0];
        opening_out[1] = open[1];
    }
    orc_prover_free(p);
    orc_verifier_free(v);
    return rc;
}
