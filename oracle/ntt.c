/* Oracle: CRT-NTT ring products in Z_q[X]/(X^n+1) (test infrastructure only).
 * Restates src/algebra/ntt.rs of the reference: a zero-padded 2n-point cyclic NTT per
 * CRT prime with `%`-based modmul, r[i] = c[i] - c[i+n], then the centred CRT lift. */
#include <pthread.h>
#include <stdlib.h>

#include "internal.h"
#include "oracle.h"

/* src/utils/config.rs:9-18, in this order */
const uint64_t ORC_PRIMES[8] = {3892314113ull, 2281701377ull, 2013265921ull, 2885681153ull,
                                2483027969ull, 1811939329ull, 469762049ull,  4194304001ull};

void orc_crt_primes(uint64_t out[8]) { memcpy(out, ORC_PRIMES, sizeof(ORC_PRIMES)); }

/* ntt.rs:332-346 */
static inline uint64_t mod_add(uint64_t a, uint64_t b, uint64_t p) {
    uint64_t s = a + b;
    return s >= p ? s - p : s;
}
static inline uint64_t mod_sub(uint64_t a, uint64_t b, uint64_t p) {
    return a >= b ? a - b : p - (b - a);
}
/* ntt.rs:338-340 */
static inline uint64_t mod_mul(uint64_t a, uint64_t b, uint64_t p) {
    return (uint64_t)(((u128)a * b) % p);
}
/* ntt.rs:342-353 */
static uint64_t mod_pow(uint64_t base, u128 exp, uint64_t p) {
    uint64_t r = 1;
    while (exp > 0) {
        if (exp & 1) r = mod_mul(r, base, p);
        base = mod_mul(base, base, p);
        exp >>= 1;
    }
    return r;
}

/* ntt.rs:300-320 — distinct prime factors by trial division */
static int factorize(uint64_t v, uint64_t *f) {
    int n = 0;
    uint64_t d = 2;
    while (d * d <= v) {
        if (v % d == 0) {
            f[n++] = d;
            while (v % d == 0) v /= d;
        }
        d += (d == 2) ? 1 : 2;
    }
    if (v > 1) f[n++] = v;
    return n;
}

/* ntt.rs:287-298 — the smallest generator of (Z/p)^* */
uint64_t orc_multiplicative_generator(uint64_t p) {
    uint64_t f[64];
    int nf = factorize(p - 1, f);
    for (uint64_t cand = 2; cand < p; ++cand) {
        int ok = 1;
        for (int i = 0; i < nf; ++i)
            if (mod_pow(cand, (p - 1) / f[i], p) == 1) {
                ok = 0;
                break;
            }
        if (ok) return cand;
    }
    return 0;
}

typedef struct {
    size_t size;
    uint64_t modulus, inv_size;
    uint64_t stage_roots[40], inv_stage_roots[40];
    int stages;
    size_t *bitrev;
} plan_t;

/* ntt.rs:21-47 NttPlan::new */
static int plan_build(plan_t *pl, size_t size, uint64_t p) {
    if (size == 0 || (size & (size - 1)) || (p - 1) % size) return ORC_ERR_ARG;
    uint64_t g = orc_multiplicative_generator(p);
    uint64_t root = mod_pow(g, (p - 1) / size, p); /* ntt.rs:275-285 */
    if (mod_pow(root, size, p) != 1) return ORC_ERR_ARG;
    if (size > 1 && mod_pow(root, size / 2, p) == 1) return ORC_ERR_ARG;
    uint64_t inv_root = mod_pow(root, p - 2, p);
    pl->size = size;
    pl->modulus = p;
    pl->stages = 0;
    for (size_t len = 2; len <= size; len <<= 1) {
        pl->stage_roots[pl->stages] = mod_pow(root, size / len, p);
        pl->inv_stage_roots[pl->stages] = mod_pow(inv_root, size / len, p);
        pl->stages++;
    }
    pl->inv_size = mod_pow(size % p, p - 2, p);
    pl->bitrev = (size_t *)malloc(size * sizeof(size_t));
    if (!pl->bitrev) return ORC_ERR_OOM;
    int bits = 0;
    while (((size_t)1 << bits) < size) ++bits;
    for (size_t i = 0; i < size; ++i) { /* ntt.rs:355-363 */
        size_t r = 0;
        for (int b = 0; b < bits; ++b)
            if (i & ((size_t)1 << b)) r |= (size_t)1 << (bits - 1 - b);
        pl->bitrev[i] = r;
    }
    return ORC_OK;
}

/* ntt.rs:256-273 cached_plan, keyed by (log2 size, prime index) */
#define MAX_LOG 26
static plan_t *plan_cache[MAX_LOG + 1][8];
static pthread_mutex_t plan_lock = PTHREAD_MUTEX_INITIALIZER;

static const plan_t *cached_plan(size_t size, int prime_idx, uint64_t p) {
    int lg = 0;
    while (((size_t)1 << lg) < size) ++lg;
    if (lg > MAX_LOG) return NULL;
    pthread_mutex_lock(&plan_lock);
    plan_t *pl = plan_cache[lg][prime_idx];
    if (!pl) {
        pl = (plan_t *)calloc(1, sizeof(plan_t));
        if (pl && plan_build(pl, size, p) != ORC_OK) {
            free(pl);
            pl = NULL;
        }
        plan_cache[lg][prime_idx] = pl;
    }
    pthread_mutex_unlock(&plan_lock);
    return pl;
}

/* ntt.rs:365-370 */
static void bit_reverse_permute(uint64_t *v, const size_t *bitrev, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        size_t o = bitrev[i];
        if (i < o) {
            uint64_t t = v[i];
            v[i] = v[o];
            v[o] = t;
        }
    }
}

/* ntt.rs:243-254 */
static void butterfly_chunk(uint64_t *chunk, size_t len, uint64_t wlen, uint64_t p) {
    size_t half = len / 2;
    uint64_t w = 1;
    for (size_t j = 0; j < half; ++j) {
        uint64_t t = mod_mul(chunk[j + half], w, p);
        uint64_t u = chunk[j];
        chunk[j] = mod_add(u, t, p);
        chunk[j + half] = mod_sub(u, t, p);
        w = mod_mul(w, wlen, p);
    }
}

/* ntt.rs:58-96 */
static void plan_forward(uint64_t *v, const plan_t *pl) {
    bit_reverse_permute(v, pl->bitrev, pl->size);
    size_t len = 2;
    for (int s = 0; s < pl->stages; ++s, len <<= 1)
        for (size_t c = 0; c < pl->size; c += len) butterfly_chunk(v + c, len, pl->stage_roots[s], pl->modulus);
}
static void plan_inverse(uint64_t *v, const plan_t *pl) {
    bit_reverse_permute(v, pl->bitrev, pl->size);
    size_t len = 2;
    for (int s = 0; s < pl->stages; ++s, len <<= 1)
        for (size_t c = 0; c < pl->size; c += len)
            butterfly_chunk(v + c, len, pl->inv_stage_roots[s], pl->modulus);
    for (size_t i = 0; i < pl->size; ++i) v[i] = mod_mul(v[i], pl->inv_size, pl->modulus);
}

static int prime_index(uint64_t p) {
    for (int i = 0; i < 8; ++i)
        if (ORC_PRIMES[i] == p) return i;
    return -1;
}

int orc_ntt_forward(uint64_t *values, size_t size, uint64_t modulus) {
    plan_t local;
    const plan_t *pl;
    int pi = prime_index(modulus);
    if (pi >= 0) {
        pl = cached_plan(size, pi, modulus);
        if (!pl) return ORC_ERR_ARG;
    } else {
        if (plan_build(&local, size, modulus) != ORC_OK) return ORC_ERR_ARG;
        pl = &local;
    }
    plan_forward(values, pl);
    if (pi < 0) free(local.bitrev);
    return ORC_OK;
}

int orc_ntt_inverse(uint64_t *values, size_t size, uint64_t modulus) {
    plan_t local;
    const plan_t *pl;
    int pi = prime_index(modulus);
    if (pi >= 0) {
        pl = cached_plan(size, pi, modulus);
        if (!pl) return ORC_ERR_ARG;
    } else {
        if (plan_build(&local, size, modulus) != ORC_OK) return ORC_ERR_ARG;
        pl = &local;
    }
    plan_inverse(values, pl);
    if (pi < 0) free(local.bitrev);
    return ORC_OK;
}

/* ---- 256-bit unsigned integers for the CRT lift (num-bigint in the reference) ---- */
typedef struct {
    uint64_t w[4];
} big_t;

static uint64_t big_mod_u64(const big_t *a, uint64_t p) {
    u128 r = 0;
    for (int i = 3; i >= 0; --i) r = ((r << 64) | a->w[i]) % p;
    return (uint64_t)r;
}
static void big_add_mul(big_t *acc, const big_t *b, uint64_t c) {
    u128 carry = 0;
    for (int i = 0; i < 4; ++i) {
        u128 t = (u128)b->w[i] * c + acc->w[i] + carry;
        acc->w[i] = (uint64_t)t;
        carry = t >> 64;
    }
}
static void big_mul_u64(big_t *a, uint64_t c) {
    u128 carry = 0;
    for (int i = 0; i < 4; ++i) {
        u128 t = (u128)a->w[i] * c + carry;
        a->w[i] = (uint64_t)t;
        carry = t >> 64;
    }
}
static int big_cmp(const big_t *a, const big_t *b) {
    for (int i = 3; i >= 0; --i) {
        if (a->w[i] != b->w[i]) return a->w[i] > b->w[i] ? 1 : -1;
    }
    return 0;
}
static void big_sub(big_t *a, const big_t *b) { /* a -= b, a >= b */
    uint64_t borrow = 0;
    for (int i = 0; i < 4; ++i) {
        u128 t = (u128)a->w[i] - b->w[i] - borrow;
        a->w[i] = (uint64_t)t;
        borrow = (uint64_t)(t >> 64) ? 1 : 0;
    }
}
static u128 big_mod_q(const big_t *a) {
    u128 r = 0;
    for (int i = 15; i >= 0; --i) {
        uint64_t digit = (a->w[i / 4] >> (16 * (i % 4))) & 0xFFFF;
        r = ((r << 16) | digit) % ZQ_Q;
    }
    return r;
}

/* ntt.rs:379-398 crt_data */
static struct {
    int ready;
    big_t prefix[8];
    uint64_t inv[8];
    big_t total, half;
} crt;
static pthread_mutex_t crt_lock = PTHREAD_MUTEX_INITIALIZER;

static void crt_init(void) {
    pthread_mutex_lock(&crt_lock);
    if (!crt.ready) {
        big_t prod = {{1, 0, 0, 0}};
        for (int i = 0; i < 8; ++i) {
            crt.prefix[i] = prod;
            uint64_t pm = big_mod_u64(&prod, ORC_PRIMES[i]);
            crt.inv[i] = mod_pow(pm, ORC_PRIMES[i] - 2, ORC_PRIMES[i]);
            big_mul_u64(&prod, ORC_PRIMES[i]);
        }
        crt.total = prod;
        for (int i = 0; i < 4; ++i)
            crt.half.w[i] = (prod.w[i] >> 1) | (i < 3 ? (prod.w[i + 1] << 63) : 0);
        __atomic_store_n(&crt.ready, 1, __ATOMIC_RELEASE);
    }
    pthread_mutex_unlock(&crt_lock);
}

/* ntt.rs:400-427 reconstruct_mod_q: Garner accumulation then the centred lift */
u128 crt_reconstruct_u128(const uint64_t residues[8]) {
    if (!__atomic_load_n(&crt.ready, __ATOMIC_ACQUIRE)) crt_init();
    big_t acc = {{0, 0, 0, 0}};
    for (int i = 0; i < 8; ++i) {
        uint64_t p = ORC_PRIMES[i];
        uint64_t cur = big_mod_u64(&acc, p);
        uint64_t delta = mod_sub(residues[i], cur, p);
        uint64_t corr = mod_mul(delta, crt.inv[i], p);
        big_add_mul(&acc, &crt.prefix[i], corr);
    }
    if (big_cmp(&acc, &crt.half) > 0) {
        big_t dist = crt.total;
        big_sub(&dist, &acc);
        u128 d = big_mod_q(&dist);
        return d == 0 ? 0 : ZQ_Q - d;
    }
    return big_mod_q(&acc);
}

void orc_crt_reconstruct(const uint64_t residues[8], uint64_t out[2]) {
    zq_store(out, crt_reconstruct_u128(residues));
}

/* ntt.rs:226-240 */
void naive_negacyclic_u128(const u128 *a, const u128 *b, size_t n, u128 *out) {
    for (size_t i = 0; i < n; ++i) out[i] = 0;
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) {
            size_t idx = i + j;
            u128 term = zq_mul(a[i], b[j]);
            if (idx >= n) {
                idx -= n;
                term = zq_neg(term);
            }
            out[idx] = zq_add(out[idx], term);
        }
}

/* ntt.rs:112-195 negacyclic_multiply, with cpu_ntt_residues (ntt.rs:197-224) */
int negacyclic_multiply_u128(const u128 *lhs, const u128 *rhs, size_t n, u128 *out) {
    if (n == 0 || (n & (n - 1))) return ORC_ERR_ARG;
    size_t S = 2 * n;
    for (int i = 0; i < 8; ++i)
        if ((ORC_PRIMES[i] - 1) % S) {
            naive_negacyclic_u128(lhs, rhs, n, out);
            return ORC_OK;
        }
    uint64_t *a = (uint64_t *)malloc(S * sizeof(uint64_t));
    uint64_t *b = (uint64_t *)malloc(S * sizeof(uint64_t));
    uint64_t *res = (uint64_t *)malloc(8 * n * sizeof(uint64_t));
    if (!a || !b || !res) {
        free(a);
        free(b);
        free(res);
        return ORC_ERR_OOM;
    }
    for (int pi = 0; pi < 8; ++pi) {
        uint64_t p = ORC_PRIMES[pi];
        const plan_t *pl = cached_plan(S, pi, p);
        for (size_t i = 0; i < S; ++i) {
            a[i] = i < n ? (uint64_t)(lhs[i] % p) : 0;
            b[i] = i < n ? (uint64_t)(rhs[i] % p) : 0;
        }
        plan_forward(a, pl);
        plan_forward(b, pl);
        for (size_t i = 0; i < S; ++i) a[i] = mod_mul(a[i], b[i], p);
        plan_inverse(a, pl);
        for (size_t i = 0; i < n; ++i) res[pi * n + i] = mod_sub(a[i], a[i + n], p);
    }
    for (size_t c = 0; c < n; ++c) {
        uint64_t r[8];
        for (int pi = 0; pi < 8; ++pi) r[pi] = res[pi * n + c];
        out[c] = crt_reconstruct_u128(r);
    }
    free(a);
    free(b);
    free(res);
    return ORC_OK;
}

/* poly.rs:134-142 Poly::mul */
int ring_mul(const u128 *a, const u128 *b, size_t n, int ntt_enabled, u128 *out) {
    if (ntt_enabled) return negacyclic_multiply_u128(a, b, n, out);
    naive_negacyclic_u128(a, b, n, out);
    return ORC_OK;
}

static u128 *load_arr(const uint64_t *w, size_t n) {
    u128 *v = (u128 *)malloc((n ? n : 1) * sizeof(u128));
    if (v)
        for (size_t i = 0; i < n; ++i) v[i] = zq_load(w + 2 * i);
    return v;
}

int orc_negacyclic_multiply(const uint64_t *lhs, const uint64_t *rhs, size_t n, uint64_t *out) {
    u128 *a = load_arr(lhs, n), *b = load_arr(rhs, n), *o = load_arr(lhs, n);
    int rc = (a && b && o) ? negacyclic_multiply_u128(a, b, n, o) : ORC_ERR_OOM;
    if (rc == ORC_OK)
        for (size_t i = 0; i < n; ++i) zq_store(out + 2 * i, o[i]);
    free(a);
    free(b);
    free(o);
    return rc;
}

int orc_naive_negacyclic(const uint64_t *lhs, const uint64_t *rhs, size_t n, uint64_t *out) {
    u128 *a = load_arr(lhs, n), *b = load_arr(rhs, n), *o = load_arr(lhs, n);
    if (!a || !b || !o) {
        free(a);
        free(b);
        free(o);
        return ORC_ERR_OOM;
    }
    naive_negacyclic_u128(a, b, n, o);
    for (size_t i = 0; i < n; ++i) zq_store(out + 2 * i, o[i]);
    free(a);
    free(b);
    free(o);
    return ORC_OK;
}

/* ntt.rs:98-110 plain convolution */
int orc_convolution(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out) {
    if (na == 0 || nb == 0) return ORC_ERR_ARG;
    size_t no = na + nb - 1;
    u128 *o = (u128 *)calloc(no, sizeof(u128));
    if (!o) return ORC_ERR_OOM;
    for (size_t i = 0; i < na; ++i)
        for (size_t j = 0; j < nb; ++j)
            o[i + j] = zq_add(o[i + j], zq_mul(zq_load(a + 2 * i), zq_load(b + 2 * j)));
    for (size_t i = 0; i < no; ++i) zq_store(out + 2 * i, o[i]);
    free(o);
    return ORC_OK;
}
