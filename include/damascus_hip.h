/*
 * damascus_hip.h — C ABI of the MI355X (gfx950) implementation of Damascus Fold's
 * commit-and-fold path.
 *
 * Conventions
 *   - Plain C: pointers and sizes only. Status is an int: 0 = ok, negative = error (never
 *     aborts). dmx_last_error() returns a message for the calling thread's last failure.
 *   - Z_q elements (q = 2^112 - 75) cross the boundary as 16-byte little-endian words
 *     (two uint64_t: lo, hi), the bytes of Fp::to_le_bytes (src/algebra/field.rs:81-83).
 *     Module elements are 8 coordinates stored coordinate-major: elem[k * ring_len + c].
 *     Generators are [vector_idx][k][coeff].
 *   - Round records and statements cross as bincode 1.3 bytes (fixint, little endian), the
 *     wire format the reference serialises RoundRecord / Statement with
 *     (src/protocol/prover.rs:14-21, src/commitment/sis.rs:23-30).
 *   - Caller owns every host buffer. Device memory is owned by the context and by the
 *     prover / verifier objects created from it.
 *   - DAMASCUS_GPU=0 in the environment, or no HIP device, makes every compute entry
 *     point return DMX_ERR_NO_DEVICE so the caller keeps its own CPU path
 *     (src/utils/gpu.rs:27-33 semantics). The library has no CPU compute path.
 *   - Threading: a context may be shared; each prover / verifier runs on its own HIP
 *     stream and must be driven by one thread at a time.
 */
#ifndef DAMASCUS_HIP_H
#define DAMASCUS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMX_OK 0
#define DMX_ERR_ARG -1          /* bad argument or shape */
#define DMX_ERR_NO_DEVICE -2    /* no HIP device, or DAMASCUS_GPU=0 */
#define DMX_ERR_HIP -3          /* HIP runtime error */
#define DMX_ERR_MEMORY_LIMIT -4 /* io.rs:46-49 "split the file into smaller statements" */
#define DMX_ERR_ROUND -5        /* round index mismatch / all rounds done (prover.rs:141-142, verifier.rs:58) */
#define DMX_ERR_INVARIANT -6    /* folding invariant failed (prover.rs:204-207, 288-291) */
#define DMX_ERR_SHAPE -7        /* ring length mismatch (poly.rs:58, module.rs:36-43) */
#define DMX_ERR_OPENING -8      /* final opening mismatch (verifier.rs:132) */
#define DMX_ERR_DECODE -9       /* bincode decode failure */
#define DMX_ERR_NOT_FINAL -10   /* opening before every round replayed (verifier.rs:123-130) */
#define DMX_ERR_ZERO_TERMS -11  /* scalar stage emitted non-zero cross terms (verifier.rs:82-85, 109-112) */
#define DMX_ERR_OOM -12         /* device allocation failed */
#define DMX_ERR_BUFFER -13      /* caller's output buffer too small; *len says how much */
#define DMX_ERR_POISONED -14    /* an earlier error left the object unusable */
#define DMX_ERR_PENDING -15     /* a served job has not finished within the wait */

typedef struct dmx_ctx dmx_ctx;
typedef struct dmx_prover dmx_prover;
typedef struct dmx_verifier dmx_verifier;

/* SystemParams (src/utils/config.rs:24-31): the dimensions are derived from the file. */
typedef struct {
    uint8_t seed_generators[32];
    uint8_t epoch_seed[32];
} dmx_system_params;

/* RuntimeConfig (src/utils/config.rs:69-87). ntt_enabled / parallel_enabled / gpu_min_elements
 * change only speed in the reference (its NTT and naive products agree bit for bit), so they
 * are accepted and recorded; gpu_enabled = 0 returns DMX_ERR_NO_DEVICE. */
typedef struct {
    uint64_t max_preprocess_bytes;
    int32_t ntt_enabled;
    int32_t parallel_enabled;
    int32_t gpu_enabled;
    int32_t reserved;
    uint64_t gpu_min_elements;
} dmx_runtime_config;

/* RoundOutput timings (src/protocol/prover.rs:25-31), nanoseconds of device time */
typedef struct {
    uint64_t vector_fold_ns;
    uint64_t poly_fold_ns;
    uint64_t total_round_ns;
} dmx_round_output;

/* ---- backend probe: replaces cuda_backend_ready / cuda_device_info (src/utils/gpu.rs:19-67) ---- */
int dmx_backend_ready(void);
int dmx_device_summary(char *buf, size_t cap);
const char *dmx_last_error(void);

/* ---- context: device, stream pool, NTT twiddle tables and CRT constants ---- */
int dmx_ctx_create(int device_index, dmx_ctx **out);
void dmx_ctx_destroy(dmx_ctx *ctx);

/* ---- reference-compatible low-level entry points ----
 * dmx_ntt_batch replaces damascus_cuda_ntt_batch (src/utils/gpu.rs:83-92, cuda/ntt_kernel.cu:106-114):
 * for each of 8 primes, the length-n cyclic convolution of host_a and host_b (prime-major,
 * 8*n words) by the reference's radix-2 DIT with the caller's stage roots. Returns 0, or 1 for
 * bad arguments (primes must be < 2^32), 2 for a device failure, like the reference's codes. */
int dmx_ntt_batch(const uint64_t *host_a, const uint64_t *host_b, uint64_t *host_out, size_t n,
                  const uint64_t *primes, const uint64_t *stage_roots, const uint64_t *inv_roots,
                  const uint64_t *inv_sizes);
/* dmx_fold_batch replaces damascus_cuda_fold_batch (src/utils/gpu.rs:71-81,
 * cuda/fold_kernel.cu:209-218): out = left + challenge * right mod q over SoA lo/hi words.
 * Returns 0, 1 for bad arguments, 2 for a device failure. */
int dmx_fold_batch(const uint64_t *left_lo, const uint64_t *left_hi, const uint64_t *right_lo,
                   const uint64_t *right_hi, uint64_t *out_lo, uint64_t *out_hi, size_t len,
                   uint64_t challenge_lo, uint64_t challenge_hi);

/* ---- ring products: negacyclic_multiply (src/algebra/ntt.rs:112-195) over a ragged batch.
 * Product i has degree 2^log_n[i]; a, b, out hold the coefficient blocks back to back. */
int dmx_ring_mul_batch(dmx_ctx *ctx, const uint64_t *a, const uint64_t *b, uint64_t *out,
                       const uint32_t *log_n, size_t count);
/* RingElement::mul(rhs, ntt_enabled, gpu_enabled) (src/algebra/poly.rs:134-140) over the
 * same ragged batch: config->ntt_enabled = 0 selects the schoolbook product the reference then
 * uses (naive_negacyclic, src/algebra/ntt.rs:226-240); gpu_enabled = 0 returns
 * DMX_ERR_NO_DEVICE. Results are identical either way. */
int dmx_ring_mul_batch_cfg(dmx_ctx *ctx, const uint64_t *a, const uint64_t *b, uint64_t *out,
                           const uint32_t *log_n, size_t count, const dmx_runtime_config *config);
/* the same on device-resident operands (d_a, d_b, d_out: device addresses of the coefficient
 * blocks; log_n in host memory). mode 0: one persistent launch over the whole ragged batch,
 * products taken largest first from a device work counter; mode 1: one launch sequence per
 * degree group; mode 2: the schoolbook product. No reference counterpart (the reference
 * multiplies one pair per call). */
int dmx_ring_mul_batch_device(dmx_ctx *ctx, const void *d_a, const void *d_b, void *d_out,
                              const uint32_t *log_n, size_t count, int mode);

/* The same products (mode 0) with no allocation of any kind: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The launch's queue lives in a fixed device array of about
 * 1.5 MiB per device, looked up when a context is created; at most 65,536 products of degree at
 * most 2^12 per call (DMX_ERR_ARG / DMX_ERR_SHAPE beyond). Calls on one device share that array:
 * calls that overlap on one device race (calls through one context are serialised). */
int dmx_ring_mul_batch_device_noalloc(dmx_ctx *ctx, const void *d_a, const void *d_b, void *d_out,
                                      const uint32_t *log_n, size_t count);

/* ---- commitment (src/commitment/sis.rs) ---- */
/* derive_subseed(seed, "g0") (sis.rs:208-210) */
int dmx_derive_subseed(const uint8_t seed[32], const uint8_t *label, size_t label_len, uint8_t out[32]);
/* derive_generators (sis.rs:212-229) into rows x 8 x ring_len words */
int dmx_derive_generators(dmx_ctx *ctx, const uint8_t g_seed[32], size_t rows, size_t ring_len,
                          uint64_t *out);
/* ModuleSisCommitter::commit (sis.rs:111-132): generators from keyed(seed, "g0");
 * rows == 0 gives zero(1) like sis.rs:154-156. com_out holds 8 * max(ring_len, 1) words. */
int dmx_commit(dmx_ctx *ctx, const uint8_t seed[32], const uint64_t *witness, size_t rows,
               size_t ring_len, uint64_t *com_out, size_t *com_ring_len);
/* commit_with_generators (sis.rs:146-167) */
int dmx_commit_with_generators(dmx_ctx *ctx, const uint64_t *witness, const uint64_t *gens,
                               size_t rows, size_t ring_len, uint64_t *com_out, size_t *com_ring_len);

/* ---- prover: DamascusProver (src/protocol/prover.rs:56-313) ---- */
int dmx_prover_new(dmx_ctx *ctx, const uint8_t *file, size_t len, const dmx_system_params *params,
                   const dmx_runtime_config *config, dmx_prover **out);
void dmx_prover_free(dmx_prover *p);
size_t dmx_prover_rounds_total(const dmx_prover *p);
size_t dmx_prover_current_round(const dmx_prover *p);
void dmx_prover_current_dimensions(const dmx_prover *p, size_t *rows, size_t *ring_len);
int dmx_prover_statement(const dmx_prover *p, uint8_t *out, size_t cap, size_t *len);
int dmx_prover_current_commitment(const dmx_prover *p, uint64_t *out, size_t cap_words, size_t *ring_len);
int dmx_prover_fold_round(dmx_prover *p, size_t round_idx, uint8_t *record_out, size_t cap,
                          size_t *len, dmx_round_output *timing);
int dmx_prover_final_opening(const dmx_prover *p, uint64_t out[2]);

/* ---- verifier: DamascusVerifier (src/protocol/verifier.rs:21-134) ---- */
int dmx_verifier_new(dmx_ctx *ctx, const dmx_system_params *params, const uint8_t *statement,
                     size_t len, dmx_verifier **out);
dmx_verifier *dmx_verifier_clone(const dmx_verifier *v);
void dmx_verifier_free(dmx_verifier *v);
size_t dmx_verifier_round(const dmx_verifier *v);
void dmx_verifier_current_dimensions(const dmx_verifier *v, size_t *rows, size_t *ring_len);
int dmx_verifier_current_commitment(const dmx_verifier *v, uint64_t *out, size_t cap_words, size_t *ring_len);
int dmx_verifier_update_commitment(dmx_verifier *v, const uint8_t *record, size_t len);
int dmx_verifier_verify_final_opening(const dmx_verifier *v, const uint64_t opening[2]);

/* ---- whole jobs (examples/full_flow.rs:47-104): initialize + d x fold_round, then replay.
 * records_out receives the d bincode records back to back. */
int dmx_prove(dmx_ctx *ctx, const uint8_t *file, size_t len, const dmx_system_params *params,
              const dmx_runtime_config *config, uint8_t *statement_out, size_t statement_cap,
              size_t *statement_len, uint8_t *records_out, size_t records_cap, size_t *records_len,
              uint64_t opening_out[2]);
int dmx_verify(dmx_ctx *ctx, const dmx_system_params *params, const uint8_t *statement,
               size_t statement_len, const uint8_t *records, size_t records_len,
               const uint64_t opening[2]);
/* prove then replay with the file already resident in device memory (d_file is a device
 * pointer); returns DMX_OK only when the verifier accepts. Output buffers may be NULL. */
int dmx_prove_verify_device(dmx_ctx *ctx, const void *d_file, size_t len,
                            const dmx_system_params *params, uint8_t *statement_out,
                            size_t statement_cap, size_t *statement_len, uint8_t *records_out,
                            size_t records_cap, size_t *records_len, uint64_t opening_out[2]);

/* One job of a batch: a file to prove and replay, and where its proof goes.
 * Inputs: file (device address, or host address when files_on_host), len, params.
 * Outputs: statement / records (bincode, as dmx_prove writes them; NULL skips), their
 * lengths, the opening m*, depth, and status (DMX_OK when the proof was produced and the
 * replay accepted it). */
typedef struct dmx_job {
    const void *file;
    uint64_t len;
    dmx_system_params params;
    uint8_t *statement;
    size_t statement_cap;
    size_t statement_len;
    uint8_t *records;
    size_t records_cap;
    size_t records_len;
    uint64_t opening[2];
    int32_t depth;
    int32_t status;
} dmx_job;

/* prove + replay of `count` independent files of any sizes (a ragged batch). Jobs of one depth
 * run side by side in shared launches (at most max_batch per launch; 0 sizes batches by
 * workspace), batches of different depths run concurrently, and every result is bit-identical
 * to proving that file alone. No reference counterpart: the reference proves one file per
 * call (examples/full_flow.rs:10-112); this is the batched form of that loop. Returns DMX_OK
 * when every job succeeded; per-job outcomes are in jobs[i].status. */
int dmx_prove_verify_many(dmx_ctx *ctx, dmx_job *jobs, size_t count, int files_on_host,
                          size_t max_batch);

/* ---- serving: continuous batching of commit+fold jobs. No counterpart in the reference
 * (it proves one file per call); the loop that a service built on examples/full_flow.rs would
 * run. Jobs are submitted as they arrive; a scheduler thread admits queued jobs whenever a
 * batch slot and device memory are free (jobs of one depth share launches, batches run
 * concurrently), so requests arriving while others run overlap on the device. Results are
 * bit-identical to proving each file alone. The dmx_job and its file and output buffers must
 * stay valid until the job has been waited for. */
typedef struct dmx_server_config {
    uint64_t hbm_budget_bytes; /* device memory for in-flight workspaces (0: 3/4 of free memory) */
    uint32_t max_batch;        /* jobs per launch (0: 64); batches are powers of two */
    uint32_t max_inflight;     /* batches on the device at once (0: 4) */
    int32_t files_on_host;     /* job files are host addresses (copied in by the server) */
    int32_t reserved;
} dmx_server_config;
typedef struct dmx_server dmx_server;
int dmx_server_create(dmx_ctx *ctx, const dmx_server_config *cfg, dmx_server **out);
int dmx_server_submit(dmx_server *srv, dmx_job *job, uint64_t *ticket);
/* Waits up to timeout_ms (negative: no limit) for a job. DMX_OK when it has finished (its
 * outcome is in job->status; latency_s = submit to verified result) and the ticket is
 * released; DMX_ERR_PENDING when it is still running. */
int dmx_server_wait(dmx_server *srv, uint64_t ticket, int64_t timeout_ms, double *latency_s);
/* counters of the scheduler: batches launched and the jobs in them, scheduler seconds spent
 * launching; workspaces (and graphs) created by the process and the seconds that took */
typedef struct dmx_server_stats_t {
    uint64_t batches;
    uint64_t jobs;
    double launch_seconds;
    uint64_t workspace_creations;
    double workspace_creation_seconds;
} dmx_server_stats_t;
int dmx_server_stats(dmx_server *srv, dmx_server_stats_t *out);
/* finishes every submitted job, then stops the scheduler */
void dmx_server_destroy(dmx_server *srv);

/* ---- device buffers for callers that keep inputs resident in HBM. No counterpart in the
 * reference (its boundary takes host buffers only); for benchmarks and serving loops that
 * hand over files already on the device. */
int dmx_device_alloc(dmx_ctx *ctx, size_t bytes, void **out);
int dmx_device_free(dmx_ctx *ctx, void *ptr);
int dmx_memcpy_htod(dmx_ctx *ctx, void *dst, const void *src, size_t bytes);
int dmx_memcpy_dtoh(dmx_ctx *ctx, void *dst, const void *src, size_t bytes);
/* device copy kernel (a calibration for HBM-traffic counters) */
int dmx_device_copy(dmx_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ---- kernel timing: with timing on, each forward-transform launch of every prover made
 * from this context is bracketed by HIP events on its own stream. dmx_kernel_timing returns
 * the summed device time (ms), the launch count and the summed algorithmic work of those
 * launches in VALU lane-cycles: with the fixed weights per butterfly (19), mod-p reduction (42),
 * Z_q fold multiply-add (204) and BLAKE3 compression (860), which price a 32-bit multiply at 4,
 * and with 32-bit multiplies at the full rate (10, 15, 108, 860); then clears the record. Any
 * output pointer may be null. No reference counterpart. */
int dmx_kernel_timing_enable(dmx_ctx *ctx, int on);
int dmx_kernel_timing(dmx_ctx *ctx, double *total_ms, uint64_t *launches, double *lane_cycles,
                      double *lane_cycles_full_rate);
/* record only launches of degree 2^log_m (negative: every degree) */
int dmx_kernel_timing_degree(dmx_ctx *ctx, int log_m);

/* ---- forward-transform autotuning: per degree 2^log_m the context holds one of *count kernel
 * variants (register blockings of E elements x TP threads per prime slice, or the per-layer
 * kernel), all computing the same residues. dmx_ctx_create times each on synthetic fold rows and
 * keeps the fastest (DAMASCUS_NTT_TUNE=0: variant 0). tuned_us (count floats, may be null)
 * receives the tuning times, 0 where none ran. Setting a variant makes the context's cached
 * graphs be captured again. No reference counterpart (its NTT has one kernel per call). */
int dmx_ctx_transform_variant(dmx_ctx *ctx, int log_m, int *variant, int *count, float *tuned_us);
int dmx_ctx_set_transform_variant(dmx_ctx *ctx, int log_m, int variant);

/* ---- MAC-pass tuning: per degree 2^log_m the context splits each NTT-domain MAC pass into row
 * segments by one of *count shapes (shape i: segments of at least 2^(3 + i / 3) rows, at most
 * 2^(17 + i % 3) threads per pass; shape 4 is the default), all giving the same sums.
 * dmx_ctx_create times each shape on synthetic residues at degrees 2^6..2^10 and switches from the
 * default only to a shape at least 2 % faster (DAMASCUS_NTT_TUNE=0: shape 4). tuned_us (count
 * floats, may be null) receives the times, 0 where none ran. Setting a shape makes the context's
 * cached graphs be captured again. No reference counterpart. */
int dmx_ctx_mac_shape(dmx_ctx *ctx, int log_m, int *shape, int *count, float *tuned_us);
int dmx_ctx_set_mac_shape(dmx_ctx *ctx, int log_m, int shape);
/* Where the context's transform variants and MAC shapes came from: 0 defaults (tuning off),
 * 1 timed at creation (and written to the tuning cache), 2 read from the tuning cache, which holds
 * the choices of this build for this kind of device (architecture name, compute units).
 * DAMASCUS_TUNE_CACHE names the cache file ("0": none); by default
 * $XDG_CACHE_HOME/damascus000_amd/tuning.txt or ~/.cache/damascus000_amd/tuning.txt. path (cap
 * bytes, may be null) receives the file name, empty when there is none. No reference counterpart. */
int dmx_ctx_tuning_source(const dmx_ctx *ctx, char *path, size_t cap);

/* sizes for buffer planning */
size_t dmx_statement_size(size_t ring_len);
size_t dmx_record_size(size_t vec_ring_len, size_t poly_ring_len);
size_t dmx_proof_records_size(uint64_t file_len);
/* device memory of one job's workspace (prover + verifier) for a file of file_len bytes: the unit
 * a server's hbm_budget_bytes admits jobs by. No reference counterpart. */
size_t dmx_job_workspace_size(uint64_t file_len);

#ifdef __cplusplus
}
#endif
#endif
