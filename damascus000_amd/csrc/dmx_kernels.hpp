// Kernel-side data structures and the launch wrappers the host runtime calls.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <type_traits>

#include "dmx_arith.hpp"

namespace dmx {

// Device-resident NTT tables (one context owns them).
// zeta[p][k] = psi^brv13(k) with psi a primitive 2^14-th root mod p; the first m entries are
// exactly the size-m negacyclic table (the tables nest), so one table serves every degree.
struct NttTables {
    const uint32_t* zeta_m;    // [NPRIMES][TW_SIZE]  Montgomery-form twiddles
    const uint32_t* zeta_mp;   // [NPRIMES][TW_SIZE]  twiddle * p^-1 mod 2^32
    const uint32_t* izeta_m;   // inverse twiddles psi^-brv13(k)
    const uint32_t* izeta_mp;
    const uint32_t* invn_m;    // [NPRIMES][MAX_LOG_N+1]  m^-1 (Montgomery form)
    const uint32_t* invn_mp;
};

// Garner / centred-lift constants for the CRT primes (src/algebra/ntt.rs:379-427)
struct CrtConsts {
    uint32_t mod_prefix[NPRIMES][NPRIMES];  // (p_0..p_{j-1}) mod p_i, j < i
    uint32_t inv[NPRIMES];                  // ((p_0..p_{i-1}) mod p_i)^-1 mod p_i
    uint32_t pre_m[NPRIMES][NPRIMES];       // mod_prefix as mulc constants (w*R mod p_i)
    uint32_t pre_mp[NPRIMES][NPRIMES];      //   and (w*R*p_i^-1 mod 2^32)
    uint32_t inv_m[NPRIMES], inv_mp[NPRIMES];
    uint32_t half_digits[NPRIMES];          // mixed-radix digits of floor(P/2)
    zq prefix_q[NPRIMES];                   // (p_0..p_{j-1}) mod q
    zq P_q;                                 // P mod q
};

// Where a forward-NTT row comes from.
enum RowMode : int {
    ROW_PLAIN = 0,   // x = a[row][c]
    ROW_VFOLD = 1,   // x = a[row][c] + s * a[row + b_off][c]            (vector fold)
    ROW_PFOLD_W = 2, // x = a[row][2c] + s * a[row][2c+1]                (poly fold, witness)
    ROW_PFOLD_G = 3, // x = a[row][2c] + s * (X * odd)[c]                 (poly fold, generators)
    ROW_GEN = 4,     // x = keyed_blake3(g_seed, "g"||i||k||c) mod q, row = i*K + k
    ROW_VFOLD_L = 5, // x = s * a[row][c] + a[row + b_off][c]            (vector fold, left scaled)
    ROW_PFOLD_GL = 6,// x = s * a[row][2c] + (X * odd)[c]                 (poly fold, even scaled)
};

struct RowSrc {
    const zq* a;            // source rows (ROW_GEN: unused)
    zq* dst;                // when non-null the Z_q row is stored here (row-major, m per row)
    zq* dst2;               // when non-null a second copy goes here (forward transforms only)
    size_t b_off;           // ROW_VFOLD partner offset in rows
    const zq* scal;         // ROW_*FOLD scalar (device)
    int mode;
    const uint32_t* gseed;  // ROW_GEN key words (device, 8 words)
    int a_shared;           // `a` is one buffer read by every job of a batch (not offset per job)
};

// one segment of a forward-transform launch: rows from `src`, residues to `out`
struct FwdSeg {
    RowSrc src;
    uint32_t* out;
    uint32_t rows;
};

// A batch of jobs of one shape. Job b's device buffers sit at byte offset b * stride from job
// 0's, so a kernel handles job blockIdx.y by offsetting every per-job pointer it was given
// (null pointers stay null). Tables and constants are shared and never offset.
struct Jobs {
    uint32_t n = 1;
    size_t stride = 0;
};

// (Byte arithmetic on a char pointer, not through an integer: the compiler then still knows a
// kernel argument's pointer is global memory and emits global loads and stores instead of flat
// ones, which would also count against the LDS wait counter.)
template <class T>
__device__ __forceinline__ T* jp(T* p, size_t jo) {
    using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return p ? reinterpret_cast<T*>(reinterpret_cast<B*>(p) + jo) : p;
}
__device__ __forceinline__ size_t job_off(size_t js) { return (size_t)blockIdx.y * js; }

// Latency-bound kernels on a proof's critical path raise their waves' issue priority above the
// default, so a concurrent throughput kernel (the verifier's generator expansion) takes only the
// issue slots they leave free.
__device__ __forceinline__ void critical_path_priority() { __builtin_amdgcn_s_setprio(2); }

void set_device_constants(const PrimeConsts* pc, const CrtConsts* crt, hipStream_t s);

// forward negacyclic NTT of `rows` rows of degree 2^log_m, all 8 primes; out[(row*8+p)*m + c]
void launch_ntt_fwd_rows(const RowSrc& src, size_t rows, int log_m, uint32_t* out,
                         const NttTables& tw, hipStream_t s, const Jobs& jb = Jobs{}, int variant = -1);
// A challenge inverse carried by a forward-transform launch: an extra workgroup 0 sets
// chal[1] = chal[0]^-1 and *u_out = *u_prev * chal[0]^-1 (u_prev null: 1) while the other
// workgroups transform rows. chal null: no task.
struct InvTask {
    zq* chal = nullptr;
    const zq* u_prev = nullptr;
    zq* u_out = nullptr;
};
// two row sets of the same degree in one launch (segment 2 may be empty)
// `variant` picks the transform kernel (fwd_variant_count(log_m) of them, all computing the same
// residues; -1: variant 0)
void launch_ntt_fwd_rows2(const RowSrc& src1, size_t rows1, uint32_t* out1, const RowSrc& src2, size_t rows2,
                          uint32_t* out2, int log_m, const NttTables& tw, hipStream_t s, const Jobs& jb = Jobs{},
                          const InvTask& it = InvTask{}, int variant = -1);
int fwd_variant_count(int log_m);
// inverse NTT in place on `slices` slices of [8][m]: data[(slice*8 + p)*m + c], times m^-1
void launch_ntt_inv(uint32_t* data, size_t slices, int log_m, const NttTables& tw, hipStream_t s,
                    const Jobs& jb = Jobs{});
// How a MAC pass splits its rows into segments (see mac_segments): segments keep at least
// 2^log_min_rows rows and the pass at most 2^log_max_threads threads. Every shape gives the same
// sums; the context picks one per degree by timing them.
struct MacShape {
    uint8_t log_min_rows = 4;
    uint8_t log_max_threads = 18;
};
constexpr int MAC_SHAPES = 9;          // min rows 8 / 16 / 32 x thread cap 2^17 / 2^18 / 2^19
constexpr int MAC_SHAPE_DEFAULT = 4;   // 16 rows, 2^18 threads
inline MacShape mac_shape_of(int i) {
    MacShape m;
    m.log_min_rows = (uint8_t)(3 + i / 3);
    m.log_max_threads = (uint8_t)(17 + i % 3);
    return m;
}
// inv = sum over all rows W[i]*G[i,k]; L = sum_{i<h} G[i+h,k]*W[i]; R = sum_{i<h} G[i,k]*W[i+h]
// (h = rows/2; L/R skipped when out_L == nullptr). Outputs [k][8][m] residues.
// `part` is scratch of mac_part_words(log_m) words: rows are split into segments whose partial
// sums a second kernel adds.
void launch_mac_pairs(const uint32_t* What, const uint32_t* Ghat, size_t rows, int log_m,
                      uint32_t* out_inv, uint32_t* out_L, uint32_t* out_R, uint32_t* part, hipStream_t s,
                      const Jobs& jb = Jobs{}, bool g_shared = false, MacShape sh = MacShape{});
// vector-stage invariant at degree m plus the poly-stage cross terms at degree m/2, from the
// NTT-domain even/odd split: out_inv [k][8][m], out_L/out_R [k][8][m/2]
void launch_mac_split(const uint32_t* What, const uint32_t* Ghat, size_t rows, int log_m,
                      uint32_t* out_inv, uint32_t* out_L, uint32_t* out_R, uint32_t* part,
                      const NttTables& tw, hipStream_t s, const Jobs& jb = Jobs{}, MacShape sh = MacShape{});
// the same passes with the partial-sum reduction and the inverse transforms fused (outputs are
// the inverse-transformed residues, scaled by 1/m); split: out_lr holds L then R
void launch_mac_pairs_inv(const uint32_t* What, const uint32_t* Ghat, size_t rows, int log_m, uint32_t* out_inv,
                          uint32_t* out_L, uint32_t* out_R, uint32_t* part, const NttTables& tw, hipStream_t s,
                          const Jobs& jb = Jobs{}, bool g_shared = false, MacShape sh = MacShape{});
void launch_mac_split_inv(const uint32_t* What, const uint32_t* Ghat, size_t rows, int log_m, uint32_t* out_inv,
                          uint32_t* out_lr, uint32_t* part, const NttTables& tw, hipStream_t s,
                          const Jobs& jb = Jobs{}, MacShape sh = MacShape{});
// scratch words the MAC passes need at degrees up to 2^max_log_m
size_t mac_part_words(int max_log_m);
// Z_q rows only (no transform): dst[row][c] = row_value(src, row, c)
void launch_fold_rows_zq(const RowSrc& src, size_t rows, int log_m, hipStream_t s, const Jobs& jb = Jobs{});
// The verifier's terminal generator g_d[0] (K values) as one weighted sum over the initial
// matrix G (rows x K x n; shared by every job when g_shared), in two parts. chal: (x, x^-1, y,
// y^-1) per round; vmask / pmask: the rounds (bit r for round r, counted from chal) that folded
// rows / coefficients.
// Early: the first masked rounds fold G (2^log_hi * 2^log_lo rows of 2^log_n) into
// H[2^log_lo][K][2^log_nt] (formed as early_chunks(log_n) partial blocks over coefficient chunks,
// then added into the first: H holds that many blocks); 2^log_hi <= EARLY_MAX_HI and
// 2^log_nt <= EARLY_MAX_NT.
constexpr uint32_t EARLY_MAX_HI = 128;
constexpr uint32_t EARLY_MAX_NT = 256;
uint32_t early_chunks(int log_n);
void launch_gen_contract_early(const zq* G, bool g_shared, const zq* chal, uint32_t vmask, uint32_t pmask,
                               int log_hi, int log_lo, int log_n, int log_nt, zq* H, hipStream_t s,
                               const Jobs& jb = Jobs{});
// Late: the remaining rounds fold H (2^log_rows x K x 2^log_n, both <= 32) to g_d[0] (gfin) and
// check the final opening: flags |= bit when g_d[0]_k * opening != com_k for some k.
void launch_gen_contract_late(const zq* H, bool h_shared, const zq* chal, uint32_t vmask, uint32_t pmask,
                              int log_rows, int log_n, const zq* com, const zq* opening, uint32_t* flags,
                              uint32_t bit, zq* gfin, hipStream_t s, const Jobs& jb = Jobs{});
// residues [slice][8][m] -> Z_q [slice][m]
void launch_crt_lift(const uint32_t* res, zq* out, size_t slices, int log_m, hipStream_t s,
                     const Jobs& jb = Jobs{});

// commitment update operands: c' = c + scal[1]*l + scal[0]*r (poly: even(c) + ...)
struct ComUpd {
    const zq* c;
    const zq* l;
    const zq* r;
    const zq* scal;
    int poly;
    const zq* unscale = nullptr;  // when set, every lifted value is multiplied by *unscale first
};
// One prover stage at degree <= 32 in one launch (see k_stage_small): fold, transforms, MAC
// pass, inverse transforms, lifts, the invariant check and the challenge inverse of `it`
// (it.chal[1], and it.u_out = *it.u_prev * chal^-1, the lifts' unscale). ctr: the job's stage
// counter, zero at the start of the proof; this launch raises it to `target`.
constexpr uint32_t STAGE_WAIT_EXPIRED = 0x20000000u;  // flags bit: a stage waited too long
struct SmallStage {
    int poly;        // 0: vector stage (split pass), 1: poly stage (pairs pass)
    int log_m;       // degree of the stage's input rows
    uint32_t rows;   // rows of the MAC pass: vector stage m/2, poly stage h
    RowSrc w, g;     // witness and generator folds (dst: the next Z_q rows)
    ComUpd cu;       // the invariant's expected update (cu.unscale unused: u is formed here)
    zq* com_out;
    zq* out_lr;      // cross terms: L slices then R slices, [2K][2^log_mb]
    int cross;       // poly stage: cross terms are emitted
    uint32_t* flags;
    uint32_t bit;
    InvTask it;
    uint32_t* ctr;
    uint32_t target;
};
void launch_stage_small(const SmallStage& st, const NttTables& tw, hipStream_t s, const Jobs& jb = Jobs{});


// lift K commitment slices (checking them against the update of `cu`, written to com_out) and
// lr_slices cross-term slices (to out_lr) in one launch; see k_lift_check
void launch_lift_check(const uint32_t* res_a, int log_ma, const ComUpd& cu, zq* com_out, zq* com_out2,
                       uint32_t* flags, uint32_t bit, const uint32_t* res_b, zq* out_lr, size_t lr_slices, int log_mb,
                       hipStream_t s, const Jobs& jb = Jobs{});

// commitment updates (tiny, K x m)
// out = c + sinv*l + s*r          (VEC; scalars at scal[0]=s, scal[1]=s^-1)
// out = even(c) + sinv*l + s*r    (POLY; c has 2m coefficients per coordinate)
// With `hint`, s^-1 is taken from *hint once s * *hint == 1 is checked (computed from s when the
// check fails) and written to scal[1]; without it scal[1] must already hold s^-1.
void launch_com_update(const zq* c, const zq* l, const zq* r, zq* scal, int poly, int log_m,
                       zq* out, hipStream_t s, const Jobs& jb = Jobs{}, const zq* hint = nullptr);
// fill n 32-bit words (graph-safe replacement for memset)
void launch_fill32(uint32_t* p, size_t n, uint32_t v, hipStream_t s, const Jobs& jb = Jobs{});
// plain 16-byte copy (HBM calibration, graph-safe device copies)
void launch_copy16(void* dst, const void* src, size_t n16, hipStream_t s, const Jobs& jb = Jobs{});
// out = a + b over n Z_q coefficients
void launch_zq_add(const zq* a, const zq* b, zq* out, size_t n, hipStream_t s);
// flags[bit] set when a != b over n coefficients
void launch_com_check(const zq* a, const zq* b, size_t n, uint32_t* flags, uint32_t bit,
                      hipStream_t s, const Jobs& jb = Jobs{});

// A job's result head (16 words): the prover's flag word, the verifier's two flag words (challenge
// errors, final check), file_id and m* (the first witness coefficient after the last fold).
constexpr size_t RESULT_HEAD_BYTES = 64;
enum : uint32_t { RESULT_HEAD_PFLAGS = 0, RESULT_HEAD_VFLAGS = 1, RESULT_HEAD_FID = 4, RESULT_HEAD_MSTAR = 12 };
void launch_result_head(uint32_t* head, const uint32_t* pflags, const uint32_t* vflags, const uint8_t* file_id,
                        const zq* mstar, hipStream_t s, const Jobs& jb);

// copy each job's file from the device address in its header (u64 len, u64 src) into dst
void launch_copy_in(const uint64_t* meta, uint8_t* dst, size_t cap, hipStream_t s, const Jobs& jb);

// witness expansion (src/utils/io.rs:51-54): 13-byte chunks into a dim x dim Z_q matrix
// With `dlen` (device, per job) the length is read there and `len` only bounds it.
void launch_expand(const uint8_t* file, size_t len, zq* out, size_t capacity, hipStream_t s,
                   const Jobs& jb = Jobs{}, const uint64_t* dlen = nullptr);

// ---- BLAKE3 tree hashing on device ----
struct TranscriptMsg {
    const uint8_t* epoch_seed;  // 32 B device
    const uint8_t* file_id;     // 32 B device
    uint32_t label_len;
    uint8_t label[4];
    uint64_t round;
    const zq* mods[3];
    uint32_t lens[3];
    // device word per job (null: 0): candidates of counters below it count as zero (a test of
    // the resampling path, transcript.rs:66-75, honoured by the oracle as well)
    const uint32_t* reject = nullptr;
};
size_t transcript_len(uint32_t label_len, const uint32_t lens[3]);
// Challenge of a transcript message (transcript.rs:42-75), hashed from its sources:
// chal_out[0] = x = LE256(BLAKE3(message)) mod q, re-hashed with counter+1 while zero, and
// chal_out[1] = x^-1, for every job. `msg` (transcript_len bytes) receives the chunks that hold
// the counter; cv_scratch holds device_hash_scratch_words(transcript_len) words.
// With inverse = false only chal_out[0] is written (the caller inverts it later, see InvTask).
void transcript_challenge(const TranscriptMsg& t, const uint32_t key[8], uint32_t flags, uint8_t* msg,
                          uint32_t* cv_scratch, zq* chal_out, uint32_t* err_flags, hipStream_t s,
                          const Jobs& jb = Jobs{}, bool inverse = true);
// Whole-buffer BLAKE3 (chunks, group merges, root). cv_scratch holds device_hash_scratch_words.
// digest_out (8 words) when non-null. For a challenge (chal_out != nullptr): LE256(digest) mod q,
// re-hashed with counter+1 while zero, then chal_out[0] = x, chal_out[1] = x^-1.
// With `dlen` (device, per job) each job's length is read there and `len` is the capacity that
// sizes the launch; at most two tree levels (lengths up to 1 GiB) are supported then.
void device_hash(const uint8_t* msg, size_t len, const uint32_t key[8], uint32_t flags,
                 uint32_t* cv_scratch, size_t cv_scratch_words, uint32_t* digest_out, zq* chal_out,
                 uint32_t* err_flags, hipStream_t s, const Jobs& jb = Jobs{}, const uint64_t* dlen = nullptr);
size_t device_hash_scratch_words(size_t len);

// The persistent tail of a proof (k_tail): every stage of degree <= 32 (the last five rounds, ten
// stages) and, optionally, the verifier's replay of them, in one launch per batch. The
// descriptor lives in device memory and holds job 0's addresses; jobs are offset as elsewhere.
constexpr int TAIL_MAX_STAGES = 10;
constexpr size_t TAIL_MAX_M = 8;  // the tail's rounds have ring degree <= 8
// One replay step of the verifier: the challenge of `tx`, then the commitment update
// out = c + s^-1 l + s r (poly: even(c) + ...) at degree 2^log_m, with the candidate inverse at
// hint + hint_off checked before use (the update computes the inverse itself when it fails).
struct TailVerStep {
    TranscriptMsg tx;
    const zq* c;
    const zq* l;
    const zq* r;
    zq* scal;         // the verifier's challenge slot: [0] = s, [1] = s^-1
    zq* out;
    const zq* hint;   // the prover's challenge slot of the same stage
    int poly;
    int log_m;
};
struct TailDesc {
    uint32_t nstages;           // 2 x rounds, at most TAIL_MAX_STAGES
    uint32_t* ctr;              // stages the prover has published (zero at the start of the job)
    uint32_t* vflags;           // the verifier's challenge error word
    SmallStage st[TAIL_MAX_STAGES];      // the prover's stages (ctr and target unused)
    TranscriptMsg ptx[TAIL_MAX_STAGES];  // the challenge message of each stage
    TailVerStep vs[TAIL_MAX_STAGES];
};
// dynamic LDS of a tail launch over the stages of `d`; *msg_zq: the Z_q values of the largest
// message the launch keeps in LDS
size_t tail_lds_bytes(const TailDesc& d, uint32_t* msg_zq);
// allows the tail kernel that much dynamic LDS (above the default limit); not during capture
hipError_t tail_set_lds_limit(size_t lds_bytes);
// grid (1 + verifier, jobs); hint_off selects the candidate inverse (1: the prover's inverse)
void launch_tail(const TailDesc* d_desc, size_t lds_bytes, uint32_t msg_zq, int verifier, int hint_off,
                 const NttTables& tw, hipStream_t s, const Jobs& jb = Jobs{});

// generator expansion into Z_q rows: g[(i*K + k)*m + c] (src/commitment/sis.rs:212-245)
// gseed: 8 key words in device memory
void launch_gen_generators(const uint32_t* gseed, size_t rows, int log_m, zq* out, hipStream_t s,
                           const Jobs& jb = Jobs{});

// the reference-compatible batch entry points
// out = left + c * right mod q; in = [l_lo | l_hi | r_lo | r_hi] (len words each), out = [lo | hi];
// a non-canonical input sets *flag
void launch_fold_soa(const uint64_t* in, uint64_t* out, size_t len, uint64_t c_lo, uint64_t c_hi, uint32_t* flag,
                     hipStream_t s);
// per prime of the cyclic product: p, p^-1 mod 2^32, R^2 mod p and inv_size as mulc pairs
struct CycPrime {
    uint32_t p, pinv, r2m, r2mp, invs_m, invs_mp, pad[2];
};
constexpr int CYC_MAX_LOG_N = 14;  // both operands of one prime in LDS: 2 x 2^14 words
// cyclic products of a and b (prime-major, 8 x 2^log_n values each) with the caller's twiddle
// tables tw (per prime: 2 (n - 1) forward then 2 (n - 1) inverse words); an input >= p sets *flag
void launch_cyc_lds(const uint64_t* a, const uint64_t* b, uint64_t* out, int log_n, const CycPrime* cp,
                    const uint32_t* tw, uint32_t* flag, hipStream_t s);
// elementwise residue products: a[i] = a[i] * b[i] (mod p of slice)
void launch_pointwise(uint32_t* a, const uint32_t* b, size_t slices, int log_m, hipStream_t s);
// one persistent launch over a ragged batch of negacyclic products a[i] * b[i] of degree
// 2^logn[i] at coefficient offset offs[i]; `order` lists the products largest first; `counter`
// is one device word (reset by the launcher); `res` holds 8 residues per coefficient
void launch_ring_mul_persistent(const zq* A, const zq* B, zq* O, const uint64_t* offs, const uint32_t* logn,
                                const uint32_t* order, uint32_t count, uint32_t* counter, uint32_t* res,
                                int max_log_n, const NttTables& tw, hipStream_t s);
// LDS-resident ragged products of degrees 2^0..2^12 in one persistent launch. The products'
// coefficient blocks lie back to back in A, B and O in input order; logn (device, `count` words)
// gives each one's degree. A one-workgroup kernel first builds the launch's queue in `sc`
// (offsets, the order by degree largest first, items of up to ring_lds_per_item products of one
// degree, and a header with the item count and a zeroed queue counter), so the launch pair needs
// no host-side preparation and no memset. Scratch: ring_scratch_bytes(count) bytes.
constexpr int RING_LDS_MAX_LOG_N = 12;
struct RingScratch {
    uint64_t* offs;    // count
    uint64_t* items;   // count (at most one item per product)
    uint32_t* order;   // count
    uint32_t* hdr;     // 4 words
};
inline size_t ring_scratch_bytes(size_t count) { return count * (8 + 8 + 4) + 16; }
inline RingScratch ring_scratch_at(void* base, size_t count) {  // base 16-byte aligned
    uint8_t* p = static_cast<uint8_t*>(base);
    RingScratch r;
    r.offs = reinterpret_cast<uint64_t*>(p);
    r.items = reinterpret_cast<uint64_t*>(p + count * 8);
    r.hdr = reinterpret_cast<uint32_t*>(p + count * 16);
    r.order = reinterpret_cast<uint32_t*>(p + count * 16 + 16);
    return r;
}
uint32_t ring_lds_per_item(int log_n);
// The allocation-free entry point's scratch (dmx_noalloc.hip): a file-scope device array holding
// the degrees (NOALLOC_MAX_PRODUCTS words, then 256 bytes of alignment) followed by the queue
// (ring_scratch_bytes). noalloc_scratch_init looks up its device address on the current device,
// once per device.
constexpr size_t NOALLOC_MAX_PRODUCTS = 65536;
constexpr size_t NOALLOC_QUEUE_OFFSET = NOALLOC_MAX_PRODUCTS * 4 + 256;
size_t noalloc_scratch_bytes();
hipError_t noalloc_scratch_init(int device, void** out);
void launch_ring_mul_lds(const zq* A, const zq* B, zq* O, const uint32_t* logn, uint32_t count, RingScratch sc,
                         uint32_t workgroups, const NttTables& tw, hipStream_t s);
// schoolbook products (the reference's ntt_enabled = false path), one workgroup per product
void launch_ring_mul_naive(const zq* A, const zq* B, zq* O, const uint64_t* offs, const uint32_t* logn,
                           uint32_t count, hipStream_t s);
// gather rows by index into a dense block (for ragged batching)
void launch_gather_rows(const zq* src, const uint64_t* offs, const uint32_t* idx, size_t count,
                        int log_m, zq* dst, hipStream_t s);
void launch_scatter_rows(const zq* src, const uint64_t* offs, const uint32_t* idx, size_t count,
                         int log_m, zq* dst, hipStream_t s);

}  // namespace dmx
