// Host runtime for the gfx950 commit-and-fold path: context (device tables), prover and
// verifier state machines, bincode wire format, and the C ABI of include/damascus_hip.h.
//
// The host never computes on protocol data: every commitment, fold, transcript hash and
// challenge runs on the device stream. The host validates shapes, orders launches and moves
// the bytes the caller asked for.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <system_error>
#include <thread>
#include <unordered_map>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <sys/stat.h>
#include <unistd.h>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/damascus_hip.h"
#include "dmx_blake3.hpp"
#include "dmx_kernels.hpp"

using namespace dmx;

// ------------------------------------------------------------------------------------------
// errors
// ------------------------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail(DMX_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

extern "C" const char* dmx_last_error(void) { return g_err.c_str(); }

// ------------------------------------------------------------------------------------------
// backend probe (src/utils/gpu.rs:19-67 semantics: DAMASCUS_GPU=0 disables)
// ------------------------------------------------------------------------------------------
static bool gpu_disabled_by_env() {
    const char* v = getenv("DAMASCUS_GPU");
    return v && strcmp(v, "0") == 0;
}

extern "C" int dmx_backend_ready(void) {
    if (gpu_disabled_by_env()) return 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n > 0 ? 1 : 0;
}

extern "C" int dmx_device_summary(char* buf, size_t cap) {
    std::string s;
    if (gpu_disabled_by_env()) {
        s = "disabled by DAMASCUS_GPU=0";
    } else {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            s = "no HIP device";
        } else {
            hipDeviceProp_t p;
            (void)hipGetDeviceProperties(&p, 0);
            char t[256];
            snprintf(t, sizeof t, "%s (%s), %d CUs, %zu MiB", p.name, p.gcnArchName,
                     p.multiProcessorCount, (size_t)(p.totalGlobalMem >> 20));
            s = t;
        }
    }
    if (buf && cap) {
        strncpy(buf, s.c_str(), cap - 1);
        buf[cap - 1] = 0;
    }
    return (int)s.size();
}

// ------------------------------------------------------------------------------------------
// layout (src/utils/io.rs:72-120)
// ------------------------------------------------------------------------------------------
static size_t coeff_count_for(uint64_t len) {
    uint64_t c = (len + BYTES_PER_COEFF - 1) / BYTES_PER_COEFF;
    return c < 1 ? 1 : (size_t)c;
}
static int depth_for(uint64_t len) {
    size_t cc = coeff_count_for(len);
    int d = MIN_FOLD_DEPTH;
    while (((size_t)1 << (2 * d)) < cc) ++d;
    return d;
}

// ------------------------------------------------------------------------------------------
// bincode 1.3 (fixint, little endian)
// ------------------------------------------------------------------------------------------
static size_t module_bytes(size_t n) { return 8 + (size_t)K * (8 + 16 * n); }
extern "C" size_t dmx_statement_size(size_t ring_len) { return 32 + 8 + 8 + module_bytes(ring_len) + 32; }
extern "C" size_t dmx_record_size(size_t lv, size_t lp) { return 8 + 2 * module_bytes(lv) + 2 * module_bytes(lp); }
extern "C" size_t dmx_proof_records_size(uint64_t file_len) {
    int d = depth_for(file_len);
    size_t t = 0;
    for (int j = 0; j < d; ++j) {
        size_t m = (size_t)1 << (d - j);
        t += dmx_record_size(m, m / 2);
    }
    return t;
}

static uint8_t* put_u64(uint8_t* p, uint64_t v) {
    for (int i = 0; i < 8; ++i) *p++ = (uint8_t)(v >> (8 * i));
    return p;
}
static uint8_t* put_module(uint8_t* p, const zq* v, size_t n) {
    p = put_u64(p, K);
    for (int k = 0; k < K; ++k) {
        p = put_u64(p, n);
        memcpy(p, v + (size_t)k * n, 16 * n);
        p += 16 * n;
    }
    return p;
}

struct Reader {
    const uint8_t* p;
    size_t left;
    bool u64(uint64_t& v) {
        if (left < 8) return false;
        v = 0;
        for (int i = 0; i < 8; ++i) v |= (uint64_t)p[i] << (8 * i);
        p += 8;
        left -= 8;
        return true;
    }
};

// ModuleElement<8> (module.rs:104-115): a sequence of exactly 8 Polys of one length; each Fp
// is read as u128 and reduced mod q (field.rs:121-129).
static int read_module(Reader& r, std::vector<zq>& out, size_t& len) {
    uint64_t n;
    if (!r.u64(n)) return fail(DMX_ERR_DECODE, "decode: truncated module length");
    bool shape_ok = true;
    len = 0;
    for (uint64_t k = 0; k < n; ++k) {
        uint64_t l;
        if (!r.u64(l) || l > r.left / 16) return fail(DMX_ERR_DECODE, "decode: truncated polynomial");
        if (k == 0) {
            len = (size_t)l;
            out.assign((size_t)K * len, zq_zero());
        } else if (l != len) {
            shape_ok = false;
        }
        for (uint64_t c = 0; c < l; ++c) {
            uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int b = 0; b < 16; ++b) w[b >> 2] |= (uint32_t)r.p[b] << (8 * (b & 3));
            r.p += 16;
            r.left -= 16;
            if (k < K && shape_ok) out[k * len + c] = zq_reduce256(w);
        }
    }
    if (n != K) return fail(DMX_ERR_DECODE, "decode: module rank mismatch (%llu)", (unsigned long long)n);
    if (!shape_ok) return fail(DMX_ERR_DECODE, "decode: module coordinate length mismatch");
    return DMX_OK;
}

struct Record {
    uint64_t round;
    std::vector<zq> m[4];  // l_vec, r_vec, l_poly, r_poly
    size_t len[4];
};

static int decode_record(const uint8_t* buf, size_t n, Record& rec) {
    Reader r{buf, n};
    if (!r.u64(rec.round)) return fail(DMX_ERR_DECODE, "decode: truncated round");
    for (int i = 0; i < 4; ++i) {
        int rc = read_module(r, rec.m[i], rec.len[i]);
        if (rc) return rc;
    }
    return DMX_OK;  // trailing bytes are allowed, as bincode::deserialize does
}

struct Statement {
    uint8_t file_id[32];
    uint64_t original_len;
    uint64_t d;
    std::vector<zq> com0;
    size_t com_len;
    uint8_t g_seed[32];
};

static int decode_statement(const uint8_t* buf, size_t n, Statement& st) {
    Reader r{buf, n};
    if (r.left < 32) return fail(DMX_ERR_DECODE, "decode: truncated statement");
    memcpy(st.file_id, r.p, 32);
    r.p += 32;
    r.left -= 32;
    if (!r.u64(st.original_len) || !r.u64(st.d)) return fail(DMX_ERR_DECODE, "decode: truncated statement");
    int rc = read_module(r, st.com0, st.com_len);
    if (rc) return rc;
    if (r.left < 32) return fail(DMX_ERR_DECODE, "decode: truncated g_0_seed");
    memcpy(st.g_seed, r.p, 32);
    return DMX_OK;
}

// ------------------------------------------------------------------------------------------
// context: constants and NTT tables
// ------------------------------------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t n) {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = n;
        if (n == 0) return DMX_OK;
        if (hipMalloc(&p, n) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return fail(DMX_ERR_OOM, "device allocation of %zu bytes failed", n);
        }
        return DMX_OK;
    }
    template <class T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

struct TimedLaunch {
    hipEvent_t a, b;
    double work;       // algorithmic lane-cycles of the launch, fixed weights
    double work_full;  // the same with 32-bit multiplies at the full issue rate
};

struct dmx_ctx {
    int device = 0;
    // forward-transform kernel per degree (launch_ntt_fwd_rows2's variant), timed at creation;
    // fwd_version counts changes so captured graphs can be retaken
    int fwd_variant[MAX_LOG_N + 1];
    float fwd_us[MAX_LOG_N + 1][4];  // tuning times per variant (0: not timed)
    // MAC-pass segmentation per degree (mac_shape_of index), timed at creation like the transforms
    int mac_idx[MAX_LOG_N + 1];
    float mac_us[MAX_LOG_N + 1][MAC_SHAPES];
    int tune_source = 0;  // 0: defaults, 1: timed at creation, 2: read from the tuning cache
    std::string tune_path;
    uint64_t fwd_version = 0;  // counts changes of either choice (captured graphs are retaken)
    DevBuf tables;
    NttTables tw{};
    PrimeConsts pc[NPRIMES];
    CrtConsts crt;
    std::mutex mu;
    bool timing = false;
    int timing_log_m = -1;  // record only this degree (-1: all)
    std::vector<TimedLaunch> launches;
    std::vector<hipEvent_t> event_pool;
    // ring-product scratch (grow-only), one batch at a time
    std::mutex ring_mu;
    DevBuf ring_offs, ring_logn, ring_order, ring_res, ring_counter, ring_ga, ring_gb, ring_ra, ring_rb, ring_scratch;
    void* ring_pinned = nullptr;  // host staging of the degrees
    size_t ring_pinned_bytes = 0;
    hipStream_t ring_stream = nullptr;  // products run on their own non-blocking stream
    hipEvent_t ring_done = nullptr;
    void* noalloc = nullptr;  // device address of the allocation-free entry point's scratch
    int cus = 256;  // compute units (one LDS-resident product workgroup each)
    ~dmx_ctx() {
        for (auto& t : launches) {
            (void)hipEventDestroy(t.a);
            (void)hipEventDestroy(t.b);
        }
        for (auto e : event_pool) (void)hipEventDestroy(e);
        if (ring_pinned) (void)hipHostFree(ring_pinned);
        if (ring_done) (void)hipEventDestroy(ring_done);
        if (ring_stream) (void)hipStreamDestroy(ring_stream);
    }
    MacShape mac(int log_m) const { return mac_shape_of(mac_idx[log_m]); }
    hipEvent_t take_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
};

// Algorithmic VALU work of one forward-transform row of degree 2^log_m in lane-cycles. Fixed
// weights: butterfly 19, mod-p reduction 42, Z_q multiply-add 204, BLAKE3 compression 860; they
// price a 32-bit multiply at 4 lane-cycles. Full-rate weights count it as 1 (gfx950 issues
// v_mul_lo_u32 / v_mul_hi_u32 at the full rate): butterfly 10 (3 multiplies), reduction 15 (9),
// multiply-add 108 (32), BLAKE3 860 (none).
struct OpWeights {
    double butterfly, reduction, muladd, blake3;
};
constexpr OpWeights FIXED_WEIGHTS{19.0, 42.0, 204.0, 860.0};
constexpr OpWeights FULL_RATE_WEIGHTS{10.0, 15.0, 108.0, 860.0};
static double fwd_row_work(int log_m, int mode, const OpWeights& w = FIXED_WEIGHTS) {
    const double m = (double)((size_t)1 << log_m);
    double v = NPRIMES * m * w.reduction + NPRIMES * (m / 2) * log_m * w.butterfly;
    if (mode == ROW_GEN) v += m * w.blake3;
    if (mode == ROW_VFOLD || mode == ROW_PFOLD_W || mode == ROW_PFOLD_G || mode == ROW_VFOLD_L ||
        mode == ROW_PFOLD_GL)
        v += m * w.muladd;
    return v;
}

// forward transforms go through here so that kernel timing can bracket them; two row sets of
// one degree share a launch
static void fwd_rows2(dmx_ctx* ctx, const RowSrc& src1, size_t rows1, uint32_t* out1, const RowSrc& src2,
                      size_t rows2, uint32_t* out2, int log_m, hipStream_t s, const Jobs& jb,
                      const InvTask& it = InvTask{}) {
    const int v = ctx->fwd_variant[log_m];
    if (!ctx->timing || rows1 + rows2 == 0 || (ctx->timing_log_m >= 0 && ctx->timing_log_m != log_m)) {
        launch_ntt_fwd_rows2(src1, rows1, out1, src2, rows2, out2, log_m, ctx->tw, s, jb, it, v);
        return;
    }
    std::lock_guard<std::mutex> g(ctx->mu);
    const double work = (double)jb.n * ((double)rows1 * fwd_row_work(log_m, src1.mode) +
                                        (double)rows2 * fwd_row_work(log_m, src2.mode));
    const double work_full = (double)jb.n * ((double)rows1 * fwd_row_work(log_m, src1.mode, FULL_RATE_WEIGHTS) +
                                             (double)rows2 * fwd_row_work(log_m, src2.mode, FULL_RATE_WEIGHTS));
    TimedLaunch t{ctx->take_event(), ctx->take_event(), work, work_full};
    (void)hipEventRecord(t.a, s);
    launch_ntt_fwd_rows2(src1, rows1, out1, src2, rows2, out2, log_m, ctx->tw, s, jb, it, v);
    (void)hipEventRecord(t.b, s);
    ctx->launches.push_back(t);
}

extern "C" int dmx_kernel_timing_enable(dmx_ctx* ctx, int on) {
    if (!ctx) return fail(DMX_ERR_ARG, "null context");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->timing = on != 0;
    return DMX_OK;
}

extern "C" int dmx_kernel_timing_degree(dmx_ctx* ctx, int log_m) {
    if (!ctx) return fail(DMX_ERR_ARG, "null context");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->timing_log_m = log_m < 0 ? -1 : log_m;
    return DMX_OK;
}

extern "C" int dmx_kernel_timing(dmx_ctx* ctx, double* total_ms, uint64_t* launches, double* lane_cycles,
                                 double* lane_cycles_full_rate) {
    if (!ctx) return fail(DMX_ERR_ARG, "null context");
    std::lock_guard<std::mutex> g(ctx->mu);
    double ms = 0, work = 0, work_full = 0;
    for (auto& t : ctx->launches) {
        HIP_TRY(hipEventSynchronize(t.b));
        float e = 0;
        HIP_TRY(hipEventElapsedTime(&e, t.a, t.b));
        ms += e;
        work += t.work;
        work_full += t.work_full;
        ctx->event_pool.push_back(t.a);
        ctx->event_pool.push_back(t.b);
    }
    if (total_ms) *total_ms = ms;
    if (launches) *launches = ctx->launches.size();
    if (lane_cycles) *lane_cycles = work;
    if (lane_cycles_full_rate) *lane_cycles_full_rate = work_full;
    ctx->launches.clear();
    return DMX_OK;
}

extern "C" int dmx_device_alloc(dmx_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return fail(DMX_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return fail(DMX_ERR_OOM, "device allocation of %zu bytes failed", bytes);
    }
    return DMX_OK;
}
extern "C" int dmx_device_free(dmx_ctx* ctx, void* ptr) {
    if (!ctx) return fail(DMX_ERR_ARG, "null context");
    if (ptr) HIP_TRY(hipFree(ptr));
    return DMX_OK;
}
extern "C" int dmx_memcpy_htod(dmx_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(DMX_ERR_ARG, "null argument");
    if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return DMX_OK;
}
extern "C" int dmx_memcpy_dtoh(dmx_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(DMX_ERR_ARG, "null argument");
    if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return DMX_OK;
}
extern "C" int dmx_device_copy(dmx_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src)) || (bytes & 15)) return fail(DMX_ERR_ARG, "bad argument");
    launch_copy16(dst, src, bytes / 16, 0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return DMX_OK;
}

static uint64_t smallest_generator(uint64_t p) {
    uint64_t m = p - 1, f[64];
    int nf = 0;
    for (uint64_t d = 2; d * d <= m; ++d) {
        if (m % d == 0) {
            f[nf++] = d;
            while (m % d == 0) m /= d;
        }
    }
    if (m > 1) f[nf++] = m;
    for (uint64_t g = 2;; ++g) {
        bool ok = true;
        for (int i = 0; i < nf; ++i)
            if (host_powmod(g, (p - 1) / f[i], p) == 1) {
                ok = false;
                break;
            }
        if (ok) return g;
    }
}

static uint32_t brv(uint32_t x, int bits) {
    uint32_t r = 0;
    for (int i = 0; i < bits; ++i)
        if (x & (1u << i)) r |= 1u << (bits - 1 - i);
    return r;
}

// u32-limb big integers for the CRT constants
typedef std::vector<uint32_t> big;
static void big_mul(big& a, uint32_t m) {
    uint64_t c = 0;
    for (auto& w : a) {
        uint64_t t = (uint64_t)w * m + c;
        w = (uint32_t)t;
        c = t >> 32;
    }
    if (c) a.push_back((uint32_t)c);
}
static uint32_t big_divmod(big& a, uint32_t d) {
    uint64_t r = 0;
    for (size_t i = a.size(); i-- > 0;) {
        uint64_t cur = (r << 32) | a[i];
        a[i] = (uint32_t)(cur / d);
        r = cur % d;
    }
    return (uint32_t)r;
}

static void build_constants(dmx_ctx* c, std::vector<uint32_t>& host_tables) {
    const size_t L = MAX_LOG_N + 1;
    host_tables.assign((size_t)NPRIMES * TW_SIZE * 4 + (size_t)NPRIMES * L * 2, 0);
    uint32_t* zm = host_tables.data();
    uint32_t* zmp = zm + NPRIMES * TW_SIZE;
    uint32_t* izm = zmp + NPRIMES * TW_SIZE;
    uint32_t* izmp = izm + NPRIMES * TW_SIZE;
    uint32_t* inm = izmp + NPRIMES * TW_SIZE;
    uint32_t* inmp = inm + NPRIMES * L;
    const zq q = zq{{Q0, Q1, Q2, Q3}};
    for (int i = 0; i < NPRIMES; ++i) {
        const uint64_t p = PRIMES[i];
        PrimeConsts& k = c->pc[i];
        memset(&k, 0, sizeof k);
        k.p = (uint32_t)p;
        k.pinv = host_inv32((uint32_t)p);
        const uint64_t R = (1ull << 32) % p;
        const uint64_t r4 = host_powmod(R, 4, p);
        k.r4_m = (uint32_t)r4;
        k.r4_mp = (uint32_t)r4 * k.pinv;
        const uint64_t r3 = host_powmod(R, 3, p);
        k.r3_m = (uint32_t)r3;
        k.r3_mp = (uint32_t)r3 * k.pinv;
        k.two64 = (uint32_t)host_powmod(R, 2, p);
        // q mod p
        uint64_t qm = 0;
        for (int l = 3; l >= 0; --l) qm = ((qm << 32) | q.w[l]) % p;
        k.q_mod_p = (uint32_t)qm;
        // twiddles: psi a primitive 2^(MAX_LOG_N+1)-th root of unity
        const uint64_t g = smallest_generator(p);
        const uint64_t psi = host_powmod(g, (p - 1) >> (MAX_LOG_N + 1), p);
        const uint64_t order = 1ull << (MAX_LOG_N + 1);
        for (uint32_t t = 0; t < (uint32_t)TW_SIZE; ++t) {
            const uint64_t e = brv(t, MAX_LOG_N);
            const uint64_t z = host_powmod(psi, e, p), iz = host_powmod(psi, (order - e) % order, p);
            const uint32_t zmv = (uint32_t)((z << 32) % p), izmv = (uint32_t)((iz << 32) % p);
            zm[i * TW_SIZE + t] = zmv;
            zmp[i * TW_SIZE + t] = zmv * k.pinv;
            izm[i * TW_SIZE + t] = izmv;
            izmp[i * TW_SIZE + t] = izmv * k.pinv;
        }
        for (size_t lg = 0; lg < L; ++lg) {
            const uint64_t inv = host_powmod((1ull << lg) % p, p - 2, p);
            const uint32_t v = (uint32_t)((inv << 32) % p);
            inm[i * L + lg] = v;
            inmp[i * L + lg] = v * k.pinv;
        }
    }
    // CRT constants (ntt.rs:379-398): prefix products in CRT_PRIMES order
    CrtConsts& cr = c->crt;
    memset(&cr, 0, sizeof cr);
    for (int i = 0; i < NPRIMES; ++i) {
        const uint64_t pi = PRIMES[i];
        uint64_t pref = 1;
        for (int j = 0; j < i; ++j) {
            cr.mod_prefix[i][j] = (uint32_t)pref;
            pref = (pref * PRIMES[j]) % pi;
        }
        cr.mod_prefix[i][i] = (uint32_t)pref;
        cr.inv[i] = i ? (uint32_t)host_powmod(pref, pi - 2, pi) : 1;
        const uint32_t pinv = host_inv32((uint32_t)pi);
        for (int j = 0; j < i; ++j) {
            const uint32_t wm = (uint32_t)(((uint64_t)cr.mod_prefix[i][j] << 32) % pi);
            cr.pre_m[i][j] = wm;
            cr.pre_mp[i][j] = wm * pinv;
        }
        const uint32_t im = (uint32_t)(((uint64_t)cr.inv[i] << 32) % pi);
        cr.inv_m[i] = im;
        cr.inv_mp[i] = im * pinv;
    }
    big P{1};
    zq pq = zq_from_u64(1);
    for (int i = 0; i < NPRIMES; ++i) {
        cr.prefix_q[i] = pq;
        pq = zq_mul(pq, zq_from_u64(PRIMES[i]));
        big_mul(P, PRIMES[i]);
    }
    cr.P_q = pq;
    // floor(P/2) in mixed radix
    big h = P;
    uint32_t carry = 0;
    for (size_t i = h.size(); i-- > 0;) {
        uint32_t nc = h[i] & 1u;
        h[i] = (h[i] >> 1) | (carry << 31);
        carry = nc;
    }
    for (int i = 0; i < NPRIMES; ++i) cr.half_digits[i] = big_divmod(h, PRIMES[i]);
}

// Forward-transform autotuning: per degree 2^6..2^11, every kernel variant transforms the same
// synthetic fold rows (a + s*b over 2^21 / m rows, the shape of the prover's folds) and the
// fastest of three timed launches each is kept. All variants give the same residues, so the
// choice changes speed only. DAMASCUS_NTT_TUNE=0 keeps variant 0 everywhere.
static int tune_transforms(dmx_ctx* c) {
    for (int l = 0; l <= MAX_LOG_N; ++l) {
        c->fwd_variant[l] = 0;
        for (int v = 0; v < 4; ++v) c->fwd_us[l][v] = 0.f;
    }
    const char* e = getenv("DAMASCUS_NTT_TUNE");
    if (e && strcmp(e, "0") == 0) return DMX_OK;
    const size_t max_rows = (size_t)1 << 15, coeffs = (size_t)1 << 21;
    DevBuf in, outb, scal;
    int rc = in.alloc(2 * coeffs * 16);
    if (!rc) rc = outb.alloc(coeffs * NPRIMES * 4);
    if (!rc) rc = scal.alloc(16);
    if (rc) return DMX_OK;  // no room for the tuning rows: keep the defaults
    launch_fill32(in.as<uint32_t>(), in.bytes / 4, 0x0000ABCDu, 0);  // every 16-byte word below q
    const uint32_t s_words[4] = {0x12345678u, 0x9ABCDEF0u, 0x0FEDCBA9u, 0x00001234u};  // < q
    HIP_TRY(hipMemcpy(scal.p, s_words, 16, hipMemcpyHostToDevice));
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    int result = DMX_OK;
    for (int l = 6; l <= 11 && result == DMX_OK; ++l) {
        const int nv = fwd_variant_count(l);
        if (nv < 2) continue;
        const size_t rows = std::min(max_rows, coeffs >> l);
        RowSrc src{};
        src.mode = ROW_VFOLD;
        src.a = in.as<zq>();
        src.b_off = rows;
        src.scal = scal.as<zq>();
        float best = 0.f;
        for (int v = 0; v < nv && result == DMX_OK; ++v) {
            float t_min = 0.f;
            for (int rep = 0; rep < 4; ++rep) {  // the first launch loads the code object
                (void)hipEventRecord(a, 0);
                launch_ntt_fwd_rows(src, rows, l, outb.as<uint32_t>(), c->tw, 0, Jobs{}, v);
                (void)hipEventRecord(b, 0);
                if (hipEventSynchronize(b) != hipSuccess || hipGetLastError() != hipSuccess) {
                    result = fail(DMX_ERR_HIP, "transform tuning launch failed (degree 2^%d, variant %d)", l, v);
                    break;
                }
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, a, b);
                if (rep && (t_min == 0.f || ms < t_min)) t_min = ms;
            }
            c->fwd_us[l][v] = t_min * 1e3f;
            if (result == DMX_OK && (v == 0 || t_min < best)) {
                best = t_min;
                c->fwd_variant[l] = v;
            }
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return result;
}

// MAC-pass tuning: per degree 2^6..2^10, every segmentation shape runs the pass pair of a round at
// that degree on synthetic residues (the pairs pass with cross terms over 2^l rows and the split
// pass over 2^(l-1) rows, each with its finish kernel), three timed launches each after one
// untimed. Every shape gives the same sums, so the choice changes speed only; a shape replaces the
// default only when it is at least 2 % faster, so timing noise does not move the choice.
static int tune_mac(dmx_ctx* c) {
    for (int l = 0; l <= MAX_LOG_N; ++l) {
        c->mac_idx[l] = MAC_SHAPE_DEFAULT;
        for (int v = 0; v < MAC_SHAPES; ++v) c->mac_us[l][v] = 0.f;
    }
    const char* e = getenv("DAMASCUS_NTT_TUNE");
    if (e && strcmp(e, "0") == 0) return DMX_OK;
    const int lmin = 6, lmax = 10;
    const size_t mmax = (size_t)1 << lmax;
    DevBuf w, g, part, out;
    int rc = w.alloc(mmax * NPRIMES * mmax * 4);
    if (!rc) rc = g.alloc(mmax * K * NPRIMES * mmax * 4);
    if (!rc) rc = part.alloc(mac_part_words(lmax) * 4);
    if (!rc) rc = out.alloc((size_t)3 * K * NPRIMES * mmax * 4);
    if (rc) {  // no room for the tuning rows: keep the defaults
        (void)hipGetLastError();
        return DMX_OK;
    }
    launch_fill32(w.as<uint32_t>(), w.bytes / 4, 0x0000ABCDu, 0);  // below every prime
    launch_fill32(g.as<uint32_t>(), g.bytes / 4, 0x00001234u, 0);
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    int result = DMX_OK;
    for (int l = lmin; l <= lmax && result == DMX_OK; ++l) {
        const size_t rows = (size_t)1 << l, m = (size_t)1 << l, slot = (size_t)K * NPRIMES * m;
        uint32_t* o = out.as<uint32_t>();
        for (int v = 0; v < MAC_SHAPES && result == DMX_OK; ++v) {
            const MacShape sh = mac_shape_of(v);
            float t_min = 0.f;
            for (int rep = 0; rep < 4; ++rep) {
                (void)hipEventRecord(a, 0);
                launch_mac_pairs_inv(w.as<uint32_t>(), g.as<uint32_t>(), rows, l, o, o + slot, o + 2 * slot,
                                     part.as<uint32_t>(), c->tw, 0, Jobs{}, false, sh);
                launch_mac_split_inv(w.as<uint32_t>(), g.as<uint32_t>(), rows / 2, l, o, o + slot, part.as<uint32_t>(),
                                     c->tw, 0, Jobs{}, sh);
                (void)hipEventRecord(b, 0);
                if (hipEventSynchronize(b) != hipSuccess || hipGetLastError() != hipSuccess) {
                    result = fail(DMX_ERR_HIP, "MAC tuning launch failed (degree 2^%d, shape %d)", l, v);
                    break;
                }
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, a, b);
                if (rep && (t_min == 0.f || ms < t_min)) t_min = ms;
            }
            c->mac_us[l][v] = t_min * 1e3f;
        }
        float best = c->mac_us[l][MAC_SHAPE_DEFAULT];
        for (int v = 0; v < MAC_SHAPES; ++v)
            if (c->mac_us[l][v] > 0.f && c->mac_us[l][v] < 0.98f * c->mac_us[l][MAC_SHAPE_DEFAULT] &&
                c->mac_us[l][v] < best) {
                best = c->mac_us[l][v];
                c->mac_idx[l] = v;
            }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return result;
}

// The tuning cache: the choices and timings of a context, keyed by this build and the device
// (architecture name, compute units), so that later contexts on the same kind of device skip the
// timing runs. DAMASCUS_TUNE_CACHE names the file ("0": none); by default
// $XDG_CACHE_HOME/damascus000_amd/tuning.txt or ~/.cache/damascus000_amd/tuning.txt. The file is
// replaced by a rename, so concurrent processes read either the old or the new one.
static const char* const TUNE_BUILD = __DATE__ "_" __TIME__;

static std::string tune_cache_path() {
    const char* e = getenv("DAMASCUS_TUNE_CACHE");
    if (e) return strcmp(e, "0") == 0 ? std::string() : std::string(e);
    const char* x = getenv("XDG_CACHE_HOME");
    const char* h = getenv("HOME");
    if (x && *x) return std::string(x) + "/damascus000_amd/tuning.txt";
    if (h && *h) return std::string(h) + "/.cache/damascus000_amd/tuning.txt";
    return std::string();
}

static std::string tune_key(const dmx_ctx* c) {
    hipDeviceProp_t prop;
    std::string arch = "unknown";
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) arch = prop.gcnArchName;
    for (char& ch : arch)
        if (ch == ' ') ch = '_';
    std::string build = TUNE_BUILD;
    for (char& ch : build)
        if (ch == ' ') ch = '_';
    return "dmx-tune 2 " + build + " " + arch + " " + std::to_string(c->cus);
}

static bool tune_load(dmx_ctx* c, const std::string& path) {
    FILE* f = fopen(path.c_str(), "r");
    if (!f) return false;
    char line[512];
    bool ok = fgets(line, sizeof line, f) != nullptr;
    if (ok) {
        std::string head(line);
        while (!head.empty() && (head.back() == '\n' || head.back() == '\r')) head.pop_back();
        ok = head == tune_key(c);
    }
    int fwd_seen = 0, mac_seen = 0;
    while (ok && fgets(line, sizeof line, f)) {
        char kind[8] = {0};
        int l = -1, v = -1, n = 0;
        if (sscanf(line, "%7s %d %d%n", kind, &l, &v, &n) < 3 || l < 0 || l > MAX_LOG_N) {
            ok = false;
            break;
        }
        const char* p = line + n;
        float us[MAC_SHAPES] = {0};
        const int want = strcmp(kind, "fwd") == 0 ? 4 : MAC_SHAPES;
        for (int i = 0; i < want; ++i) {
            int used = 0;
            if (sscanf(p, "%f%n", &us[i], &used) != 1) {
                ok = false;
                break;
            }
            p += used;
        }
        if (!ok) break;
        if (strcmp(kind, "fwd") == 0 && v >= 0 && v < std::max(1, fwd_variant_count(l))) {
            c->fwd_variant[l] = v;
            for (int i = 0; i < 4; ++i) c->fwd_us[l][i] = us[i];
            ++fwd_seen;
        } else if (strcmp(kind, "mac") == 0 && v >= 0 && v < MAC_SHAPES) {
            c->mac_idx[l] = v;
            for (int i = 0; i < MAC_SHAPES; ++i) c->mac_us[l][i] = us[i];
            ++mac_seen;
        } else {
            ok = false;
        }
    }
    fclose(f);
    return ok && fwd_seen == MAX_LOG_N + 1 && mac_seen == MAX_LOG_N + 1;
}

static void tune_save(const dmx_ctx* c, const std::string& path) {
    const size_t slash = path.rfind('/');
    if (slash != std::string::npos && slash > 0) {  // create the directory (and one parent), if missing
        const std::string dir = path.substr(0, slash);
        const size_t s2 = dir.rfind('/');
        if (s2 != std::string::npos && s2 > 0) (void)mkdir(dir.substr(0, s2).c_str(), 0755);
        (void)mkdir(dir.c_str(), 0755);
    }
    const std::string tmp = path + ".tmp" + std::to_string((long)getpid());
    FILE* f = fopen(tmp.c_str(), "w");
    if (!f) return;
    fprintf(f, "%s\n", tune_key(c).c_str());
    for (int l = 0; l <= MAX_LOG_N; ++l) {
        fprintf(f, "fwd %d %d", l, c->fwd_variant[l]);
        for (int i = 0; i < 4; ++i) fprintf(f, " %.3f", c->fwd_us[l][i]);
        fprintf(f, "\n");
    }
    for (int l = 0; l <= MAX_LOG_N; ++l) {
        fprintf(f, "mac %d %d", l, c->mac_idx[l]);
        for (int i = 0; i < MAC_SHAPES; ++i) fprintf(f, " %.3f", c->mac_us[l][i]);
        fprintf(f, "\n");
    }
    const bool good = fclose(f) == 0;
    if (!good || rename(tmp.c_str(), path.c_str()) != 0) (void)remove(tmp.c_str());
}

// Transform variants and MAC shapes: from the cache when it holds this build's choices for this
// kind of device, else timed now and saved.
static int tune_context(dmx_ctx* c) {
    const char* e = getenv("DAMASCUS_NTT_TUNE");
    const bool off = e && strcmp(e, "0") == 0;
    c->tune_path = off ? std::string() : tune_cache_path();
    if (!c->tune_path.empty()) {
        for (int l = 0; l <= MAX_LOG_N; ++l) {
            c->fwd_variant[l] = 0;
            c->mac_idx[l] = MAC_SHAPE_DEFAULT;
        }
        if (tune_load(c, c->tune_path)) {
            c->tune_source = 2;
            return DMX_OK;
        }
    }
    int rc = tune_transforms(c);
    if (!rc) rc = tune_mac(c);
    if (rc) return rc;
    c->tune_source = off ? 0 : 1;
    if (!off && !c->tune_path.empty()) tune_save(c, c->tune_path);
    return DMX_OK;
}

extern "C" int dmx_ctx_mac_shape(dmx_ctx* ctx, int log_m, int* shape, int* count, float* tuned_us) {
    if (!ctx || log_m < 0 || log_m > MAX_LOG_N) return fail(DMX_ERR_ARG, "bad context or degree");
    if (shape) *shape = ctx->mac_idx[log_m];
    if (count) *count = MAC_SHAPES;
    if (tuned_us)
        for (int v = 0; v < MAC_SHAPES; ++v) tuned_us[v] = ctx->mac_us[log_m][v];
    return DMX_OK;
}

extern "C" int dmx_ctx_set_mac_shape(dmx_ctx* ctx, int log_m, int shape) {
    if (!ctx || log_m < 0 || log_m > MAX_LOG_N) return fail(DMX_ERR_ARG, "bad context or degree");
    if (shape < 0 || shape >= MAC_SHAPES) return fail(DMX_ERR_ARG, "MAC shape %d is not one of %d", shape, MAC_SHAPES);
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->mac_idx[log_m] != shape) {
        ctx->mac_idx[log_m] = shape;
        ++ctx->fwd_version;
    }
    return DMX_OK;
}

extern "C" int dmx_ctx_tuning_source(const dmx_ctx* ctx, char* path, size_t cap) {
    if (!ctx) return -1;
    if (path && cap) snprintf(path, cap, "%s", ctx->tune_path.c_str());
    return ctx->tune_source;
}

extern "C" int dmx_ctx_transform_variant(dmx_ctx* ctx, int log_m, int* variant, int* count, float* tuned_us) {
    if (!ctx || log_m < 0 || log_m > MAX_LOG_N) return fail(DMX_ERR_ARG, "bad context or degree");
    const int nv = fwd_variant_count(log_m);
    if (variant) *variant = ctx->fwd_variant[log_m];
    if (count) *count = nv;
    if (tuned_us)
        for (int v = 0; v < nv; ++v) tuned_us[v] = ctx->fwd_us[log_m][v];
    return DMX_OK;
}

extern "C" int dmx_ctx_set_transform_variant(dmx_ctx* ctx, int log_m, int variant) {
    if (!ctx || log_m < 0 || log_m > MAX_LOG_N) return fail(DMX_ERR_ARG, "bad context or degree");
    if (variant < 0 || variant >= fwd_variant_count(log_m))
        return fail(DMX_ERR_ARG, "degree 2^%d has %d transform variants", log_m, fwd_variant_count(log_m));
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->fwd_variant[log_m] != variant) {
        ctx->fwd_variant[log_m] = variant;
        ++ctx->fwd_version;
    }
    return DMX_OK;
}

extern "C" int dmx_ctx_create(int device_index, dmx_ctx** out) {
    if (!out) return fail(DMX_ERR_ARG, "null output");
    if (!dmx_backend_ready()) return fail(DMX_ERR_NO_DEVICE, "no usable HIP device (or DAMASCUS_GPU=0)");
    HIP_TRY(hipSetDevice(device_index));
    std::unique_ptr<dmx_ctx> c(new dmx_ctx);
    c->device = device_index;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device_index) == hipSuccess && prop.multiProcessorCount > 0)
            c->cus = prop.multiProcessorCount;
    }
    std::vector<uint32_t> host;
    build_constants(c.get(), host);
    int rc = c->tables.alloc(host.size() * 4);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(c->tables.p, host.data(), host.size() * 4, hipMemcpyHostToDevice));
    uint32_t* base = c->tables.as<uint32_t>();
    c->tw.zeta_m = base;
    c->tw.zeta_mp = base + NPRIMES * TW_SIZE;
    c->tw.izeta_m = base + 2 * NPRIMES * TW_SIZE;
    c->tw.izeta_mp = base + 3 * NPRIMES * TW_SIZE;
    c->tw.invn_m = base + 4 * NPRIMES * TW_SIZE;
    c->tw.invn_mp = base + 4 * NPRIMES * TW_SIZE + NPRIMES * (MAX_LOG_N + 1);
    set_device_constants(c->pc, &c->crt, 0);
    {
        const hipError_t e = noalloc_scratch_init(device_index, &c->noalloc);
        if (e != hipSuccess) return fail(DMX_ERR_HIP, "allocation-free scratch lookup: %s", hipGetErrorString(e));
    }
    HIP_TRY(hipStreamCreateWithFlags(&c->ring_stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&c->ring_done, hipEventDisableTiming));
    HIP_TRY(hipDeviceSynchronize());
    rc = tune_context(c.get());
    if (rc) return rc;
    *out = c.release();
    return DMX_OK;
}

static void batch_forget_context(dmx_ctx* ctx);
extern "C" void dmx_ctx_destroy(dmx_ctx* ctx) {
    if (!ctx) return;
    batch_forget_context(ctx);  // cached batch workspaces hold the context's tables
    delete ctx;
}

// ------------------------------------------------------------------------------------------
// job workspaces. Every device buffer of one job lives in one block; a batch of jobs of one
// depth is n such blocks at a fixed stride, so one launch covers every job (Jobs{n, stride}).
// ------------------------------------------------------------------------------------------
struct Slice {
    void* p = nullptr;
    size_t bytes = 0;
    template <class T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

struct Layout {
    size_t size = 0;
    std::vector<std::pair<Slice*, size_t>> items;
    void add(Slice& s, size_t bytes) {
        size = (size + 255) & ~(size_t)255;
        items.emplace_back(&s, size);
        s.bytes = bytes;
        size += std::max<size_t>(bytes, 16);
    }
    size_t stride() const { return (size + 255) & ~(size_t)255; }
    int commit(DevBuf& store, uint32_t jobs) {
        int rc = store.alloc(stride() * jobs);
        if (rc) return rc;
        for (auto& it : items) it.first->p = store.as<uint8_t>() + it.second;
        return DMX_OK;
    }
};

// Per-job header in device memory, written by the host before a job runs.
struct JobMeta {
    uint64_t len;        // file length in bytes
    uint64_t src;        // device address the file is copied in from (0: already in place)
    uint32_t reject;     // challenge candidates rejected per challenge (test knob, 0 normally)
    uint32_t pad0;
    uint64_t pad1;
    uint32_t gseed[8];   // generator key words: g_0_seed (sis.rs:70)
    uint8_t epoch[32];   // epoch_seed
    uint8_t file_id[32]; // written by the device
};
static_assert(sizeof(JobMeta) == 128, "job header layout");
constexpr size_t META_REJECT = 16, META_GSEED = 32, META_SEEDS = 64;

template <class T>
static T* job_ptr(T* p, const Jobs& jb, uint32_t b) {
    return reinterpret_cast<T*>(reinterpret_cast<uintptr_t>(p) + (size_t)b * jb.stride);
}

// n rows of `width` bytes between pitched buffers; one 2-D copy while the pitches are modest
static hipError_t copy_rows(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t n,
                            hipMemcpyKind kind, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (dpitch < ((size_t)1 << 30) && spitch < ((size_t)1 << 30))
        return hipMemcpy2DAsync(dst, dpitch, src, spitch, width, n, kind, s);
    for (size_t i = 0; i < n; ++i) {
        hipError_t e = hipMemcpyAsync(static_cast<uint8_t*>(dst) + i * dpitch,
                                      static_cast<const uint8_t*>(src) + i * spitch, width, kind, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

static const uint32_t IVW[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};

static void gseed_words(const uint8_t g_seed[32], uint32_t w[8]) { b3::words_le(g_seed, w, 8); }

// Test knob DAMASCUS_TEST_REJECT: "n" makes every job's challenges reject the candidates of
// counters 0..n-1 (transcript.rs:66-75's resampling branch, which real digests reach with
// probability about 2^-112); "n@len" does so only for jobs whose file is `len` bytes long. Read
// when a job's header is written, so captured graphs replay with it. n >= 64 exhausts the
// device's bounded resampling and fails the job.
static uint32_t test_reject_for(uint64_t len) {
    const char* e = getenv("DAMASCUS_TEST_REJECT");
    if (!e || !*e) return 0;
    char* end = nullptr;
    const unsigned long n = strtoul(e, &end, 10);
    if (end && *end == '@') return strtoull(end + 1, nullptr, 10) == len ? (uint32_t)n : 0;
    return (uint32_t)n;
}

static JobMeta make_meta(uint64_t len, const void* src, const uint8_t g_seed[32], const uint8_t epoch[32]) {
    JobMeta m;
    memset(&m, 0, sizeof m);
    m.len = len;
    m.src = reinterpret_cast<uint64_t>(src);
    m.reject = test_reject_for(len);
    gseed_words(g_seed, m.gseed);
    memcpy(m.epoch, epoch, 32);
    return m;
}

struct Transcriber {
    Slice msg, cvs;
    uint8_t* seeds = nullptr;  // epoch_seed[32] | file_id[32], inside the job header
    const uint32_t* reject = nullptr;  // the job header's rejection count
    size_t msg_cap = 0, cv_words = 0;
    void plan(Layout& L, size_t max_ring, size_t extra_hash_len) {
        const uint32_t lens[3] = {(uint32_t)max_ring, (uint32_t)max_ring, (uint32_t)max_ring};
        msg_cap = transcript_len(4, lens);
        cv_words = device_hash_scratch_words(std::max(msg_cap, extra_hash_len));
        L.add(msg, msg_cap);
        L.add(cvs, cv_words * 4);
    }
    const uint8_t* epoch() const { return seeds; }
    uint8_t* file_id() const { return seeds + 32; }
    // the challenge message (transcript.rs:50-61, 77-90) over the modules (c, l, r)
    TranscriptMsg message(const char* label, uint64_t round, const zq* c, size_t lc, const zq* l, size_t ll,
                          const zq* r, size_t lr) const {
        TranscriptMsg t;
        t.epoch_seed = epoch();
        t.file_id = file_id();
        t.label_len = (uint32_t)strlen(label);
        memset(t.label, 0, 4);
        memcpy(t.label, label, t.label_len);
        t.round = round;
        t.mods[0] = c;
        t.mods[1] = l;
        t.mods[2] = r;
        t.lens[0] = (uint32_t)lc;
        t.lens[1] = (uint32_t)ll;
        t.lens[2] = (uint32_t)lr;
        t.reject = reject;
        return t;
    }
    // challenge (transcript.rs:42-63) written as (x, x^-1) to out, for every job
    void challenge(const char* label, uint64_t round, const zq* c, size_t lc, const zq* l, size_t ll,
                   const zq* r, size_t lr, zq* out, uint32_t* flags, hipStream_t s, const Jobs& jb,
                   bool inverse = true) {
        const TranscriptMsg t = message(label, round, c, lc, l, ll, r, lr);
        transcript_challenge(t, IVW, 0, msg.as<uint8_t>(), cvs.as<uint32_t>(), out, flags, s, jb, inverse);
    }
};

// ------------------------------------------------------------------------------------------
// prover (src/protocol/prover.rs:56-313)
// ------------------------------------------------------------------------------------------
enum : uint32_t { FLAG_ZERO_CHALLENGE = 0xC0000000u };
static uint32_t vec_bit(size_t j) { return 1u << (2 * j); }
static uint32_t poly_bit(size_t j) { return 1u << (2 * j + 1); }

struct dmx_prover {
    dmx_ctx* ctx = nullptr;
    hipStream_t s = nullptr;
    Jobs jb;                                      // jobs of this depth proved side by side
    std::vector<dmx_system_params> params;        // per job
    std::vector<std::array<uint8_t, 32>> g_seed;  // per job: g_0_seed (sis.rs:70)
    std::vector<uint64_t> orig_len;               // per job
    int d = 0;
    size_t N = 0;
    size_t round = 0;
    bool poisoned = false;
    bool capturing = false;  // inside hipGraph capture: no timing events
    size_t hash_cap = 0;     // bound on every job's file length (sizes the file hash launch)
    DevBuf store;            // the workspace, unless a batch owns it
    // out_head, com0 and recs are consecutive: a job's result is one contiguous range
    Slice meta, file, W_a, W_b, G_a, G_b, What_a, What_b, Ghat_a, Ghat_b, res, part, lifted, C, Ct, out_head, com0,
        recs, scal, unsc, flags;
    // Challenge inverses are off the critical path: the folds take the challenge itself and
    // the generator state is kept scaled (x g_L + g_R = x (g_L + x^-1 g_R), y g_e + X g_o =
    // y (g_e + y^-1 X g_o)); stage t's lifts multiply by unsc[t+1], the product of every
    // inverse so far, which a workgroup of the stage's forward-transform launch computes while
    // the others transform (InvTask).
    std::vector<size_t> rec_off;  // offset (coefficients) of round j's record block
    Transcriber tr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    const uint8_t* d_file = nullptr;
    std::vector<JobMeta> stage;  // host copy of the job headers
    // Resident generator matrix shared by every job (one seed): Z_q rows and their transforms.
    // When set, no job expands or transforms its own generators.
    const zq* G0_src = nullptr;
    const uint32_t* Ghat0_src = nullptr;

    ~dmx_prover() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (s) (void)hipStreamDestroy(s);
    }
    size_t m_at(size_t j) const { return N >> j; }
    zq* rec_lvec(size_t j) const { return recs.as<zq>() + rec_off[j]; }
    zq* rec_rvec(size_t j) const { return rec_lvec(j) + (size_t)K * m_at(j); }
    zq* rec_lpoly(size_t j) const { return rec_lvec(j) + (size_t)2 * K * m_at(j); }
    zq* rec_rpoly(size_t j) const { return rec_lpoly(j) + (size_t)K * (m_at(j) / 2); }
    uint32_t* res_at(size_t words) const { return res.as<uint32_t>() + words; }
    zq* sc(size_t i) const { return scal.as<zq>() + i; }
    zq* un(size_t t) const { return unsc.as<zq>() + t; }
    const uint64_t* dlen() const { return meta.as<uint64_t>(); }
    const uint32_t* gseed() const { return reinterpret_cast<const uint32_t*>(meta.as<uint8_t>() + META_GSEED); }
};

// file_cap = 0: the file stays where the caller keeps it (one job)
static void prover_plan(dmx_prover* p, Layout& L, size_t file_cap) {
    const size_t N = p->N;
    const size_t NN = N * N;
    L.add(p->meta, sizeof(JobMeta));
    if (file_cap) L.add(p->file, file_cap);
    L.add(p->W_a, NN * 16);
    L.add(p->W_b, NN / 2 * 16);
    L.add(p->G_a, NN * K * 16);
    L.add(p->G_b, NN / 2 * K * 16);
    L.add(p->What_a, NN * NPRIMES * 4);
    L.add(p->What_b, NN / 2 * NPRIMES * 4);
    L.add(p->Ghat_a, NN * K * NPRIMES * 4);
    L.add(p->Ghat_b, NN / 2 * K * NPRIMES * 4);
    L.add(p->res, (size_t)3 * K * NPRIMES * N * 4);
    L.add(p->part, mac_part_words(p->d) * 4);
    L.add(p->lifted, (size_t)K * N * 16);
    L.add(p->C, (size_t)K * N * 16);
    L.add(p->Ct, (size_t)K * N * 16);
    L.add(p->out_head, RESULT_HEAD_BYTES);
    L.add(p->com0, (size_t)K * N * 16);
    size_t total = 0;
    p->rec_off.resize(p->d + 1);
    for (int j = 0; j <= p->d; ++j) {
        p->rec_off[j] = total;
        size_t m = N >> j;
        total += (size_t)K * (2 * m + m);
    }
    L.add(p->recs, total * 16);
    L.add(p->scal, (size_t)4 * (p->d + 1) * 16);
    L.add(p->unsc, (size_t)2 * (p->d + 1) * 16);
    L.add(p->flags, 256);
    p->tr.plan(L, N, p->hash_cap);
}

static void prover_bind(dmx_prover* p) {
    p->tr.seeds = p->meta.as<uint8_t>() + META_SEEDS;
    p->tr.reject = reinterpret_cast<const uint32_t*>(p->meta.as<uint8_t>() + META_REJECT);
    if (p->file.p) p->d_file = p->file.as<uint8_t>();
}

// initialize_with_config (prover.rs:61-111), first part: file_id and the witness.
static int prover_init_file(dmx_prover* p) {
    hipStream_t s = p->s;
    const Jobs& jb = p->jb;
    launch_fill32(p->flags.as<uint32_t>(), 64, 0, s, jb);
    device_hash(p->d_file, p->hash_cap, IVW, 0, p->tr.cvs.as<uint32_t>(), p->tr.cv_words,
                reinterpret_cast<uint32_t*>(p->tr.file_id()), nullptr, p->flags.as<uint32_t>(), s, jb, p->dlen());
    launch_expand(p->d_file, p->hash_cap, p->W_a.as<zq>(), p->N * p->N, s, jb, p->dlen());
    HIP_TRY(hipGetLastError());
    return DMX_OK;
}

// second part: generators, com_0 and the round-0 vector cross terms, all from one pass over
// the generator transforms.
// `g_copy`: where the generator rows also go (the fused job's verifier matrix), or null
static int prover_init_commit(dmx_prover* p, zq* g_copy = nullptr) {
    hipStream_t s = p->s;
    const Jobs& jb = p->jb;
    const size_t N = p->N;
    const int d = p->d;
    const NttTables& tw = p->ctx->tw;
    RowSrc pl{};
    pl.mode = ROW_PLAIN;
    pl.a = p->W_a.as<zq>();
    RowSrc gen{};
    gen.mode = ROW_GEN;
    gen.dst = p->G_a.as<zq>();
    gen.dst2 = g_copy;
    gen.gseed = p->gseed();
    // witness rows, and the generator rows unless a resident matrix serves every job
    fwd_rows2(p->ctx, pl, N, p->What_a.as<uint32_t>(), gen, p->Ghat0_src ? 0 : (size_t)K * N,
              p->Ghat_a.as<uint32_t>(), d, s, jb);
    const size_t slot = (size_t)K * NPRIMES * N;
    const bool shared = p->Ghat0_src != nullptr;
    launch_mac_pairs_inv(p->What_a.as<uint32_t>(), shared ? p->Ghat0_src : p->Ghat_a.as<uint32_t>(), N, d,
                         p->res_at(0), p->res_at(slot), p->res_at(2 * slot), p->part.as<uint32_t>(), tw, s, jb, shared,
                         p->ctx->mac(d));
    // com_0 (to C and com0) and L_vec, R_vec (adjacent) in one lift
    launch_lift_check(p->res_at(0), d, ComUpd{}, p->C.as<zq>(), p->com0.as<zq>(), nullptr, 0, p->res_at(slot),
                      p->rec_lvec(0), 2 * K, d, s, jb);
    HIP_TRY(hipGetLastError());
    return DMX_OK;
}

static int prover_initialize(dmx_prover* p) {
    int rc = prover_init_file(p);
    return rc ? rc : prover_init_commit(p);
}

// Stages of degree <= 32 take one launch each (launch_stage_small); DAMASCUS_SMALL_STAGE=0 keeps
// them on the multi-launch chain (for comparisons). Their stage counter is a word of the job's
// flags block.
constexpr size_t SMALL_STAGE_MAX_M = 32;
constexpr size_t STAGE_CTR_WORD = 32;
static bool small_stages_enabled() {
    static const bool on = [] {
        const char* e = getenv("DAMASCUS_SMALL_STAGE");
        return !(e && strcmp(e, "0") == 0);
    }();
    return on;
}

// The operands of one stage of fold_round (prover.rs:140-313): the witness and generator folds,
// the challenge-inverse task, the invariant's expected update and where the lifts go.
struct StageArgs {
    RowSrc w, g;
    InvTask it;
    ComUpd cu;
    zq* com_out;
    zq* out_lr;
    int cross;  // the stage emits cross terms
    uint32_t bit;
};

// vector stage of round j: m_L + x m_R and x g_L + g_R = x (g_L + x^-1 g_R); the lift checks the
// invariant against C_t = C + x^-1 L + x R (written to Ct) and emits L/R_poly
static StageArgs vec_stage_args(const dmx_prover* p, size_t j) {
    const size_t h = p->m_at(j) / 2;
    StageArgs a{};
    a.w.mode = ROW_VFOLD;
    a.w.a = p->W_a.as<zq>();
    a.w.b_off = h;
    a.w.scal = p->sc(4 * j);
    a.w.dst = p->W_b.as<zq>();
    a.g.mode = ROW_VFOLD_L;
    a.g.a = (j == 0 && p->G0_src) ? p->G0_src : p->G_a.as<zq>();
    a.g.a_shared = j == 0 && p->G0_src;
    a.g.b_off = h * K;
    a.g.scal = p->sc(4 * j);
    a.g.dst = p->G_b.as<zq>();
    a.it = InvTask{p->sc(4 * j), j ? p->un(2 * j) : nullptr, p->un(2 * j + 1)};
    a.cu = ComUpd{p->C.as<zq>(), p->rec_lvec(j), p->rec_rvec(j), p->sc(4 * j), 0, p->un(2 * j + 1)};
    a.com_out = p->Ct.as<zq>();
    a.out_lr = p->rec_lpoly(j);
    a.cross = 1;
    a.bit = vec_bit(j);
    return a;
}

// poly stage of round j: e + y o and y g_e + X g_o = y (g_e + y^-1 X g_o); the lift checks the
// invariant against C' = even(C_t) + y^-1 L + y R (written to C) and emits the next round's
// vector cross terms
static StageArgs poly_stage_args(const dmx_prover* p, size_t j) {
    const size_t h = p->m_at(j) / 2;
    StageArgs a{};
    a.w.mode = ROW_PFOLD_W;
    a.w.a = p->W_b.as<zq>();
    a.w.scal = p->sc(4 * j + 2);
    a.w.dst = p->W_a.as<zq>();
    a.g.mode = ROW_PFOLD_GL;
    a.g.a = p->G_b.as<zq>();
    a.g.scal = p->sc(4 * j + 2);
    a.g.dst = p->G_a.as<zq>();
    a.it = InvTask{p->sc(4 * j + 2), p->un(2 * j + 1), p->un(2 * j + 2)};
    a.cu = ComUpd{p->Ct.as<zq>(), p->rec_lpoly(j), p->rec_rpoly(j), p->sc(4 * j + 2), 1, p->un(2 * j + 2)};
    a.com_out = p->C.as<zq>();
    a.out_lr = p->rec_lvec(j + 1);
    a.cross = h > 1 ? 1 : 0;
    a.bit = poly_bit(j);
    return a;
}

// one stage of degree <= 32 as a single launch (k_stage_small) or a step of the tail (k_tail)
static SmallStage small_stage_of(const dmx_prover* p, size_t j, int poly, const StageArgs& a) {
    SmallStage st{};
    st.poly = poly;
    st.log_m = p->d - (int)j;
    st.rows = (uint32_t)(p->m_at(j) / 2);
    st.w = a.w;
    st.g = a.g;
    st.cu = a.cu;
    st.com_out = a.com_out;
    st.out_lr = a.out_lr;
    st.cross = a.cross;
    st.flags = p->flags.as<uint32_t>();
    st.bit = a.bit;
    st.it = a.it;
    return st;
}

static size_t first_small_round(const dmx_prover* p) {
    size_t j = 0;
    while ((p->N >> j) > SMALL_STAGE_MAX_M) ++j;
    return j;
}

// fold_round (prover.rs:140-313) for round j, m = N >> j: both stages and both invariants.
// `vec_done`, when given, is recorded after the vector stage: the round's poly cross terms are
// then in place for a verifier that follows the proof half a round behind.
static int prover_round(dmx_prover* p, size_t j, hipEvent_t vec_done = nullptr) {
    hipStream_t s = p->s;
    const Jobs& jb = p->jb;
    const NttTables& tw = p->ctx->tw;
    const size_t m = p->m_at(j), h = m / 2;
    const int lm = p->d - (int)j;
    uint32_t* flags = p->flags.as<uint32_t>();
    if (!p->capturing) HIP_TRY(hipEventRecord(p->ev[0], s));
    const bool small = small_stages_enabled() && m <= SMALL_STAGE_MAX_M;
    // the job's small stages count up on flags[STAGE_CTR_WORD] (zeroed with the flags)
    const uint32_t small_ord = small ? (uint32_t)(2 * (j - first_small_round(p))) : 0;
    // ---- vector stage ----
    p->tr.challenge("vec", j, p->C.as<zq>(), m, p->rec_lvec(j), m, p->rec_rvec(j), m, p->sc(4 * j), flags, s, jb,
                    false);
    const StageArgs va = vec_stage_args(p, j);
    // invariant at degree m and the poly cross terms at degree m/2 from one pass
    const size_t s0 = (size_t)K * NPRIMES * m;
    if (small) {
        SmallStage st = small_stage_of(p, j, 0, va);
        st.ctr = flags + STAGE_CTR_WORD;
        st.target = small_ord + 1;
        launch_stage_small(st, tw, s, jb);
    } else {
        fwd_rows2(p->ctx, va.w, h, p->What_b.as<uint32_t>(), va.g, h * K, p->Ghat_b.as<uint32_t>(), lm, s, jb, va.it);
        launch_mac_split_inv(p->What_b.as<uint32_t>(), p->Ghat_b.as<uint32_t>(), h, lm, p->res_at(0), p->res_at(s0),
                             p->part.as<uint32_t>(), tw, s, jb, p->ctx->mac(lm));
        launch_lift_check(p->res_at(0), lm, va.cu, va.com_out, nullptr, flags, va.bit, p->res_at(s0), va.out_lr,
                          2 * K, lm - 1, s, jb);
    }
    if (!p->capturing) HIP_TRY(hipEventRecord(p->ev[1], s));
    if (vec_done) HIP_TRY(hipEventRecord(vec_done, s));
    // ---- poly stage ----
    p->tr.challenge("poly", j, p->Ct.as<zq>(), m, p->rec_lpoly(j), h, p->rec_rpoly(j), h, p->sc(4 * j + 2), flags, s,
                    jb, false);
    const StageArgs pa = poly_stage_args(p, j);
    // post-poly invariant, plus the next round's vector cross terms from the same pass
    const size_t s1 = (size_t)K * NPRIMES * h;
    const bool next = pa.cross != 0;
    if (small) {
        SmallStage st = small_stage_of(p, j, 1, pa);
        st.ctr = flags + STAGE_CTR_WORD;
        st.target = small_ord + 2;
        launch_stage_small(st, tw, s, jb);
    } else {
        fwd_rows2(p->ctx, pa.w, h, p->What_a.as<uint32_t>(), pa.g, h * K, p->Ghat_a.as<uint32_t>(), lm - 1, s, jb,
                  pa.it);
        launch_mac_pairs_inv(p->What_a.as<uint32_t>(), p->Ghat_a.as<uint32_t>(), h, lm - 1, p->res_at(0),
                             next ? p->res_at(s1) : nullptr, next ? p->res_at(2 * s1) : nullptr,
                             p->part.as<uint32_t>(), tw, s, jb, false, p->ctx->mac(lm - 1));
        launch_lift_check(p->res_at(0), lm - 1, pa.cu, pa.com_out, nullptr, flags, pa.bit,
                          next ? p->res_at(s1) : nullptr, pa.out_lr, next ? 2 * K : 0, lm - 1, s, jb);
    }
    if (!p->capturing) HIP_TRY(hipEventRecord(p->ev[2], s));
    HIP_TRY(hipGetLastError());
    return DMX_OK;
}

static int check_round_flags(size_t j, uint32_t f) {
    if (f & FLAG_ZERO_CHALLENGE) return fail(DMX_ERR_HIP, "challenge resampling exhausted");
    if (f & STAGE_WAIT_EXPIRED) return fail(DMX_ERR_HIP, "a stage waited too long for its challenge inverse");
    if (f & vec_bit(j)) return fail(DMX_ERR_INVARIANT, "vector folding invariant failed (round %zu)", j);
    if (f & poly_bit(j)) return fail(DMX_ERR_INVARIANT, "round folding invariant failed (round %zu)", j);
    return DMX_OK;
}

static int encode_record(dmx_prover* p, size_t j, uint8_t* out, size_t cap, size_t* len) {
    const size_t m = p->m_at(j), h = m / 2;
    const size_t need = dmx_record_size(m, h);
    if (len) *len = need;
    if (!out) return DMX_OK;
    if (cap < need) return fail(DMX_ERR_BUFFER, "record buffer too small (%zu < %zu)", cap, need);
    std::vector<zq> host((size_t)K * (2 * m + m));
    HIP_TRY(hipMemcpyAsync(host.data(), p->rec_lvec(j), host.size() * 16, hipMemcpyDeviceToHost, p->s));
    HIP_TRY(hipStreamSynchronize(p->s));
    uint8_t* q = put_u64(out, j);
    q = put_module(q, host.data(), m);
    q = put_module(q, host.data() + (size_t)K * m, m);
    q = put_module(q, host.data() + (size_t)2 * K * m, h);
    put_module(q, host.data() + (size_t)2 * K * m + (size_t)K * h, h);
    return DMX_OK;
}

// every round's record from the host copy of one job's record blocks
static void encode_records_from(const dmx_prover* p, const zq* host, uint8_t* out) {
    uint8_t* q = out;
    for (size_t j = 0; j < (size_t)p->d; ++j) {
        const size_t m = p->m_at(j), h = m / 2;
        const zq* b = host + p->rec_off[j];
        q = put_u64(q, j);
        q = put_module(q, b, m);
        q = put_module(q, b + (size_t)K * m, m);
        q = put_module(q, b + (size_t)2 * K * m, h);
        q = put_module(q, b + (size_t)2 * K * m + (size_t)K * h, h);
    }
}

static void encode_statement_from(const dmx_prover* p, uint32_t b, const uint8_t fid[32], const zq* com,
                                  uint8_t* out) {
    uint8_t* q = out;
    memcpy(q, fid, 32);
    q = put_u64(q + 32, p->orig_len[b]);
    q = put_u64(q, (uint64_t)p->d);
    q = put_module(q, com, p->N);
    memcpy(q, p->g_seed[b].data(), 32);
}

static int prover_new_object(dmx_ctx* ctx, int d, uint32_t n, dmx_prover** out) {
    std::unique_ptr<dmx_prover> p(new dmx_prover);
    p->ctx = ctx;
    p->d = d;
    p->N = (size_t)1 << d;
    p->jb = Jobs{n, 0};
    p->params.resize(n);
    p->g_seed.resize(n);
    p->orig_len.assign(n, 0);
    p->stage.resize(n);
    HIP_TRY(hipStreamCreate(&p->s));
    for (auto& e : p->ev) HIP_TRY(hipEventCreate(&e));
    *out = p.release();
    return DMX_OK;
}

static void prover_set_job(dmx_prover* p, uint32_t b, const dmx_system_params* params, uint64_t len,
                           const void* src) {
    p->params[b] = *params;
    p->orig_len[b] = len;
    b3::host_hash_short(params->seed_generators, (const uint8_t*)"g0", 2, p->g_seed[b].data());  // sis.rs:70
    p->stage[b] = make_meta(len, src, p->g_seed[b].data(), params->epoch_seed);
}

static int prover_create(dmx_ctx* ctx, const uint8_t* file, const void* d_file, size_t len,
                         const dmx_system_params* params, const dmx_runtime_config* cfg, dmx_prover** out) {
    if (!ctx || !params || !out || (!file && !d_file && len)) return fail(DMX_ERR_ARG, "null argument");
    if (cfg && !cfg->gpu_enabled) return fail(DMX_ERR_NO_DEVICE, "gpu_enabled = false");
    if (gpu_disabled_by_env()) return fail(DMX_ERR_NO_DEVICE, "disabled by DAMASCUS_GPU=0");
    const int d = depth_for(len);
    const size_t N = (size_t)1 << d;
    const uint64_t limit = cfg ? cfg->max_preprocess_bytes : (512ull << 20);
    if (N * N > UINT64_MAX / 32 || (uint64_t)(N * N * 32) > limit)
        return fail(DMX_ERR_MEMORY_LIMIT,
                    "full witness expansion requires %llu bytes, above max_preprocess_bytes=%llu; split the file into smaller statements",
                    (unsigned long long)(N * N * 32), (unsigned long long)limit);
    if (d > 12) return fail(DMX_ERR_ARG, "depth %d exceeds the supported 12", d);
    HIP_TRY(hipSetDevice(ctx->device));
    dmx_prover* raw = nullptr;
    int rc = prover_new_object(ctx, d, 1, &raw);
    if (rc) return rc;
    std::unique_ptr<dmx_prover> p(raw);
    p->hash_cap = len;
    Layout L;
    prover_plan(p.get(), L, d_file ? 0 : std::max<size_t>(len, 16));
    rc = L.commit(p->store, 1);
    if (rc) return rc;
    prover_bind(p.get());
    if (d_file) p->d_file = static_cast<const uint8_t*>(d_file);
    prover_set_job(p.get(), 0, params, len, nullptr);
    HIP_TRY(hipMemcpyAsync(p->meta.p, p->stage.data(), sizeof(JobMeta), hipMemcpyHostToDevice, p->s));
    if (!d_file && len) HIP_TRY(hipMemcpyAsync(p->file.p, file, len, hipMemcpyHostToDevice, p->s));
    rc = prover_initialize(p.get());
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(p->s));
    uint32_t f = 0;
    HIP_TRY(hipMemcpy(&f, p->flags.p, 4, hipMemcpyDeviceToHost));
    if (f & FLAG_ZERO_CHALLENGE) return fail(DMX_ERR_HIP, "hashing failed");
    *out = p.release();
    return DMX_OK;
}

extern "C" int dmx_prover_new(dmx_ctx* ctx, const uint8_t* file, size_t len, const dmx_system_params* params,
                              const dmx_runtime_config* cfg, dmx_prover** out) {
    return prover_create(ctx, file, nullptr, len, params, cfg, out);
}

extern "C" void dmx_prover_free(dmx_prover* p) { delete p; }
extern "C" size_t dmx_prover_rounds_total(const dmx_prover* p) { return p ? (size_t)p->d : 0; }
extern "C" size_t dmx_prover_current_round(const dmx_prover* p) { return p ? p->round : 0; }
extern "C" void dmx_prover_current_dimensions(const dmx_prover* p, size_t* rows, size_t* ring_len) {
    const size_t m = p->N >> p->round;
    if (rows) *rows = m;
    if (ring_len) *ring_len = m;
}

extern "C" int dmx_prover_statement(const dmx_prover* p, uint8_t* out, size_t cap, size_t* len) {
    if (!p) return fail(DMX_ERR_ARG, "null prover");
    const size_t need = dmx_statement_size(p->N);
    if (len) *len = need;
    if (!out) return DMX_OK;
    if (cap < need) return fail(DMX_ERR_BUFFER, "statement buffer too small");
    std::vector<zq> com((size_t)K * p->N);
    uint8_t fid[32];
    HIP_TRY(hipMemcpyAsync(com.data(), p->com0.p, com.size() * 16, hipMemcpyDeviceToHost, p->s));
    HIP_TRY(hipMemcpyAsync(fid, p->tr.file_id(), 32, hipMemcpyDeviceToHost, p->s));
    HIP_TRY(hipStreamSynchronize(p->s));
    encode_statement_from(p, 0, fid, com.data(), out);
    return DMX_OK;
}

extern "C" int dmx_prover_current_commitment(const dmx_prover* p, uint64_t* out, size_t cap_words, size_t* ring_len) {
    const size_t m = p->N >> p->round;
    if (ring_len) *ring_len = m;
    if (!out) return DMX_OK;
    if (cap_words < (size_t)2 * K * m) return fail(DMX_ERR_BUFFER, "commitment buffer too small");
    HIP_TRY(hipMemcpyAsync(out, p->C.p, (size_t)K * m * 16, hipMemcpyDeviceToHost, p->s));
    HIP_TRY(hipStreamSynchronize(p->s));
    return DMX_OK;
}

extern "C" int dmx_prover_fold_round(dmx_prover* p, size_t j, uint8_t* record_out, size_t cap, size_t* len,
                                     dmx_round_output* timing) {
    if (!p) return fail(DMX_ERR_ARG, "null prover");
    if (p->poisoned) return fail(DMX_ERR_POISONED, "prover unusable after an earlier failure");
    if (j != p->round) return fail(DMX_ERR_ROUND, "round index mismatch");
    if (p->round >= (size_t)p->d) return fail(DMX_ERR_ROUND, "all rounds have finished");
    const size_t m = p->m_at(j);
    if (record_out && cap < dmx_record_size(m, m / 2)) {
        if (len) *len = dmx_record_size(m, m / 2);
        return fail(DMX_ERR_BUFFER, "record buffer too small");
    }
    int rc = prover_round(p, j);
    if (!rc) rc = hipStreamSynchronize(p->s) == hipSuccess ? DMX_OK : fail(DMX_ERR_HIP, "stream sync failed");
    uint32_t f = 0;
    if (!rc && hipMemcpy(&f, p->flags.p, 4, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(DMX_ERR_HIP, "flags");
    if (!rc) rc = check_round_flags(j, f);
    if (rc) {
        p->poisoned = true;
        return rc;
    }
    if (timing) {
        float a = 0, b = 0;
        (void)hipEventElapsedTime(&a, p->ev[0], p->ev[1]);
        (void)hipEventElapsedTime(&b, p->ev[1], p->ev[2]);
        timing->vector_fold_ns = (uint64_t)(a * 1e6);
        timing->poly_fold_ns = (uint64_t)(b * 1e6);
        timing->total_round_ns = (uint64_t)((a + b) * 1e6);
    }
    rc = encode_record(p, j, record_out, cap, len);
    if (rc) return rc;
    p->round += 1;
    return DMX_OK;
}

extern "C" int dmx_prover_final_opening(const dmx_prover* p, uint64_t out[2]) {
    if (!p || p->round != (size_t)p->d) return fail(DMX_ERR_NOT_FINAL, "rounds remain before the final opening");
    HIP_TRY(hipMemcpyAsync(out, p->W_a.p, 16, hipMemcpyDeviceToHost, p->s));
    HIP_TRY(hipStreamSynchronize(p->s));
    return DMX_OK;
}

// ------------------------------------------------------------------------------------------
// verifier (src/protocol/verifier.rs:21-134)
// ------------------------------------------------------------------------------------------
static int log2_exact(size_t n) {
    int l = 0;
    while (((size_t)1 << l) < n) ++l;
    return l;
}

struct dmx_verifier {
    dmx_ctx* ctx = nullptr;
    hipStream_t s = nullptr;
    Jobs jb;
    size_t rounds = 0, round = 0;
    size_t rows = 0, glen = 0, clen = 0;  // current dimensions
    size_t rows0 = 0, n0 = 0;             // the initial generator matrix: rows0 x K x n0
    size_t g_cap = 0, c_cap = 0;
    uint32_t vec_mask = 0, poly_mask = 0;  // rounds whose vector / poly stage ran
    DevBuf store;  // the workspace, unless a batch owns it
    // G0: the generator matrix, never folded round by round; chal: (x, x^-1, y, y^-1) per
    // round; H: the matrix folded by the early rounds' weights (gen_split); gfin: g_d[0]
    Slice meta, G0, C, Ct, rec, chal, H, gfin, flags, opening;
    Transcriber tr;
    std::vector<JobMeta> stage;
    const zq* G0_src = nullptr;  // resident generator matrix shared by every job (see dmx_prover)
    ~dmx_verifier() {
        if (s) (void)hipStreamDestroy(s);
    }
    const zq* G() const { return G0_src ? G0_src : G0.as<zq>(); }
    zq* round_chal(size_t r) const { return chal.as<zq>() + 4 * r; }
    const uint32_t* gseed() const { return reinterpret_cast<const uint32_t*>(meta.as<uint8_t>() + META_GSEED); }
};

// The terminal generator is one weighted sum over the matrix, split at round T: the rounds
// before T fold the matrix into H as soon as their challenges are known, the last (at most
// LATE_ROUNDS) rounds fold H at the final opening. A round folds rows while more than one is left
// and coefficients while the ring length is above one, so rounds 0 .. log2(rows0) - 1 are the
// vector folds and 0 .. log2(n0) - 1 the poly folds.
constexpr size_t LATE_ROUNDS = 5;
struct GenSplit {
    size_t T = 0;
    int log_hi = 0, log_lo = 0, log_n = 0, log_nt = 0;  // early rows folded into one, rows left, ...
    uint32_t vm = 0, pm = 0;                           // the early rounds' vector / poly folds
};
static GenSplit gen_split(const dmx_verifier* v) {
    GenSplit g;
    g.T = v->rounds > LATE_ROUNDS ? v->rounds - LATE_ROUNDS : 0;
    const int lr = log2_exact(v->rows0), ln = log2_exact(v->n0);
    g.log_hi = std::min<int>((int)g.T, lr);
    g.log_lo = lr - g.log_hi;
    g.log_n = ln;
    g.log_nt = ln - std::min<int>((int)g.T, ln);
    g.vm = (1u << g.log_hi) - 1;
    g.pm = (1u << (ln - g.log_nt)) - 1;
    return g;
}

static void verifier_plan(dmx_verifier* v, Layout& L) {
    L.add(v->meta, sizeof(JobMeta));
    L.add(v->G0, v->g_cap * 16);
    L.add(v->C, (size_t)K * v->c_cap * 16);
    L.add(v->Ct, (size_t)K * v->c_cap * 16);
    L.add(v->rec, (size_t)4 * K * v->c_cap * 16);
    L.add(v->chal, std::max<size_t>(v->rounds, 1) * 4 * 16);
    const GenSplit g = gen_split(v);
    L.add(v->H, g.T ? ((size_t)K << (g.log_lo + g.log_nt)) * early_chunks(g.log_n) * 16 : 16);
    L.add(v->gfin, K * 16);
    L.add(v->flags, 256);
    L.add(v->opening, 16);
    v->tr.plan(L, v->c_cap, 0);
}

static void verifier_bind(dmx_verifier* v) {
    v->tr.seeds = v->meta.as<uint8_t>() + META_SEEDS;
    v->tr.reject = reinterpret_cast<const uint32_t*>(v->meta.as<uint8_t>() + META_REJECT);
}

// Generator derivation from g_0_seed (verifier.rs:31-33) for a standalone verifier (a fused job
// takes the rows from the prover's first transform instead). Kernels only.
static int verifier_enqueue_gen(dmx_verifier* v) {
    launch_fill32(v->flags.as<uint32_t>(), 64, 0, v->s, v->jb);
    if (!v->G0_src) launch_gen_generators(v->gseed(), v->rows0 * K, log2_exact(v->n0), v->G0.as<zq>(), v->s, v->jb);
    HIP_TRY(hipGetLastError());
    return DMX_OK;
}

// the statement's file_id and com_0 from device memory (the fused path)
static int verifier_enqueue_statement(dmx_verifier* v, const uint8_t* d_file_id, const zq* d_com0) {
    launch_copy16(v->tr.file_id(), d_file_id, 2, v->s, v->jb);
    launch_copy16(v->C.p, d_com0, (size_t)K * v->clen, v->s, v->jb);
    HIP_TRY(hipGetLastError());
    return DMX_OK;
}

static int verifier_new_object(dmx_ctx* ctx, size_t d, size_t rows, size_t ring_len, uint32_t n,
                               dmx_verifier** out) {
    std::unique_ptr<dmx_verifier> v(new dmx_verifier);
    v->ctx = ctx;
    v->jb = Jobs{n, 0};
    v->rounds = d;
    v->rows = v->rows0 = rows;
    v->glen = v->clen = v->n0 = ring_len;
    v->g_cap = v->rows * K * v->glen;
    v->c_cap = v->clen;
    v->stage.resize(n);
    HIP_TRY(hipStreamCreate(&v->s));
    *out = v.release();
    return DMX_OK;
}

// DamascusVerifier::new (verifier.rs:22-43) from a host statement: generators re-derived from
// the statement's g_0_seed on the device.
static int verifier_make(dmx_ctx* ctx, const dmx_system_params* params, const uint8_t* file_id, uint64_t orig_len,
                         uint64_t d, const zq* com0, size_t com_len, const uint8_t g_seed[32], dmx_verifier** out) {
    if (gpu_disabled_by_env()) return fail(DMX_ERR_NO_DEVICE, "disabled by DAMASCUS_GPU=0");
    if (com_len == 0 || (com_len & (com_len - 1)) || com_len > ((size_t)1 << MAX_LOG_N))
        return fail(DMX_ERR_SHAPE, "statement ring length %zu is not a supported power of two", com_len);
    const int ld = depth_for(orig_len);
    if (ld > 12) return fail(DMX_ERR_ARG, "depth %d exceeds the supported 12", ld);
    HIP_TRY(hipSetDevice(ctx->device));
    dmx_verifier* raw = nullptr;
    int rc = verifier_new_object(ctx, (size_t)d, (size_t)1 << ld, com_len, 1, &raw);
    if (rc) return rc;
    std::unique_ptr<dmx_verifier> v(raw);
    Layout L;
    verifier_plan(v.get(), L);
    rc = L.commit(v->store, 1);
    if (rc) return rc;
    verifier_bind(v.get());
    v->stage[0] = make_meta(orig_len, nullptr, g_seed, params->epoch_seed);
    memcpy(v->stage[0].file_id, file_id, 32);
    HIP_TRY(hipMemcpyAsync(v->meta.p, v->stage.data(), sizeof(JobMeta), hipMemcpyHostToDevice, v->s));
    HIP_TRY(hipMemcpyAsync(v->C.p, com0, (size_t)K * v->clen * 16, hipMemcpyHostToDevice, v->s));
    rc = verifier_enqueue_gen(v.get());
    if (rc) return rc;
    *out = v.release();
    return DMX_OK;
}

extern "C" int dmx_verifier_new(dmx_ctx* ctx, const dmx_system_params* params, const uint8_t* statement, size_t len,
                                dmx_verifier** out) {
    if (!ctx || !params || !statement || !out) return fail(DMX_ERR_ARG, "null argument");
    Statement st;
    int rc = decode_statement(statement, len, st);
    if (rc) return rc;
    dmx_verifier* v = nullptr;
    rc = verifier_make(ctx, params, st.file_id, st.original_len, st.d, st.com0.data(), st.com_len, st.g_seed, &v);
    if (rc) return rc;
    if (hipStreamSynchronize(v->s) != hipSuccess) {
        delete v;
        return fail(DMX_ERR_HIP, "verifier setup failed");
    }
    *out = v;
    return DMX_OK;
}

extern "C" dmx_verifier* dmx_verifier_clone(const dmx_verifier* v) {
    if (!v) return nullptr;
    dmx_verifier* raw = nullptr;
    if (verifier_new_object(v->ctx, v->rounds, v->rows, v->clen, 1, &raw)) return nullptr;
    std::unique_ptr<dmx_verifier> c(raw);
    c->round = v->round;
    c->rows = v->rows;
    c->glen = v->glen;
    c->rows0 = v->rows0;
    c->n0 = v->n0;
    c->g_cap = v->g_cap;
    c->vec_mask = v->vec_mask;
    c->poly_mask = v->poly_mask;
    Layout L;
    verifier_plan(c.get(), L);
    if (L.commit(c->store, 1)) return nullptr;
    verifier_bind(c.get());
    (void)hipStreamSynchronize(v->s);
    if (hipMemcpyAsync(c->G0.p, v->G(), c->g_cap * 16, hipMemcpyDeviceToDevice, c->s) != hipSuccess) return nullptr;
    if (hipMemcpyAsync(c->chal.p, v->chal.p, std::max<size_t>(v->rounds, 1) * 4 * 16, hipMemcpyDeviceToDevice,
                       c->s) != hipSuccess)
        return nullptr;
    if (hipMemcpyAsync(c->C.p, v->C.p, (size_t)K * v->clen * 16, hipMemcpyDeviceToDevice, c->s) != hipSuccess)
        return nullptr;
    if (hipMemcpyAsync(c->meta.p, v->meta.p, sizeof(JobMeta), hipMemcpyDeviceToDevice, c->s) != hipSuccess)
        return nullptr;
    launch_fill32(c->flags.as<uint32_t>(), 64, 0, c->s);
    if (hipStreamSynchronize(c->s) != hipSuccess) return nullptr;
    return c.release();
}

extern "C" void dmx_verifier_free(dmx_verifier* v) { delete v; }
extern "C" size_t dmx_verifier_round(const dmx_verifier* v) { return v ? v->round : 0; }
extern "C" void dmx_verifier_current_dimensions(const dmx_verifier* v, size_t* rows, size_t* ring_len) {
    if (rows) *rows = v->rows;
    if (ring_len) *ring_len = v->clen;
}
extern "C" int dmx_verifier_current_commitment(const dmx_verifier* v, uint64_t* out, size_t cap_words, size_t* ring_len) {
    if (ring_len) *ring_len = v->clen;
    if (!out) return DMX_OK;
    if (cap_words < (size_t)2 * K * v->clen) return fail(DMX_ERR_BUFFER, "commitment buffer too small");
    HIP_TRY(hipMemcpyAsync(out, v->C.p, (size_t)K * v->clen * 16, hipMemcpyDeviceToHost, v->s));
    HIP_TRY(hipStreamSynchronize(v->s));
    return DMX_OK;
}

static bool all_zero(const std::vector<zq>& m) {
    for (const zq& x : m)
        if (!zq_is_zero(x)) return false;
    return true;
}

// A replay round in four steps: the vector challenge x from C and the vector cross terms; the
// update C_t = C + x^-1 L + x R into Ct; the poly challenge y from C_t and the poly cross terms;
// the update C = even(C_t) + y^-1 L + y R, which also advances the round. A round folds rows
// while more than one is left and coefficients while the ring length is above one.
//
// `inverse`: the challenge kernel also computes the inverse. Otherwise the update takes it from
// `hint`, a candidate the verifier accepts only after checking x * hint == 1 (the update computes
// the inverse itself when the check fails), so the result never depends on where the candidate
// came from. The fused prove-and-verify job passes the prover's own inverse of the same
// challenge: the verifier's chain then carries no inversion.
static int verifier_vec_challenge(dmx_verifier* v, const zq* d_l, const zq* d_r, bool inverse) {
    if (v->rows <= 1) return DMX_OK;
    const size_t n = v->clen;
    v->tr.challenge("vec", v->round, v->C.as<zq>(), n, d_l, n, d_r, n, v->round_chal(v->round),
                    v->flags.as<uint32_t>(), v->s, v->jb, inverse);
    v->vec_mask |= 1u << v->round;
    HIP_TRY(hipGetLastError());
    return DMX_OK;
}

static int verifier_vec_update(dmx_verifier* v, const zq* d_l, const zq* d_r, const zq* hint) {
    if (v->rows <= 1) return DMX_OK;
    launch_com_update(v->C.as<zq>(), d_l, d_r, v->round_chal(v->round), 0, log2_exact(v->clen), v->Ct.as<zq>(), v->s,
                      v->jb, hint);
    HIP_TRY(hipGetLastError());
    return DMX_OK;
}

static int verifier_poly_challenge(dmx_verifier* v, const zq* d_lp, const zq* d_rp, bool inverse) {
    if (v->glen <= 1) return DMX_OK;
    const size_t n = v->clen;
    const zq* Ct = v->rows > 1 ? v->Ct.as<zq>() : v->C.as<zq>();
    v->tr.challenge("poly", v->round, Ct, n, d_lp, n / 2, d_rp, n / 2, v->round_chal(v->round) + 2,
                    v->flags.as<uint32_t>(), v->s, v->jb, inverse);
    v->poly_mask |= 1u << v->round;
    HIP_TRY(hipGetLastError());
    return DMX_OK;
}

static int verifier_poly_update(dmx_verifier* v, const zq* d_lp, const zq* d_rp, const zq* hint) {
    const size_t n = v->clen;
    const bool vec = v->rows > 1, poly = v->glen > 1;
    const zq* Ct = vec ? v->Ct.as<zq>() : v->C.as<zq>();
    size_t nclen = n, nglen = v->glen;
    if (poly) {
        launch_com_update(Ct, d_lp, d_rp, v->round_chal(v->round) + 2, 1, log2_exact(n / 2), v->C.as<zq>(), v->s,
                          v->jb, hint);
        nclen = n / 2;
        nglen = v->glen / 2;
    } else if (vec) {
        launch_copy16(v->C.p, v->Ct.p, (size_t)K * n, v->s, v->jb);
    }
    HIP_TRY(hipGetLastError());
    if (vec) v->rows /= 2;
    v->clen = nclen;
    v->glen = nglen;
    v->round += 1;
    return DMX_OK;
}

// One replay round on the device (verifier.rs:57-120) with the round's four cross-term
// modules already in device memory and their shapes already checked. No host sync. The
// generator folds of the round (verifier.rs:75, 102) are not materialised: the round's challenge
// inverses are kept and the terminal generator is formed once, at the final opening.
static int verifier_round_launch(dmx_verifier* v, const zq* d_l, const zq* d_r, const zq* d_lp, const zq* d_rp) {
    int rc = verifier_vec_challenge(v, d_l, d_r, true);
    if (!rc) rc = verifier_vec_update(v, d_l, d_r, nullptr);
    if (!rc) rc = verifier_poly_challenge(v, d_lp, d_rp, true);
    return rc ? rc : verifier_poly_update(v, d_lp, d_rp, nullptr);
}

// update_commitment (verifier.rs:57-120); every check runs before any device work, so a
// rejected record leaves the state unchanged, as in the reference.
extern "C" int dmx_verifier_update_commitment(dmx_verifier* v, const uint8_t* record, size_t len) {
    if (!v || !record) return fail(DMX_ERR_ARG, "null argument");
    Record rec;
    int rc = decode_record(record, len, rec);
    if (rc) return rc;
    if (rec.round != v->round) return fail(DMX_ERR_ROUND, "round mismatch in round record");
    const size_t n = v->clen;
    const bool vec = v->rows > 1, poly = v->glen > 1;
    if (vec) {
        if (rec.len[0] != n || rec.len[1] != n) return fail(DMX_ERR_SHAPE, "module ring length mismatch");
    } else {
        if (rec.len[0] != n || rec.len[1] != n || !all_zero(rec.m[0]) || !all_zero(rec.m[1]))
            return fail(DMX_ERR_ZERO_TERMS, "scalar vector stage must emit zero vec cross-terms");
    }
    if (poly) {
        if (n != v->glen || n < 2) return fail(DMX_ERR_SHAPE, "module ring length mismatch");
        if (rec.len[2] != n / 2 || rec.len[3] != n / 2) return fail(DMX_ERR_SHAPE, "module ring length mismatch");
    } else {
        if (rec.len[2] != 1 || rec.len[3] != 1 || !all_zero(rec.m[2]) || !all_zero(rec.m[3]))
            return fail(DMX_ERR_ZERO_TERMS, "scalar stage must emit zero poly cross-terms");
    }
    hipStream_t s = v->s;
    zq* d_l = v->rec.as<zq>();
    zq* d_r = d_l + (size_t)K * n;
    zq* d_lp = d_l + (size_t)2 * K * n;
    zq* d_rp = d_lp + (size_t)K * (n / 2 ? n / 2 : 1);
    if (vec) {
        HIP_TRY(hipMemcpyAsync(d_l, rec.m[0].data(), (size_t)K * n * 16, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_r, rec.m[1].data(), (size_t)K * n * 16, hipMemcpyHostToDevice, s));
    }
    if (poly) {
        HIP_TRY(hipMemcpyAsync(d_lp, rec.m[2].data(), (size_t)K * (n / 2) * 16, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_rp, rec.m[3].data(), (size_t)K * (n / 2) * 16, hipMemcpyHostToDevice, s));
    }
    // state (rows/lens/round) only advances after the device work is known good
    const size_t rows0 = v->rows, clen0 = v->clen, glen0 = v->glen, round0 = v->round;
    const uint32_t vmask0 = v->vec_mask, pmask0 = v->poly_mask;
    rc = verifier_round_launch(v, d_l, d_r, d_lp, d_rp);
    uint32_t f = 0;
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = fail(DMX_ERR_HIP, "stream sync failed");
    if (!rc && hipMemcpy(&f, v->flags.p, 4, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(DMX_ERR_HIP, "flags");
    if (!rc && f) rc = fail(DMX_ERR_HIP, "challenge resampling exhausted");
    if (rc) {
        v->rows = rows0;
        v->clen = clen0;
        v->glen = glen0;
        v->round = round0;
        v->vec_mask = vmask0;
        v->poly_mask = pmask0;
    }
    return rc;
}

// The early fold of the terminal generator (k_gen_contract_early) on stream `s`; needs the
// challenges of the rounds before gen_split's T and the generator matrix.
static int verifier_enqueue_early(dmx_verifier* v, hipStream_t s) {
    const GenSplit g = gen_split(v);
    if (!g.T) return DMX_OK;
    if ((v->vec_mask & ((1u << g.T) - 1)) != g.vm || (v->poly_mask & ((1u << g.T) - 1)) != g.pm)
        return fail(DMX_ERR_NOT_FINAL, "the early rounds have not been replayed");
    if ((1u << g.log_hi) > EARLY_MAX_HI || (1u << g.log_nt) > EARLY_MAX_NT)
        return fail(DMX_ERR_SHAPE, "generator matrix shape outside the early fold's limits");
    launch_gen_contract_early(v->G(), v->G0_src != nullptr, v->chal.as<zq>(), g.vm, g.pm, g.log_hi, g.log_lo,
                              g.log_n, g.log_nt, v->H.as<zq>(), s, v->jb);
    HIP_TRY(hipGetLastError());
    return DMX_OK;
}

// The final check's result bit lives in its own flag word (word 0 collects the challenge kernels'
// errors), cleared here unless `fresh_flags` (a fused job's flags were zeroed at its start).
constexpr size_t FINAL_WORD = 1;

// `early_done`: the early fold was already enqueued (a fused job runs it on its side stream)
static int verifier_final_launch(dmx_verifier* v, const zq* d_opening, bool early_done = false,
                                 bool fresh_flags = false) {
    if (v->round != v->rounds) return fail(DMX_ERR_NOT_FINAL, "cannot verify final opening before all rounds replay");
    if (v->rows != 1) return fail(DMX_ERR_NOT_FINAL, "terminal generator family must have length one");
    if (v->glen != v->clen) return fail(DMX_ERR_OPENING, "final opening mismatch");
    if (v->glen != 1 || (size_t)__builtin_popcount(v->vec_mask) != (size_t)log2_exact(v->rows0) ||
        (size_t)__builtin_popcount(v->poly_mask) != (size_t)log2_exact(v->n0))
        return fail(DMX_ERR_NOT_FINAL, "the generator family has not been folded to scalars");
    const GenSplit g = gen_split(v);
    if (g.log_lo > 5 || g.log_nt > 5) return fail(DMX_ERR_SHAPE, "generator matrix shape outside the late fold's limits");
    if (!early_done) {
        int rc = verifier_enqueue_early(v, v->s);
        if (rc) return rc;
    }
    const zq* H = g.T ? v->H.as<zq>() : v->G();
    const bool h_shared = !g.T && v->G0_src != nullptr;
    if (!fresh_flags) launch_fill32(v->flags.as<uint32_t>() + FINAL_WORD, 1, 0, v->s, v->jb);
    launch_gen_contract_late(H, h_shared, v->chal.as<zq>() + 4 * g.T, v->vec_mask >> g.T, v->poly_mask >> g.T, g.log_lo,
                             g.log_nt, v->C.as<zq>(), d_opening, v->flags.as<uint32_t>() + FINAL_WORD, 1,
                             v->gfin.as<zq>(), v->s, v->jb);
    HIP_TRY(hipGetLastError());
    return DMX_OK;
}

static int verifier_final_result(dmx_verifier* v) {
    uint32_t f = 0;
    HIP_TRY(hipMemcpyAsync(&f, v->flags.as<uint32_t>() + FINAL_WORD, 4, hipMemcpyDeviceToHost, v->s));
    HIP_TRY(hipStreamSynchronize(v->s));
    if (f) return fail(DMX_ERR_OPENING, "final opening mismatch");
    return DMX_OK;
}

// verify_final_opening (verifier.rs:122-134): g_d[0] * opening == C_d
extern "C" int dmx_verifier_verify_final_opening(const dmx_verifier* v_, const uint64_t opening[2]) {
    dmx_verifier* v = const_cast<dmx_verifier*>(v_);
    if (!v || !opening) return fail(DMX_ERR_ARG, "null argument");
    uint32_t w[8] = {(uint32_t)opening[0], (uint32_t)(opening[0] >> 32), (uint32_t)opening[1],
                     (uint32_t)(opening[1] >> 32), 0, 0, 0, 0};
    zq o = zq_reduce256(w);  // an Fp is always canonical
    HIP_TRY(hipMemcpyAsync(v->opening.p, &o, 16, hipMemcpyHostToDevice, v->s));
    int rc = verifier_final_launch(v, v->opening.as<zq>());
    if (rc) return rc;
    return verifier_final_result(v);
}

// ------------------------------------------------------------------------------------------
// whole jobs
// ------------------------------------------------------------------------------------------
extern "C" int dmx_prove(dmx_ctx* ctx, const uint8_t* file, size_t len, const dmx_system_params* params,
                         const dmx_runtime_config* cfg, uint8_t* st_out, size_t st_cap, size_t* st_len,
                         uint8_t* rec_out, size_t rec_cap, size_t* rec_len, uint64_t opening_out[2]) {
    dmx_prover* p = nullptr;
    int rc = dmx_prover_new(ctx, file, len, params, cfg, &p);
    if (rc) return rc;
    std::unique_ptr<dmx_prover> guard(p);
    rc = dmx_prover_statement(p, st_out, st_cap, st_len);
    size_t pos = 0;
    for (size_t j = 0; !rc && j < (size_t)p->d; ++j) {
        size_t got = 0;
        const size_t need = dmx_record_size(p->N >> j, p->N >> (j + 1));
        uint8_t* dst = (rec_out && pos + need <= rec_cap) ? rec_out + pos : nullptr;
        rc = dmx_prover_fold_round(p, j, dst, need, &got, nullptr);
        pos += need;
    }
    if (rec_len) *rec_len = pos;
    if (!rc && rec_out && pos > rec_cap) rc = fail(DMX_ERR_BUFFER, "records buffer too small");
    if (!rc && opening_out) rc = dmx_prover_final_opening(p, opening_out);
    return rc;
}

extern "C" int dmx_verify(dmx_ctx* ctx, const dmx_system_params* params, const uint8_t* st, size_t st_len,
                          const uint8_t* recs, size_t rec_len, const uint64_t opening[2]) {
    dmx_verifier* v = nullptr;
    int rc = dmx_verifier_new(ctx, params, st, st_len, &v);
    if (rc) return rc;
    std::unique_ptr<dmx_verifier> guard(v);
    size_t pos = 0;
    while (!rc && pos < rec_len) {
        Record r;
        rc = decode_record(recs + pos, rec_len - pos, r);
        if (rc) break;
        size_t used = 8;
        for (int i = 0; i < 4; ++i) used += module_bytes(r.len[i]);
        rc = dmx_verifier_update_commitment(v, recs + pos, used);
        pos += used;
    }
    if (!rc) rc = dmx_verifier_verify_final_opening(v, opening);
    return rc;
}

// ------------------------------------------------------------------------------------------
// Batched device jobs: prove + replay of n files of one depth side by side.
// The prover and the verifier each run on their own stream over one shared workspace (n job
// blocks); every launch covers all n jobs. The verifier's generator derivation starts with the
// job (it needs only the seed), and from then on it consumes the prover's device-resident
// records round by round, recomputing every challenge itself. File lengths, seeds and source
// addresses live in the per-job headers, so one captured hipGraph per (depth, n) replays for
// any files of that depth: one launch instead of ~400 per batch.
// ------------------------------------------------------------------------------------------
// Resident generator matrices per (context, g_0_seed, depth): the Z_q rows g[i][k] (sis.rs:212-229)
// and their forward transforms, derived once and read by every later job under that seed. A
// served stream of requests under one SystemParams then spends no work on generator expansion.
// DAMASCUS_GEN_CACHE sets the number of matrices kept (default 8; 0 disables).
struct GenEntry {
    dmx_ctx* ctx = nullptr;
    int d = 0;
    uint8_t seed[32];
    uint32_t kw[8];
    DevBuf G, Ghat, key;
    hipEvent_t ready = nullptr;
    int refs = 0;  // batch graphs captured against this matrix
    uint64_t last_use = 0;
    ~GenEntry() {
        if (ready) (void)hipEventDestroy(ready);
    }
};
static std::mutex g_gen_mu;
static std::vector<std::unique_ptr<GenEntry>> g_gen;
static uint64_t g_gen_clock = 0;

static size_t gen_cache_slots() {
    const char* e = getenv("DAMASCUS_GEN_CACHE");
    return e ? (size_t)atoi(e) : 8;
}

// Admission: a (seed, depth) gets a matrix on its second miss. A seed met once, such as a request
// that carries its own generator seed, expands its generators inside its own job: no allocation,
// no eviction (whose free would wait for the device) and no extra launch. The last GEN_SEEN_MAX
// misses are remembered, oldest first.
struct GenSeen {
    const dmx_ctx* ctx;
    int d;
    uint8_t seed[32];
};
static std::vector<GenSeen> g_gen_seen;
constexpr size_t GEN_SEEN_MAX = 256;

// true when (ctx, seed, d) missed before (and forgets it); records the miss otherwise. Called
// with g_gen_mu held.
static bool gen_missed_before(const dmx_ctx* ctx, const uint8_t seed[32], int d) {
    for (auto it = g_gen_seen.begin(); it != g_gen_seen.end(); ++it)
        if (it->ctx == ctx && it->d == d && !memcmp(it->seed, seed, 32)) {
            g_gen_seen.erase(it);
            return true;
        }
    if (g_gen_seen.size() >= GEN_SEEN_MAX) g_gen_seen.erase(g_gen_seen.begin());
    GenSeen m{ctx, d, {}};
    memcpy(m.seed, seed, 32);
    g_gen_seen.push_back(m);
    return false;
}

// The matrix for (seed, d), derived on stream s when new; nullptr when the cache is off, the
// pair is met for the first time, or the cache is full of matrices still in use. The caller
// waits on ->ready before reading it.
static int gen_acquire(dmx_ctx* ctx, const uint8_t seed[32], int d, hipStream_t s, GenEntry** out) {
    *out = nullptr;
    const size_t slots = gen_cache_slots();
    if (!slots) return DMX_OK;
    std::lock_guard<std::mutex> g(g_gen_mu);
    for (auto& e : g_gen)
        if (e->ctx == ctx && e->d == d && !memcmp(e->seed, seed, 32)) {
            e->refs++;
            e->last_use = ++g_gen_clock;
            *out = e.get();
            return DMX_OK;
        }
    if (!gen_missed_before(ctx, seed, d)) return DMX_OK;  // first request: derive per job
    while (g_gen.size() >= slots) {
        size_t victim = SIZE_MAX;
        for (size_t i = 0; i < g_gen.size(); ++i)
            if (!g_gen[i]->refs && (victim == SIZE_MAX || g_gen[i]->last_use < g_gen[victim]->last_use)) victim = i;
        if (victim == SIZE_MAX) return DMX_OK;  // every matrix is in use: derive per job instead
        g_gen.erase(g_gen.begin() + victim);
    }
    std::unique_ptr<GenEntry> e(new GenEntry);
    e->ctx = ctx;
    e->d = d;
    memcpy(e->seed, seed, 32);
    gseed_words(seed, e->kw);
    const size_t N = (size_t)1 << d;
    int rc = e->G.alloc(K * N * N * 16);
    if (!rc) rc = e->Ghat.alloc(K * N * N * NPRIMES * 4);
    if (!rc) rc = e->key.alloc(32);
    if (rc) {
        (void)hipGetLastError();
        return DMX_OK;  // no room: derive per job
    }
    HIP_TRY(hipEventCreateWithFlags(&e->ready, hipEventDisableTiming));
    HIP_TRY(hipMemcpyAsync(e->key.p, e->kw, 32, hipMemcpyHostToDevice, s));
    RowSrc gen{};
    gen.mode = ROW_GEN;
    gen.dst = e->G.as<zq>();
    gen.gseed = e->key.as<uint32_t>();
    launch_ntt_fwd_rows(gen, K * N, d, e->Ghat.as<uint32_t>(), ctx->tw, s, Jobs{}, ctx->fwd_variant[d]);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(e->ready, s));
    e->refs = 1;
    e->last_use = ++g_gen_clock;
    *out = e.get();
    g_gen.push_back(std::move(e));
    return DMX_OK;
}

static void gen_release(GenEntry* e) {
    std::lock_guard<std::mutex> g(g_gen_mu);
    e->refs--;
}

struct BatchJob {
    dmx_ctx* ctx = nullptr;
    int d = 0;
    uint32_t n = 0;
    DevBuf store;
    std::unique_ptr<dmx_prover> p;
    std::unique_ptr<dmx_verifier> v;
    std::vector<hipEvent_t> evs;
    hipEvent_t done = nullptr;
    // the early fold of the verifier's generator matrix runs on its own stream while the proof's
    // last rounds are in flight; the replay needs the folded matrix only at the final opening
    hipStream_t gs = nullptr;
    hipEvent_t side_done = nullptr;
    // Captured job graphs by [stream shape][generators]. Shape 0: three streams (prover,
    // verifier, side), so one job's verifier overlaps its prover (the plain and ragged-batch
    // paths). Shape 1: every step on the prover's stream, a single chain that the runtime
    // launches in tens of microseconds and that needs no waits across hardware queues (the
    // server, whose concurrency comes from several batches in flight). Generators 0: the jobs
    // expand their own; 1: they read the resident matrix `gen`.
    hipGraph_t graphs[2][2] = {};
    hipGraphExec_t execs[2][2] = {};
    uint64_t graph_version = 0;  // the context's fwd_version the graphs were captured with
    bool one_stream = false;     // the shape of the next enqueue
    GenEntry* gen = nullptr;
    // the persistent tail (k_tail) of every job: stage descriptors for the prover's stages of
    // degree <= 32 and the verifier's replay of them, built once from the workspace layout
    DevBuf tail_desc;
    size_t tail_lds = 0;
    uint32_t tail_msg = 0;
    size_t tail_round = 0;  // the first round of the tail (0: no tail)
    bool busy = false;
    uint64_t last_use = 0;
    // pinned staging for files handed over in host memory: the host copies each file in, and
    // the upload is an asynchronous DMA on the batch's stream, overlapping other batches' work
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    // pinned staging for the results (flags, file ids, com_0, records, openings): DMA copies,
    // not the runtime's pageable path through a bounce buffer
    void* pinned_out = nullptr;
    size_t pinned_out_bytes = 0;
    void drop_graph(int shape, int r) {
        if (execs[shape][r]) (void)hipGraphExecDestroy(execs[shape][r]);
        if (graphs[shape][r]) (void)hipGraphDestroy(graphs[shape][r]);
        execs[shape][r] = nullptr;
        graphs[shape][r] = nullptr;
    }
    void drop_graphs() {
        for (int shape = 0; shape < 2; ++shape)
            for (int r = 0; r < 2; ++r) drop_graph(shape, r);
    }
    void drop_resident() {
        drop_graph(0, 1);
        drop_graph(1, 1);
        if (gen) gen_release(gen);
        gen = nullptr;
    }
    ~BatchJob() {
        if (pinned) (void)hipHostFree(pinned);  // idle batch: its uploads have completed
        if (pinned_out) (void)hipHostFree(pinned_out);
        drop_resident();
        drop_graphs();
        for (auto e : evs) (void)hipEventDestroy(e);
        if (done) (void)hipEventDestroy(done);
        if (side_done) (void)hipEventDestroy(side_done);
        if (gs) (void)hipStreamDestroy(gs);
    }
    size_t bytes() const { return store.bytes; }
};

static size_t max_file_len(int d) { return (size_t)BYTES_PER_COEFF << (2 * d); }

// The tail of a batch (k_tail): the prover's stages from the first round of degree <= TAIL_MAX_M
// on and the verifier's replay of those rounds, one launch. Above that degree a stage's work is
// large enough to want more than the tail's one workgroup per job, and those stages keep their
// per-stage launches. DAMASCUS_TAIL=0 keeps every stage on them (for comparisons).
constexpr size_t TAIL_CTR_WORD = 33;  // flags word: the tail's stage counter (zeroed with the flags)
static bool tail_enabled() {
    static const bool on = [] {
        const char* e = getenv("DAMASCUS_TAIL");
        return !(e && strcmp(e, "0") == 0);
    }();
    return on;
}

static int tail_build(BatchJob* J) {
    dmx_prover* p = J->p.get();
    dmx_verifier* v = J->v.get();
    const size_t d = (size_t)p->d;
    size_t T0 = 0;
    while ((p->N >> T0) > TAIL_MAX_M) ++T0;
    J->tail_round = 0;
    if (T0 == 0 || 2 * (d - T0) > (size_t)TAIL_MAX_STAGES || T0 < gen_split(v).T) return DMX_OK;
    std::unique_ptr<TailDesc> td(new TailDesc());
    memset(td.get(), 0, sizeof(TailDesc));
    td->nstages = (uint32_t)(2 * (d - T0));
    td->ctr = p->flags.as<uint32_t>() + TAIL_CTR_WORD;
    td->vflags = v->flags.as<uint32_t>();
    for (size_t j = T0; j < d; ++j) {
        const size_t s = 2 * (j - T0), m = p->m_at(j), h = m / 2;
        td->st[s] = small_stage_of(p, j, 0, vec_stage_args(p, j));
        td->st[s + 1] = small_stage_of(p, j, 1, poly_stage_args(p, j));
        td->ptx[s] = p->tr.message("vec", j, p->C.as<zq>(), m, p->rec_lvec(j), m, p->rec_rvec(j), m);
        td->ptx[s + 1] = p->tr.message("poly", j, p->Ct.as<zq>(), m, p->rec_lpoly(j), h, p->rec_rpoly(j), h);
        TailVerStep& a = td->vs[s];  // vector challenge x, then C_t = C + x^-1 L + x R
        a.tx = v->tr.message("vec", j, v->C.as<zq>(), m, p->rec_lvec(j), m, p->rec_rvec(j), m);
        a.c = v->C.as<zq>();
        a.l = p->rec_lvec(j);
        a.r = p->rec_rvec(j);
        a.scal = v->round_chal(j);
        a.out = v->Ct.as<zq>();
        a.hint = p->sc(4 * j);
        a.poly = 0;
        a.log_m = log2_exact(m);
        TailVerStep& b = td->vs[s + 1];  // poly challenge y, then C = even(C_t) + y^-1 L + y R
        b.tx = v->tr.message("poly", j, v->Ct.as<zq>(), m, p->rec_lpoly(j), h, p->rec_rpoly(j), h);
        b.c = v->Ct.as<zq>();
        b.l = p->rec_lpoly(j);
        b.r = p->rec_rpoly(j);
        b.scal = v->round_chal(j) + 2;
        b.out = v->C.as<zq>();
        b.hint = p->sc(4 * j + 2);
        b.poly = 1;
        b.log_m = log2_exact(h);
    }
    int rc = J->tail_desc.alloc(sizeof(TailDesc));
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(J->tail_desc.p, td.get(), sizeof(TailDesc), hipMemcpyHostToDevice, p->s));
    HIP_TRY(hipStreamSynchronize(p->s));
    J->tail_lds = tail_lds_bytes(*td, &J->tail_msg);
    HIP_TRY(tail_set_lds_limit(J->tail_lds));
    J->tail_round = T0;
    return DMX_OK;
}

static int batch_create(dmx_ctx* ctx, int d, uint32_t n, BatchJob** out) {
    std::unique_ptr<BatchJob> J(new BatchJob);
    J->ctx = ctx;
    J->d = d;
    J->n = n;
    dmx_prover* p = nullptr;
    int rc = prover_new_object(ctx, d, n, &p);
    if (rc) return rc;
    J->p.reset(p);
    p->hash_cap = max_file_len(d);
    dmx_verifier* v = nullptr;
    rc = verifier_new_object(ctx, (size_t)d, p->N, p->N, n, &v);
    if (rc) return rc;
    J->v.reset(v);
    Layout L;
    prover_plan(p, L, p->hash_cap);
    verifier_plan(v, L);
    rc = L.commit(J->store, n);
    if (rc) return rc;
    p->jb = v->jb = Jobs{n, L.stride()};
    prover_bind(p);
    verifier_bind(v);
    J->evs.resize(2 * (size_t)d + 5);
    for (auto& e : J->evs) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&J->done, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&J->side_done, hipEventDisableTiming));
    HIP_TRY(hipStreamCreate(&J->gs));
    rc = tail_build(J.get());
    if (rc) return rc;
    *out = J.release();
    return DMX_OK;
}

// all device work of one batch, enqueued on p->s (forking to v->s); no host syncs
static int batch_enqueue_streams(BatchJob* J);

// The job on its streams, or with J->one_stream every step on the prover's stream (the
// verifier's and the side stream's work then follows the prover's in order; every event wait
// becomes a no-op).
static int batch_enqueue(BatchJob* J) {
    if (!J->one_stream) return batch_enqueue_streams(J);
    const hipStream_t vs = J->v->s, gs = J->gs;
    J->v->s = J->p->s;
    J->gs = J->p->s;
    const int rc = batch_enqueue_streams(J);
    J->v->s = vs;
    J->gs = gs;
    return rc;
}

static int batch_enqueue_streams(BatchJob* J) {
    dmx_prover* p = J->p.get();
    dmx_verifier* v = J->v.get();
    p->round = 0;
    v->round = 0;
    v->rows = p->N;
    v->glen = v->clen = p->N;
    v->vec_mask = v->poly_mask = 0;
    launch_copy_in(p->meta.as<uint64_t>(), p->file.as<uint8_t>(), p->hash_cap, p->s, p->jb);
    int rc = prover_init_file(p);
    if (rc) return rc;
    // fork the verifier stream
    HIP_TRY(hipEventRecord(J->evs[0], p->s));
    HIP_TRY(hipStreamWaitEvent(v->s, J->evs[0], 0));
    launch_fill32(v->flags.as<uint32_t>(), 64, 0, v->s, v->jb);
    // The generator matrix is a function of the statement's seed alone (verifier.rs:31-33).
    // Without a resident copy, the prover's first transform, which expands every generator row
    // for com_0 anyway, also stores the rows into the verifier's matrix: one expansion per job.
    const bool own_gen = v->G0_src == nullptr;
    rc = prover_init_commit(p, own_gen ? v->G0.as<zq>() : nullptr);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(J->evs[1], p->s));
    HIP_TRY(hipStreamWaitEvent(v->s, J->evs[1], 0));
    rc = verifier_enqueue_statement(v, p->tr.file_id(), p->com0.as<zq>());
    if (rc) return rc;
    // The verifier follows the prover stage by stage. Its vector challenge of round j needs the
    // prover's round j-1 (the vector cross terms of round j); the update with x^-1 and the poly
    // challenge need the prover's vector stage of round j (the poly cross terms, and the inverse
    // it offers as a hint); the poly update needs the prover's poly stage, so it runs at the start
    // of the next round. The side stream gs folds the generator matrix by the early rounds'
    // challenges once the verifier has them all (at the start of round T), so the final opening
    // only folds what is left.
    const size_t d = (size_t)p->d;
    const size_t T = gen_split(v).T;
    // The inverse candidates offered to the verifier: the prover's x^-1 and y^-1 slots.
    // DAMASCUS_VERIFIER_HINT=wrong offers x and y themselves (a test of the verifier's own
    // inversion when the check x * candidate == 1 fails).
    const char* he = getenv("DAMASCUS_VERIFIER_HINT");
    const size_t hint_off = he && strcmp(he, "wrong") == 0 ? 0 : 1;
    // The last stages (degree <= TAIL_MAX_M) and their replay run as one persistent launch (k_tail)
    // from round T0 on: the loop stops there, and the launch follows the verifier's poly update of
    // the round before it.
    const size_t T0 = (tail_enabled() && small_stages_enabled()) ? J->tail_round : 0;
    const size_t jend = T0 ? T0 : d;
    for (size_t j = 0; j < jend; ++j) {
        rc = prover_round(p, j, J->evs[2 + d + j]);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(J->evs[2 + j], p->s));
        if (j > 0) {
            rc = verifier_poly_update(v, p->rec_lpoly(j - 1), p->rec_rpoly(j - 1), p->sc(4 * (j - 1) + 2 + hint_off));
            if (rc) return rc;
        }
        if (T && j == T) {
            HIP_TRY(hipEventRecord(J->evs[2 + 2 * d], v->s));
            HIP_TRY(hipStreamWaitEvent(J->gs, J->evs[2 + 2 * d], 0));
            rc = verifier_enqueue_early(v, J->gs);
            if (rc) return rc;
        }
        rc = verifier_vec_challenge(v, p->rec_lvec(j), p->rec_rvec(j), false);
        if (rc) return rc;
        HIP_TRY(hipStreamWaitEvent(v->s, J->evs[2 + d + j], 0));
        rc = verifier_vec_update(v, p->rec_lvec(j), p->rec_rvec(j), p->sc(4 * j + hint_off));
        if (!rc) rc = verifier_poly_challenge(v, p->rec_lpoly(j), p->rec_rpoly(j), false);
        if (rc) return rc;
        HIP_TRY(hipStreamWaitEvent(v->s, J->evs[2 + j], 0));
    }
    if (T0) {
        rc = verifier_poly_update(v, p->rec_lpoly(T0 - 1), p->rec_rpoly(T0 - 1), p->sc(4 * (T0 - 1) + 2 + hint_off));
        if (rc) return rc;
        // the tail's verifier workgroup continues from the verifier's state after round T0 - 1
        hipEvent_t vready = J->evs[2 + 2 * d + 1];
        HIP_TRY(hipEventRecord(vready, v->s));
        HIP_TRY(hipStreamWaitEvent(p->s, vready, 0));
        launch_tail(J->tail_desc.as<TailDesc>(), J->tail_lds, J->tail_msg, 1, (int)hint_off, p->ctx->tw, p->s,
                    p->jb);
        HIP_TRY(hipGetLastError());
        if (T && T == T0) {  // the early generator fold needs the rounds before T only: it overlaps the tail
            HIP_TRY(hipStreamWaitEvent(J->gs, vready, 0));
            rc = verifier_enqueue_early(v, J->gs);
            if (rc) return rc;
        }  // (for T < T0 the loop enqueued it at round T)
        for (size_t j = T0; j < d; ++j) {  // what the replayed rounds leave on the host side
            v->vec_mask |= 1u << j;
            v->poly_mask |= 1u << j;
        }
        v->rows = 1;
        v->clen = v->glen = 1;
        v->round = d;
        p->round = d;
        if (T) {
            HIP_TRY(hipEventRecord(J->side_done, J->gs));
            HIP_TRY(hipStreamWaitEvent(p->s, J->side_done, 0));
        }
        const char* oe = getenv("DAMASCUS_TEST_OPENING_OFFSET");
        const size_t open_off = oe ? std::min<size_t>(strtoull(oe, nullptr, 10), p->N - 1) : 0;
        const hipStream_t vs = v->s;  // the final opening follows the tail on the prover's stream
        v->s = p->s;
        rc = verifier_final_launch(v, p->W_a.as<zq>() + open_off, T != 0, true);
        v->s = vs;
        if (rc) return rc;
        launch_result_head(p->out_head.as<uint32_t>(), p->flags.as<uint32_t>(), v->flags.as<uint32_t>(),
                           p->tr.file_id(), p->W_a.as<zq>(), p->s, p->jb);
        HIP_TRY(hipGetLastError());
        return DMX_OK;
    }
    rc = verifier_poly_update(v, p->rec_lpoly(d - 1), p->rec_rpoly(d - 1), p->sc(4 * (d - 1) + 2 + hint_off));
    if (rc) return rc;
    p->round = (size_t)p->d;
    if (T) {
        HIP_TRY(hipEventRecord(J->side_done, J->gs));
        HIP_TRY(hipStreamWaitEvent(v->s, J->side_done, 0));
    }
    // m* = message[0].coeffs[0] (prover.rs:133-138)
    // DAMASCUS_TEST_OPENING_OFFSET=k checks the final opening against witness coefficient k
    // instead of 0 (a test that the fused job's verifier rejects a wrong opening)
    const char* oe = getenv("DAMASCUS_TEST_OPENING_OFFSET");
    const size_t open_off = oe ? std::min<size_t>(strtoull(oe, nullptr, 10), p->N - 1) : 0;
    rc = verifier_final_launch(v, p->W_a.as<zq>() + open_off, T != 0, true);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(J->evs[J->evs.size() - 1], v->s));
    HIP_TRY(hipStreamWaitEvent(p->s, J->evs[J->evs.size() - 1], 0));  // join
    // the job's flags, file_id and m* next to com_0 and the records: one copy brings a result home
    launch_result_head(p->out_head.as<uint32_t>(), p->flags.as<uint32_t>(), v->flags.as<uint32_t>(), p->tr.file_id(),
                       p->W_a.as<zq>(), p->s, p->jb);
    HIP_TRY(hipGetLastError());
    return DMX_OK;
}

static bool graphs_enabled() {
    const char* e = getenv("DAMASCUS_GRAPH");
    return !(e && strcmp(e, "0") == 0);
}

// A file handed over as a device address is read by the copy-in kernel: it must lie inside one
// allocation the device can read. One handed over as a host address is copied by the host into
// pinned staging: it must not be device memory. Checked before any work is enqueued, so a
// mismatched pointer is an argument error instead of a GPU memory fault or a host crash.
static int check_device_file(const void* p, uint64_t len, size_t i) {
    if (!len) return DMX_OK;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return fail(DMX_ERR_ARG, "job %zu: the file is not device memory (files are device addresses)", i);
    }
    if (a.type == hipMemoryTypeUnregistered)
        return fail(DMX_ERR_ARG, "job %zu: the file is not device memory (files are device addresses)", i);
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(DMX_ERR_ARG, "job %zu: the file is not inside a device allocation", i);
    }
    const uintptr_t b = reinterpret_cast<uintptr_t>(base), q = reinterpret_cast<uintptr_t>(p);
    if (q < b || q - b > size || len > size - (q - b))
        return fail(DMX_ERR_ARG, "job %zu: %llu file bytes run past the end of their device allocation", i,
                    (unsigned long long)len);
    return DMX_OK;
}
static int check_host_file(const void* p, uint64_t len, size_t i) {
    if (!len) return DMX_OK;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {  // unknown to HIP: pageable host memory
        (void)hipGetLastError();
        return DMX_OK;
    }
    if (a.type == hipMemoryTypeDevice)
        return fail(DMX_ERR_ARG, "job %zu: the file is device memory, but the files are host addresses", i);
    return DMX_OK;
}
static int check_file(const void* p, uint64_t len, bool on_host, size_t i) {
    return on_host ? check_host_file(p, len, i) : check_device_file(p, len, i);
}

// Writes the job headers (lengths, copy-in sources, seeds) and launches the batch. Files are
// device addresses (copied in by the first kernel) or host addresses (copied here).
struct JobIn {
    const void* file;
    uint64_t len;
    const dmx_system_params* params;
};

// A host file into pinned staging. One core copies some 6-8 GB/s, so a large file is split
// over up to eight threads; its upload then starts (asynchronously) while the next file copies.
static void staged_copy(uint8_t* dst, const uint8_t* src, size_t len) {
    constexpr size_t PIECE = (size_t)2 << 20;
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>(std::min<size_t>(8, hw), (len + PIECE - 1) / PIECE);
    if (nt <= 1) {
        memcpy(dst, src, len);
        return;
    }
    const size_t per = (len + nt - 1) / nt;
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    size_t t = 1;
    try {
        for (; t < nt; ++t) {
            const size_t off = t * per, n = off < len ? std::min(per, len - off) : 0;
            th.emplace_back([=] { memcpy(dst + off, src + off, n); });
        }
    } catch (const std::system_error&) {  // no thread to spare: copy the remaining pieces here
        for (; t < nt; ++t) {
            const size_t off = t * per;
            if (off < len) memcpy(dst + off, src + off, std::min(per, len - off));
        }
    }
    memcpy(dst, src, std::min(per, len));
    for (auto& x : th) x.join();
}

// With `resident`, a batch whose jobs share one generator seed reads the resident matrix for
// that seed instead of expanding generators per job.
static int batch_launch(BatchJob* J, const JobIn* jobs, bool host_files, bool resident, bool one_stream = false) {
    dmx_prover* p = J->p.get();
    dmx_verifier* v = J->v.get();
    bool same_seed = true;
    for (uint32_t b = 0; b < J->n; ++b) {
        prover_set_job(p, b, jobs[b].params, jobs[b].len, host_files ? nullptr : jobs[b].file);
        v->stage[b] = p->stage[b];
        same_seed = same_seed && p->g_seed[b] == p->g_seed[0];
    }
    GenEntry* E = nullptr;
    if (resident && same_seed) {
        if (J->gen && !memcmp(J->gen->seed, p->g_seed[0].data(), 32)) {
            E = J->gen;
        } else {
            int rc = gen_acquire(J->ctx, p->g_seed[0].data(), J->d, p->s, &E);
            if (rc) return rc;
            if (E) {
                J->drop_resident();
                J->gen = E;
            }
        }
    }
    if (E) HIP_TRY(hipStreamWaitEvent(p->s, E->ready, 0));
    p->G0_src = E ? E->G.as<zq>() : nullptr;
    p->Ghat0_src = E ? E->Ghat.as<uint32_t>() : nullptr;
    v->G0_src = p->G0_src;
    const size_t st = p->jb.stride;
    HIP_TRY(copy_rows(p->meta.p, st, p->stage.data(), sizeof(JobMeta), sizeof(JobMeta), J->n,
                             hipMemcpyHostToDevice, p->s));
    HIP_TRY(copy_rows(v->meta.p, st, v->stage.data(), sizeof(JobMeta), sizeof(JobMeta), J->n,
                             hipMemcpyHostToDevice, p->s));
    if (host_files) {
        size_t need = 0;
        for (uint32_t b = 0; b < J->n; ++b) need += (jobs[b].len + 255) & ~(size_t)255;
        if (need > J->pinned_bytes) {
            HIP_TRY(hipStreamSynchronize(p->s));  // the previous uploads from the old staging are done
            if (J->pinned) (void)hipHostFree(J->pinned);
            J->pinned = nullptr;
            J->pinned_bytes = 0;
            if (hipHostMalloc(&J->pinned, need, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                J->pinned = nullptr;
                return fail(DMX_ERR_OOM, "pinned staging of %zu bytes failed", need);
            }
            J->pinned_bytes = need;
        }
        size_t off = 0;
        for (uint32_t b = 0; b < J->n; ++b) {
            if (!jobs[b].len) continue;
            uint8_t* st = static_cast<uint8_t*>(J->pinned) + off;
            staged_copy(st, static_cast<const uint8_t*>(jobs[b].file), jobs[b].len);
            HIP_TRY(hipMemcpyAsync(job_ptr(p->file.p, p->jb, b), st, jobs[b].len, hipMemcpyHostToDevice, p->s));
            off += (jobs[b].len + 255) & ~(size_t)255;
        }
    }
    bool timing;
    uint64_t fwd_version;
    {
        std::lock_guard<std::mutex> g(J->ctx->mu);
        timing = J->ctx->timing;
        fwd_version = J->ctx->fwd_version;
    }
    if (J->graph_version != fwd_version) {  // the transform kernels changed: capture again
        J->drop_graphs();
        J->graph_version = fwd_version;
    }
    J->one_stream = one_stream;
    if (!graphs_enabled() || timing) {  // eager: kernel timing needs per-launch events
        int rc = batch_enqueue(J);
        if (rc) return rc;
    } else {
        hipGraph_t& graph_slot = J->graphs[one_stream][E ? 1 : 0];
        hipGraphExec_t& exec_slot = J->execs[one_stream][E ? 1 : 0];
        if (!exec_slot) {
            p->capturing = true;
            HIP_TRY(hipStreamBeginCapture(p->s, hipStreamCaptureModeThreadLocal));
            int rc = batch_enqueue(J);
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(p->s, &graph);
            p->capturing = false;
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (e != hipSuccess) return fail(DMX_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
            graph_slot = graph;
            HIP_TRY(hipGraphInstantiate(&exec_slot, graph_slot, nullptr, nullptr, 0));
        }
        HIP_TRY(hipGraphLaunch(exec_slot, p->s));
    }
    HIP_TRY(hipEventRecord(J->done, p->s));
    return DMX_OK;
}

// Output of one job of a finished batch.
struct JobOut {
    uint8_t* statement;
    size_t statement_cap;
    size_t* statement_len;
    uint8_t* records;
    size_t records_cap;
    size_t* records_len;
    uint64_t* opening;
    int* status;
};

// Waits for the batch, checks every job's flags and encodes the proofs.
static int batch_collect(BatchJob* J, const JobOut* outs) {
    dmx_prover* p = J->p.get();
    const uint32_t n = J->n;
    const size_t st = p->jb.stride;
    HIP_TRY(hipEventSynchronize(J->done));
    const size_t com_b = (size_t)K * p->N * 16, rb = p->rec_off[p->d] * 16;
    // one job's result: the head (launch_result_head), com_0 and the records, contiguous
    const size_t com_at = p->com0.as<uint8_t>() - p->out_head.as<uint8_t>();
    const size_t rec_at = p->recs.as<uint8_t>() - p->out_head.as<uint8_t>();
    if (com_at < RESULT_HEAD_BYTES || rec_at != com_at + com_b || com_at % 16)
        return fail(DMX_ERR_HIP, "result layout is not contiguous");
    const size_t blk = rec_at + rb;
    const size_t need = (size_t)n * blk;
    if (need > J->pinned_out_bytes) {
        if (J->pinned_out) (void)hipHostFree(J->pinned_out);
        J->pinned_out = nullptr;
        J->pinned_out_bytes = 0;
        if (hipHostMalloc(&J->pinned_out, need, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            J->pinned_out = nullptr;
            return fail(DMX_ERR_OOM, "pinned result staging of %zu bytes failed", need);
        }
        J->pinned_out_bytes = need;
    }
    uint8_t* base = static_cast<uint8_t*>(J->pinned_out);
    hipStream_t s = p->s;
    HIP_TRY(copy_rows(base, blk, p->out_head.p, st, blk, n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    int first = DMX_OK;
    for (uint32_t b = 0; b < n; ++b) {
        const JobOut& o = outs[b];
        uint8_t* blk_b = base + (size_t)b * blk;
        const uint32_t* head = reinterpret_cast<const uint32_t*>(blk_b);
        const uint8_t* fid = blk_b + 4 * RESULT_HEAD_FID;
        const zq* mstar = reinterpret_cast<const zq*>(blk_b + 4 * RESULT_HEAD_MSTAR);
        const zq* com = reinterpret_cast<const zq*>(blk_b + com_at);
        const zq* recs = reinterpret_cast<const zq*>(blk_b + rec_at);
        int rc = DMX_OK;
        for (size_t j = 0; !rc && j < (size_t)p->d; ++j) rc = check_round_flags(j, head[RESULT_HEAD_PFLAGS]);
        if (!rc && head[RESULT_HEAD_VFLAGS]) rc = fail(DMX_ERR_HIP, "challenge resampling exhausted");
        if (!rc && head[RESULT_HEAD_VFLAGS + 1]) rc = fail(DMX_ERR_OPENING, "final opening mismatch");
        const size_t st_need = dmx_statement_size(p->N);
        if (o.statement_len) *o.statement_len = st_need;
        if (!rc && o.statement) {
            if (o.statement_cap < st_need)
                rc = fail(DMX_ERR_BUFFER, "statement buffer too small");
            else
                encode_statement_from(p, b, fid, com, o.statement);
        }
        const size_t rec_need = dmx_proof_records_size(p->orig_len[b]);
        if (o.records_len) *o.records_len = rec_need;
        if (!rc && o.records) {
            if (o.records_cap < rec_need)
                rc = fail(DMX_ERR_BUFFER, "records buffer too small");
            else
                encode_records_from(p, recs, o.records);
        }
        if (!rc && o.opening) {
            o.opening[0] = (uint64_t)mstar->w[0] | ((uint64_t)mstar->w[1] << 32);
            o.opening[1] = (uint64_t)mstar->w[2] | ((uint64_t)mstar->w[3] << 32);
        }
        if (o.status) *o.status = rc;
        if (rc && !first) first = rc;
    }
    return first;
}

// device memory of one job's block in a batch of depth d
static size_t job_workspace_bytes(int d) {
    dmx_prover tp;
    tp.d = d;
    tp.N = (size_t)1 << d;
    tp.hash_cap = max_file_len(d);
    dmx_verifier tv;
    tv.rows = tv.glen = tv.clen = tv.c_cap = tv.rows0 = tv.n0 = tp.N;
    tv.rounds = (size_t)d;
    tv.g_cap = tp.N * K * tp.N;
    Layout L;
    prover_plan(&tp, L, tp.hash_cap);
    verifier_plan(&tv, L);
    return L.stride();
}

extern "C" size_t dmx_job_workspace_size(uint64_t file_len) {
    const int d = depth_for(file_len);
    return d > 12 ? 0 : job_workspace_bytes(d);
}

// Cache of batch workspaces and their graphs, per (context, depth, n). The cache holds at most
// DAMASCUS_CACHE_GB (default 60 % of device memory); idle workspaces are freed, least recently
// used first, only when a new one would not fit (freeing stalls the device, so never on
// release).
static std::mutex g_batches_mu;
static uint64_t g_batch_creations = 0;
static double g_batch_create_s = 0;
static std::vector<std::unique_ptr<BatchJob>> g_batches;
static uint64_t g_batch_clock = 0;

static size_t cache_budget() {
    const char* e = getenv("DAMASCUS_CACHE_GB");
    if (e) return (size_t)(atof(e) * (double)(1ull << 30));
    static size_t half = [] {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
            (void)hipGetLastError();
            return (size_t)32 << 30;
        }
        return total_b / 10 * 6;
    }();
    return half;
}

// device memory held by cached batch workspaces, busy or idle
static size_t batch_cached_bytes_locked() {
    size_t t = 0;
    for (auto& b : g_batches) t += b->bytes();
    return t;
}

static bool batch_evict_lru_locked() {
    size_t victim = SIZE_MAX;
    for (size_t i = 0; i < g_batches.size(); ++i)
        if (!g_batches[i]->busy && (victim == SIZE_MAX || g_batches[i]->last_use < g_batches[victim]->last_use))
            victim = i;
    if (victim == SIZE_MAX) return false;
    g_batches.erase(g_batches.begin() + victim);
    return true;
}

// An idle cached workspace for (ctx, d, n), else a new one when `may_create` and the cache can
// hold it within its budget after evicting idle workspaces (least recently used first). Returns
// DMX_ERR_PENDING when nothing is free: the caller waits for a batch to finish.
static int batch_try_acquire(dmx_ctx* ctx, int d, uint32_t n, bool may_create, BatchJob** out) {
    {
        std::lock_guard<std::mutex> g(g_batches_mu);
        for (auto& b : g_batches)
            if (b->ctx == ctx && b->d == d && b->n == n && !b->busy) {
                b->busy = true;
                b->last_use = ++g_batch_clock;
                *out = b.get();
                return DMX_OK;
            }
        if (!may_create) return DMX_ERR_PENDING;
        const size_t need = job_workspace_bytes(d) * n, budget = cache_budget();
        while (batch_cached_bytes_locked() + need > budget)
            if (!batch_evict_lru_locked()) return DMX_ERR_PENDING;
    }
    BatchJob* J = nullptr;
    int rc;
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        rc = batch_create(ctx, d, n, &J);
        if (rc != DMX_ERR_OOM) break;
        std::lock_guard<std::mutex> g(g_batches_mu);
        if (!batch_evict_lru_locked()) return DMX_ERR_PENDING;
    }
    if (rc) return rc;
    std::lock_guard<std::mutex> g(g_batches_mu);
    g_batch_creations++;
    g_batch_create_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    J->busy = true;
    J->last_use = ++g_batch_clock;
    g_batches.emplace_back(J);
    *out = J;
    return DMX_OK;
}

// the unmanaged callers (one-shot batches): a workspace even beyond the cache budget
static int batch_acquire(dmx_ctx* ctx, int d, uint32_t n, BatchJob** out) {
    int rc = batch_try_acquire(ctx, d, n, true, out);
    if (rc != DMX_ERR_PENDING) return rc;
    BatchJob* J = nullptr;
    rc = batch_create(ctx, d, n, &J);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(g_batches_mu);
    J->busy = true;
    J->last_use = ++g_batch_clock;
    g_batches.emplace_back(J);
    *out = J;
    return DMX_OK;
}

static void batch_release(BatchJob* J) {
    std::lock_guard<std::mutex> g(g_batches_mu);
    J->busy = false;
}

static void batch_forget_context(dmx_ctx* ctx) {
    {
        std::lock_guard<std::mutex> g(g_batches_mu);
        for (size_t i = 0; i < g_batches.size();)
            if (g_batches[i]->ctx == ctx)
                g_batches.erase(g_batches.begin() + i);
            else
                ++i;
    }
    std::lock_guard<std::mutex> g(g_gen_mu);
    for (size_t i = 0; i < g_gen.size();)
        if (g_gen[i]->ctx == ctx)
            g_gen.erase(g_gen.begin() + i);
        else
            ++i;
}

extern "C" int dmx_prove_verify_device(dmx_ctx* ctx, const void* d_file, size_t len, const dmx_system_params* params,
                                       uint8_t* st_out, size_t st_cap, size_t* st_len, uint8_t* rec_out,
                                       size_t rec_cap, size_t* rec_len, uint64_t opening_out[2]) {
    if (!ctx || !params || (!d_file && len)) return fail(DMX_ERR_ARG, "null argument");
    if (gpu_disabled_by_env()) return fail(DMX_ERR_NO_DEVICE, "disabled by DAMASCUS_GPU=0");
    const int d = depth_for(len);
    if (d > 12) return fail(DMX_ERR_ARG, "depth %d exceeds the supported 12", d);
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = check_device_file(d_file, len, 0);
    if (rc) return rc;
    BatchJob* J = nullptr;
    rc = batch_acquire(ctx, d, 1, &J);
    if (rc) return rc;
    const JobIn in{d_file, len, params};
    int status = DMX_OK;
    const JobOut out{st_out, st_cap, st_len, rec_out, rec_cap, rec_len, opening_out, &status};
    rc = batch_launch(J, &in, false, false);  // the plain path derives generators per call
    if (!rc) rc = batch_collect(J, &out);
    batch_release(J);
    return rc;
}


// Ragged batch: jobs of any depths. Jobs are grouped by depth into batches of at most
// `max_batch` (0: sized by workspace, about 16 GiB per batch); every batch is launched before
// any is awaited, so batches of different depths run concurrently on their own streams.
extern "C" int dmx_prove_verify_many(dmx_ctx* ctx, dmx_job* jobs, size_t count, int files_on_host,
                                     size_t max_batch) {
    if (!ctx || (count && !jobs)) return fail(DMX_ERR_ARG, "null argument");
    if (gpu_disabled_by_env()) return fail(DMX_ERR_NO_DEVICE, "disabled by DAMASCUS_GPU=0");
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<std::vector<size_t>> by_d(13);
    for (size_t i = 0; i < count; ++i) {
        if (!jobs[i].file && jobs[i].len) return fail(DMX_ERR_ARG, "job %zu: null file", i);
        const int d = depth_for(jobs[i].len);
        if (d > 12) return fail(DMX_ERR_ARG, "job %zu: depth %d exceeds the supported 12", i, d);
        const int crc = check_file(jobs[i].file, jobs[i].len, files_on_host != 0, i);
        if (crc) return crc;
        jobs[i].depth = d;
        jobs[i].status = DMX_ERR_HIP;
        by_d[d].push_back(i);
    }
    struct Launched {
        BatchJob* J;
        std::vector<size_t> idx;
    };
    std::vector<Launched> live;
    int rc = DMX_OK;
    for (int d = 12; d >= 0 && !rc; --d) {  // deepest first: their chains are longest
        const std::vector<size_t>& ids = by_d[d];
        if (ids.empty()) continue;
        const size_t per_job = job_workspace_bytes(d);
        size_t cap = max_batch ? max_batch : std::max<size_t>(1, ((size_t)16 << 30) / per_job);
        cap = std::min<size_t>(cap, 65535);
        for (size_t at = 0; at < ids.size() && !rc; at += cap) {
            const size_t nb = std::min(cap, ids.size() - at);
            Launched L{nullptr, std::vector<size_t>(ids.begin() + at, ids.begin() + at + nb)};
            rc = batch_acquire(ctx, d, (uint32_t)nb, &L.J);
            if (rc) break;
            std::vector<JobIn> in(nb);
            for (size_t k = 0; k < nb; ++k) {
                const dmx_job& jj = jobs[L.idx[k]];
                in[k] = JobIn{jj.file, jj.len, &jj.params};
            }
            rc = batch_launch(L.J, in.data(), files_on_host != 0, true);
            live.push_back(std::move(L));
        }
    }
    for (auto& L : live) {
        std::vector<JobOut> outs(L.idx.size());
        for (size_t k = 0; k < L.idx.size(); ++k) {
            dmx_job& jj = jobs[L.idx[k]];
            outs[k] = JobOut{jj.statement, jj.statement_cap, &jj.statement_len, jj.records, jj.records_cap,
                             &jj.records_len, jj.opening, &jj.status};
        }
        if (!rc) {
            const int r2 = batch_collect(L.J, outs.data());
            (void)r2;  // per-job status carries the outcome
        } else {
            (void)hipEventSynchronize(L.J->done);
        }
        batch_release(L.J);
    }
    if (rc) return rc;
    for (size_t i = 0; i < count; ++i)
        if (jobs[i].status != DMX_OK) return fail(jobs[i].status, "job %zu failed", i);
    return DMX_OK;
}

// ------------------------------------------------------------------------------------------
// reference-compatible batch entry points (return codes follow cuda/*.cu: 0 ok, 1 args, 2 device)
// ------------------------------------------------------------------------------------------
// The reference's entry points take no context: their state lives per device, created on the
// first call with that device current. Device buffers and pinned staging grow and are kept, the
// twiddle tables of the last few plans (n, primes, roots) are kept, and every call is one upload,
// one launch and one readback on the state's own stream. Calls are serialised per device.
struct DropIn {
    std::mutex mu;
    hipStream_t s = nullptr;
    hipEvent_t done = nullptr;
    DevBuf dev;
    void* pin = nullptr;
    size_t pin_bytes = 0;
    struct Plan {
        std::vector<uint64_t> key;  // n, primes, stage roots, inverse roots, inverse sizes
        DevBuf tables;              // CycPrime[8] then the twiddle pairs
        uint64_t last_use = 0;
    };
    std::vector<std::unique_ptr<Plan>> plans;
    uint64_t clock = 0;
    ~DropIn() {
        if (pin) (void)hipHostFree(pin);
        if (done) (void)hipEventDestroy(done);
        if (s) (void)hipStreamDestroy(s);
    }
    int stage(size_t bytes) {  // pinned staging of at least `bytes`
        if (bytes <= pin_bytes) return DMX_OK;
        if (pin) (void)hipHostFree(pin);
        pin = nullptr;
        pin_bytes = 0;
        if (hipHostMalloc(&pin, bytes, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            pin = nullptr;
            return DMX_ERR_OOM;
        }
        pin_bytes = bytes;
        return DMX_OK;
    }
    int wait() {
        if (hipEventRecord(done, s) != hipSuccess) return DMX_ERR_HIP;
        for (;;) {
            const hipError_t e = hipEventQuery(done);
            if (e == hipSuccess) return DMX_OK;
            if (e != hipErrorNotReady) return DMX_ERR_HIP;
            std::this_thread::yield();
        }
    }
};
static std::mutex g_dropin_mu;
static std::unordered_map<int, std::unique_ptr<DropIn>> g_dropin;

static DropIn* dropin_state() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> g(g_dropin_mu);
    auto& slot = g_dropin[dev];
    if (!slot) {
        std::unique_ptr<DropIn> d(new DropIn);
        if (hipStreamCreateWithFlags(&d->s, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&d->done, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        slot = std::move(d);
    }
    return slot.get();
}

// damascus_cuda_fold_batch (cuda/fold_kernel.cu:209-334, src/utils/gpu.rs:95-138): out = left +
// challenge * right mod q over SoA lo/hi words. Returns 0; 1 for bad arguments, including values
// that are not canonical elements of Z_q (every Fp is; the call is then declined and the
// caller's CPU path runs); 2 for a device failure.
extern "C" int dmx_fold_batch(const uint64_t* l_lo, const uint64_t* l_hi, const uint64_t* r_lo, const uint64_t* r_hi,
                              uint64_t* o_lo, uint64_t* o_hi, size_t len, uint64_t c_lo, uint64_t c_hi) {
    if (!l_lo || !l_hi || !r_lo || !r_hi || !o_lo || !o_hi || len == 0) return 1;
    if (c_hi > 0xFFFFFFFFFFFFull || (c_hi == 0xFFFFFFFFFFFFull && c_lo >= 0xFFFFFFFFFFFFFFB5ull)) return 1;
    if (!dmx_backend_ready()) return 2;
    DropIn* D = dropin_state();
    if (!D) return 2;
    std::lock_guard<std::mutex> g(D->mu);
    const size_t by = len * 8;
    if (D->stage(6 * by + 256)) return 2;
    if (D->dev.bytes < 6 * by + 256 && D->dev.alloc(6 * by + 256)) return 2;
    uint8_t* h = static_cast<uint8_t*>(D->pin);
    memcpy(h, l_lo, by);
    memcpy(h + by, l_hi, by);
    memcpy(h + 2 * by, r_lo, by);
    memcpy(h + 3 * by, r_hi, by);
    uint64_t* d = D->dev.as<uint64_t>();
    uint32_t* dflag = reinterpret_cast<uint32_t*>(d + 6 * len);
    if (hipMemcpyAsync(d, h, 4 * by, hipMemcpyHostToDevice, D->s) != hipSuccess) return 2;
    launch_fill32(dflag, 1, 0, D->s);
    launch_fold_soa(d, d + 4 * len, len, c_lo, c_hi, dflag, D->s);
    if (hipGetLastError() != hipSuccess) return 2;
    if (hipMemcpyAsync(h + 4 * by, d + 4 * len, 2 * by + 4, hipMemcpyDeviceToHost, D->s) != hipSuccess) return 2;
    if (D->wait()) return 2;
    uint32_t flag = 0;
    memcpy(&flag, h + 6 * by, 4);
    if (flag) return 1;
    memcpy(o_lo, h + 4 * by, by);
    memcpy(o_hi, h + 5 * by, by);
    return 0;
}

static uint64_t mulmod_u64(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)((unsigned __int128)a * b % p); }
static uint64_t powmod_u64(uint64_t b, uint64_t e, uint64_t p) {
    uint64_t r = 1 % p;
    for (; e; e >>= 1, b = mulmod_u64(b, b, p))
        if (e & 1) r = mulmod_u64(r, b, p);
    return r;
}

// The device tables of a plan: CycPrime per prime, then per prime the Montgomery pairs of
// root_s^j (j < 2^s) for every forward then every inverse stage. The roots are checked to give
// the reference's stage spans: each stage-s root (forward and inverse) has order exactly
// 2^(s+1), i.e. root^(2^s) = p - 1 (stage_len_from_root, ntt_kernel.cu:33-41); roots and
// inverse sizes are below p. DMX_ERR_ARG when they are not.
static int dropin_plan(DropIn* D, size_t n, int log_n, const uint64_t* primes, const uint64_t* roots,
                       const uint64_t* iroots, const uint64_t* inv_sizes, const DropIn::Plan** out) {
    std::vector<uint64_t> key;
    key.reserve(1 + 2 * NPRIMES + 2 * NPRIMES * log_n);
    key.push_back(n);
    key.insert(key.end(), primes, primes + NPRIMES);
    key.insert(key.end(), roots, roots + (size_t)NPRIMES * log_n);
    key.insert(key.end(), iroots, iroots + (size_t)NPRIMES * log_n);
    key.insert(key.end(), inv_sizes, inv_sizes + NPRIMES);
    for (auto& pl : D->plans)
        if (pl->key == key) {
            pl->last_use = ++D->clock;
            *out = pl.get();
            return DMX_OK;
        }
    const size_t per = 4 * (n - 1);
    std::vector<uint32_t> host(sizeof(CycPrime) / 4 * NPRIMES + per * NPRIMES);
    CycPrime* cps = reinterpret_cast<CycPrime*>(host.data());
    uint32_t* tw = host.data() + sizeof(CycPrime) / 4 * NPRIMES;
    for (int q = 0; q < NPRIMES; ++q) {
        const uint64_t p = primes[q];
        if (inv_sizes[q] >= p) return DMX_ERR_ARG;
        const uint32_t pinv = host_inv32((uint32_t)p);
        auto pair = [&](uint64_t w, uint32_t* dst) {  // (w R mod p, that * p^-1 mod 2^32)
            const uint32_t wm = (uint32_t)(((unsigned __int128)w << 32) % p);
            dst[0] = wm;
            dst[1] = wm * pinv;
        };
        CycPrime& c = cps[q];
        memset(&c, 0, sizeof c);
        c.p = (uint32_t)p;
        c.pinv = pinv;
        // the pointwise product is redc(a b) = a b / R, then times R: mulc(t, wm, wmp) = t w for
        // wm = w R mod p, so w = R takes wm = R^2 mod p
        const uint32_t r2 = (uint32_t)(((unsigned __int128)1 << 64) % p);
        c.r2m = r2;
        c.r2mp = r2 * pinv;
        uint32_t inv_pair[2];
        pair(inv_sizes[q], inv_pair);
        c.invs_m = inv_pair[0];
        c.invs_mp = inv_pair[1];
        for (int dir = 0; dir < 2; ++dir) {
            const uint64_t* rs = (dir ? iroots : roots) + (size_t)q * log_n;
            uint32_t* t = tw + (size_t)q * per + (size_t)dir * 2 * (n - 1);
            for (int st = 0; st < log_n; ++st) {
                const uint64_t r = rs[st];
                if (r >= p || powmod_u64(r, (uint64_t)1 << st, p) != p - 1) return DMX_ERR_ARG;
                uint64_t w = 1;
                for (size_t j = 0; j < ((size_t)1 << st); ++j) {
                    pair(w, t + 2 * (((size_t)1 << st) - 1 + j));
                    w = mulmod_u64(w, r, p);
                }
            }
        }
    }
    if (D->plans.size() >= 4) {  // keep the most recently used plans
        size_t victim = 0;
        for (size_t i = 1; i < D->plans.size(); ++i)
            if (D->plans[i]->last_use < D->plans[victim]->last_use) victim = i;
        D->plans.erase(D->plans.begin() + victim);
    }
    std::unique_ptr<DropIn::Plan> pl(new DropIn::Plan);
    pl->key = std::move(key);
    if (pl->tables.alloc(host.size() * 4)) return DMX_ERR_OOM;
    if (hipMemcpy(pl->tables.p, host.data(), host.size() * 4, hipMemcpyHostToDevice) != hipSuccess) return DMX_ERR_HIP;
    pl->last_use = ++D->clock;
    *out = pl.get();
    D->plans.push_back(std::move(pl));
    return DMX_OK;
}

// damascus_cuda_ntt_batch (cuda/ntt_kernel.cu:106-293, src/utils/gpu.rs:140-207): for each of 8
// primes, the length-n cyclic product of host_a and host_b (prime-major, 8 n words) computed as
// the reference does (bit reversal, radix-2 DIT with the caller's stage roots, pointwise product,
// the inverse with the caller's inverse roots, times inv_sizes). Returns 0; 1 for bad arguments,
// which include a plan the reference would not run as radix-2 stages (a stage root whose order is
// not 2^(s+1), a root or inverse size >= p) and inputs that are not residues below p: the call is
// declined and the caller's CPU path runs, as for any non-zero code (gpu.rs:191-193); n above
// 2^14 is declined the same way; 2 for a device failure.
extern "C" int dmx_ntt_batch(const uint64_t* a, const uint64_t* b, uint64_t* out, size_t n, const uint64_t* primes,
                             const uint64_t* stage_roots, const uint64_t* inv_roots, const uint64_t* inv_sizes) {
    if (!a || !b || !out || !primes || !stage_roots || !inv_roots || !inv_sizes || n == 0 || (n & (n - 1))) return 1;
    for (int i = 0; i < NPRIMES; ++i)
        if (primes[i] < 3 || primes[i] >= (1ull << 32) || !(primes[i] & 1)) return 1;
    int log_n = 0;
    while (((size_t)1 << log_n) < n) ++log_n;
    if (log_n > CYC_MAX_LOG_N) return 1;
    if (!dmx_backend_ready()) return 2;
    DropIn* D = dropin_state();
    if (!D) return 2;
    std::lock_guard<std::mutex> g(D->mu);
    const DropIn::Plan* plan = nullptr;
    const int prc = dropin_plan(D, n, log_n, primes, stage_roots, inv_roots, inv_sizes, &plan);
    if (prc == DMX_ERR_ARG) return 1;
    if (prc) return 2;
    const size_t words = (size_t)NPRIMES * n, by = words * 8;
    if (D->stage(3 * by + 256)) return 2;
    if (D->dev.bytes < 3 * by + 256 && D->dev.alloc(3 * by + 256)) return 2;
    uint8_t* h = static_cast<uint8_t*>(D->pin);
    memcpy(h, a, by);
    memcpy(h + by, b, by);
    uint64_t* d = D->dev.as<uint64_t>();
    uint32_t* dflag = reinterpret_cast<uint32_t*>(d + 3 * words);
    if (hipMemcpyAsync(d, h, 2 * by, hipMemcpyHostToDevice, D->s) != hipSuccess) return 2;
    launch_fill32(dflag, 1, 0, D->s);
    const CycPrime* cp = plan->tables.as<CycPrime>();
    const uint32_t* tw = plan->tables.as<uint32_t>() + sizeof(CycPrime) / 4 * NPRIMES;
    launch_cyc_lds(d, d + words, d + 2 * words, log_n, cp, tw, dflag, D->s);
    if (hipGetLastError() != hipSuccess) return 2;
    if (hipMemcpyAsync(h + 2 * by, d + 2 * words, by + 4, hipMemcpyDeviceToHost, D->s) != hipSuccess) return 2;
    if (D->wait()) return 2;
    uint32_t flag = 0;
    memcpy(&flag, h + 3 * by, 4);
    if (flag) return 1;
    memcpy(out, h + 2 * by, by);
    return 0;
}

// ------------------------------------------------------------------------------------------
// commitments and products through the context
// ------------------------------------------------------------------------------------------
extern "C" int dmx_derive_subseed(const uint8_t seed[32], const uint8_t* label, size_t label_len, uint8_t out[32]) {
    if (!seed || !out || (label_len && !label) || label_len > 64) return fail(DMX_ERR_ARG, "bad argument");
    b3::host_hash_short(seed, label, label_len, out);
    return DMX_OK;
}

extern "C" int dmx_derive_generators(dmx_ctx* ctx, const uint8_t g_seed[32], size_t rows, size_t ring_len,
                                     uint64_t* out) {
    if (!ctx || !g_seed || !out) return fail(DMX_ERR_ARG, "null argument");
    if (!ring_len || (ring_len & (ring_len - 1)) || ring_len > ((size_t)1 << MAX_LOG_N))
        return fail(DMX_ERR_SHAPE, "ring length must be a power of two <= 8192");
    if (!rows) return DMX_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    DevBuf g, key;
    int rc = g.alloc(rows * K * ring_len * 16);
    if (!rc) rc = key.alloc(32);
    if (rc) return rc;
    uint32_t gw[8];
    gseed_words(g_seed, gw);
    HIP_TRY(hipMemcpy(key.p, gw, 32, hipMemcpyHostToDevice));
    launch_gen_generators(key.as<uint32_t>(), rows * K, log2_exact(ring_len), g.as<zq>(), 0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, g.p, rows * K * ring_len * 16, hipMemcpyDeviceToHost));
    return DMX_OK;
}

// commit_with_generators_ntt (sis.rs:146-167) on device; gens == nullptr derives them from
// g_seed. The sum over rows is split so that each partial stays inside the CRT range.
static int commit_impl(dmx_ctx* ctx, const uint64_t* witness, const uint64_t* gens, const uint8_t* g_seed,
                       size_t rows, size_t ring_len, uint64_t* com_out, size_t* com_len) {
    if (!ctx || !com_out || (rows && !witness)) return fail(DMX_ERR_ARG, "null argument");
    if (gpu_disabled_by_env()) return fail(DMX_ERR_NO_DEVICE, "disabled by DAMASCUS_GPU=0");
    if (rows == 0) {
        memset(com_out, 0, (size_t)K * 16);
        if (com_len) *com_len = 1;
        return DMX_OK;
    }
    if (!ring_len || (ring_len & (ring_len - 1)) || ring_len > ((size_t)1 << MAX_LOG_N))
        return fail(DMX_ERR_SHAPE, "ring length must be a power of two <= 8192");
    HIP_TRY(hipSetDevice(ctx->device));
    const int lg = log2_exact(ring_len);
    const size_t m = ring_len;
    // rows per partial sum: rows * m <= 2^24 keeps |V| <= rows*m*(q/2)^2 < P/2
    const size_t chunk = std::max<size_t>(1, ((size_t)1 << 24) / m);
    const size_t crow = std::min(rows, chunk);
    DevBuf W, G, What, Ghat, res, part, acc, mpart;
    int rc = 0;
    rc = rc ? rc : W.alloc(crow * m * 16);
    rc = rc ? rc : G.alloc(crow * K * m * 16);
    rc = rc ? rc : What.alloc(crow * NPRIMES * m * 4);
    rc = rc ? rc : Ghat.alloc(crow * K * NPRIMES * m * 4);
    rc = rc ? rc : res.alloc((size_t)K * NPRIMES * m * 4);
    rc = rc ? rc : part.alloc((size_t)K * m * 16);
    rc = rc ? rc : acc.alloc((size_t)K * m * 16);
    rc = rc ? rc : mpart.alloc(mac_part_words(lg) * 4);
    if (rc) return rc;
    hipStream_t s = 0;
    HIP_TRY(hipMemsetAsync(acc.p, 0, (size_t)K * m * 16, s));
    uint32_t gw[8] = {0};
    if (g_seed) gseed_words(g_seed, gw);
    DevBuf key;
    rc = key.alloc(32);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(key.p, gw, 32, hipMemcpyHostToDevice));
    for (size_t r0 = 0; r0 < rows; r0 += crow) {
        const size_t nr = std::min(crow, rows - r0);
        HIP_TRY(hipMemcpyAsync(W.p, witness + 2 * r0 * m, nr * m * 16, hipMemcpyHostToDevice, s));
        RowSrc pw{};
        pw.mode = ROW_PLAIN;
        pw.a = W.as<zq>();
        launch_ntt_fwd_rows(pw, nr, lg, What.as<uint32_t>(), ctx->tw, s, Jobs{}, ctx->fwd_variant[lg]);
        RowSrc pg{};
        if (gens) {
            HIP_TRY(hipMemcpyAsync(G.p, gens + 2 * r0 * K * m, nr * K * m * 16, hipMemcpyHostToDevice, s));
            pg.mode = ROW_PLAIN;
            pg.a = G.as<zq>();
            launch_ntt_fwd_rows(pg, nr * K, lg, Ghat.as<uint32_t>(), ctx->tw, s, Jobs{}, ctx->fwd_variant[lg]);
        } else {
            // generator rows r0.. of the family: generate the full family prefix, keep the slice
            DevBuf full;
            rc = full.alloc((r0 + nr) * K * m * 16);
            if (rc) return rc;
            launch_gen_generators(key.as<uint32_t>(), (r0 + nr) * K, lg, full.as<zq>(), s);
            pg.mode = ROW_PLAIN;
            pg.a = full.as<zq>() + r0 * K * m;
            launch_ntt_fwd_rows(pg, nr * K, lg, Ghat.as<uint32_t>(), ctx->tw, s, Jobs{}, ctx->fwd_variant[lg]);
            HIP_TRY(hipStreamSynchronize(s));
        }
        launch_mac_pairs_inv(What.as<uint32_t>(), Ghat.as<uint32_t>(), nr, lg, res.as<uint32_t>(), nullptr, nullptr,
                             mpart.as<uint32_t>(), ctx->tw, s, Jobs{}, false, ctx->mac(lg));
        launch_crt_lift(res.as<uint32_t>(), part.as<zq>(), K, lg, s);
        launch_zq_add(acc.as<zq>(), part.as<zq>(), acc.as<zq>(), (size_t)K * m, s);
        HIP_TRY(hipStreamSynchronize(s));
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(com_out, acc.p, (size_t)K * m * 16, hipMemcpyDeviceToHost));
    if (com_len) *com_len = m;
    return DMX_OK;
}

extern "C" int dmx_commit(dmx_ctx* ctx, const uint8_t seed[32], const uint64_t* witness, size_t rows, size_t ring_len,
                          uint64_t* com_out, size_t* com_len) {
    if (!seed) return fail(DMX_ERR_ARG, "null seed");
    uint8_t gs[32];
    b3::host_hash_short(seed, (const uint8_t*)"g0", 2, gs);
    return commit_impl(ctx, witness, nullptr, gs, rows, ring_len, com_out, com_len);
}

extern "C" int dmx_commit_with_generators(dmx_ctx* ctx, const uint64_t* witness, const uint64_t* gens, size_t rows,
                                          size_t ring_len, uint64_t* com_out, size_t* com_len) {
    if (rows && !gens) return fail(DMX_ERR_ARG, "null generators");
    return commit_impl(ctx, witness, gens, nullptr, rows, ring_len, com_out, com_len);
}

// Waits for the product stream by polling an event (a blocking synchronize adds tens of
// microseconds of wake-up latency to a call that is itself well under a millisecond).
static int ring_wait(dmx_ctx* ctx, hipStream_t s) {
    HIP_TRY(hipEventRecord(ctx->ring_done, s));
    for (;;) {
        const hipError_t e = hipEventQuery(ctx->ring_done);
        if (e == hipSuccess) return DMX_OK;
        if (e != hipErrorNotReady) return fail(DMX_ERR_HIP, "product stream: %s", hipGetErrorString(e));
        std::this_thread::yield();
    }
}

// ragged negacyclic products: group by degree, transform both operands, multiply, invert, lift
// Ragged negacyclic products on device-resident operands. mode 0: one persistent launch over
// the whole batch; mode 1: one launch sequence per degree group (gather, transforms, pointwise,
// inverse, lift, scatter). Scratch buffers are kept in the context and grow on demand.
static int ring_mul_device(dmx_ctx* ctx, const zq* dA, const zq* dB, zq* dO, const uint32_t* log_n, size_t count,
                           int mode, hipStream_t s) {
    std::vector<uint64_t> offs(count);
    size_t total = 0;
    int max_lg = 0;
    for (size_t i = 0; i < count; ++i) {
        if (log_n[i] > (uint32_t)MAX_LOG_N) return fail(DMX_ERR_SHAPE, "degree 2^%u exceeds 2^13", log_n[i]);
        offs[i] = total;
        total += (size_t)1 << log_n[i];
        max_lg = std::max(max_lg, (int)log_n[i]);
    }
    std::vector<uint32_t> order(count);
    for (size_t i = 0; i < count; ++i) order[i] = (uint32_t)i;
    std::lock_guard<std::mutex> g(ctx->ring_mu);
    auto grow = [](DevBuf& b, size_t bytes) { return b.bytes >= bytes ? DMX_OK : b.alloc(bytes); };
    int rc = DMX_OK;
    if (mode == 0 && max_lg <= RING_LDS_MAX_LOG_N) {
        // LDS-resident products: the degrees go up (one copy from pinned staging) and the device
        // builds the launch's queue itself
        if (count * 4 > ctx->ring_pinned_bytes) {
            if (ctx->ring_pinned) (void)hipHostFree(ctx->ring_pinned);
            ctx->ring_pinned = nullptr;
            ctx->ring_pinned_bytes = 0;
            if (hipHostMalloc(&ctx->ring_pinned, count * 4, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                ctx->ring_pinned = nullptr;
                return fail(DMX_ERR_OOM, "pinned staging of %zu bytes failed", count * 4);
            }
            ctx->ring_pinned_bytes = count * 4;
        }
        memcpy(ctx->ring_pinned, log_n, count * 4);
        rc = grow(ctx->ring_logn, count * 4);
        rc = rc ? rc : grow(ctx->ring_scratch, ring_scratch_bytes(count));
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(ctx->ring_logn.p, ctx->ring_pinned, count * 4, hipMemcpyHostToDevice, s));
        launch_ring_mul_lds(dA, dB, dO, ctx->ring_logn.as<uint32_t>(), (uint32_t)count,
                            ring_scratch_at(ctx->ring_scratch.p, count), (uint32_t)ctx->cus, ctx->tw, s);
        HIP_TRY(hipGetLastError());
        return ring_wait(ctx, s);  // the pinned staging is reused by the next call
    }
    rc = grow(ctx->ring_offs, count * 8);
    rc = rc ? rc : grow(ctx->ring_logn, count * 4);
    rc = rc ? rc : grow(ctx->ring_order, count * 4);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(ctx->ring_offs.p, offs.data(), count * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(ctx->ring_logn.p, log_n, count * 4, hipMemcpyHostToDevice, s));
    if (mode == 2) {  // schoolbook (ntt_enabled = false)
        launch_ring_mul_naive(dA, dB, dO, ctx->ring_offs.as<uint64_t>(), ctx->ring_logn.as<uint32_t>(),
                              (uint32_t)count, s);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(s));
        return DMX_OK;
    }
    if (mode == 0) {  // degree 2^13 present: the layered persistent kernel over the whole batch
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return log_n[x] > log_n[y]; });
        HIP_TRY(hipMemcpyAsync(ctx->ring_order.p, order.data(), count * 4, hipMemcpyHostToDevice, s));
        rc = grow(ctx->ring_res, total * NPRIMES * 4);
        rc = rc ? rc : grow(ctx->ring_counter, 256);
        if (rc) return rc;
        launch_ring_mul_persistent(dA, dB, dO, ctx->ring_offs.as<uint64_t>(), ctx->ring_logn.as<uint32_t>(),
                                   ctx->ring_order.as<uint32_t>(), (uint32_t)count, ctx->ring_counter.as<uint32_t>(),
                                   ctx->ring_res.as<uint32_t>(), max_lg, ctx->tw, s);
        HIP_TRY(hipGetLastError());
        return ring_wait(ctx, s);  // host vectors above outlive the copies
    }
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return log_n[x] < log_n[y]; });
    HIP_TRY(hipMemcpyAsync(ctx->ring_order.p, order.data(), count * 4, hipMemcpyHostToDevice, s));
    size_t maxw = 0;
    for (size_t i = 0; i < count;) {
        size_t j = i;
        while (j < count && log_n[order[j]] == log_n[order[i]]) ++j;
        maxw = std::max(maxw, (j - i) << log_n[order[i]]);
        i = j;
    }
    rc = grow(ctx->ring_ga, maxw * 16);
    rc = rc ? rc : grow(ctx->ring_gb, maxw * 16);
    rc = rc ? rc : grow(ctx->ring_ra, maxw * NPRIMES * 4);
    rc = rc ? rc : grow(ctx->ring_rb, maxw * NPRIMES * 4);
    if (rc) return rc;
    for (size_t i = 0; i < count;) {
        size_t j = i;
        while (j < count && log_n[order[j]] == log_n[order[i]]) ++j;
        const int lg = (int)log_n[order[i]];
        const size_t cnt = j - i;
        const uint32_t* idx = ctx->ring_order.as<uint32_t>() + i;
        const uint64_t* doffs = ctx->ring_offs.as<uint64_t>();
        launch_gather_rows(dA, doffs, idx, cnt, lg, ctx->ring_ga.as<zq>(), s);
        launch_gather_rows(dB, doffs, idx, cnt, lg, ctx->ring_gb.as<zq>(), s);
        RowSrc sa{}, sb{};
        sa.mode = sb.mode = ROW_PLAIN;
        sa.a = ctx->ring_ga.as<zq>();
        sb.a = ctx->ring_gb.as<zq>();
        launch_ntt_fwd_rows(sa, cnt, lg, ctx->ring_ra.as<uint32_t>(), ctx->tw, s, Jobs{}, ctx->fwd_variant[lg]);
        launch_ntt_fwd_rows(sb, cnt, lg, ctx->ring_rb.as<uint32_t>(), ctx->tw, s, Jobs{}, ctx->fwd_variant[lg]);
        launch_pointwise(ctx->ring_ra.as<uint32_t>(), ctx->ring_rb.as<uint32_t>(), cnt, lg, s);
        launch_ntt_inv(ctx->ring_ra.as<uint32_t>(), cnt, lg, ctx->tw, s);
        launch_crt_lift(ctx->ring_ra.as<uint32_t>(), ctx->ring_ga.as<zq>(), cnt, lg, s);
        launch_scatter_rows(ctx->ring_ga.as<zq>(), doffs, idx, cnt, lg, dO, s);
        i = j;
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return DMX_OK;
}

extern "C" int dmx_ring_mul_batch(dmx_ctx* ctx, const uint64_t* a, const uint64_t* b, uint64_t* out,
                                  const uint32_t* log_n, size_t count) {
    if (!ctx || (count && (!a || !b || !out || !log_n))) return fail(DMX_ERR_ARG, "null argument");
    if (gpu_disabled_by_env()) return fail(DMX_ERR_NO_DEVICE, "disabled by DAMASCUS_GPU=0");
    if (!count) return DMX_OK;
    size_t total = 0;
    for (size_t i = 0; i < count; ++i) {
        if (log_n[i] > (uint32_t)MAX_LOG_N) return fail(DMX_ERR_SHAPE, "degree 2^%u exceeds 2^13", log_n[i]);
        total += (size_t)1 << log_n[i];
    }
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = ctx->ring_stream;
    DevBuf dA, dB, dO;
    int rc = 0;
    rc = rc ? rc : dA.alloc(total * 16);
    rc = rc ? rc : dB.alloc(total * 16);
    rc = rc ? rc : dO.alloc(total * 16);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(dA.p, a, total * 16, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dB.p, b, total * 16, hipMemcpyHostToDevice, s));
    rc = ring_mul_device(ctx, dA.as<zq>(), dB.as<zq>(), dO.as<zq>(), log_n, count, 0, s);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(out, dO.p, total * 16, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return DMX_OK;
}

// RingElement::mul(rhs, ntt_enabled, gpu_enabled) (src/algebra/poly.rs:134-140) over a ragged
// batch: NTT products, or the schoolbook product when cfg->ntt_enabled == 0; gpu_enabled == 0
// returns DMX_ERR_NO_DEVICE (the caller's CPU path). Bit-identical either way.
extern "C" int dmx_ring_mul_batch_cfg(dmx_ctx* ctx, const uint64_t* a, const uint64_t* b, uint64_t* out,
                                      const uint32_t* log_n, size_t count, const dmx_runtime_config* cfg) {
    if (cfg && !cfg->gpu_enabled) return fail(DMX_ERR_NO_DEVICE, "gpu_enabled = false");
    if (!cfg || cfg->ntt_enabled) return dmx_ring_mul_batch(ctx, a, b, out, log_n, count);
    if (!ctx || (count && (!a || !b || !out || !log_n))) return fail(DMX_ERR_ARG, "null argument");
    if (gpu_disabled_by_env()) return fail(DMX_ERR_NO_DEVICE, "disabled by DAMASCUS_GPU=0");
    if (!count) return DMX_OK;
    size_t total = 0;
    for (size_t i = 0; i < count; ++i) {
        if (log_n[i] > (uint32_t)MAX_LOG_N) return fail(DMX_ERR_SHAPE, "degree 2^%u exceeds 2^13", log_n[i]);
        total += (size_t)1 << log_n[i];
    }
    HIP_TRY(hipSetDevice(ctx->device));
    DevBuf dA, dB, dO;
    int rc = dA.alloc(total * 16);
    rc = rc ? rc : dB.alloc(total * 16);
    rc = rc ? rc : dO.alloc(total * 16);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(dA.p, a, total * 16, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dB.p, b, total * 16, hipMemcpyHostToDevice));
    rc = ring_mul_device(ctx, dA.as<zq>(), dB.as<zq>(), dO.as<zq>(), log_n, count, 2, ctx->ring_stream);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(out, dO.p, total * 16, hipMemcpyDeviceToHost));
    return DMX_OK;
}

extern "C" int dmx_ring_mul_batch_device(dmx_ctx* ctx, const void* d_a, const void* d_b, void* d_out,
                                         const uint32_t* log_n, size_t count, int mode) {
    if (!ctx || (count && (!d_a || !d_b || !d_out || !log_n))) return fail(DMX_ERR_ARG, "null argument");
    if (mode < 0 || mode > 2) return fail(DMX_ERR_ARG, "mode must be 0 (persistent), 1 (per degree) or 2 (schoolbook)");
    if (gpu_disabled_by_env()) return fail(DMX_ERR_NO_DEVICE, "disabled by DAMASCUS_GPU=0");
    if (!count) return DMX_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    return ring_mul_device(ctx, static_cast<const zq*>(d_a), static_cast<const zq*>(d_b), static_cast<zq*>(d_out),
                           log_n, count, mode, ctx->ring_stream);
}

// The ragged products of dmx_ring_mul_batch_device (mode 0) without any allocation: the degrees and
// the launch's queue go to the device array looked up at context creation (ctx->noalloc). Up to
// NOALLOC_MAX_PRODUCTS products of degree at most 2^12. Calls on one device share that array, so
// calls that overlap on one device (other contexts, other threads) race; calls through one
// context are serialised by its product lock.
extern "C" int dmx_ring_mul_batch_device_noalloc(dmx_ctx* ctx, const void* d_a, const void* d_b, void* d_out,
                                                 const uint32_t* log_n, size_t count) {
    if (!ctx || (count && (!d_a || !d_b || !d_out || !log_n))) return fail(DMX_ERR_ARG, "null argument");
    if (gpu_disabled_by_env()) return fail(DMX_ERR_NO_DEVICE, "disabled by DAMASCUS_GPU=0");
    if (count > NOALLOC_MAX_PRODUCTS)
        return fail(DMX_ERR_ARG, "%zu products exceed the allocation-free limit of %zu", count,
                    (size_t)NOALLOC_MAX_PRODUCTS);
    for (size_t i = 0; i < count; ++i)
        if (log_n[i] > (uint32_t)RING_LDS_MAX_LOG_N)
            return fail(DMX_ERR_SHAPE, "degree 2^%u exceeds the allocation-free limit of 2^%d", log_n[i],
                        RING_LDS_MAX_LOG_N);
    if (!count) return DMX_OK;
    if (!ctx->noalloc) return fail(DMX_ERR_HIP, "the allocation-free scratch was not looked up");
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> g(ctx->ring_mu);
    hipStream_t s = ctx->ring_stream;
    uint8_t* base = static_cast<uint8_t*>(ctx->noalloc);
    uint32_t* d_logn = reinterpret_cast<uint32_t*>(base);
    HIP_TRY(hipMemcpyAsync(d_logn, log_n, count * 4, hipMemcpyHostToDevice, s));
    launch_ring_mul_lds(static_cast<const zq*>(d_a), static_cast<const zq*>(d_b), static_cast<zq*>(d_out), d_logn,
                        (uint32_t)count, ring_scratch_at(base + NOALLOC_QUEUE_OFFSET, count), (uint32_t)ctx->cus,
                        ctx->tw, s);
    HIP_TRY(hipGetLastError());
    return ring_wait(ctx, s);  // the caller's degrees array is read by the copy before this returns
}

// ------------------------------------------------------------------------------------------
// Serving loop: continuous batching. A scheduler thread owns the device side: it launches
// batches from the arrival queue whenever a batch slot and device memory are free, and polls
// the in-flight batches' completion events, so jobs that arrive while others run are admitted
// at once and overlap them on the device.
// ------------------------------------------------------------------------------------------
struct ServedJob {
    dmx_job* job = nullptr;
    int d = 0;
    double t_submit = 0, t_done = 0;
    bool done = false;
};

struct dmx_server {
    dmx_ctx* ctx = nullptr;
    dmx_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> queue;  // tickets in arrival order
    std::unordered_map<uint64_t, ServedJob> jobs;
    struct Flight {
        BatchJob* J;
        std::vector<uint64_t> tickets;
        size_t bytes;
    };
    std::vector<Flight> flights;  // running on the device
    std::deque<Flight> done_q;     // finished on the device, results not yet collected
    std::condition_variable cv_fin;
    size_t hbm_used = 0;
    uint64_t next_ticket = 1;
    uint64_t batches = 0, batched_jobs = 0;
    double launch_s = 0;  // scheduler time spent acquiring workspaces and launching
    bool stop = false, fin_stop = false;
    std::thread th, fth;  // scheduler; finisher (result copies and encoding, off the scheduler)
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

static void server_finish(dmx_server* S, dmx_server::Flight& f, int launch_rc) {
    std::vector<JobOut> outs(f.tickets.size());
    {
        std::lock_guard<std::mutex> g(S->mu);
        for (size_t k = 0; k < f.tickets.size(); ++k) {
            dmx_job& jj = *S->jobs[f.tickets[k]].job;
            outs[k] = JobOut{jj.statement, jj.statement_cap, &jj.statement_len, jj.records, jj.records_cap,
                             &jj.records_len, jj.opening, &jj.status};
        }
    }
    if (!launch_rc) {
        (void)batch_collect(f.J, outs.data());
    } else {
        (void)hipEventSynchronize(f.J->done);
        for (auto& o : outs) *o.status = launch_rc;
    }
    batch_release(f.J);
    std::lock_guard<std::mutex> g(S->mu);
    const double t = S->now();
    for (uint64_t tk : f.tickets) {
        ServedJob& sj = S->jobs[tk];
        sj.done = true;
        sj.t_done = t;
    }
    S->hbm_used -= f.bytes;
    S->cv_done.notify_all();
}

// Collects finished batches (device-to-host copies, proof encoding, workspace release) so the
// scheduler never stalls on them and keeps the device fed.
static void server_finisher(dmx_server* S) {
    (void)hipSetDevice(S->ctx->device);
    for (;;) {
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv_fin.wait(lk, [S] { return S->fin_stop || !S->done_q.empty(); });
        if (S->done_q.empty()) break;  // stopping and drained
        dmx_server::Flight f = std::move(S->done_q.front());
        S->done_q.pop_front();
        lk.unlock();
        server_finish(S, f, DMX_OK);
    }
}

static void server_loop(dmx_server* S) {
    (void)hipSetDevice(S->ctx->device);
    const uint32_t max_batch = S->cfg.max_batch ? S->cfg.max_batch : 64;
    const uint32_t max_inflight = S->cfg.max_inflight ? S->cfg.max_inflight : 4;
    for (;;) {
        // completions, in any order, handed to the finisher
        for (size_t i = 0; i < S->flights.size();) {
            if (hipEventQuery(S->flights[i].J->done) == hipSuccess) {
                {
                    std::lock_guard<std::mutex> g(S->mu);
                    S->done_q.push_back(std::move(S->flights[i]));
                }
                S->cv_fin.notify_one();
                S->flights.erase(S->flights.begin() + i);
            } else {
                (void)hipGetLastError();  // "not ready" must not surface at the next launch check
                ++i;
            }
        }
        // admission: the oldest waiting job's depth goes first; a batch takes the waiting jobs
        // of that depth in arrival order, a power of two of them, as many as memory allows
        for (;;) {
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->queue.empty() || S->flights.size() >= max_inflight) break;
            const int d = S->jobs[S->queue.front()].d;
            std::vector<size_t> pos;
            for (size_t i = 0; i < S->queue.size() && pos.size() < max_batch; ++i)
                if (S->jobs[S->queue[i]].d == d) pos.push_back(i);
            size_t nb = 1;
            while (nb * 2 <= pos.size()) nb *= 2;
            const size_t per = job_workspace_bytes(d);
            // one batch's workspace stays within 1/8 of the budget (at least one job), so a few
            // large batches cannot pin all memory
            while (nb > 1 && (S->hbm_used + nb * per > S->cfg.hbm_budget_bytes || nb * per > S->cfg.hbm_budget_bytes / 8))
                nb /= 2;
            // batches still running or still being collected will return memory
            const bool pending = !S->flights.empty() || !S->done_q.empty();
            if (S->hbm_used + nb * per > S->cfg.hbm_budget_bytes && pending) break;  // wait for memory
            // a cached workspace of this size, else a new one if the cache has room, else a
            // smaller cached one; otherwise wait for a batch to finish
            lk.unlock();
            BatchJob* J = nullptr;
            int arc = batch_try_acquire(S->ctx, d, (uint32_t)nb, false, &J);
            if (arc == DMX_ERR_PENDING) arc = batch_try_acquire(S->ctx, d, (uint32_t)nb, true, &J);
            for (size_t m = nb / 2; arc == DMX_ERR_PENDING && m >= 1; m /= 2) {
                arc = batch_try_acquire(S->ctx, d, (uint32_t)m, false, &J);
                if (!arc) nb = m;
            }
            if (arc == DMX_ERR_PENDING && !pending) {  // nothing will free up: create anyway
                arc = batch_acquire(S->ctx, d, (uint32_t)nb, &J);
            }
            lk.lock();
            if (arc == DMX_ERR_PENDING) break;
            dmx_server::Flight f{J, {}, nb * per};
            for (size_t k = 0; k < nb; ++k) f.tickets.push_back(S->queue[pos[k]]);
            for (size_t k = nb; k-- > 0;) S->queue.erase(S->queue.begin() + pos[k]);
            std::vector<JobIn> in(nb);
            for (size_t k = 0; k < nb; ++k) {
                const dmx_job& jj = *S->jobs[f.tickets[k]].job;
                in[k] = JobIn{jj.file, jj.len, &jj.params};
            }
            S->hbm_used += f.bytes;
            lk.unlock();
            int rc = arc;
            const auto tl = std::chrono::steady_clock::now();
            if (!rc) rc = batch_launch(f.J, in.data(), S->cfg.files_on_host != 0, true, true);
            S->launch_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - tl).count();
            S->batches++;
            S->batched_jobs += nb;
            if (rc) {
                if (f.J) {
                    server_finish(S, f, rc);
                } else {  // no workspace: fail these jobs
                    std::lock_guard<std::mutex> g(S->mu);
                    const double t = S->now();
                    for (uint64_t tk : f.tickets) {
                        ServedJob& sj = S->jobs[tk];
                        sj.job->status = rc;
                        sj.done = true;
                        sj.t_done = t;
                    }
                    S->hbm_used -= f.bytes;
                    S->cv_done.notify_all();
                }
                continue;
            }
            S->flights.push_back(std::move(f));
        }
        std::unique_lock<std::mutex> lk(S->mu);
        if (S->stop && S->queue.empty() && S->flights.empty()) break;
        if (S->flights.empty() && S->queue.empty())
            S->cv_work.wait(lk, [S] { return S->stop || !S->queue.empty(); });
        else  // poll the in-flight batches
            S->cv_work.wait_for(lk, std::chrono::microseconds(20));
    }
}

extern "C" int dmx_server_create(dmx_ctx* ctx, const dmx_server_config* cfg, dmx_server** out) {
    if (!ctx || !out) return fail(DMX_ERR_ARG, "null argument");
    if (gpu_disabled_by_env()) return fail(DMX_ERR_NO_DEVICE, "disabled by DAMASCUS_GPU=0");
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<dmx_server> S(new dmx_server);
    S->ctx = ctx;
    if (cfg) S->cfg = *cfg;
    if (!S->cfg.hbm_budget_bytes) {
        size_t free_b = 0, total_b = 0;
        HIP_TRY(hipMemGetInfo(&free_b, &total_b));
        S->cfg.hbm_budget_bytes = free_b / 4 * 3;
    }
    S->fth = std::thread(server_finisher, S.get());
    S->th = std::thread(server_loop, S.get());
    *out = S.release();
    return DMX_OK;
}

extern "C" int dmx_server_submit(dmx_server* S, dmx_job* job, uint64_t* ticket) {
    if (!S || !job || !ticket) return fail(DMX_ERR_ARG, "null argument");
    if (!job->file && job->len) return fail(DMX_ERR_ARG, "null file");
    const int d = depth_for(job->len);
    if (d > 12) return fail(DMX_ERR_ARG, "depth %d exceeds the supported 12", d);
    if (job_workspace_bytes(d) > S->cfg.hbm_budget_bytes)
        return fail(DMX_ERR_MEMORY_LIMIT, "job workspace exceeds the server's device memory budget");
    const int crc = check_file(job->file, job->len, S->cfg.files_on_host != 0, 0);
    if (crc) return crc;
    job->depth = d;
    job->status = DMX_ERR_PENDING;
    std::lock_guard<std::mutex> g(S->mu);
    const uint64_t t = S->next_ticket++;
    ServedJob sj;
    sj.job = job;
    sj.d = d;
    sj.t_submit = S->now();
    S->jobs[t] = sj;
    S->queue.push_back(t);
    *ticket = t;
    S->cv_work.notify_one();
    return DMX_OK;
}

extern "C" int dmx_server_wait(dmx_server* S, uint64_t ticket, int64_t timeout_ms, double* latency_s) {
    if (!S) return fail(DMX_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->jobs.find(ticket) == S->jobs.end())
        return fail(DMX_ERR_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    // find(), not operator[]: a ticket released by another waiter must not be re-created
    auto ready = [&] {
        auto f = S->jobs.find(ticket);
        return f == S->jobs.end() || f->second.done;
    };
    if (timeout_ms < 0)
        S->cv_done.wait(lk, ready);
    else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return DMX_ERR_PENDING;
    auto it = S->jobs.find(ticket);
    if (it == S->jobs.end())
        return fail(DMX_ERR_ARG, "ticket %llu was released by another waiter", (unsigned long long)ticket);
    if (latency_s) *latency_s = it->second.t_done - it->second.t_submit;
    S->jobs.erase(it);
    return DMX_OK;
}

extern "C" int dmx_server_stats(dmx_server* S, dmx_server_stats_t* out) {
    if (!S || !out) return fail(DMX_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(S->mu);
    out->batches = S->batches;
    out->jobs = S->batched_jobs;
    out->launch_seconds = S->launch_s;
    {
        std::lock_guard<std::mutex> g2(g_batches_mu);
        out->workspace_creations = g_batch_creations;
        out->workspace_creation_seconds = g_batch_create_s;
    }
    return DMX_OK;
}

extern "C" void dmx_server_destroy(dmx_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
        S->cv_work.notify_all();
    }
    if (S->th.joinable()) S->th.join();
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->fin_stop = true;
    }
    S->cv_fin.notify_all();
    if (S->fth.joinable()) S->fth.join();
    delete S;
}
