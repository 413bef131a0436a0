"""damascus000_amd — Damascus Fold's commit-and-fold path on MI355X (gfx950).

A Python mirror of the reference's public surface (src/lib.rs:6-10: DamascusProver,
DamascusVerifier, ModuleSisCommitter, SystemParams, RuntimeConfig, Statement / RoundRecord
bytes) over the C ABI of include/damascus_hip.h. All compute runs in libdamascus_hip.so on the
GPU; there is no CPU path. Importing fails loudly when the library has not been built.

Z_q values are Python ints or numpy uint64 arrays whose last axis is (lo, hi), the 16-byte
little-endian encoding of Fp (src/algebra/field.rs:81-83).
"""
import ctypes
import os

import numpy as np

__all__ = [
    "Q", "K", "PRIMES", "DamascusError", "SystemParams", "RuntimeConfig", "Context", "backend_ready",
    "device_summary", "ModuleSisCommitter", "DamascusProver", "DamascusVerifier", "ntt_batch", "fold_batch",
    "ring_mul_batch", "ring_mul_batch_device", "prove", "verify", "prove_verify_device", "prove_verify_many", "Server", "map_file", "derive_subseed", "lib", "LIB_PATH",
    "DeviceBuffer",
]

Q = (1 << 112) - 75
K = 8
PRIMES = (3892314113, 2281701377, 2013265921, 2885681153, 2483027969, 1811939329, 469762049, 4194304001)
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdamascus_hip.so")

DMX_OK = 0
ERRORS = {
    -1: "ERR_ARG", -2: "ERR_NO_DEVICE", -3: "ERR_HIP", -4: "ERR_MEMORY_LIMIT", -5: "ERR_ROUND",
    -6: "ERR_INVARIANT", -7: "ERR_SHAPE", -8: "ERR_OPENING", -9: "ERR_DECODE", -10: "ERR_NOT_FINAL",
    -11: "ERR_ZERO_TERMS", -12: "ERR_OOM", -13: "ERR_BUFFER", -14: "ERR_POISONED", -15: "ERR_PENDING",
}
ERR_CODE = {v: k for k, v in ERRORS.items()}


class DamascusError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"{ERRORS.get(code, code)}: {message}")
        self.code = code
        self.name = ERRORS.get(code, str(code))


_lib = None
_c = ctypes
_VP = ctypes.c_void_p

# The register-blocked forward-transform kernels per degree 2^log_m, variants 0 and 1, as
# launch_ntt_fwd_rows2 (csrc/dmx_ntt.hip) launches them; the last variant is k_ntt_fwd_rows.
FWD_KERNELS = {
    6: ("k_ntt_fwd_reg<8, 8>", "k_ntt_fwd_reg<16, 4>"),
    7: ("k_ntt_fwd_reg<16, 8>", "k_ntt_fwd_reg<32, 4>"),
    8: ("k_ntt_fwd_reg<16, 16>", "k_ntt_fwd_reg3<8, 32>"),
    9: ("k_ntt_fwd_reg<32, 16>", "k_ntt_fwd_reg3<9, 64>"),
    10: ("k_ntt_fwd_reg<32, 32>", "k_ntt_fwd_reg3<10, 64>"),
    11: ("k_ntt_fwd_reg3<11, 128>", "k_ntt_fwd_reg3<11, 64>"),
}
_SZ = ctypes.c_size_t
_SZP = ctypes.POINTER(ctypes.c_size_t)


class SystemParams(ctypes.Structure):
    """SystemParams (src/utils/config.rs:24-44): seeds only, the dimensions follow the file."""
    _fields_ = [("seed_generators", ctypes.c_uint8 * 32), ("epoch_seed", ctypes.c_uint8 * 32)]

    def __init__(self, seed_generators=bytes([7] * 32), epoch_seed=bytes([11] * 32)):
        super().__init__()
        self.seed_generators[:] = list(bytes(seed_generators))
        self.epoch_seed[:] = list(bytes(epoch_seed))


class RuntimeConfig(ctypes.Structure):
    """RuntimeConfig (src/utils/config.rs:69-87)"""
    _fields_ = [("max_preprocess_bytes", ctypes.c_uint64), ("ntt_enabled", ctypes.c_int32),
                ("parallel_enabled", ctypes.c_int32), ("gpu_enabled", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("gpu_min_elements", ctypes.c_uint64)]

    def __init__(self, max_preprocess_bytes=512 << 20, ntt_enabled=True, parallel_enabled=True, gpu_enabled=True,
                 gpu_min_elements=4096):
        super().__init__()
        self.max_preprocess_bytes = max_preprocess_bytes
        self.ntt_enabled = int(ntt_enabled)
        self.parallel_enabled = int(parallel_enabled)
        self.gpu_enabled = int(gpu_enabled)
        self.gpu_min_elements = gpu_min_elements


class RoundOutput(ctypes.Structure):
    _fields_ = [("vector_fold_ns", ctypes.c_uint64), ("poly_fold_ns", ctypes.c_uint64),
                ("total_round_ns", ctypes.c_uint64)]


class Job(ctypes.Structure):
    """dmx_job: one file of a ragged batch and where its proof goes"""
    _fields_ = [("file", _VP), ("len", ctypes.c_uint64), ("params", SystemParams), ("statement", _VP),
                ("statement_cap", _SZ), ("statement_len", _SZ), ("records", _VP), ("records_cap", _SZ),
                ("records_len", _SZ), ("opening", ctypes.c_uint64 * 2), ("depth", ctypes.c_int32),
                ("status", ctypes.c_int32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_uint64), ("jobs", ctypes.c_uint64), ("launch_seconds", ctypes.c_double),
                ("workspace_creations", ctypes.c_uint64), ("workspace_creation_seconds", ctypes.c_double)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_batch", ctypes.c_uint32),
                ("max_inflight", ctypes.c_uint32), ("files_on_host", ctypes.c_int32), ("reserved", ctypes.c_int32)]


# name -> (restype, argtypes)
_SIGS = {
    "dmx_backend_ready": (ctypes.c_int, []),
    "dmx_device_summary": (ctypes.c_int, [ctypes.c_char_p, _SZ]),
    "dmx_last_error": (ctypes.c_char_p, []),
    "dmx_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_VP)]),
    "dmx_ctx_destroy": (None, [_VP]),
    "dmx_ntt_batch": (ctypes.c_int, [_VP] * 3 + [_SZ] + [_VP] * 4),
    "dmx_fold_batch": (ctypes.c_int, [_VP] * 6 + [_SZ, ctypes.c_uint64, ctypes.c_uint64]),
    "dmx_ring_mul_batch": (ctypes.c_int, [_VP] * 5 + [_SZ]),
    "dmx_ring_mul_batch_device": (ctypes.c_int, [_VP] * 5 + [_SZ, ctypes.c_int]),
    "dmx_ring_mul_batch_device_noalloc": (ctypes.c_int, [_VP] * 5 + [_SZ]),
    "dmx_ring_mul_batch_cfg": (ctypes.c_int, [_VP] * 5 + [_SZ, _VP]),
    "dmx_derive_subseed": (ctypes.c_int, [_VP, _VP, _SZ, _VP]),
    "dmx_derive_generators": (ctypes.c_int, [_VP, _VP, _SZ, _SZ, _VP]),
    "dmx_commit": (ctypes.c_int, [_VP, _VP, _VP, _SZ, _SZ, _VP, _SZP]),
    "dmx_commit_with_generators": (ctypes.c_int, [_VP, _VP, _VP, _SZ, _SZ, _VP, _SZP]),
    "dmx_prover_new": (ctypes.c_int, [_VP, _VP, _SZ, _VP, _VP, ctypes.POINTER(_VP)]),
    "dmx_prover_free": (None, [_VP]),
    "dmx_prover_rounds_total": (_SZ, [_VP]),
    "dmx_prover_current_round": (_SZ, [_VP]),
    "dmx_prover_current_dimensions": (None, [_VP, _SZP, _SZP]),
    "dmx_prover_statement": (ctypes.c_int, [_VP, _VP, _SZ, _SZP]),
    "dmx_prover_current_commitment": (ctypes.c_int, [_VP, _VP, _SZ, _SZP]),
    "dmx_prover_fold_round": (ctypes.c_int, [_VP, _SZ, _VP, _SZ, _SZP, _VP]),
    "dmx_prover_final_opening": (ctypes.c_int, [_VP, _VP]),
    "dmx_verifier_new": (ctypes.c_int, [_VP, _VP, _VP, _SZ, ctypes.POINTER(_VP)]),
    "dmx_verifier_clone": (_VP, [_VP]),
    "dmx_verifier_free": (None, [_VP]),
    "dmx_verifier_round": (_SZ, [_VP]),
    "dmx_verifier_current_dimensions": (None, [_VP, _SZP, _SZP]),
    "dmx_verifier_current_commitment": (ctypes.c_int, [_VP, _VP, _SZ, _SZP]),
    "dmx_verifier_update_commitment": (ctypes.c_int, [_VP, _VP, _SZ]),
    "dmx_verifier_verify_final_opening": (ctypes.c_int, [_VP, _VP]),
    "dmx_prove": (ctypes.c_int, [_VP, _VP, _SZ, _VP, _VP, _VP, _SZ, _SZP, _VP, _SZ, _SZP, _VP]),
    "dmx_verify": (ctypes.c_int, [_VP, _VP, _VP, _SZ, _VP, _SZ, _VP]),
    "dmx_prove_verify_device": (ctypes.c_int, [_VP, _VP, _SZ, _VP, _VP, _SZ, _SZP, _VP, _SZ, _SZP, _VP]),
    "dmx_prove_verify_many": (ctypes.c_int, [_VP, _VP, _SZ, ctypes.c_int, _SZ]),
    "dmx_server_create": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(_VP)]),
    "dmx_server_submit": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(ctypes.c_uint64)]),
    "dmx_server_wait": (ctypes.c_int, [_VP, ctypes.c_uint64, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]),
    "dmx_server_destroy": (None, [_VP]),
    "dmx_server_stats": (ctypes.c_int, [_VP, _VP]),
    "dmx_device_alloc": (ctypes.c_int, [_VP, _SZ, ctypes.POINTER(_VP)]),
    "dmx_device_free": (ctypes.c_int, [_VP, _VP]),
    "dmx_memcpy_htod": (ctypes.c_int, [_VP, _VP, _VP, _SZ]),
    "dmx_memcpy_dtoh": (ctypes.c_int, [_VP, _VP, _VP, _SZ]),
    "dmx_device_copy": (ctypes.c_int, [_VP, _VP, _VP, _SZ]),
    "dmx_kernel_timing_enable": (ctypes.c_int, [_VP, ctypes.c_int]),
    "dmx_kernel_timing_degree": (ctypes.c_int, [_VP, ctypes.c_int]),
    "dmx_ctx_transform_variant": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                 ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)]),
    "dmx_ctx_set_transform_variant": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int]),
    "dmx_ctx_mac_shape": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)]),
    "dmx_ctx_set_mac_shape": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int]),
    "dmx_ctx_tuning_source": (ctypes.c_int, [_VP, ctypes.c_char_p, _SZ]),
    "dmx_kernel_timing": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64),
                                         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "dmx_statement_size": (_SZ, [_SZ]),
    "dmx_record_size": (_SZ, [_SZ, _SZ]),
    "dmx_proof_records_size": (_SZ, [ctypes.c_uint64]),
    "dmx_job_workspace_size": (_SZ, [ctypes.c_uint64]),
}


def lib():
    """The loaded libdamascus_hip.so. Raises ImportError when it was never built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C damascus000_amd/csrc` "
                              "(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_VP)


def _check(rc):
    if rc != DMX_OK:
        raise DamascusError(rc, lib().dmx_last_error().decode(errors="replace"))


def _bytes_ptr(b):
    buf = ctypes.create_string_buffer(bytes(b), max(1, len(b)))
    return buf


def to_words(vals):
    vals = [int(v) for v in vals]
    out = np.zeros((len(vals), 2), dtype=np.uint64)
    for i, v in enumerate(vals):
        out[i, 0] = v & 0xFFFFFFFFFFFFFFFF
        out[i, 1] = v >> 64
    return out


def from_words(w):
    w = np.asarray(w, dtype=np.uint64).reshape(-1, 2)
    return [int(lo) | (int(hi) << 64) for lo, hi in w]


def backend_ready():
    """cuda_backend_ready (src/utils/gpu.rs:23-25): library built, device present, DAMASCUS_GPU != 0"""
    return bool(lib().dmx_backend_ready())


def device_summary():
    buf = ctypes.create_string_buffer(256)
    lib().dmx_device_summary(buf, 256)
    return buf.value.decode()


class Context:
    """Device context: NTT tables and CRT constants resident in HBM, shared by every object."""
    _default = None

    def __init__(self, device=0):
        self.h = _VP()
        _check(lib().dmx_ctx_create(device, ctypes.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.dmx_ctx_destroy(self.h)
            self.h = None

    @classmethod
    def default(cls):
        if cls._default is None:
            cls._default = Context()
        return cls._default

    def upload(self, data):
        """copy host bytes into a new device buffer; returns a DeviceBuffer"""
        buf = DeviceBuffer(self, len(data))
        _check(lib().dmx_memcpy_htod(self.h, _VP(buf.ptr), _bytes_ptr(data), len(data)))
        return buf

    def kernel_timing(self, on=None, log_m=-1, full_rate=False):
        """enable (on=True/False; log_m >= 0 records only degree 2^log_m) or read the
        forward-transform timing record: returns (total_ms, launches, lane_cycles) and clears it;
        with full_rate, (total_ms, launches, lane_cycles, lane_cycles at full-rate multiplies)"""
        if on is not None:
            _check(lib().dmx_kernel_timing_degree(self.h, log_m if on else -1))
            _check(lib().dmx_kernel_timing_enable(self.h, int(bool(on))))
            return None
        ms, n, w, wf = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_double(), ctypes.c_double()
        _check(lib().dmx_kernel_timing(self.h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(w), ctypes.byref(wf)))
        if full_rate:
            return ms.value, n.value, w.value, wf.value
        return ms.value, n.value, w.value

    def transform_variant(self, log_m):
        """(chosen variant, variant count, tuning times in us) of the degree-2^log_m forward
        transform"""
        v, n = ctypes.c_int(), ctypes.c_int()
        us = (ctypes.c_float * 4)()
        _check(lib().dmx_ctx_transform_variant(self.h, log_m, ctypes.byref(v), ctypes.byref(n), us))
        return v.value, n.value, [round(us[i], 2) for i in range(n.value)]

    def set_transform_variant(self, log_m, variant):
        _check(lib().dmx_ctx_set_transform_variant(self.h, log_m, variant))

    def transform_kernel(self, log_m):
        """name of the kernel that runs the degree-2^log_m forward transform (the chosen variant)"""
        v, n, _ = self.transform_variant(log_m)
        return FWD_KERNELS[log_m][v] if v < n - 1 and log_m in FWD_KERNELS else "k_ntt_fwd_rows"

    MAC_SHAPES = 9

    def mac_shape(self, log_m):
        """(chosen shape, shape count, tuning times in us) of the degree-2^log_m MAC passes; shape
        i splits rows into segments of at least 2^(3 + i // 3) rows, at most 2^(17 + i % 3) threads"""
        v, n = ctypes.c_int(), ctypes.c_int()
        us = (ctypes.c_float * self.MAC_SHAPES)()
        _check(lib().dmx_ctx_mac_shape(self.h, log_m, ctypes.byref(v), ctypes.byref(n), us))
        return v.value, n.value, [round(us[i], 2) for i in range(n.value)]

    def set_mac_shape(self, log_m, shape):
        _check(lib().dmx_ctx_set_mac_shape(self.h, log_m, shape))

    def tuning_source(self):
        """("defaults" | "timed" | "cache", cache file or "")"""
        buf = ctypes.create_string_buffer(1024)
        src = lib().dmx_ctx_tuning_source(self.h, buf, len(buf))
        return {0: "defaults", 1: "timed", 2: "cache"}.get(src, "unknown"), buf.value.decode()


class DeviceBuffer:
    """device memory owned through the context (no torch: one HIP runtime per process)"""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = nbytes
        p = _VP()
        _check(lib().dmx_device_alloc(ctx.h, nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def copy_from(self, other):
        _check(lib().dmx_device_copy(self.ctx.h, _VP(self.ptr), _VP(other.ptr), min(self.nbytes, other.nbytes)))

    def read(self, nbytes=None):
        n = self.nbytes if nbytes is None else nbytes
        out = ctypes.create_string_buffer(max(n, 1))
        _check(lib().dmx_memcpy_dtoh(self.ctx.h, out, _VP(self.ptr), n))
        return out.raw[:n]

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.dmx_device_free(self.ctx.h, _VP(self.ptr))
            self.ptr = None


def derive_subseed(seed, label):
    """derive_subseed (src/commitment/sis.rs:208-210): keyed BLAKE3(seed, label)"""
    out = ctypes.create_string_buffer(32)
    _check(lib().dmx_derive_subseed(_bytes_ptr(seed), _bytes_ptr(label), len(label), out))
    return out.raw


class ModuleSisCommitter:
    """ModuleSisCommitter (src/commitment/sis.rs:67-196) on the GPU"""

    def __init__(self, seed=bytes([7] * 32), ctx=None):
        self.ctx = ctx or Context.default()
        self.seed = bytes(seed)
        self.g_seed = derive_subseed(self.seed, b"g0")

    def generator_seed(self):
        return self.g_seed

    def generators_for(self, vector_len, ring_len):
        out = np.zeros((vector_len, K, ring_len, 2), dtype=np.uint64)
        _check(lib().dmx_derive_generators(self.ctx.h, _bytes_ptr(self.g_seed), vector_len, ring_len, _ptr(out)))
        return out

    def commit(self, witness):
        """witness: (rows, ring_len, 2) words -> (8, ring_len_out, 2)"""
        w = np.ascontiguousarray(witness, dtype=np.uint64)
        rows = w.shape[0]
        n = w.shape[1] if rows else 1
        out = np.zeros((K, max(n, 1), 2), dtype=np.uint64)
        ol = _SZ(0)
        _check(lib().dmx_commit(self.ctx.h, _bytes_ptr(self.seed), _ptr(w), rows, n, _ptr(out), ctypes.byref(ol)))
        return out[:, : ol.value]

    def commit_with_generators(self, witness, gens):
        w = np.ascontiguousarray(witness, dtype=np.uint64)
        g = np.ascontiguousarray(gens, dtype=np.uint64)
        rows = w.shape[0]
        n = w.shape[1] if rows else 1
        out = np.zeros((K, max(n, 1), 2), dtype=np.uint64)
        ol = _SZ(0)
        _check(lib().dmx_commit_with_generators(self.ctx.h, _ptr(w), _ptr(g), rows, n, _ptr(out), ctypes.byref(ol)))
        return out[:, : ol.value]


class DamascusProver:
    """DamascusProver (src/protocol/prover.rs:41-313)"""

    def __init__(self, handle, ctx):
        self.h = handle
        self.ctx = ctx

    @classmethod
    def initialize(cls, data, params=None, config=None, ctx=None):
        """initialize / initialize_with_config (prover.rs:57-111) from the file's bytes"""
        ctx = ctx or Context.default()
        params = params or SystemParams()
        config = config or RuntimeConfig()
        h = _VP()
        buf = _bytes_ptr(data)
        _check(lib().dmx_prover_new(ctx.h, buf, len(data), ctypes.byref(params), ctypes.byref(config), ctypes.byref(h)))
        return cls(h, ctx)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.dmx_prover_free(self.h)
            self.h = None

    def rounds_total(self):
        return lib().dmx_prover_rounds_total(self.h)

    def current_round(self):
        return lib().dmx_prover_current_round(self.h)

    def current_dimensions(self):
        r, n = _SZ(), _SZ()
        lib().dmx_prover_current_dimensions(self.h, ctypes.byref(r), ctypes.byref(n))
        return r.value, n.value

    def statement(self):
        """bincode bytes of the Statement (src/commitment/sis.rs:23-30)"""
        n = _SZ()
        _check(lib().dmx_prover_statement(self.h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(lib().dmx_prover_statement(self.h, buf, n.value, ctypes.byref(n)))
        return buf.raw

    def current_commitment(self):
        _, n = self.current_dimensions()
        out = np.zeros((K, n, 2), dtype=np.uint64)
        ln = _SZ()
        _check(lib().dmx_prover_current_commitment(self.h, _ptr(out), out.size, ctypes.byref(ln)))
        return out

    def fold_round(self, round_idx, with_timing=False):
        """fold_round (prover.rs:140-313): returns the bincode RoundRecord"""
        _, n = self.current_dimensions()
        cap = lib().dmx_record_size(max(n, 1), max(n // 2, 1))
        buf = ctypes.create_string_buffer(cap)
        got = _SZ()
        t = RoundOutput()
        _check(lib().dmx_prover_fold_round(self.h, round_idx, buf, cap, ctypes.byref(got), ctypes.byref(t)))
        rec = buf.raw[: got.value]
        if with_timing:
            return rec, {"vector_fold_ns": t.vector_fold_ns, "poly_fold_ns": t.poly_fold_ns,
                         "total_round_ns": t.total_round_ns}
        return rec

    def final_opening(self):
        """final_opening (prover.rs:133-138): None until every round has run"""
        out = np.zeros(2, dtype=np.uint64)
        rc = lib().dmx_prover_final_opening(self.h, _ptr(out))
        return None if rc != DMX_OK else from_words(out)[0]


class DamascusVerifier:
    """DamascusVerifier (src/protocol/verifier.rs:11-139)"""

    def __init__(self, params, statement, ctx=None, _handle=None):
        self.ctx = ctx or Context.default()
        if _handle is not None:
            self.h = _VP(_handle)
            return
        self.h = _VP()
        params = params or SystemParams()
        st = bytes(statement)
        _check(lib().dmx_verifier_new(self.ctx.h, ctypes.byref(params), _bytes_ptr(st), len(st), ctypes.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.dmx_verifier_free(self.h)
            self.h = None

    def clone(self):
        h = lib().dmx_verifier_clone(self.h)
        if not h:
            raise DamascusError(-3, "verifier clone failed")
        return DamascusVerifier(None, None, ctx=self.ctx, _handle=h)

    def round(self):
        return lib().dmx_verifier_round(self.h)

    def current_dimensions(self):
        r, n = _SZ(), _SZ()
        lib().dmx_verifier_current_dimensions(self.h, ctypes.byref(r), ctypes.byref(n))
        return r.value, n.value

    def current_commitment(self):
        _, n = self.current_dimensions()
        out = np.zeros((K, n, 2), dtype=np.uint64)
        ln = _SZ()
        _check(lib().dmx_verifier_current_commitment(self.h, _ptr(out), out.size, ctypes.byref(ln)))
        return out

    def update_commitment(self, record):
        rec = bytes(record)
        _check(lib().dmx_verifier_update_commitment(self.h, _bytes_ptr(rec), len(rec)))

    def verify_final_opening(self, opening):
        _check(lib().dmx_verifier_verify_final_opening(self.h, _ptr(to_words([opening]))))


def job_workspace_size(nbytes):
    """device bytes of one job's workspace for a file of nbytes (a server budget's unit)"""
    return int(lib().dmx_job_workspace_size(int(nbytes)))


def _ring_len(nbytes):
    """N = 2^d of a file of nbytes: c = max(1, ceil(len / 13)) coefficients and the smallest
    d >= 6 with 4^d >= c (src/utils/io.rs:72-120); sizes the statement buffer."""
    c = max(1, -(-int(nbytes) // 13))
    d = 6
    while 4**d < c:
        d += 1
    return 1 << d


def _out_buffer(n):
    """an uninitialised host buffer of at least one byte (the library writes what it reports)"""
    return np.empty(max(int(n), 1), dtype=np.uint8)


def prove(data, params=None, config=None, ctx=None):
    """initialize + every fold_round; returns (statement, records_blob, opening)"""
    ctx = ctx or Context.default()
    params = params or SystemParams()
    config = config or RuntimeConfig(max_preprocess_bytes=(1 << 64) - 1)
    st = _out_buffer(lib().dmx_statement_size(_ring_len(len(data))))
    recs = _out_buffer(lib().dmx_proof_records_size(len(data)))
    sl, rl = _SZ(), _SZ()
    op = np.zeros(2, dtype=np.uint64)
    _check(lib().dmx_prove(ctx.h, _bytes_ptr(data), len(data), ctypes.byref(params), ctypes.byref(config),
                           st.ctypes.data, st.size, ctypes.byref(sl), recs.ctypes.data, recs.size, ctypes.byref(rl),
                           _ptr(op)))
    return st[: sl.value].tobytes(), recs[: rl.value].tobytes(), from_words(op)[0]


def verify(statement, records, opening, params=None, ctx=None):
    ctx = ctx or Context.default()
    params = params or SystemParams()
    st, rb = bytes(statement), bytes(records)
    _check(lib().dmx_verify(ctx.h, ctypes.byref(params), _bytes_ptr(st), len(st), _bytes_ptr(rb), len(rb),
                            _ptr(to_words([opening]))))


def prove_verify_device(d_file_ptr, length, params=None, ctx=None, want_proof=True):
    """prove + replay with the file already in device memory (a device pointer as int).
    Returns (statement, records_blob, opening) when want_proof, else None."""
    ctx = ctx or Context.default()
    params = params or SystemParams()
    if want_proof:
        st = _out_buffer(lib().dmx_statement_size(_ring_len(length)))
        recs = _out_buffer(lib().dmx_proof_records_size(length))
        sl, rl = _SZ(), _SZ()
        op = np.zeros(2, dtype=np.uint64)
        _check(lib().dmx_prove_verify_device(ctx.h, _VP(d_file_ptr), length, ctypes.byref(params), st.ctypes.data,
                                             st.size, ctypes.byref(sl), recs.ctypes.data, recs.size, ctypes.byref(rl),
                                             _ptr(op)))
        return st[: sl.value].tobytes(), recs[: rl.value].tobytes(), from_words(op)[0]
    _check(lib().dmx_prove_verify_device(ctx.h, _VP(d_file_ptr), length, ctypes.byref(params), None, 0, None, None, 0,
                                         None, None))
    return None


def prove_verify_many(files, params=None, ctx=None, max_batch=0, want_proof=True, raise_on_error=True,
                      files_on_host=None):
    """prove + replay of many files in one ragged batch (dmx_prove_verify_many).

    files: bytes objects (host memory) or (pointer, length) pairs: device pointers, or host
    pointers (e.g. MappedFile.job()) with files_on_host=True.
    params: one SystemParams for all, or a list with one per file.
    Returns a list of (statement, records_blob, opening) per file, or of status codes when
    raise_on_error is False and a job failed (then the entry is the negative code)."""
    ctx = ctx or Context.default()
    n = len(files)
    if n == 0:
        return []
    kinds = {isinstance(f, tuple) for f in files}
    if len(kinds) > 1:
        raise TypeError("prove_verify_many: every file must be of one kind, bytes or (pointer, length) pairs")
    on_host = (not isinstance(files[0], tuple)) if files_on_host is None else bool(files_on_host)
    if not on_host and not isinstance(files[0], tuple):
        raise TypeError("prove_verify_many: bytes are host memory; pass files_on_host=True or device (ptr, len) pairs")
    plist = params if isinstance(params, (list, tuple)) else [params or SystemParams()] * n
    jobs = (Job * n)()
    keep = []
    for i, f in enumerate(files):
        if not isinstance(f, tuple):
            buf = ctypes.create_string_buffer(bytes(f), max(len(f), 1))
            keep.append(buf)
            jobs[i].file = ctypes.cast(buf, _VP)
            jobs[i].len = len(f)
        else:
            jobs[i].file = _VP(f[0])
            jobs[i].len = f[1]
        jobs[i].params = plist[i]
        if want_proof:
            st = _out_buffer(lib().dmx_statement_size(_ring_len(jobs[i].len)))
            rc = _out_buffer(lib().dmx_proof_records_size(jobs[i].len))
            keep += [st, rc]
            jobs[i].statement, jobs[i].statement_cap = st.ctypes.data, st.size
            jobs[i].records, jobs[i].records_cap = rc.ctypes.data, rc.size
    rc = lib().dmx_prove_verify_many(ctx.h, jobs, n, int(on_host), max_batch)
    if rc != DMX_OK and raise_on_error:
        _check(rc)
    out = []
    for i in range(n):
        j = jobs[i]
        if j.status != DMX_OK:
            out.append(int(j.status))
        elif want_proof:
            st = ctypes.string_at(j.statement, j.statement_len)
            recs = ctypes.string_at(j.records, j.records_len)
            out.append((st, recs, int(j.opening[0]) | (int(j.opening[1]) << 64)))
        else:
            out.append(None)
    return out


def _job_for(f, params, want_proof, keep):
    j = Job()
    if isinstance(f, tuple):
        j.file, j.len = _VP(f[0]), f[1]
    else:
        buf = ctypes.create_string_buffer(bytes(f), max(len(f), 1))
        keep.append(buf)
        j.file, j.len = ctypes.cast(buf, _VP), len(f)
    j.params = params
    if want_proof:
        st = _out_buffer(lib().dmx_statement_size(_ring_len(j.len)))
        rc = _out_buffer(lib().dmx_proof_records_size(j.len))
        keep += [st, rc]
        j.statement, j.statement_cap = st.ctypes.data, st.size
        j.records, j.records_cap = rc.ctypes.data, rc.size
    return j


def _job_result(j, want_proof):
    if j.status != DMX_OK:
        return int(j.status)
    if not want_proof:
        return None
    return (ctypes.string_at(j.statement, j.statement_len), ctypes.string_at(j.records, j.records_len),
            int(j.opening[0]) | (int(j.opening[1]) << 64))


class Server:
    """Continuous-batching prove+verify service (dmx_server_*). Files are (device_pointer,
    length) pairs, or bytes when files_on_host. submit() returns a ticket; wait(ticket) returns
    (result, latency_seconds) where result is (statement, records, opening) or an error code."""

    def __init__(self, ctx=None, hbm_budget=0, max_batch=0, max_inflight=0, files_on_host=False):
        self.ctx = ctx or Context.default()
        self.files_on_host = bool(files_on_host)
        cfg = ServerConfig(hbm_budget, max_batch, max_inflight, int(files_on_host), 0)
        h = _VP()
        _check(lib().dmx_server_create(self.ctx.h, ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self._live = {}

    def submit(self, f, params=None, want_proof=True):
        if not isinstance(f, tuple) and not self.files_on_host:
            raise TypeError("Server.submit: bytes are host memory, but this server takes device (ptr, len) pairs; "
                            "create it with files_on_host=True or upload the file first")
        keep = []
        j = _job_for(f, params or SystemParams(), want_proof, keep)
        t = ctypes.c_uint64()
        _check(lib().dmx_server_submit(self.h, ctypes.byref(j), ctypes.byref(t)))
        self._live[t.value] = (j, keep, want_proof)
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        lat = ctypes.c_double()
        rc = lib().dmx_server_wait(self.h, ticket, timeout_ms, ctypes.byref(lat))
        if rc == ERR_CODE["ERR_PENDING"]:
            return None
        _check(rc)
        j, keep, want = self._live.pop(ticket)
        return _job_result(j, want), lat.value

    def stats(self):
        st = ServerStats()
        _check(lib().dmx_server_stats(self.h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            lib().dmx_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ntt_batch(a, b, n, primes, stage_roots, inv_roots, inv_sizes):
    """damascus_cuda_ntt_batch replacement (src/utils/gpu.rs:140-207): returns (rc, out)"""
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.zeros_like(a)
    arrs = [np.ascontiguousarray(x, dtype=np.uint64) for x in (primes, stage_roots, inv_roots, inv_sizes)]
    rc = lib().dmx_ntt_batch(_ptr(a), _ptr(b), _ptr(out), n, *[_ptr(x) for x in arrs])
    return rc, out


def fold_batch(left, right, challenge):
    """damascus_cuda_fold_batch replacement (src/utils/gpu.rs:95-138) on lists of ints"""
    n = len(left)
    lw, rw = to_words(left), to_words(right)
    l_lo, l_hi = np.ascontiguousarray(lw[:, 0]), np.ascontiguousarray(lw[:, 1])
    r_lo, r_hi = np.ascontiguousarray(rw[:, 0]), np.ascontiguousarray(rw[:, 1])
    o_lo, o_hi = np.zeros(n, np.uint64), np.zeros(n, np.uint64)
    rc = lib().dmx_fold_batch(_ptr(l_lo), _ptr(l_hi), _ptr(r_lo), _ptr(r_hi), _ptr(o_lo), _ptr(o_hi), n,
                              challenge & 0xFFFFFFFFFFFFFFFF, challenge >> 64)
    if rc != 0:
        return rc, None
    return rc, [int(a) | (int(b) << 64) for a, b in zip(o_lo, o_hi)]


def ring_mul_batch(a_list, b_list, ctx=None, config=None):
    """negacyclic products a_i * b_i of mixed power-of-two degrees; lists of (n_i, 2) words.
    config.ntt_enabled = False takes the schoolbook product (RingElement::mul, poly.rs:134-140)."""
    ctx = ctx or Context.default()
    logs = np.array([int(x.shape[0]).bit_length() - 1 for x in a_list], dtype=np.uint32)
    A = np.ascontiguousarray(np.concatenate(a_list), dtype=np.uint64)
    B = np.ascontiguousarray(np.concatenate(b_list), dtype=np.uint64)
    out = np.zeros_like(A)
    if config is None:
        _check(lib().dmx_ring_mul_batch(ctx.h, _ptr(A), _ptr(B), _ptr(out), _ptr(logs), len(a_list)))
    else:
        _check(lib().dmx_ring_mul_batch_cfg(ctx.h, _ptr(A), _ptr(B), _ptr(out), _ptr(logs), len(a_list),
                                            ctypes.byref(config)))
    res, pos = [], 0
    for x in a_list:
        res.append(out[pos: pos + x.shape[0]])
        pos += x.shape[0]
    return res


def ring_mul_batch_device(d_a, d_b, d_out, log_n, mode=0, ctx=None):
    """ragged products on device-resident coefficient blocks (device pointers as ints; log_n a
    list of degrees' log2). mode 0: one persistent launch; mode 1: per-degree launches."""
    ctx = ctx or Context.default()
    logs = np.ascontiguousarray(np.asarray(log_n, dtype=np.uint32))
    _check(lib().dmx_ring_mul_batch_device(ctx.h, _VP(d_a), _VP(d_b), _VP(d_out), _ptr(logs), len(logs), mode))


class MappedFile:
    """a read-only memory map of a file, the way the reference reads its input
    (src/utils/io.rs:21-70 maps the file): .ptr / .len describe the bytes in host memory, for
    a files_on_host Server or prove_verify_many with (ptr, len) pairs."""

    def __init__(self, path):
        self.len = os.path.getsize(path)
        if self.len:
            self._map = np.memmap(path, dtype=np.uint8, mode="r")
            self.ptr = self._map.ctypes.data
        else:
            self._map = np.zeros(1, dtype=np.uint8)
            self.ptr = self._map.ctypes.data

    def job(self):
        return (self.ptr, self.len)


def map_file(path):
    return MappedFile(path)
