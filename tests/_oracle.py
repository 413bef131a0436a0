"""ctypes bindings for the oracle (oracle/liboracle.so), the CPU restatement of the reference.

Test infrastructure only: the oracle is the checker, never the thing measured or shipped.
Z_q arrays are numpy uint64 arrays whose last axis holds (lo, hi) words, the 16-byte
little-endian layout of Fp::to_le_bytes (src/algebra/field.rs:81-83).
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
Q = (1 << 112) - 75
K = 8
PRIMES = [3892314113, 2281701377, 2013265921, 2885681153, 2483027969, 1811939329, 469762049,
          4194304001]

ORC_OK = 0
ORC_ERR_MEMORY_LIMIT = -2
ORC_ERR_ROUND = -3
ORC_ERR_INVARIANT = -4
ORC_ERR_SHAPE = -5
ORC_ERR_OPENING = -6
ORC_ERR_DECODE = -7
ORC_ERR_NOT_FINAL = -8
ORC_ERR_ZERO_TERMS = -9

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                           stdout=subprocess.DEVNULL)
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.orc_prover_rounds.restype = ctypes.c_size_t
        _lib.orc_prover_round.restype = ctypes.c_size_t
        _lib.orc_prover_statement.restype = ctypes.c_size_t
        _lib.orc_verifier_round.restype = ctypes.c_size_t
        _lib.orc_record_size.restype = ctypes.c_size_t
        _lib.orc_coeff_count_for_byte_len.restype = ctypes.c_size_t
        _lib.orc_multiplicative_generator.restype = ctypes.c_uint64
        _lib.orc_multiplicative_generator.argtypes = [ctypes.c_uint64]
        _lib.orc_verifier_clone.restype = ctypes.c_void_p
        _lib.orc_get_threads.restype = ctypes.c_int
    return _lib


def P(a):
    """pointer to a contiguous numpy array"""
    return a.ctypes.data_as(ctypes.c_void_p)


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


def random_zq(rng, n):
    """n uniform Z_q values as words (rejection sampling on 112 bits)"""
    out = np.zeros((n, 2), dtype=np.uint64)
    lo = rng.integers(0, 2**64, size=n, dtype=np.uint64)
    hi = rng.integers(0, 2**48, size=n, dtype=np.uint64)
    out[:, 0] = lo
    out[:, 1] = hi
    # values >= q (probability ~2^-105 each) are folded down; keeps arrays canonical
    for i in range(n):
        if (int(hi[i]) << 64 | int(lo[i])) >= Q:
            out[i, 1] = 0
    return out


def set_threads(n):
    lib().orc_set_threads(ctypes.c_int(n))


def set_challenge_reject(n):
    """test knob: every challenge rejects the candidates of counters below n (0: the reference)"""
    lib().orc_set_challenge_reject(ctypes.c_uint32(n))


def sample_nonzero_mock(zero_below, value, cap=64):
    """sample_nonzero_challenge (transcript.rs:66-75) on the mock candidate of transcript.rs:130-143:
    returns (sampled value, counters seen)"""
    seen = (ctypes.c_uint32 * cap)()
    n = ctypes.c_size_t(0)
    lib().orc_sample_nonzero_mock.restype = ctypes.c_uint64
    v = lib().orc_sample_nonzero_mock(ctypes.c_uint32(zero_below), ctypes.c_uint64(value), seen,
                                      ctypes.c_size_t(cap), ctypes.byref(n))
    return int(v), [int(seen[i]) for i in range(min(n.value, cap))]


# ---- Z_q ----
def zq_binop(name, a, b):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.zeros_like(a)
    getattr(lib(), "orc_zq_" + name)(P(a), P(b), P(out), ctypes.c_size_t(a.shape[0]))
    return out


def zq_unop(name, a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    out = np.zeros_like(a)
    getattr(lib(), "orc_zq_" + name)(P(a), P(out), ctypes.c_size_t(a.shape[0]))
    return out


def from_le_bytes_mod_order(raw):
    out = np.zeros(2, dtype=np.uint64)
    buf = ctypes.create_string_buffer(bytes(raw), max(1, len(raw)))
    lib().orc_zq_from_le_bytes_mod_order(buf, ctypes.c_size_t(len(raw)), P(out))
    return from_words(out)[0]


def from_le_chunk(raw):
    out = np.zeros(2, dtype=np.uint64)
    buf = ctypes.create_string_buffer(bytes(raw), max(1, len(raw)))
    lib().orc_zq_from_le_chunk(buf, ctypes.c_size_t(len(raw)), P(out))
    return from_words(out)[0]


# ---- BLAKE3 ----
def blake3(data, key=None, n=32):
    out = ctypes.create_string_buffer(n)
    buf = ctypes.create_string_buffer(bytes(data), max(1, len(data)))
    kb = ctypes.create_string_buffer(bytes(key), 32) if key is not None else None
    lib().orc_blake3_hash_xof(kb, buf, ctypes.c_size_t(len(data)), out, ctypes.c_size_t(n))
    return out.raw[:n]


# ---- NTT / products ----
def ntt_forward(values, prime):
    v = np.ascontiguousarray(values, dtype=np.uint64).copy()
    rc = lib().orc_ntt_forward(P(v), ctypes.c_size_t(v.shape[0]), ctypes.c_uint64(prime))
    assert rc == 0
    return v


def ntt_inverse(values, prime):
    v = np.ascontiguousarray(values, dtype=np.uint64).copy()
    rc = lib().orc_ntt_inverse(P(v), ctypes.c_size_t(v.shape[0]), ctypes.c_uint64(prime))
    assert rc == 0
    return v


def negacyclic_multiply(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.zeros_like(a)
    rc = lib().orc_negacyclic_multiply(P(a), P(b), ctypes.c_size_t(a.shape[0]), P(out))
    assert rc == 0
    return out


def naive_negacyclic(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.zeros_like(a)
    rc = lib().orc_naive_negacyclic(P(a), P(b), ctypes.c_size_t(a.shape[0]), P(out))
    assert rc == 0
    return out


def convolution(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.zeros((a.shape[0] + b.shape[0] - 1, 2), dtype=np.uint64)
    rc = lib().orc_convolution(P(a), ctypes.c_size_t(a.shape[0]), P(b), ctypes.c_size_t(b.shape[0]), P(out))
    assert rc == 0
    return out


def crt_reconstruct(residues):
    r = np.ascontiguousarray(residues, dtype=np.uint64)
    out = np.zeros(2, dtype=np.uint64)
    lib().orc_crt_reconstruct(P(r), P(out))
    return from_words(out)[0]


def mul_by_x(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    out = np.zeros_like(a)
    lib().orc_poly_mul_by_x(P(a), ctypes.c_size_t(a.shape[0]), P(out))
    return out


# ---- commitment ----
def derive_subseed(seed, label):
    out = ctypes.create_string_buffer(32)
    lb = ctypes.create_string_buffer(label, max(1, len(label)))
    lib().orc_derive_subseed(ctypes.create_string_buffer(bytes(seed), 32), lb,
                             ctypes.c_size_t(len(label)), out)
    return out.raw


def derive_generators(g_seed, rows, ring_len):
    out = np.zeros((rows, K, ring_len, 2), dtype=np.uint64)
    lib().orc_derive_generators(ctypes.create_string_buffer(bytes(g_seed), 32), ctypes.c_size_t(rows),
                                ctypes.c_size_t(ring_len), P(out))
    return out


def commit_with_generators(witness, gens, ntt_enabled=True):
    """witness (rows, m, 2); gens (rows, K, m, 2) -> (K, m_out, 2)"""
    witness = np.ascontiguousarray(witness, dtype=np.uint64)
    gens = np.ascontiguousarray(gens, dtype=np.uint64)
    rows = witness.shape[0]
    m = witness.shape[1] if rows else 1
    out = np.zeros((K, max(m, 1), 2), dtype=np.uint64)
    ol = ctypes.c_size_t(0)
    rc = lib().orc_commit_with_generators(P(witness), P(gens), ctypes.c_size_t(rows), ctypes.c_size_t(m),
                                          ctypes.c_int(1 if ntt_enabled else 0), P(out), ctypes.byref(ol))
    assert rc == 0
    return out[:, : ol.value]


def commit(seed, witness, ntt_enabled=True):
    witness = np.ascontiguousarray(witness, dtype=np.uint64)
    rows = witness.shape[0]
    m = witness.shape[1] if rows else 1
    out = np.zeros((K, max(m, 1), 2), dtype=np.uint64)
    ol = ctypes.c_size_t(0)
    rc = lib().orc_commit(ctypes.create_string_buffer(bytes(seed), 32), P(witness), ctypes.c_size_t(rows),
                          ctypes.c_size_t(m), ctypes.c_int(1 if ntt_enabled else 0), P(out), ctypes.byref(ol))
    assert rc == 0
    return out[:, : ol.value]


def challenge(epoch_seed, label, file_id, rnd, cur, lhs, rhs):
    cur = np.ascontiguousarray(cur, dtype=np.uint64)
    lhs = np.ascontiguousarray(lhs, dtype=np.uint64)
    rhs = np.ascontiguousarray(rhs, dtype=np.uint64)
    out = np.zeros(2, dtype=np.uint64)
    lib().orc_challenge(ctypes.create_string_buffer(bytes(epoch_seed), 32), label.encode(),
                        ctypes.create_string_buffer(bytes(file_id), 32), ctypes.c_uint64(rnd),
                        P(cur), ctypes.c_size_t(cur.shape[1]), P(lhs), ctypes.c_size_t(lhs.shape[1]),
                        P(rhs), ctypes.c_size_t(rhs.shape[1]), P(out))
    return from_words(out)[0]


# ---- io ----
def layout_for_byte_len(n):
    d, dim, cc, cap = (ctypes.c_size_t() for _ in range(4))
    rc = lib().orc_layout_for_byte_len(ctypes.c_uint64(n), ctypes.byref(d), ctypes.byref(dim),
                                       ctypes.byref(cc), ctypes.byref(cap))
    assert rc == 0
    return {"depth": d.value, "dimension": dim.value, "coeff_count": cc.value, "capacity": cap.value}


def expand(data, max_bytes=2**62):
    lay = layout_for_byte_len(len(data))
    out = np.zeros((lay["capacity"], 2), dtype=np.uint64)
    fid = ctypes.create_string_buffer(32)
    buf = ctypes.create_string_buffer(bytes(data), max(1, len(data)))
    rc = lib().orc_expand(buf, ctypes.c_size_t(len(data)), ctypes.c_size_t(max_bytes), fid, P(out),
                          ctypes.c_size_t(lay["capacity"]))
    return rc, fid.raw, out.reshape(lay["dimension"], lay["dimension"], 2), lay


# ---- prover / verifier ----
class OracleError(Exception):
    def __init__(self, rc, what):
        super().__init__(f"{what} failed with oracle status {rc}")
        self.rc = rc


class Prover:
    """DamascusProver (src/protocol/prover.rs) restated in the oracle"""

    def __init__(self, data, seed_generators=bytes([7] * 32), epoch_seed=bytes([11] * 32),
                 max_preprocess_bytes=512 << 20, ntt_enabled=True):
        self.h = ctypes.c_void_p()
        self._buf = ctypes.create_string_buffer(bytes(data), max(1, len(data)))
        rc = lib().orc_prover_new(self._buf, ctypes.c_size_t(len(data)),
                                  ctypes.create_string_buffer(bytes(seed_generators), 32),
                                  ctypes.create_string_buffer(bytes(epoch_seed), 32),
                                  ctypes.c_size_t(max_preprocess_bytes), ctypes.c_int(int(ntt_enabled)),
                                  ctypes.byref(self.h))
        if rc != 0:
            self.h = None
            raise OracleError(rc, "initialize")
        self.epoch_seed = bytes(epoch_seed)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_prover_free(self.h)

    def rounds_total(self):
        return lib().orc_prover_rounds(self.h)

    def current_round(self):
        return lib().orc_prover_round(self.h)

    def dims(self):
        r, n = ctypes.c_size_t(), ctypes.c_size_t()
        lib().orc_prover_dims(self.h, ctypes.byref(r), ctypes.byref(n))
        return r.value, n.value

    def statement(self):
        n = lib().orc_prover_statement(self.h, None, ctypes.c_size_t(0))
        buf = ctypes.create_string_buffer(n)
        lib().orc_prover_statement(self.h, buf, ctypes.c_size_t(n))
        return buf.raw

    def current_commitment(self):
        _, n = self.dims()
        out = np.zeros((K, n, 2), dtype=np.uint64)
        lib().orc_prover_current_commitment(self.h, P(out))
        return out

    def message(self):
        r, n = self.dims()
        out = np.zeros((r, n, 2), dtype=np.uint64)
        lib().orc_prover_message(self.h, P(out))
        return out

    def generators(self):
        r, n = self.dims()
        out = np.zeros((r, K, n, 2), dtype=np.uint64)
        lib().orc_prover_generators(self.h, P(out))
        return out

    def fold_round(self, j):
        _, n = self.dims()
        cap = lib().orc_record_size(ctypes.c_size_t(max(n, 1)), ctypes.c_size_t(max(n // 2, 1)))
        buf = ctypes.create_string_buffer(cap)
        got = ctypes.c_size_t(0)
        rc = lib().orc_prover_fold_round(self.h, ctypes.c_size_t(j), buf, ctypes.c_size_t(cap), ctypes.byref(got))
        if rc != 0:
            raise OracleError(rc, f"fold_round({j})")
        return buf.raw[: got.value]

    def final_opening(self):
        out = np.zeros(2, dtype=np.uint64)
        rc = lib().orc_prover_final_opening(self.h, P(out))
        return None if rc != 0 else from_words(out)[0]


class Verifier:
    """DamascusVerifier (src/protocol/verifier.rs) restated in the oracle"""

    def __init__(self, statement, epoch_seed=bytes([11] * 32), _handle=None):
        if _handle is not None:
            self.h = ctypes.c_void_p(_handle)
            return
        self.h = ctypes.c_void_p()
        rc = lib().orc_verifier_new(ctypes.create_string_buffer(bytes(epoch_seed), 32),
                                    ctypes.create_string_buffer(bytes(statement), max(1, len(statement))),
                                    ctypes.c_size_t(len(statement)), ctypes.byref(self.h))
        if rc != 0:
            self.h = None
            raise OracleError(rc, "verifier new")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_verifier_free(self.h)

    def clone(self):
        return Verifier(None, _handle=lib().orc_verifier_clone(self.h))

    def dims(self):
        r, n = ctypes.c_size_t(), ctypes.c_size_t()
        lib().orc_verifier_dims(self.h, ctypes.byref(r), ctypes.byref(n))
        return r.value, n.value

    def current_commitment(self):
        _, n = self.dims()
        out = np.zeros((K, n, 2), dtype=np.uint64)
        lib().orc_verifier_current_commitment(self.h, P(out))
        return out

    def update_commitment(self, record):
        return lib().orc_verifier_update(self.h, ctypes.create_string_buffer(bytes(record), max(1, len(record))),
                                         ctypes.c_size_t(len(record)))

    def verify_final_opening(self, opening):
        return lib().orc_verifier_verify_final_opening(self.h, P(to_words([opening])))


def full_flow(data, seed_generators=bytes([7] * 32), epoch_seed=bytes([11] * 32)):
    """prove + replay + verify; returns (rc, statement, records_blob, opening)"""
    lay = layout_for_byte_len(len(data))
    d, n = lay["depth"], lay["dimension"]
    st_cap = 32 + 16 + 8 + K * (8 + 16 * n) + 32
    rec_cap = sum(lib().orc_record_size(ctypes.c_size_t(n >> j), ctypes.c_size_t(max(1, n >> (j + 1))))
                  for j in range(d))
    st = ctypes.create_string_buffer(st_cap)
    recs = ctypes.create_string_buffer(rec_cap)
    sl, rl = ctypes.c_size_t(), ctypes.c_size_t()
    op = np.zeros(2, dtype=np.uint64)
    buf = ctypes.create_string_buffer(bytes(data), max(1, len(data)))
    rc = lib().orc_full_flow(buf, ctypes.c_size_t(len(data)), ctypes.create_string_buffer(bytes(seed_generators), 32),
                             ctypes.create_string_buffer(bytes(epoch_seed), 32), st, ctypes.c_size_t(st_cap),
                             ctypes.byref(sl), recs, ctypes.c_size_t(rec_cap), ctypes.byref(rl), P(op))
    return rc, st.raw[: sl.value], recs.raw[: rl.value], from_words(op)[0]
