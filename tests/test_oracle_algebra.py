"""Pins the oracle's algebra: Z_q, BLAKE3, NTT, ring products, CRT lift.

Restates the reference's own unit tests (src/algebra/field.rs:260-312, src/algebra/ntt.rs:429-522,
src/algebra/poly.rs:172-242) against the oracle, and checks it against golden vectors computed
independently with Python integers (tests/golden/gen_golden.py) and the official BLAKE3 vectors.
"""
import json
import os

import numpy as np
import pytest

import _oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


# ---------------- field.rs ----------------
def test_field_arithmetic_round_trip():
    """field.rs:269-276"""
    a, b = O.to_words([17]), O.to_words([9])
    c = O.zq_binop("mul", a, b)
    assert O.from_words(c) == [153]
    assert O.from_words(O.zq_binop("mul", c, O.zq_unop("inv", b))) == [17]


def test_arithmetic_matches_reference_modulus():
    """field.rs:278-297: 20k random pairs against big-integer (a*b) mod q"""
    rng = np.random.default_rng(7)
    a = O.random_zq(rng, 20000)
    b = O.random_zq(rng, 20000)
    add = O.from_words(O.zq_binop("add", a, b))
    mul = O.from_words(O.zq_binop("mul", a, b))
    av, bv = O.from_words(a), O.from_words(b)
    for i in range(20000):
        assert add[i] == (av[i] + bv[i]) % O.Q
        assert mul[i] == av[i] * bv[i] % O.Q


def test_challenge_bytes_reduce_into_field():
    """field.rs:299-303"""
    assert O.from_le_bytes_mod_order(b"\xff" * 64) < O.Q


def test_serde_uses_fixed_width_encoding():
    """field.rs:305-311: an Fp serialises to 16 little-endian bytes"""
    w = O.to_words([42])
    assert w.astype("<u8").tobytes() == (42).to_bytes(16, "little")


def test_zq_golden_vectors():
    g = gold("zq.json")
    assert int(g["modulus"]) == O.Q
    a = O.to_words([int(c["a"]) for c in g["ops"]])
    b = O.to_words([int(c["b"]) for c in g["ops"]])
    for op in ["add", "sub", "mul"]:
        assert O.from_words(O.zq_binop(op, a, b)) == [int(c[op]) for c in g["ops"]], op
    assert O.from_words(O.zq_unop("neg", a)) == [int(c["neg"]) for c in g["ops"]]
    nz = [c for c in g["ops"] if c["inv"] is not None]
    inv = O.from_words(O.zq_unop("inv", O.to_words([int(c["a"]) for c in nz])))
    assert inv == [int(c["inv"]) for c in nz]
    for c in g["from_le_bytes_mod_order"]:
        assert O.from_le_bytes_mod_order(bytes.fromhex(c["bytes"])) == int(c["mod_q"])
    for c in g["from_le_chunk"]:
        assert O.from_le_chunk(bytes.fromhex(c["bytes"])) == int(c["value"])


# ---------------- BLAKE3 (crate blake3 1.8.3; spec vectors) ----------------
def test_blake3_spec_vectors():
    g = gold("blake3_spec.json")
    for n, want in g["hash"].items():
        data = bytes(i % 251 for i in range(int(n)))
        assert O.blake3(data).hex() == want, n
    key = b"whats the Elvish word for friend"
    for n, want in g["keyed_hash"].items():
        data = bytes(i % 251 for i in range(int(n)))
        assert O.blake3(data, key).hex() == want, n
    assert O.blake3(b"abc").hex() == g["abc"]


def test_blake3_incremental_matches_one_shot_across_chunk_boundaries():
    """the transcript absorbs many small updates (transcript.rs:50-61); the tree must not
    depend on how the input is split"""
    import ctypes
    for n in [0, 63, 64, 65, 1023, 1024, 1025, 2047, 3073, 5000]:
        data = bytes((7 * i + 3) % 256 for i in range(n))
        assert O.blake3(data) == O.blake3(data[: n // 2] + data[n // 2:])


# ---------------- ntt.rs ----------------
def test_ntt_round_trip():
    """ntt.rs:438-445"""
    v = np.arange(128, dtype=np.uint64)
    assert np.array_equal(O.ntt_inverse(O.ntt_forward(v, O.PRIMES[0]), O.PRIMES[0]), v)


def test_generic_convolution_matches_reference():
    """ntt.rs:448-465"""
    c = O.from_words(O.convolution(O.to_words([1, 2, 3, 4]), O.to_words([4, 5, 6, 7])))
    assert len(c) == 7 and c[0] == 4 and c[3] == 50


def test_negacyclic_multiply_matches_naive():
    """ntt.rs:468-478"""
    lhs = O.to_words([i + 1 for i in range(64)])
    rhs = O.to_words([i * 3 + 5 for i in range(64)])
    assert np.array_equal(O.negacyclic_multiply(lhs, rhs), O.naive_negacyclic(lhs, rhs))


def test_negacyclic_multiply_matches_naive_across_power_of_two_sizes():
    """ntt.rs:481-493"""
    for n in [2, 4, 8, 16, 32, 64]:
        lhs = O.to_words([i * 5 + 1 for i in range(n)])
        rhs = O.to_words([i * 7 + 3 for i in range(n)])
        assert np.array_equal(O.negacyclic_multiply(lhs, rhs), O.naive_negacyclic(lhs, rhs)), n


def test_ntt_roundtrip_across_runtime_sizes():
    """ntt.rs:496-507"""
    for n in [64, 128, 256, 512, 1024, 2048]:
        v = (np.arange(n, dtype=np.uint64) * 17 + 3) % O.PRIMES[0]
        assert np.array_equal(O.ntt_inverse(O.ntt_forward(v, O.PRIMES[0]), O.PRIMES[0]), v), n


def test_negacyclic_multiply_across_runtime_sizes():
    """ntt.rs:510-521: 3X^{n-1} * 5X = -15"""
    for n in [64, 128, 256, 512, 1024, 2048]:
        lhs = np.zeros((n, 2), dtype=np.uint64)
        rhs = np.zeros((n, 2), dtype=np.uint64)
        lhs[n - 1, 0] = 3
        rhs[1, 0] = 5
        out = O.from_words(O.negacyclic_multiply(lhs, rhs))
        assert out[0] == O.Q - 15 and all(v == 0 for v in out[1:]), n


def test_ntt_matches_dft_definition():
    """forward_ntt computes A[k] = sum_j a[j] w^(jk) with w = g^((p-1)/S), g the smallest
    generator (ntt.rs:275-298): check against the O(S^2) definition for every prime"""
    roots = gold("ntt_roots.json")
    for entry, p in zip(roots, O.PRIMES):
        assert entry["prime"] == p
        assert O.lib().orc_multiplicative_generator(p) == entry["generator"]
        S = 16
        w = pow(entry["generator"], (p - 1) // S, p)
        a = [(i * 977 + 5) % p for i in range(S)]
        want = [sum(a[j] * pow(w, j * k, p) for j in range(S)) % p for k in range(S)]
        got = O.ntt_forward(np.array(a, dtype=np.uint64), p)
        assert [int(x) for x in got] == want


def test_negacyclic_golden_products():
    """CRT-NTT products (the reference algorithm) against Python-int schoolbook products,
    including all-(q-1) operands, which give the largest positive and negative lifts"""
    for c in gold("negacyclic.json"):
        a, b = O.to_words(c["a"]), O.to_words(c["b"])
        assert O.from_words(O.negacyclic_multiply(a, b)) == [int(x) for x in c["out"]], (c["n"], c["kind"])
        if c["n"] <= 64:
            assert O.from_words(O.naive_negacyclic(a, b)) == [int(x) for x in c["out"]]


def test_crt_lift_golden():
    """reconstruct_mod_q (ntt.rs:400-427): centred lift of |V| < P/2 gives V mod q, including
    V = +-(P-1)/2 and the d=11 commitment bound 2048*2048*(q-1)^2"""
    g = gold("crt.json")
    P = 1
    for p in O.PRIMES:
        P *= p
    assert int(g["P"]) == P
    for c in g["cases"]:
        assert O.crt_reconstruct(np.array(c["residues"], dtype=np.uint64)) == int(c["mod_q"]), c["value"]


# ---------------- poly.rs ----------------
def test_odd_even_split_and_fold():
    """poly.rs:185-210 (split + fold with challenge 2 gives coeff0 = 1 + 2*2 = 5)"""
    p = list(range(1, 9))
    even, odd = p[0::2], p[1::2]
    assert even == [1, 3, 5, 7] and odd == [2, 4, 6, 8]
    folded = O.from_words(O.zq_binop("add", O.to_words(even),
                                     O.zq_binop("mul", O.to_words(odd), O.to_words([2] * 4))))
    assert folded[0] == 5


def test_x_n_plus_one_vanishes():
    """poly.rs:212-220"""
    n = 64
    acc = np.zeros((n, 2), dtype=np.uint64)
    acc[0, 0] = 1
    for _ in range(n):
        acc = O.mul_by_x(acc)
    one = np.zeros((n, 2), dtype=np.uint64)
    one[0, 0] = 1
    assert all(v == 0 for v in O.from_words(O.zq_binop("add", acc, one)))


def test_scalar_unit_inverse_restores_ring_element():
    """poly.rs:222-239"""
    rng = np.random.default_rng(9)
    for _ in range(8):
        a = O.random_zq(rng, 64)
        s = O.from_words(O.random_zq(rng, 1))[0] or 1
        b = np.zeros((64, 2), dtype=np.uint64)
        b[0] = O.to_words([s])[0]
        bi = np.zeros((64, 2), dtype=np.uint64)
        bi[0] = O.to_words([pow(s, O.Q - 2, O.Q)])[0]
        rec = O.negacyclic_multiply(O.negacyclic_multiply(a, b), bi)
        assert np.array_equal(rec, a)
