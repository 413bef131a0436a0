"""Host build of the device Z_q inverse (damascus000_amd/csrc/dmx_arith.hpp, zq_inv_divsteps).

The challenge kernels invert x = LE256(digest) mod q on the device (transcript.rs:66-75 feeds
x; prover.rs:195-207 and verifier.rs use x^-1). The device routine is batched divsteps, not
the reference's x^(q-2) (field.rs:106-109); the inverse is unique, so both must agree on every
input. Compiled here for the host from the same header, it is checked against the golden
inverses (tests/golden/zq.json, Python pow), against x^(q-2) on seeded random inputs and edge
cases, and by x * x^-1 = 1.
"""
import json
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "damascus000_amd", "csrc")
Q = 2**112 - 75


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("host_arith") / "host_arith_check")
    subprocess.run([hipcc, "--offload-host-only", "-O2", "-std=c++17", "-I", CSRC, "-o", exe,
                    os.path.join(HERE, "host_arith_check.cpp")], check=True, capture_output=True)
    return exe


def _inverses(exe, xs):
    out = subprocess.run([exe, "inv"], input="".join(f"{x:x}\n" for x in xs), capture_output=True,
                         text=True, check=True).stdout.split()
    return [int(v, 16) for v in out]


def test_inverse_matches_golden_vectors(checker):
    g = json.load(open(os.path.join(HERE, "golden", "zq.json")))
    nz = [c for c in g["ops"] if c["inv"] is not None]
    assert _inverses(checker, [int(c["a"]) for c in nz]) == [int(c["inv"]) for c in nz]


def test_inverse_closed_form_cases(checker):
    xs = [1, 2, 3, 75, 76, Q - 1, Q - 2, 2**111, 2**111 - 38, 2**111 - 37] + [pow(3, k, Q) for k in range(1, 200, 7)]
    assert _inverses(checker, xs) == [pow(x, Q - 2, Q) for x in xs]
    assert _inverses(checker, [0]) == [0]  # as x^(q-2)


def test_inverse_random_against_fermat(checker):
    r = subprocess.run([checker, "random", "200000", "20261203"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout


def _tx_message(exe, label, rnd, epoch, file_id, mods):
    ns = [len(m) // 8 for m in mods]
    inp = f"{label} {rnd} {epoch.hex()} {file_id.hex()} {ns[0]} {ns[1]} {ns[2]}\n"
    inp += "".join(f"{c:x}\n" for m in mods for c in m)
    r = subprocess.run([exe, "tx"], input=inp, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return bytes.fromhex(r.stdout.strip())


@pytest.mark.parametrize("label,rnd,ns", [("vec", 0, (1, 1, 1)), ("poly", 0, (2, 1, 1)), ("vec", 3, (4, 4, 4)),
                                          ("poly", 5, (32, 16, 16)), ("vec", 2**40 + 7, (64, 64, 64)),
                                          ("poly", 1, (8, 64, 2))])
def test_transcript_words_hash_to_the_oracle_challenge(checker, label, rnd, ns):
    """The hash kernels build the challenge message (transcript.rs:50-61, 77-90) from aligned
    8-byte words (dmx_transcript.hpp, tx_words4). BLAKE3 of the assembled bytes, reduced mod q,
    must equal the oracle's challenge, which absorbs the same fields field by field."""
    import numpy as np

    import _oracle as O

    rng = np.random.default_rng(sum(ns) + rnd % 1000)
    epoch = bytes(rng.integers(0, 256, 32, dtype=np.uint8))
    file_id = bytes(rng.integers(0, 256, 32, dtype=np.uint8))
    mods = [[int.from_bytes(rng.bytes(14), "little") % Q for _ in range(8 * n)] for n in ns]
    msg = _tx_message(checker, label, rnd, epoch, file_id, mods)
    assert len(msg) == 96 + len(label) + sum(8 + 8 * (8 + 16 * n) for n in ns) + 4
    assert msg[:32] == epoch and msg[-4:] == b"\0\0\0\0"
    x = O.from_le_bytes_mod_order(O.blake3(msg))
    arrs = [O.to_words(m).reshape(8, n, 2) for m, n in zip(mods, ns)]
    assert x == O.challenge(epoch, label, file_id, rnd, *arrs)


PRIMES = [3892314113, 2281701377, 2013265921, 2885681153, 2483027969, 1811939329, 469762049, 4194304001]


def test_mod_p_routines_against_exact_integers(checker):
    """The Montgomery routines every transform, MAC pass and lift is built on: addp/subp
    (canonical sums of residues, which carry out of 32 bits for the primes above 2^31),
    mulc (b*w mod p for any 32-bit b), redc64 ((hi*2^32 + lo)*2^-32 mod p, below 2^32, below p
    when hi < p), zq_mod_p and its centred form (the residue of x, or of x - q when
    x > (q-1)/2), and reduce_u96 (the u96 MAC accumulators), checked against Python integers on
    edge and seeded cases for all eight primes."""
    out = subprocess.run([checker, "modp", "3000", "7"], capture_output=True, text=True, check=True).stdout
    R = 1 << 32
    counts = {}
    for line in out.splitlines():
        f = line.split()
        op, p = f[0], PRIMES[int(f[1])]
        v = [int(t, 16) for t in f[2:]]
        counts[op] = counts.get(op, 0) + 1
        if op == "add":
            assert v[2] == (v[0] + v[1]) % p, line
        elif op == "sub":
            assert v[2] == (v[0] - v[1]) % p, line
        elif op == "mulc":
            assert v[2] == v[0] * v[1] % p, line
        elif op == "redc":
            hi, lo, r = v
            assert r < R and (r * R - (hi * R + lo)) % p == 0, line
            if hi < p:
                assert r < p, line
        elif op == "modp":
            x, r, rc = v
            assert x < Q
            assert r == x % p, line
            assert rc == ((x - Q) % p if x > (Q - 1) // 2 else x % p), line
        elif op == "u96":
            lo, hi, r = v
            assert r == (hi * 2**64 + lo) % p, line
    assert all(counts.get(k, 0) >= 8 * 3000 for k in ("add", "sub", "mulc", "redc", "modp", "u96")), counts
