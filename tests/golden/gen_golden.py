"""Writes the golden fixtures under tests/golden/ from closed forms in Python ints.

Nothing here imports or runs the reference: every expected value is computed from the
mathematics the reference implements (src/algebra/field.rs, src/algebra/ntt.rs,
src/utils/io.rs), with Python's arbitrary-precision integers as the independent engine.
The BLAKE3 vectors are the official BLAKE3 test vectors (input bytes i % 251, key
"whats the Elvish word for friend"), first 32 bytes of output, kept as data.

Run:  python tests/golden/gen_golden.py
"""
import json
import os
import random

Q = (1 << 112) - 75
PRIMES = [3892314113, 2281701377, 2013265921, 2885681153, 2483027969, 1811939329, 469762049,
          4194304001]
HERE = os.path.dirname(os.path.abspath(__file__))


def naive_negacyclic(a, b):
    n = len(a)
    out = [0] * n
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            k = i + j
            if k >= n:
                out[k - n] -= x * y
            else:
                out[k] += x * y
    return [v % Q for v in out]


def zq_vectors(rng):
    cases = []
    edge = [0, 1, 2, Q - 1, Q - 2, (1 << 111), (1 << 112) - 76, 75, 1 << 56, (1 << 56) - 1]
    vals = edge + [rng.randrange(Q) for _ in range(54)]
    for i in range(len(vals)):
        a = vals[i]
        b = vals[(i * 7 + 3) % len(vals)]
        cases.append({
            "a": str(a), "b": str(b),
            "add": str((a + b) % Q), "sub": str((a - b) % Q), "neg": str((-a) % Q),
            "mul": str(a * b % Q),
            "inv": str(pow(a, Q - 2, Q)) if a else None,
        })
    byte_cases = []
    for raw in [b"\xff" * 64, b"\xff" * 32, bytes(range(32)), b"\x01", b"",
                bytes(rng.randrange(256) for _ in range(32)),
                ((Q).to_bytes(16, "little") + b"\x00" * 16),
                ((Q - 1).to_bytes(16, "little") + b"\x00" * 16)]:
        byte_cases.append({"bytes": raw.hex(), "mod_q": str(int.from_bytes(raw, "little") % Q)})
    chunk_cases = []
    for n in range(0, 14):
        raw = bytes(rng.randrange(256) for _ in range(n))
        chunk_cases.append({"bytes": raw.hex(), "value": str(int.from_bytes(raw, "little") % Q)})
    return {"modulus": str(Q), "ops": cases, "from_le_bytes_mod_order": byte_cases,
            "from_le_chunk": chunk_cases}


def product_vectors(rng):
    cases = []
    for n in [1, 2, 4, 8, 16, 32, 64, 128, 256]:
        for kind in ["uniform", "top", "sparse"]:
            if kind == "uniform":
                a = [rng.randrange(Q) for _ in range(n)]
                b = [rng.randrange(Q) for _ in range(n)]
            elif kind == "top":  # every coefficient q-1: the largest positive and negative sums
                a = [Q - 1] * n
                b = [Q - 1] * n
            else:
                a = [0] * n
                b = [0] * n
                a[n - 1] = 3
                b[min(1, n - 1)] = 5
            cases.append({"n": n, "kind": kind, "a": [str(x) for x in a], "b": [str(x) for x in b],
                          "out": [str(x) for x in naive_negacyclic(a, b)]})
    return cases


def crt_vectors(rng):
    """Centred CRT lift: any integer V with |V| < P/2 must come back as V mod q."""
    P = 1
    for p in PRIMES:
        P *= p
    cases = []
    specials = [0, 1, -1, (P - 1) // 2, -((P - 1) // 2), Q, -Q, 64 * (Q - 1) ** 2,
                -64 * (Q - 1) ** 2, 2048 * 2048 * (Q - 1) ** 2, -2048 * 2048 * (Q - 1) ** 2]
    vals = specials + [rng.randrange(-(P // 2), P // 2) for _ in range(40)]
    for v in vals:
        cases.append({"value": str(v), "residues": [v % p for p in PRIMES], "mod_q": str(v % Q)})
    return {"P": str(P), "cases": cases}


def layout_vectors():
    cases = []
    for nbytes in [0, 1, 12, 13, 14, 44, 53248, 53249, 8192, 8 << 20, 1 << 20, 65536, 16 << 20,
                   128 << 20, 1 << 30]:
        cc = max(1, -(-nbytes // 13))
        d = 6
        while (1 << d) * (1 << d) < cc:
            d += 1
        cases.append({"bytes": nbytes, "coeff_count": cc, "depth": d, "dimension": 1 << d})
    return cases


def ntt_root_vectors():
    """Smallest multiplicative generator per prime (ntt.rs:287-298), by brute force."""
    out = []
    for p in PRIMES:
        m = p - 1
        fac = []
        d = 2
        while d * d <= m:
            if m % d == 0:
                fac.append(d)
                while m % d == 0:
                    m //= d
            d += 1
        if m > 1:
            fac.append(m)
        g = 2
        while any(pow(g, (p - 1) // f, p) == 1 for f in fac):
            g += 1
        out.append({"prime": p, "generator": g, "two_adicity": ((p - 1) & -(p - 1)).bit_length() - 1})
    return out


BLAKE3_SPEC = {
    "source": "official BLAKE3 test vectors (input i % 251; keyed key 'whats the Elvish word for friend')",
    "hash": {
        "0": "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262",
        "1": "2d3adedff11b61f14c886e35afa036736dcd87a74d27b5c1510225d0f592e213",
        "1023": "10108970eeda3eb932baac1428c7a2163b0e924c9a9e25b35bba72b28f70bd11",
        "1024": "42214739f095a406f3fc83deb889744ac00df831c10daa55189b5d121c855af7",
        "1025": "d00278ae47eb27b34faecf67b4fe263f82d5412916c1ffd97c8cb7fb814b8444",
        "2048": "e776b6028c7cd22a4d0ba182a8bf62205d2ef576467e838ed6f2529b85fba24a",
        "2049": "5f4d72f40d7a5f82b15ca2b2e44b1de3c2ef86c426c95c1af0b6879522563030",
        "3072": "b98cb0ff3623be03326b373de6b9095218513e64f1ee2edd2525c7ad1e5cffd2",
        "4096": "015094013f57a5277b59d8475c0501042c0b642e531b0a1c8f58d2163229e969",
        "8192": "aae792484c8efe4f19e2ca7d371d8c467ffb10748d8a5a1ae579948f718a2a63",
        "31744": "62b6960e1a44bcc1eb1a611a8d6235b6b4b78f32e7abc4fb4c6cdcce94895c47",
    },
    "keyed_hash": {
        "0": "92b2b75604ed3c761f9d6f62392c8a9227ad0ea3f09573e783f1498a4ed60d26",
        "1": "6d7878dfff2f485635d39013278ae14f1454b8c0a3a2d34bc1ab38228a80c95b",
    },
    "abc": "6437b3ac38465133ffb63b75273a8db548c558465d79db03fd359c6cd5bd9d85",
}


def main():
    rng = random.Random(20261203)
    fixtures = {
        "zq.json": zq_vectors(rng),
        "negacyclic.json": product_vectors(rng),
        "crt.json": crt_vectors(rng),
        "layout.json": layout_vectors(),
        "ntt_roots.json": ntt_root_vectors(),
        "blake3_spec.json": BLAKE3_SPEC,
    }
    for name, obj in fixtures.items():
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(obj, f, indent=0)
        print("wrote", name)


if __name__ == "__main__":
    main()
