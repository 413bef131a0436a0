"""Writes the C1 end-to-end golden fixture: the CPU-runnable case of the commit-and-fold path.

Input: 53,248 bytes (4,096 coefficients at 13 bytes each, a 64 x 64 witness, d = 6) taken from
SplitMix64 seeded with C1_SEED, each 64-bit output little-endian. Default SystemParams seeds
(src/utils/config.rs:19-20): seed_generators = [7; 32], epoch_seed = [11; 32].

Output, produced by the oracle (the C restatement of the reference, examples/full_flow.rs:47-104):
  c1_statement.bin  bincode Statement (src/commitment/sis.rs:23-30)
  c1_records.bin    the 6 bincode RoundRecords back to back (src/protocol/prover.rs:14-21)
  c1_full_flow.json m*, sizes, file_id and the SHA-256 of the input and of both files

Run from the repository root after building the oracle: python tests/golden/gen_c1_fixture.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

C1_SEED = 20261203
C1_BYTES = 53248
MASK = (1 << 64) - 1


def splitmix64_bytes(seed, n):
    """n bytes of the SplitMix64 stream (Steele, Lea, Flood 2014), outputs little-endian"""
    out = bytearray()
    x = seed & MASK
    while len(out) < n:
        x = (x + 0x9E3779B97F4A7C15) & MASK
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        z ^= z >> 31
        out += z.to_bytes(8, "little")
    return bytes(out[:n])


def c1_input():
    return splitmix64_bytes(C1_SEED, C1_BYTES)


def load():
    """(input bytes, statement, records blob, m*) of the committed fixture"""
    with open(os.path.join(HERE, "c1_full_flow.json")) as f:
        meta = json.load(f)
    with open(os.path.join(HERE, "c1_statement.bin"), "rb") as f:
        st = f.read()
    with open(os.path.join(HERE, "c1_records.bin"), "rb") as f:
        recs = f.read()
    data = c1_input()
    assert hashlib.sha256(data).hexdigest() == meta["input_sha256"], "C1 input stream changed"
    return data, st, recs, int(meta["m_star"]), meta


def main():
    import _oracle as O

    data = c1_input()
    rc, st, recs, op = O.full_flow(data)
    if rc != 0:
        raise SystemExit(f"oracle full flow failed: {rc}")
    with open(os.path.join(HERE, "c1_statement.bin"), "wb") as f:
        f.write(st)
    with open(os.path.join(HERE, "c1_records.bin"), "wb") as f:
        f.write(recs)
    meta = {
        "what": "C1: 53,248 SplitMix64 bytes, d = 6, default seeds; oracle prove + verify",
        "prng": "splitmix64", "seed": C1_SEED, "input_bytes": C1_BYTES,
        "input_sha256": hashlib.sha256(data).hexdigest(),
        "seed_generators": [7] * 32, "epoch_seed": [11] * 32, "depth": 6,
        "file_id": st[:32].hex(),
        "statement_len": len(st), "statement_sha256": hashlib.sha256(st).hexdigest(),
        "records_len": len(recs), "records_sha256": hashlib.sha256(recs).hexdigest(),
        "m_star": str(op),
    }
    with open(os.path.join(HERE, "c1_full_flow.json"), "w") as f:
        json.dump(meta, f, indent=1)
        f.write("\n")
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main()
