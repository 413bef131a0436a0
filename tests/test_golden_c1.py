"""The C1 end-to-end golden fixture (tests/golden/c1_*): the oracle must keep producing it, and the
fixture must hold together on its own terms (file_id is BLAKE3 of the input, the statement and
records decode with the reference's bincode layout, the oracle verifier accepts it, one changed
byte is rejected). The GPU path is checked against the same files in test_gpu_deep.py."""
import os
import sys

import numpy as np

import _oracle as O
import _wire as W

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import gen_c1_fixture as G  # noqa: E402


def test_splitmix64_known_answer():
    """SplitMix64 from seed 0: first output 0xE220A8397B1DCDAF (the published reference stream)"""
    assert G.splitmix64_bytes(0, 8) == (0xE220A8397B1DCDAF).to_bytes(8, "little")


def test_oracle_reproduces_the_c1_fixture():
    data, st, recs, op, meta = G.load()
    rc, ost, orecs, oop = O.full_flow(data)
    assert rc == 0
    assert ost == st, "statement differs from the committed fixture"
    assert orecs == recs, "records differ from the committed fixture"
    assert oop == op, "m* differs from the committed fixture"


def test_c1_fixture_is_self_consistent():
    data, st, recs, op, meta = G.load()
    s = W.decode_statement(st)
    assert s["file_id"] == O.blake3(data) and s["file_id"].hex() == meta["file_id"]
    assert s["original_len_bytes"] == len(data) == 53248 and s["d"] == 6
    assert s["g_0_seed"] == O.derive_subseed(bytes([7] * 32), b"g0")
    blobs = W.split_records(recs)
    assert len(blobs) == 6
    for j, b in enumerate(blobs):
        rnd, mods, used = W.decode_record(b)
        m = 64 >> j
        assert rnd == j and [x.shape[1] for x in mods] == [m, m, m // 2, m // 2]
    v = O.Verifier(st)
    for b in blobs:
        assert v.update_commitment(b) == 0
    assert v.verify_final_opening(op) == 0
    v = O.Verifier(st)
    for b in blobs:
        assert v.update_commitment(b) == 0
    assert v.verify_final_opening((op + 1) % O.Q) != 0
