"""Pins the oracle's commitment, transcript, witness layout, prover and verifier.

Restates the reference's tests in src/commitment/sis.rs:247-327, src/protocol/transcript.rs:92-143,
src/utils/io.rs:122-166, src/utils/config.rs:100-115 and src/protocol/prover.rs:459-878 against the
oracle, and adds closed-form checks (Python-int commitments, fold identities).
"""
import json
import os
import struct

import numpy as np
import pytest

import _oracle as O
import _wire as W

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SEED = bytes([7] * 32)
EPOCH = bytes([11] * 32)


# ---------------- sis.rs ----------------
def test_commitment_is_deterministic():
    """sis.rs:277-284"""
    w = np.zeros((2, 64, 2), dtype=np.uint64)
    assert np.array_equal(O.commit(bytes([42] * 32), w), O.commit(bytes([42] * 32), w))


def _ring_mul_rows(rows, a):
    return np.stack([O.negacyclic_multiply(r, a) for r in rows])


def test_commitment_preserves_ring_linearity():
    """sis.rs:286-313 (K=8, the only module rank the protocol uses)"""
    rng = np.random.default_rng(12)
    seed = bytes([9] * 32)
    wa = np.stack([O.random_zq(rng, 64) for _ in range(4)])
    wb = np.stack([O.random_zq(rng, 64) for _ in range(4)])
    ra, rb = O.random_zq(rng, 64), O.random_zq(rng, 64)
    ca, cb = O.commit(seed, wa), O.commit(seed, wb)
    lin = np.stack([O.zq_binop("add", O.negacyclic_multiply(x, ra), O.negacyclic_multiply(y, rb))
                    for x, y in zip(wa, wb)])
    clin = O.commit(seed, lin)
    rhs = np.stack([O.zq_binop("add", O.negacyclic_multiply(ca[k], ra), O.negacyclic_multiply(cb[k], rb))
                    for k in range(8)])
    assert np.array_equal(clin, rhs)


def test_generator_family_is_module_valued():
    """sis.rs:316-326"""
    gs = O.derive_subseed(bytes([1] * 32), b"g0")
    g = O.derive_generators(gs, 2, 64)
    assert g.shape == (2, 8, 64, 2)
    assert all(v < O.Q for v in O.from_words(g))


def test_generators_match_keyed_blake3_closed_form():
    """g[i][k][c] = LE256(keyed_blake3(g_0_seed, "g"||u64(i)||u64(k)||u64(c))) mod q (sis.rs:231-245)"""
    gs = O.derive_subseed(SEED, b"g0")
    assert gs == O.blake3(b"g0", SEED)
    g = O.derive_generators(gs, 3, 4)
    for i in range(3):
        for k in range(8):
            for c in range(4):
                msg = b"g" + struct.pack("<QQQ", i, k, c)
                want = int.from_bytes(O.blake3(msg, gs), "little") % O.Q
                assert O.from_words(g[i, k, c])[0] == want


def test_commit_matches_python_schoolbook():
    """com_k = sum_i g[i]_k * m[i] in Z_q[X]/(X^m+1), with Python-int schoolbook products"""
    rng = np.random.default_rng(3)
    rows, m = 3, 8
    w = np.stack([O.random_zq(rng, m) for _ in range(rows)])
    gs = O.derive_subseed(SEED, b"g0")
    g = O.derive_generators(gs, rows, m)
    got = O.commit(SEED, w)
    wv = [O.from_words(w[i]) for i in range(rows)]
    for k in range(8):
        acc = [0] * m
        for i in range(rows):
            gv = O.from_words(g[i, k])
            for a in range(m):
                for b in range(m):
                    t = gv[a] * wv[i][b]
                    if a + b >= m:
                        acc[a + b - m] -= t
                    else:
                        acc[a + b] += t
        assert O.from_words(got[k]) == [x % O.Q for x in acc]


def test_commit_ntt_off_matches_ntt_on():
    """ntt_enabled=false uses the naive product (poly.rs:134-142); results are identical"""
    rng = np.random.default_rng(5)
    w = np.stack([O.random_zq(rng, 16) for _ in range(4)])
    assert np.array_equal(O.commit(SEED, w, True), O.commit(SEED, w, False))


def test_empty_witness_commits_to_zero_of_length_one():
    """sis.rs:154-156"""
    out = O.commit_with_generators(np.zeros((0, 1, 2), np.uint64), np.zeros((0, 8, 1, 2), np.uint64))
    assert out.shape == (8, 1, 2) and not out.any()


# ---------------- transcript.rs ----------------
def _tagged(tag):
    m = W.zero_module(64)
    m[0, 0] = O.to_words([tag])[0]
    m[1, 1] = O.to_words([tag + 1])[0]
    return m


def test_vec_and_poly_challenges_are_domain_separated():
    """transcript.rs:117-128"""
    fid = bytes([3] * 32)
    cur, lhs, rhs = _tagged(11), _tagged(17), _tagged(23)
    x = O.challenge(EPOCH, "vec", fid, 0, cur, lhs, rhs)
    y = O.challenge(EPOCH, "poly", fid, 0, cur, lhs, rhs)
    assert x != y and 0 < x < O.Q and 0 < y < O.Q


def test_challenge_matches_byte_encoding():
    """transcript.rs:42-63, 77-90: the challenge is LE256(BLAKE3(msg)) mod q for the documented
    message; rebuilt here byte by byte and hashed with the spec-pinned BLAKE3"""
    fid = bytes(range(32))
    cur, lhs, rhs = _tagged(5), _tagged(6)[:, :32], _tagged(7)[:, :32]

    def enc(m):
        out = struct.pack("<Q", 8)
        for k in range(8):
            out += struct.pack("<Q", m.shape[1]) + m[k].astype("<u8").tobytes()
        return out

    msg = (EPOCH + struct.pack("<Q", 4) + b"poly" + struct.pack("<Q", 32) + fid + struct.pack("<QQ", 0, 9)
           + enc(cur) + enc(lhs) + enc(rhs) + struct.pack("<I", 0))
    want = int.from_bytes(O.blake3(msg), "little") % O.Q
    assert O.challenge(EPOCH, "poly", fid, 9, cur, lhs, rhs) == want


# ---------------- io.rs / config.rs ----------------
def test_expand_preserves_injective_chunks():
    """io.rs:129-141"""
    rc, fid, w, lay = O.expand(bytes([0xA5] * (13 * 3 + 5)))
    assert rc == 0 and lay["depth"] == 6 and w.shape[:2] == (64, 64)
    vals = O.from_words(w.reshape(-1, 2))
    assert vals[0] == int.from_bytes(bytes([0xA5] * 13), "little")
    assert vals[3] == int.from_bytes(bytes([0xA5] * 5), "little")
    assert all(v == 0 for v in vals[4:])
    assert fid == O.blake3(bytes([0xA5] * 44))


def test_square_layout_expands_both_dimensions_for_large_inputs():
    """io.rs:144-155"""
    for nbytes in [128 << 20, 1 << 30]:
        lay = O.layout_for_byte_len(nbytes)
        assert lay["dimension"] == 1 << lay["depth"] and lay["depth"] >= 6
        assert lay["capacity"] >= lay["coeff_count"] and lay["capacity"] * 13 >= nbytes


def test_layout_golden():
    with open(os.path.join(GOLD, "layout.json")) as f:
        for c in json.load(f):
            lay = O.layout_for_byte_len(c["bytes"])
            assert (lay["coeff_count"], lay["depth"], lay["dimension"]) == (c["coeff_count"], c["depth"], c["dimension"])


def test_expand_rejects_memory_overcommit():
    """io.rs:158-165"""
    rc, _, _, _ = O.expand(bytes([0x55] * 4096), max_bytes=64)
    assert rc == O.ORC_ERR_MEMORY_LIMIT


def test_hard_spec_constants_are_wired():
    """config.rs:106-114"""
    assert O.K == 8 and len(O.PRIMES) == 8 and O.layout_for_byte_len(1)["depth"] == 6


# ---------------- prover.rs ----------------
def test_distinct_files_produce_distinct_initial_commitments():
    """prover.rs:512-523 (64 KiB files instead of 1 MiB: same property, d=7 keeps the oracle fast)"""
    a = O.Prover(W.pattern_bytes(1 << 16, 3))
    b = O.Prover(W.pattern_bytes(1 << 16, 4))
    assert not np.array_equal(a.current_commitment(), b.current_commitment())
    assert W.decode_statement(a.statement())["file_id"] != W.decode_statement(b.statement())["file_id"]


def test_initialize_errors_when_full_expansion_exceeds_memory_limit():
    """prover.rs:526-544"""
    with pytest.raises(O.OracleError) as e:
        O.Prover(W.pattern_bytes(1 << 20, 7), max_preprocess_bytes=1024)
    assert e.value.rc == O.ORC_ERR_MEMORY_LIMIT


def _prove(data):
    p = O.Prover(data)
    st = p.statement()
    recs = [p.fold_round(j) for j in range(p.rounds_total())]
    return p, st, recs, p.final_opening()


def test_honest_transcript_replays_and_tampering_is_rejected():
    """prover.rs:547-588"""
    p, st, recs, opening = _prove(W.pattern_bytes(8192, 9))
    v = O.Verifier(st)
    for r in recs:
        assert v.update_commitment(r) == 0
    assert v.verify_final_opening(opening) == 0
    rnd, mods, _ = W.decode_record(recs[0])
    mods[0][0, 0] = O.to_words([(O.from_words(mods[0][0, 0])[0] + 1) % O.Q])[0]
    bad = W.encode_record(rnd, *mods)
    v2 = O.Verifier(st)
    assert v2.update_commitment(bad) == 0
    for r in recs[1:]:
        assert v2.update_commitment(r) == 0
    assert v2.verify_final_opening(opening) == O.ORC_ERR_OPENING


def test_round_record_serialization_matches_module_payload_scale():
    """prover.rs:591-625; also the exact bincode size 24,872 B for round 0 at d=6"""
    p = O.Prover(W.pattern_bytes(8192, 13))
    rec = p.fold_round(0)
    assert len(rec) == 24872
    assert len(rec) >= 4 * 8 * 32 * 16
    rnd, mods, used = W.decode_record(rec)
    assert used == len(rec) and rnd == 0
    assert [m.shape[1] for m in mods] == [64, 64, 32, 32]


def test_single_round_preserves_commitment_through_vector_and_poly_stages():
    """prover.rs:628-692: re-derives x, y and both invariants from outside the prover"""
    p = O.Prover(W.pattern_bytes(8192, 31))
    st = W.decode_statement(p.statement())
    cur, msg, g = p.current_commitment(), p.message(), p.generators()
    rec = p.fold_round(0)
    _, (l_vec, r_vec, l_poly, r_poly), _ = W.decode_record(rec)
    mid = msg.shape[0] // 2
    x = O.challenge(EPOCH, "vec", st["file_id"], 0, cur, l_vec, r_vec)
    xi = pow(x, O.Q - 2, O.Q)

    def scale(m, s):
        flat = m.reshape(-1, 2)
        return O.zq_binop("mul", flat, O.to_words([s] * flat.shape[0])).reshape(m.shape)

    def add(a, b):
        return O.zq_binop("add", a.reshape(-1, 2), b.reshape(-1, 2)).reshape(a.shape)

    vfc = add(add(cur, scale(l_vec, xi)), scale(r_vec, x))
    fm = add(msg[:mid], scale(msg[mid:], x))
    fg = add(g[:mid], scale(g[mid:], xi))
    assert np.array_equal(O.commit_with_generators(fm, fg), vfc)
    e, o = fm[:, 0::2], fm[:, 1::2]
    ge = fg[:, :, 0::2]
    gox = np.stack([np.stack([O.mul_by_x(fg[i, k, 1::2]) for k in range(8)]) for i in range(mid)])
    y = O.challenge(EPOCH, "poly", st["file_id"], 0, vfc, l_poly, r_poly)
    yi = pow(y, O.Q - 2, O.Q)
    nm = add(np.ascontiguousarray(e), scale(np.ascontiguousarray(o), y))
    ng = add(np.ascontiguousarray(ge), scale(gox, yi))
    ncom = add(add(np.ascontiguousarray(vfc[:, 0::2]), scale(l_poly, yi)), scale(r_poly, y))
    assert np.array_equal(O.commit_with_generators(nm, ng), ncom)
    assert np.array_equal(ncom, p.current_commitment())


def test_serialized_round_record_byte_flip_is_rejected():
    """prover.rs:695-738"""
    p, st, recs, opening = _prove(W.pattern_bytes(8192, 21))
    enc = bytearray(recs[0])
    enc[len(enc) - 8] ^= 1
    v = O.Verifier(st)
    rc = v.update_commitment(bytes(enc))
    if rc == 0:
        for r in recs[1:]:
            assert v.update_commitment(r) == 0
        assert v.verify_final_opening(opening) == O.ORC_ERR_OPENING


def _zero_statement(d):
    return W.encode_statement(bytes([d] * 32), W.byte_len_for_depth(d), d, W.zero_module(1 << d), bytes([17] * 32))


def _zero_record(d, j):
    cur, nxt = 1 << (d - j), 1 << (d - j - 1)
    return W.encode_record(j, W.zero_module(cur), W.zero_module(cur), W.zero_module(nxt), W.zero_module(nxt))


def test_dual_dimension_fold_halves_each_round_and_round_record_size_tracks_ring_degree():
    """prover.rs:741-788 (d in {6, 8}; d=10 derives 8.4M generators in the single-threaded oracle)"""
    for d in [6, 8]:
        assert O.layout_for_byte_len(W.byte_len_for_depth(d))["depth"] == d
        v = O.Verifier(_zero_statement(d))
        total = (1 << d) ** 2
        for j in range(d):
            assert v.dims() == (1 << (d - j), 1 << (d - j))
            rec = _zero_record(d, j)
            assert v.update_commitment(rec) == 0
            assert len(rec) >= 2 * 8 * (1 << (d - j)) * 16
            nxt = v.dims()
            assert nxt == (1 << (d - j - 1), 1 << (d - j - 1))
            assert nxt[0] * nxt[1] * 4 == total
            total = nxt[0] * nxt[1]
        assert v.verify_final_opening(0) == 0


def test_e2e_succeeds_across_depth_sweep():
    """prover.rs:791-808 (d = 6..8)"""
    for d in [6, 7, 8]:
        v = O.Verifier(_zero_statement(d))
        for j in range(d):
            assert v.update_commitment(_zero_record(d, j)) == 0
        assert v.verify_final_opening(0) == 0


def test_round_mismatch_and_early_opening_are_errors():
    """verifier.rs:58, 123-126; prover.rs:141-142"""
    p = O.Prover(W.pattern_bytes(8192, 1))
    with pytest.raises(O.OracleError) as e:
        p.fold_round(1)
    assert e.value.rc == O.ORC_ERR_ROUND
    assert p.final_opening() is None
    v = O.Verifier(p.statement())
    assert v.update_commitment(_zero_record(6, 1)) == O.ORC_ERR_ROUND
    assert v.verify_final_opening(0) == O.ORC_ERR_NOT_FINAL


@pytest.mark.slow
def test_every_serialized_micro_block_byte_flip_is_rejected():
    """prover.rs:811-877 on a sample of byte positions per record (every 97th byte plus all
    length prefixes and round fields); the exhaustive sweep runs against the HIP verifier"""
    p, st, recs, opening = _prove(W.pattern_bytes(8192, 41))
    v = O.Verifier(st)
    prefixes = []
    for r in recs:
        prefixes.append(v.clone())
        assert v.update_commitment(r) == 0
    for ri, enc in enumerate(recs):
        positions = set(range(0, len(enc), 97)) | set(range(0, 16)) | {len(enc) - 1}
        for bi in sorted(positions):
            t = bytearray(enc)
            t[bi] ^= 1
            vv = prefixes[ri].clone()
            rejected = vv.update_commitment(bytes(t)) != 0
            if not rejected:
                for r in recs[ri + 1:]:
                    if vv.update_commitment(r) != 0:
                        rejected = True
                        break
            if not rejected:
                rejected = vv.verify_final_opening(opening) != 0
            assert rejected, (ri, bi)


def test_full_flow_accepts_and_is_deterministic():
    """examples/full_flow.rs:47-104 on a small file: the verifier accepts and a rerun reproduces
    every byte"""
    data = W.pattern_bytes(20000, 5)
    rc, st, recs, op = O.full_flow(data)
    assert rc == 0
    rc2, st2, recs2, op2 = O.full_flow(data)
    assert (st, recs, op) == (st2, recs2, op2)
    assert len(W.split_records(recs)) == 6


def test_challenge_sampler_rejects_zero_candidates():
    """transcript.rs:130-143: the sampler asks for counters 0, 1, 2 and returns the first
    non-zero candidate"""
    v, seen = O.sample_nonzero_mock(2, 9)
    assert v == 9 and seen == [0, 1, 2]
    v, seen = O.sample_nonzero_mock(0, 5)
    assert v == 5 and seen == [0]


def test_rejected_candidate_resamples_with_the_next_counter():
    """With the first candidate rejected (the test knob), the challenge is the counter-1
    candidate: the same message with u32le(1) as its last field, rebuilt byte by byte."""
    fid = bytes(range(32))
    cur, lhs, rhs = _tagged(5), _tagged(6)[:, :32], _tagged(7)[:, :32]

    def enc(m):
        out = struct.pack("<Q", 8)
        for k in range(8):
            out += struct.pack("<Q", m.shape[1]) + m[k].astype("<u8").tobytes()
        return out

    head = (EPOCH + struct.pack("<Q", 3) + b"vec" + struct.pack("<Q", 32) + fid + struct.pack("<QQ", 0, 4)
            + enc(cur) + enc(lhs) + enc(rhs))
    c0 = int.from_bytes(O.blake3(head + struct.pack("<I", 0)), "little") % O.Q
    c1 = int.from_bytes(O.blake3(head + struct.pack("<I", 1)), "little") % O.Q
    c3 = int.from_bytes(O.blake3(head + struct.pack("<I", 3)), "little") % O.Q
    assert c0 != c1
    try:
        assert O.challenge(EPOCH, "vec", fid, 4, cur, lhs, rhs) == c0
        O.set_challenge_reject(1)
        assert O.challenge(EPOCH, "vec", fid, 4, cur, lhs, rhs) == c1
        O.set_challenge_reject(3)
        assert O.challenge(EPOCH, "vec", fid, 4, cur, lhs, rhs) == c3
    finally:
        O.set_challenge_reject(0)
    assert O.challenge(EPOCH, "vec", fid, 4, cur, lhs, rhs) == c0


def test_full_flow_with_rejected_candidates_replays():
    """every challenge of a d = 6 proof resampled once: a different proof that the oracle's own
    verifier replays and accepts under the same knob"""
    data = W.pattern_bytes(8192, 5)
    rc0, st0, recs0, op0 = O.full_flow(data)
    try:
        O.set_challenge_reject(1)
        rc, st, recs, op = O.full_flow(data)
    finally:
        O.set_challenge_reject(0)
    assert rc0 == 0 and rc == 0
    assert st == st0 and recs != recs0
