"""GPU parity: every HIP result through the C ABI against the oracle (the CPU restatement of the
reference), bit for bit. Integer work, so the bar is exact equality."""
import os
import struct
import sys

import numpy as np
import pytest

import _oracle as O
import _wire as W

pytestmark = pytest.mark.gpu

dm = pytest.importorskip("damascus000_amd")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    if not dm.backend_ready():
        pytest.fail("HIP backend not ready on a GPU test run: " + dm.device_summary())
    say("device:", dm.device_summary())


# ---------------- reference-compatible entry points ----------------
def test_fold_batch_matches_field_arithmetic():
    rng = np.random.default_rng(11)
    n = 4099
    left = O.from_words(O.random_zq(rng, n))
    right = O.from_words(O.random_zq(rng, n))
    c = O.from_words(O.random_zq(rng, 1))[0]
    say("fold_batch n", n)
    rc, out = dm.fold_batch(left, right, c)
    assert rc == 0
    assert out == [(l + c * r) % O.Q for l, r in zip(left, right)]


def _plans(S):
    """stage roots as the reference builds them (ntt.rs:21-47), from the oracle's generator"""
    roots, iroots, inv_sizes = [], [], []
    for p in O.PRIMES:
        g = O.lib().orc_multiplicative_generator(p)
        w = pow(g, (p - 1) // S, p)
        wi = pow(w, p - 2, p)
        ln = 2
        while ln <= S:
            roots.append(pow(w, S // ln, p))
            iroots.append(pow(wi, S // ln, p))
            ln *= 2
        inv_sizes.append(pow(S, p - 2, p))
    return roots, iroots, inv_sizes


def test_ntt_batch_matches_reference_cyclic_product():
    """dmx_ntt_batch = the reference's cyclic NTT product (ntt.rs:197-224) per prime"""
    rng = np.random.default_rng(5)
    for ring_len in [1, 8, 64, 512]:
        S = 2 * ring_len
        a = np.zeros((8, S), dtype=np.uint64)
        b = np.zeros((8, S), dtype=np.uint64)
        for i, p in enumerate(O.PRIMES):
            a[i, :ring_len] = rng.integers(0, p, size=ring_len, dtype=np.uint64)
            b[i, :ring_len] = rng.integers(0, p, size=ring_len, dtype=np.uint64)
        roots, iroots, inv_sizes = _plans(S)
        say("ntt_batch S", S)
        rc, out = dm.ntt_batch(a, b, S, O.PRIMES, roots, iroots, inv_sizes)
        assert rc == 0
        for i, p in enumerate(O.PRIMES):
            fa = O.ntt_forward(a[i], p)
            fb = O.ntt_forward(b[i], p)
            want = O.ntt_inverse((fa.astype(object) * fb.astype(object) % p).astype(np.uint64), p)
            assert np.array_equal(out[i], want), (S, p)


def test_ntt_batch_large_plans_and_repeat_calls():
    """n up to 2^14 (ring length 8192), each plan called twice (its twiddle tables are kept), a
    second plan between the calls: every product equals the oracle's cyclic NTT product"""
    rng = np.random.default_rng(6)
    for S in [4096, 16384, 4096]:
        a = np.stack([rng.integers(0, p, size=S, dtype=np.uint64) for p in O.PRIMES])
        b = np.stack([rng.integers(0, p, size=S, dtype=np.uint64) for p in O.PRIMES])
        roots, iroots, inv_sizes = _plans(S)
        say("ntt_batch S", S)
        rc, out = dm.ntt_batch(a, b, S, O.PRIMES, roots, iroots, inv_sizes)
        assert rc == 0
        for i, p in enumerate(O.PRIMES[:3]):
            fa = O.ntt_forward(a[i], p)
            fb = O.ntt_forward(b[i], p)
            want = O.ntt_inverse((fa.astype(object) * fb.astype(object) % p).astype(np.uint64), p)
            assert np.array_equal(out[i], want), (S, p)
        rc2, out2 = dm.ntt_batch(a, b, S, O.PRIMES, roots, iroots, inv_sizes)
        assert rc2 == 0 and np.array_equal(out2, out)


def test_drop_ins_decline_what_the_reference_would_not_compute_the_same_way():
    """Return code 1 (the caller's CPU path runs, gpu.rs:191-193) for a stage root whose order is
    not 2^(s+1) (the reference's stage span would differ), a root or input not below p, a size past
    2^14, and fold values that are not canonical elements of Z_q"""
    S = 64
    a = np.zeros((8, S), dtype=np.uint64)
    b = np.zeros((8, S), dtype=np.uint64)
    roots, iroots, inv_sizes = _plans(S)
    bad = list(roots)
    bad[1] = 1  # stage 1 of prime 0 with a root of order 1
    assert dm.ntt_batch(a, b, S, O.PRIMES, bad, iroots, inv_sizes)[0] == 1
    big = list(iroots)
    big[2] += O.PRIMES[0]  # a root not below p
    assert dm.ntt_batch(a, b, S, O.PRIMES, roots, big, inv_sizes)[0] == 1
    a2 = a.copy()
    a2[3, 5] = O.PRIMES[3]  # an input that is not a residue below p
    assert dm.ntt_batch(a2, b, S, O.PRIMES, roots, iroots, inv_sizes)[0] == 1
    assert dm.ntt_batch(a, b, S, O.PRIMES, roots, iroots, inv_sizes)[0] == 0  # the same plan still runs
    n = 1 << 15
    r15, i15, s15 = _plans(n)
    z = np.zeros((8, n), dtype=np.uint64)
    assert dm.ntt_batch(z, z, n, O.PRIMES, r15, i15, s15)[0] == 1
    assert dm.fold_batch([O.Q], [1], 3)[0] == 1
    assert dm.fold_batch([1], [2], O.Q)[0] == 1
    rc, out = dm.fold_batch([O.Q - 1, 0], [O.Q - 1, 5], O.Q - 1)
    assert rc == 0 and out == [(O.Q - 1 + (O.Q - 1) ** 2) % O.Q, (5 * (O.Q - 1)) % O.Q]


# ---------------- ring products ----------------
def test_ring_products_match_golden_schoolbook():
    import json
    import os

    with open(os.path.join(os.path.dirname(__file__), "golden", "negacyclic.json")) as f:
        cases = json.load(f)
    a = [O.to_words(c["a"]) for c in cases]
    b = [O.to_words(c["b"]) for c in cases]
    say("ring_mul golden", len(cases))
    outs = dm.ring_mul_batch(a, b)
    for c, o in zip(cases, outs):
        assert O.from_words(o) == [int(x) for x in c["out"]], (c["n"], c["kind"])


def test_ragged_ring_products_match_oracle():
    """C3 shape at reduced count: n drawn from 2^6..2^12 per pair, uniform coefficients"""
    rng = np.random.default_rng(33)
    logs = rng.integers(6, 13, size=40)
    a = [O.random_zq(rng, 1 << int(l)) for l in logs]
    b = [O.random_zq(rng, 1 << int(l)) for l in logs]
    say("ragged ring_mul", len(logs))
    outs = dm.ring_mul_batch(a, b)
    for x, y, o in zip(a, b, outs):
        assert np.array_equal(o, O.negacyclic_multiply(x, y))


def test_ring_products_small_degrees_and_extremes():
    rng = np.random.default_rng(2)
    a, b = [], []
    for lg in range(0, 14):
        n = 1 << lg
        a.append(O.random_zq(rng, n))
        b.append(O.random_zq(rng, n))
        top = O.to_words([O.Q - 1] * n)
        a.append(top)
        b.append(top.copy())
    say("ring_mul degrees 1..8192")
    outs = dm.ring_mul_batch(a, b)
    for x, y, o in zip(a, b, outs):
        if x.shape[0] <= 2048:
            assert np.array_equal(o, O.negacyclic_multiply(x, y)), x.shape[0]
        else:  # the oracle's big-int lift is slow at 8192; check X^n+1 linearity instead
            s = O.random_zq(np.random.default_rng(x.shape[0]), 1)
            lhs = dm.ring_mul_batch([O.zq_binop("mul", x, np.repeat(s, x.shape[0], 0))], [y])[0]
            rhs = O.zq_binop("mul", o, np.repeat(s, x.shape[0], 0))
            assert np.array_equal(lhs, rhs)


# ---------------- commitment ----------------
def test_generators_match_oracle():
    gs = O.derive_subseed(bytes([7] * 32), b"g0")
    assert dm.derive_subseed(bytes([7] * 32), b"g0") == gs
    say("generators 5x8x64")
    c = dm.ModuleSisCommitter(bytes([7] * 32))
    assert np.array_equal(c.generators_for(5, 64), O.derive_generators(gs, 5, 64))


@pytest.mark.parametrize("rows,n", [(1, 1), (3, 2), (4, 64), (7, 128), (64, 64), (2, 2048)])
def test_commit_matches_oracle(rows, n):
    rng = np.random.default_rng(rows * 1000 + n)
    w = np.stack([O.random_zq(rng, n) for _ in range(rows)])
    say("commit", rows, n)
    got = dm.ModuleSisCommitter(bytes([9] * 32)).commit(w)
    assert np.array_equal(got, O.commit(bytes([9] * 32), w))


def test_commit_empty_witness_is_zero_of_length_one():
    got = dm.ModuleSisCommitter().commit(np.zeros((0, 1, 2), np.uint64))
    assert got.shape == (8, 1, 2) and not got.any()


def test_commit_with_generators_all_top_coefficients():
    """every coefficient q-1: the largest centred sums the lift must handle"""
    rows, n = 16, 256
    w = np.stack([O.to_words([O.Q - 1] * n) for _ in range(rows)])
    g = np.stack([np.stack([O.to_words([O.Q - 1] * n) for _ in range(8)]) for _ in range(rows)])
    say("commit top", rows, n)
    got = dm.ModuleSisCommitter().commit_with_generators(w, g)
    assert np.array_equal(got, O.commit_with_generators(w, g))


# ---------------- protocol ----------------
def _gpu_prove(data, params=None):
    p = dm.DamascusProver.initialize(data, params)
    st = p.statement()
    recs = [p.fold_round(j) for j in range(p.rounds_total())]
    return p, st, recs, p.final_opening()


@pytest.mark.parametrize("size,tweak", [(8192, 9), (53248, 0), (60000, 3)])
def test_prover_bit_exact_with_oracle(size, tweak):
    data = W.pattern_bytes(size, tweak)
    say("prove", size)
    p, st, recs, op = _gpu_prove(data)
    orc = O.Prover(data)
    assert st == orc.statement()
    for j, r in enumerate(recs):
        assert r == orc.fold_round(j), j
    assert op == orc.final_opening()


def test_c1_random_file_full_flow_matches_oracle():
    """the CPU-runnable case: 53,248 seeded uniform bytes, d=6, default seeds: statement, records, m*"""
    data = np.random.default_rng(20261203).integers(0, 256, size=53248, dtype=np.uint8).tobytes()
    say("C1 full flow")
    st, recs, op = dm.prove(data)
    rc, ost, orecs, oop = O.full_flow(data)
    assert rc == 0 and (st, recs, op) == (ost, orecs, oop)
    dm.verify(st, recs, op)


def test_edge_file_sizes_match_oracle():
    """io.rs:72-120 edges: an empty file (one zero coefficient, d = 6), one byte, exactly 13 * 4^6
    bytes (the largest d = 6 file) and one byte more (the smallest d = 7 file), through the
    host-bytes path and the fused device path, each equal to the oracle's proof."""
    ctx = dm.Context.default()
    for n in [0, 1, 13 * 4 ** 6, 13 * 4 ** 6 + 1]:
        data = np.random.default_rng(n + 3).integers(0, 256, size=n, dtype=np.uint8).tobytes()
        say("edge size", n)
        rc, ost, orecs, oop = O.full_flow(data)
        assert rc == 0
        assert dm.prove(data) == (ost, orecs, oop), n
        if n:
            buf = ctx.upload(data)
            assert dm.prove_verify_device(buf.ptr, n) == (ost, orecs, oop), n
        dm.verify(ost, orecs, oop)


def test_deepest_supported_file_proves_and_verifies():
    """d = 12, the largest depth (4096 x 4096 witness, 2^24 coefficients, a 2 GiB generator
    matrix): the smallest such file proves and its fused verifier replay accepts; the host-bytes
    prover gives the same proof bytes. One byte past the largest d = 12 file is refused."""
    n = 13 * 4 ** 11 + 1
    data = np.random.default_rng(12).integers(0, 256, size=n, dtype=np.uint8).tobytes()
    buf = dm.Context.default().upload(data)
    say("d = 12 fused job")
    fused = dm.prove_verify_device(buf.ptr, n)
    del buf
    say("d = 12 host-bytes prover")
    assert dm.prove(data) == fused
    with pytest.raises(dm.DamascusError) as e:
        dm.DamascusProver.initialize(b"\0" * (13 * 4 ** 12 + 1))
    assert e.value.name in ("ERR_ARG", "ERR_MEMORY_LIMIT")


def test_d7_random_file_matches_oracle_with_other_seeds():
    params_seed, epoch = bytes(range(32)), bytes(range(100, 132))
    data = np.random.default_rng(7).integers(0, 256, size=13 * 4096 + 1, dtype=np.uint8).tobytes()  # d = 7
    say("d7 full flow")
    st, recs, op = dm.prove(data, dm.SystemParams(params_seed, epoch))
    rc, ost, orecs, oop = O.full_flow(data, params_seed, epoch)
    assert rc == 0 and (st, recs, op) == (ost, orecs, oop)


def test_gpu_verifier_accepts_and_rejects_tampering():
    """prover.rs:547-588 on the GPU verifier"""
    data = W.pattern_bytes(8192, 9)
    p, st, recs, op = _gpu_prove(data)
    v = dm.DamascusVerifier(None, st)
    for r in recs:
        v.update_commitment(r)
    v.verify_final_opening(op)
    rnd, mods, _ = W.decode_record(recs[0])
    mods[0][0, 0] = O.to_words([(O.from_words(mods[0][0, 0])[0] + 1) % O.Q])[0]
    bad = W.encode_record(rnd, *mods)
    v2 = dm.DamascusVerifier(None, st)
    v2.update_commitment(bad)
    for r in recs[1:]:
        v2.update_commitment(r)
    with pytest.raises(dm.DamascusError) as e:
        v2.verify_final_opening(op)
    assert e.value.name == "ERR_OPENING"


def test_gpu_verifier_matches_oracle_verifier_state():
    data = W.pattern_bytes(20000, 4)
    p, st, recs, op = _gpu_prove(data)
    v = dm.DamascusVerifier(None, st)
    ov = O.Verifier(st)
    for r in recs:
        v.update_commitment(r)
        assert ov.update_commitment(r) == 0
        assert v.current_dimensions() == ov.dims()
        assert np.array_equal(v.current_commitment(), ov.current_commitment())


def test_gpu_verifier_zero_statement_depth_sweep():
    """prover.rs:741-808: all-zero statements and records replay and accept for d = 6..9"""
    for d in [6, 7, 8, 9]:
        st = W.encode_statement(bytes([d] * 32), W.byte_len_for_depth(d), d, W.zero_module(1 << d), bytes([17] * 32))
        v = dm.DamascusVerifier(None, st)
        for j in range(d):
            assert v.current_dimensions() == (1 << (d - j), 1 << (d - j))
            cur, nxt = 1 << (d - j), 1 << (d - j - 1)
            v.update_commitment(W.encode_record(j, W.zero_module(cur), W.zero_module(cur), W.zero_module(nxt),
                                                W.zero_module(nxt)))
        v.verify_final_opening(0)


def test_gpu_verifier_split_generator_fold_accepts_and_rejects():
    """The terminal generator is folded in two parts, the rounds before the last five and then
    those five. A real proof replayed record by record accepts, and the same replay with the
    opening changed by one is rejected, at d = 7 and 8 (2 and 3 early rounds)."""
    for d in [7, 8]:
        data = W.pattern_bytes(W.byte_len_for_depth(d), d)
        say("split generator fold, d =", d)
        p, st, recs, op = _gpu_prove(data)
        rc, ost, orecs, oop = O.full_flow(data)
        assert rc == 0 and st == ost and b"".join(recs) == orecs and op == oop
        for opening, ok in [(op, True), ((op + 1) % O.Q, False)]:
            v = dm.DamascusVerifier(None, st)
            for r in recs:
                v.update_commitment(r)
            if ok:
                v.verify_final_opening(opening)
            else:
                with pytest.raises(dm.DamascusError) as e:
                    v.verify_final_opening(opening)
                assert e.value.name == "ERR_OPENING"


def test_gpu_verifier_error_paths():
    data = W.pattern_bytes(8192, 2)
    p, st, recs, op = _gpu_prove(data)
    v = dm.DamascusVerifier(None, st)
    with pytest.raises(dm.DamascusError) as e:
        v.update_commitment(recs[1])
    assert e.value.name == "ERR_ROUND"
    with pytest.raises(dm.DamascusError) as e:
        v.verify_final_opening(op)
    assert e.value.name == "ERR_NOT_FINAL"
    with pytest.raises(dm.DamascusError) as e:
        v.update_commitment(recs[0][:100])
    assert e.value.name == "ERR_DECODE"
    rnd, mods, _ = W.decode_record(recs[0])
    short = W.encode_record(0, mods[0][:, :32], mods[1][:, :32], mods[2], mods[3])
    with pytest.raises(dm.DamascusError) as e:
        v.update_commitment(short)
    assert e.value.name == "ERR_SHAPE"
    v.update_commitment(recs[0])  # state unchanged by the rejected records
    assert v.round() == 1


def test_prover_round_mismatch_and_memory_limit():
    with pytest.raises(dm.DamascusError) as e:
        dm.DamascusProver.initialize(W.pattern_bytes(1 << 20, 7), config=dm.RuntimeConfig(max_preprocess_bytes=1024))
    assert e.value.name == "ERR_MEMORY_LIMIT"
    p = dm.DamascusProver.initialize(W.pattern_bytes(8192, 1))
    with pytest.raises(dm.DamascusError) as e:
        p.fold_round(1)
    assert e.value.name == "ERR_ROUND"
    assert p.final_opening() is None


def test_prove_verify_device_matches_host_api():
    data = np.random.default_rng(99).integers(0, 256, size=150000, dtype=np.uint8).tobytes()  # d = 7
    ctx = dm.Context.default()
    buf = ctx.upload(data)
    assert buf.read() == data
    say("prove_verify_device", len(data))
    st, recs, op = dm.prove_verify_device(buf.ptr, len(data))
    st2, recs2, op2 = dm.prove(data)
    assert (st, recs, op) == (st2, recs2, op2)
    rc, ost, orecs, oop = O.full_flow(data)
    assert rc == 0 and (st, recs, op) == (ost, orecs, oop)
    assert dm.prove_verify_device(buf.ptr, len(data), want_proof=False) is None


def test_graph_replays_match_eager_runs():
    """the fused job is captured into a hipGraph on first use and replayed afterwards: every
    replay, with eager work and synchronisations in between, equals the eager result"""
    import os

    data = np.random.default_rng(123).integers(0, 256, size=13 * 5000, dtype=np.uint8).tobytes()  # d = 7
    ctx = dm.Context.default()
    buf = ctx.upload(data)
    eager = dm.prove(data)
    os.environ["DAMASCUS_GRAPH"] = "0"
    try:
        assert dm.prove_verify_device(buf.ptr, len(data)) == eager
    finally:
        os.environ.pop("DAMASCUS_GRAPH")
    for rep in range(4):
        say("graph replay", rep)
        assert dm.prove_verify_device(buf.ptr, len(data)) == eager, rep
        other = W.pattern_bytes(8192, rep)  # eager work between replays
        st, recs, op = dm.prove(other)
        dm.verify(st, recs, op)
    buf2 = ctx.upload(bytes(reversed(data)))  # same shape, different file: replays with new bytes
    assert dm.prove_verify_device(buf2.ptr, len(data)) == dm.prove(bytes(reversed(data)))


def _rand_file(rng, n):
    return rng.integers(0, 256, size=n, dtype=np.uint8).tobytes()


def test_verifier_computes_inverses_when_the_candidate_fails_its_check():
    """The fused job's verifier takes the prover's challenge inverses as candidates it checks
    (x * candidate == 1); offered x itself instead, it inverts on its own and still accepts the
    proof, whose bytes do not change. Eager launches, so no graph captured earlier is replayed."""
    data = np.random.default_rng(31).integers(0, 256, size=13 * 4096 + 7, dtype=np.uint8).tobytes()  # d = 7
    buf = dm.Context.default().upload(data)
    want = dm.prove_verify_device(buf.ptr, len(data))
    os.environ["DAMASCUS_GRAPH"] = "0"
    os.environ["DAMASCUS_VERIFIER_HINT"] = "wrong"
    try:
        say("verifier with wrong inverse candidates")
        got = dm.prove_verify_device(buf.ptr, len(data))
    finally:
        del os.environ["DAMASCUS_VERIFIER_HINT"]
        del os.environ["DAMASCUS_GRAPH"]
    assert got == want
    rc, ost, orecs, oop = O.full_flow(data)
    assert rc == 0 and got == (ost, orecs, oop)


def test_fused_job_rejects_a_wrong_final_opening():
    """The fused job's verifier checks g_d[0] * m* == C_d: checked against another witness
    coefficient instead of m*, the job fails with ERR_OPENING (eager launches, so no graph
    captured earlier is replayed); with the right opening the same file proves and verifies."""
    data = np.random.default_rng(41).integers(0, 256, size=13 * 4096 + 9, dtype=np.uint8).tobytes()  # d = 7
    buf = dm.Context.default().upload(data)
    os.environ["DAMASCUS_GRAPH"] = "0"
    os.environ["DAMASCUS_TEST_OPENING_OFFSET"] = "1"
    try:
        say("fused job with a wrong final opening")
        with pytest.raises(dm.DamascusError) as e:
            dm.prove_verify_device(buf.ptr, len(data))
        assert e.value.name == "ERR_OPENING"
    finally:
        del os.environ["DAMASCUS_TEST_OPENING_OFFSET"]
        del os.environ["DAMASCUS_GRAPH"]
    rc, ost, orecs, oop = O.full_flow(data)
    assert rc == 0 and dm.prove_verify_device(buf.ptr, len(data)) == (ost, orecs, oop)


def test_transform_variants_give_identical_proofs():
    """Every forward-transform variant (register blockings of the prime slices, the three-pass
    kernels at 2^8..2^11 and the per-layer kernel) gives the same proof: at d = 7 each against the
    oracle, at d = 8..11 (generator rows at 2^8..2^11, fold rows down to 2^6) across variants
    through the fused job, whose captured graph is retaken after a change."""
    ctx = dm.Context.default()
    saved = {lg: ctx.transform_variant(lg)[0] for lg in range(6, 12)}
    say("tuned transform variants (variant, count, us):", {lg: ctx.transform_variant(lg) for lg in range(6, 12)})
    small = W.pattern_bytes(W.byte_len_for_depth(7), 5)
    d8 = np.random.default_rng(8).integers(0, 256, size=512 << 10, dtype=np.uint8).tobytes()  # d = 8
    d9 = np.random.default_rng(9).integers(0, 256, size=2 << 20, dtype=np.uint8).tobytes()  # d = 9
    mid = np.random.default_rng(10).integers(0, 256, size=8 << 20, dtype=np.uint8).tobytes()  # d = 10
    big = np.random.default_rng(11).integers(0, 256, size=16 << 20, dtype=np.uint8).tobytes()  # d = 11
    rc, ost, orecs, oop = O.full_flow(small)
    assert rc == 0
    b8, b9, mbuf, buf = ctx.upload(d8), ctx.upload(d9), ctx.upload(mid), ctx.upload(big)
    outs, mids, d9s, d8s = [], [], [], []
    try:
        for v in range(3):
            for lg in range(6, 12):
                ctx.set_transform_variant(lg, min(v, ctx.transform_variant(lg)[1] - 1))
            say("all degrees at variant", v)
            assert dm.prove(small) == (ost, orecs, oop)
            d8s.append(dm.prove_verify_device(b8.ptr, len(d8)))
            d9s.append(dm.prove_verify_device(b9.ptr, len(d9)))
            mids.append(dm.prove_verify_device(mbuf.ptr, len(mid)))
            outs.append(dm.prove_verify_device(buf.ptr, len(big)))
        assert d8s[0] == d8s[1] == d8s[2]
        assert d9s[0] == d9s[1] == d9s[2]
        assert mids[0] == mids[1] == mids[2]
        assert outs[0] == outs[1] == outs[2]
        with pytest.raises(dm.DamascusError):
            ctx.set_transform_variant(11, 3)  # degree 2^11 has three variants
    finally:
        for lg, v in saved.items():
            ctx.set_transform_variant(lg, v)



def test_mac_shapes_give_identical_proofs():
    """Every MAC-pass segmentation shape gives the same proof: at d = 7 each against the oracle
    (the host-API prover, per-stage launches), at d = 10 across the fewest, default and most
    segments through the fused job, whose captured graph is retaken after a change."""
    ctx = dm.Context.default()
    saved = {lg: ctx.mac_shape(lg)[0] for lg in range(6, 12)}
    say("tuned MAC shapes (shape, count, us):", {lg: ctx.mac_shape(lg) for lg in range(6, 11)})
    small = W.pattern_bytes(W.byte_len_for_depth(7), 7)
    d8 = np.random.default_rng(8).integers(0, 256, size=512 << 10, dtype=np.uint8).tobytes()  # d = 8
    d9 = np.random.default_rng(9).integers(0, 256, size=2 << 20, dtype=np.uint8).tobytes()  # d = 9
    mid = np.random.default_rng(10).integers(0, 256, size=8 << 20, dtype=np.uint8).tobytes()  # d = 10
    rc, ost, orecs, oop = O.full_flow(small)
    assert rc == 0
    buf = ctx.upload(mid)
    outs = []
    try:
        for v in range(ctx.MAC_SHAPES):
            for lg in range(6, 12):
                ctx.set_mac_shape(lg, v)
            assert dm.prove(small) == (ost, orecs, oop), v
            if v in (0, 4, 8):
                say("fused d = 10 job with MAC shape", v)
                outs.append(dm.prove_verify_device(buf.ptr, len(mid)))
        assert outs[0] == outs[1] == outs[2]
        with pytest.raises(dm.DamascusError):
            ctx.set_mac_shape(10, ctx.MAC_SHAPES)
    finally:
        for lg, v in saved.items():
            ctx.set_mac_shape(lg, v)


def test_tuning_cache_round_trip(tmp_path):
    """A context created with a tuning cache file times its choices and writes them there; a
    second context reads them back instead of timing (same transform variants and MAC shapes)."""
    path = str(tmp_path / "tuning.txt")
    old = os.environ.get("DAMASCUS_TUNE_CACHE")
    os.environ["DAMASCUS_TUNE_CACHE"] = path
    try:
        a = dm.Context()
        assert a.tuning_source() == ("timed", path) and os.path.exists(path)
        b = dm.Context()
        assert b.tuning_source() == ("cache", path)
        # the cache keeps times to 0.001 us and the mirror rounds them to 0.01 us, so a time on a
        # rounding boundary may come back one step of 0.01 away
        for lg in range(0, 14):
            va, na, ta = a.transform_variant(lg)
            vb, nb, tb = b.transform_variant(lg)
            assert (va, na) == (vb, nb) and all(abs(x - y) < 0.015 for x, y in zip(ta, tb)), (lg, ta, tb)
            sa, ca, ua = a.mac_shape(lg)
            sb, cb, ub = b.mac_shape(lg)
            assert (sa, ca) == (sb, cb) and all(abs(x - y) < 0.015 for x, y in zip(ua, ub)), (lg, ua, ub)
        with open(path) as f:
            assert f.readline().split()[:2] == ["dmx-tune", "2"]
        say("tuning cache", path, "MAC shapes", {lg: b.mac_shape(lg)[0] for lg in range(6, 11)})
    finally:
        if old is None:
            del os.environ["DAMASCUS_TUNE_CACHE"]
        else:
            os.environ["DAMASCUS_TUNE_CACHE"] = old

def test_batch_of_one_depth_equals_solo_runs():
    """five d=6 files of different lengths and seeds proved side by side in one launch sequence:
    each proof equals proving that file alone and the oracle's, so no job reads another's data"""
    rng = np.random.default_rng(2024)
    sizes = [0, 1, 4096, 13 * 4096, 50000]  # all d = 6 (capacity 53,248 bytes)
    files = [_rand_file(rng, n) for n in sizes]
    params = [dm.SystemParams(bytes([i + 1] * 32), bytes([2 * i + 3] * 32)) for i in range(len(files))]
    say("batch d6", sizes)
    out = dm.prove_verify_many(files, params)
    for f, p, o in zip(files, params, out):
        assert o == dm.prove(f, p), len(f)
    rc, ost, orecs, oop = O.full_flow(files[3], bytes([4] * 32), bytes([9] * 32))
    assert rc == 0 and out[3] == (ost, orecs, oop)


def test_ragged_batch_mixed_depths_device_files():
    """a ragged batch over d = 6..8 with device-resident files, split into launches of at most
    two jobs, run twice with new contents of the same lengths (the cached graphs replay):
    every proof equals its solo run"""
    rng = np.random.default_rng(77)
    sizes = [100, 60000, 200000, 53248, 300000, 13 * 4 ** 7, 700000, 5]  # d = 6,7,7,6,8,7,8,6
    ctx = dm.Context.default()
    for rep in range(2):
        files = [_rand_file(rng, n) for n in sizes]
        bufs = [ctx.upload(f) for f in files]
        say("ragged batch rep", rep, sizes)
        out = dm.prove_verify_many([(b.ptr, len(f)) for b, f in zip(bufs, files)], max_batch=2)
        for f, o in zip(files, out):
            assert o == dm.prove(f), len(f)
    rc, ost, orecs, oop = O.full_flow(files[2])
    assert rc == 0 and out[2] == (ost, orecs, oop)


def test_kernel_timing_counts_forward_transforms():
    ctx = dm.Context.default()
    data = W.pattern_bytes(8192, 3)
    buf = ctx.upload(data)
    ctx.kernel_timing(True)
    dm.prove_verify_device(buf.ptr, len(data), want_proof=False)
    ms, n, work = ctx.kernel_timing()
    ctx.kernel_timing(False)
    # init (witness + generators) + 2 per round; stages of degree <= 32 run as single launches
    # (k_stage_small) whose transforms are not timed separately
    small = os.environ.get("DAMASCUS_SMALL_STAGE", "1") != "0"
    rounds_timed = sum(1 for j in range(6) if not small or (64 >> j) > 32)
    assert n == 1 + 2 * rounds_timed and ms > 0 and work > 0


def test_every_byte_flip_of_first_two_records_is_rejected_on_gpu():
    """prover.rs:811-877, exhaustive over records 0..1 of a d=6 proof on positions of every
    structural field and every 13th coefficient byte"""
    data = W.pattern_bytes(8192, 41)
    p, st, recs, op = _gpu_prove(data)
    v = dm.DamascusVerifier(None, st)
    prefixes = []
    for r in recs:
        prefixes.append(v.clone())
        v.update_commitment(r)
    v.verify_final_opening(op)
    for ri in range(2):
        enc = recs[ri]
        positions = sorted(set(range(0, len(enc), 13)) | set(range(16)) | {len(enc) - 1})
        say("byte flips record", ri, len(positions))
        for bi in positions:
            t = bytearray(enc)
            t[bi] ^= 1
            vv = prefixes[ri].clone()
            rejected = False
            try:
                vv.update_commitment(bytes(t))
                for r in recs[ri + 1:]:
                    vv.update_commitment(r)
                vv.verify_final_opening(op)
            except dm.DamascusError:
                rejected = True
            assert rejected, (ri, bi)


def test_server_results_equal_solo_runs_in_the_mix():
    """jobs submitted to the continuous-batching server while earlier ones run (depths 6..8,
    staggered arrivals, batches of up to 4, three batches in flight) give exactly the proofs of
    solo runs, and the oracle's"""
    import time

    rng = np.random.default_rng(515)
    sizes = [70000, 100, 300000, 53248, 250000, 9000, 800000, 60000, 13, 200000]
    files = [_rand_file(rng, n) for n in sizes]
    params = [dm.SystemParams(bytes([i + 3] * 32), bytes([i + 5] * 32)) for i in range(len(files))]
    ctx = dm.Context.default()
    bufs = [ctx.upload(f) for f in files]
    srv = dm.Server(ctx, max_batch=4, max_inflight=3)
    try:
        tickets = []
        for i, (b, f) in enumerate(zip(bufs, files)):
            say("server submit", i, len(f))
            tickets.append(srv.submit((b.ptr, len(f)), params[i]))
            if i % 3 == 2:
                time.sleep(0.002)  # later arrivals land while earlier batches run
        res = [srv.wait(t) for t in tickets]
    finally:
        srv.close()
    for f, p, (r, lat) in zip(files, params, res):
        assert r == dm.prove(f, p), len(f)
        assert lat > 0
    rc, ost, orecs, oop = O.full_flow(files[1], bytes([4] * 32), bytes([6] * 32))
    assert rc == 0 and res[1][0] == (ost, orecs, oop)


def test_server_host_files_and_pending_wait():
    rng = np.random.default_rng(516)
    files = [_rand_file(rng, n) for n in (5000, 120000, 40000)]
    srv = dm.Server(files_on_host=True, max_batch=2)
    try:
        tickets = [srv.submit(f) for f in files]
        res = [srv.wait(t) for t in tickets]
        with pytest.raises(dm.DamascusError):
            srv.wait(tickets[0])  # a ticket is released once waited for
    finally:
        srv.close()
    for f, (r, lat) in zip(files, res):
        assert r == dm.prove(f)


def test_resident_generators_across_seeds_and_batches():
    """batches whose jobs share a generator seed read one resident matrix per (seed, depth) from
    the seed's second batch on (its first expands per job, as a request with a seed of its own
    does); a second seed gets its own matrix, the first is reused later, and a batch mixing seeds
    expands per job: every proof equals its solo run"""
    rng = np.random.default_rng(808)
    pa = dm.SystemParams(bytes([21] * 32), bytes([22] * 32))
    pb = dm.SystemParams(bytes([23] * 32), bytes([22] * 32))
    for rep, plist in enumerate([[pa] * 3, [pb] * 3, [pa] * 3, [pb] * 3, [pa] * 3, [pa, pb, pa]]):
        files = [_rand_file(rng, n) for n in (60000, 150000, 200000)]  # d = 7
        say("resident generators rep", rep)
        out = dm.prove_verify_many(files, plist)
        for f, p, o in zip(files, plist, out):
            assert o == dm.prove(f, p)


def test_persistent_ragged_products_equal_per_degree_launches_and_oracle():
    """the C3 shape at reduced count on device-resident operands: one persistent launch and the
    per-degree launch sequences give the same products, equal to the oracle on every pair"""
    rng = np.random.default_rng(3030)
    logs = [int(x) for x in rng.integers(6, 13, size=60)] + [0, 1, 13]
    a = [O.random_zq(rng, 1 << l) for l in logs]
    b = [O.random_zq(rng, 1 << l) for l in logs]
    A = np.ascontiguousarray(np.concatenate(a), dtype=np.uint64)
    B = np.ascontiguousarray(np.concatenate(b), dtype=np.uint64)
    ctx = dm.Context.default()
    da, db = ctx.upload(A.tobytes()), ctx.upload(B.tobytes())
    outs = []
    for mode in (0, 1):
        do = dm.DeviceBuffer(ctx, A.nbytes)
        say("ring_mul_batch_device mode", mode, len(logs))
        dm.ring_mul_batch_device(da.ptr, db.ptr, do.ptr, logs, mode)
        outs.append(np.frombuffer(do.read(), dtype=np.uint64).reshape(-1, 2))
    assert np.array_equal(outs[0], outs[1])
    pos = 0
    for x, y in zip(a, b):
        n = x.shape[0]
        if n <= 2048:
            assert np.array_equal(outs[0][pos: pos + n], O.negacyclic_multiply(x, y)), n
        pos += n


def test_ntt_off_schoolbook_products_match_ntt_products_and_oracle():
    """RuntimeConfig.ntt_enabled = false takes the reference's schoolbook product
    (naive_negacyclic); results are identical to the NTT path and to the oracle"""
    rng = np.random.default_rng(4040)
    logs = [0, 1, 3, 6, 9, 10, 12]
    a = [O.random_zq(rng, 1 << l) for l in logs]
    b = [O.random_zq(rng, 1 << l) for l in logs]
    a.append(O.to_words([O.Q - 1] * 256))
    b.append(O.to_words([O.Q - 1] * 256))
    say("schoolbook products", logs)
    naive = dm.ring_mul_batch(a, b, config=dm.RuntimeConfig(ntt_enabled=False))
    fast = dm.ring_mul_batch(a, b)
    for x, y, p, q in zip(a, b, naive, fast):
        assert np.array_equal(p, q), x.shape[0]
        if x.shape[0] <= 1024:
            assert np.array_equal(p, O.negacyclic_multiply(x, y))
    with pytest.raises(dm.DamascusError):
        dm.ring_mul_batch(a[:1], b[:1], config=dm.RuntimeConfig(gpu_enabled=False))


def test_server_ingests_mapped_host_files_through_pinned_staging(tmp_path):
    """files handed over as read-only memory maps of files on disk (the reference maps its
    input): the server copies them through pinned staging with asynchronous uploads"""
    rng = np.random.default_rng(517)
    files = [_rand_file(rng, n) for n in (70000, 1000, 300000, 53248)]
    maps = []
    for i, f in enumerate(files):
        path = tmp_path / f"in{i}.bin"
        path.write_bytes(f)
        maps.append(dm.map_file(str(path)))
    srv = dm.Server(files_on_host=True, max_batch=2, max_inflight=2)
    try:
        tickets = [srv.submit(m.job()) for m in maps]
        res = [srv.wait(t) for t in tickets]
    finally:
        srv.close()
    for f, (r, lat) in zip(files, res):
        assert r == dm.prove(f), len(f)
    many = dm.prove_verify_many([m.job() for m in maps], files_on_host=True)
    assert many == [r for r, _ in res]
