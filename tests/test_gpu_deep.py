"""GPU parity at full sizes and at the edges of the serving path, against the oracle (the CPU
restatement of the reference) bit for bit:

* the 8 MiB file (d = 10) and a 16 MiB file (d = 11): whole proofs;
* file_id of multi-MiB files, whose BLAKE3 tree has a group level (src/utils/io.rs:33);
* a batch of d = 6..10 files with their own generator seeds, and a served mix with d = 9..11
  requests: every job equal to its solo run and to the oracle;
* the files on either side of the depth boundaries 7/8, 8/9 and 9/10 in one batch, each equal to
  the oracle;
* the committed C1 golden fixture (tests/golden/c1_*);
* the resampling branch of sample_nonzero_challenge (src/protocol/transcript.rs:66-75), forced
  by a test knob the oracle honours too, at the chunk level and at the group level (d = 12);
* server admission within a small device-memory budget, a failing job isolated from the others
  in its batch (src/utils/gpu.rs:122-131: a failure is per call), and file pointers of the wrong
  kind refused before any launch.
"""
import os
import sys
import time

import numpy as np
import pytest

import _oracle as O
import _wire as W

pytestmark = pytest.mark.gpu

dm = pytest.importorskip("damascus000_amd")
sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import gen_c1_fixture as G  # noqa: E402

THREADS = min(16, os.cpu_count() or 1)
DEFAULT_SEEDS = (bytes([7] * 32), bytes([11] * 32))


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _device():
    if not dm.backend_ready():
        pytest.fail("HIP backend not ready on a GPU test run: " + dm.device_summary())
    O.set_threads(THREADS)
    say("device:", dm.device_summary(), "oracle threads:", THREADS)


def rand(n, seed):
    return np.random.default_rng(seed).integers(0, 256, size=int(n), dtype=np.uint8).tobytes()


_orc = {}


def oracle(data, seeds=DEFAULT_SEEDS):
    """the oracle's (statement, records, m*), computed once per (file, seeds)"""
    key = (len(data), hash(data), seeds)
    if key not in _orc:
        t0 = time.time()
        rc, st, recs, op = O.full_flow(data, *seeds)
        assert rc == 0, rc
        _orc[key] = (st, recs, op)
        say(f"  oracle d={W.decode_statement(st)['d']} {len(data)} B in {time.time() - t0:.1f} s")
    return _orc[key]


def params_of(seeds):
    return dm.SystemParams(*seeds)


# ---------------- full-size single files ----------------
def test_c2_8mib_file_matches_oracle():
    """the 8 MiB seeded file (645,278 coefficients in a 1024 x 1024 witness, d = 10): the fused
    job and the host-bytes prover both equal the oracle's proof"""
    data = rand(8 << 20, 2)
    buf = dm.Context.default().upload(data)
    say("C2 fused job")
    fused = dm.prove_verify_device(buf.ptr, len(data))
    assert fused[0][:32] == O.blake3(data)
    assert fused == oracle(data)
    say("C2 host-bytes prover")
    assert dm.prove(data) == fused


def test_d11_file_matches_oracle():
    """a 16 MiB file (d = 11, 2048 x 2048 witness, the largest served depth)"""
    data = rand(16 << 20, 13)
    buf = dm.Context.default().upload(data)
    say("d = 11 fused job")
    fused = dm.prove_verify_device(buf.ptr, len(data))
    assert fused == oracle(data)


def test_file_id_of_multi_mib_files_matches_blake3():
    """file_id = BLAKE3(file) over files whose chunk tree has a group level (more than 1024
    chunks): exactly one group, one group and one chunk, full groups, a partial last group, and
    the largest d = 10 file; through the fused job and in one ragged batch (per-job lengths)"""
    ctx = dm.Context.default()
    sizes = [1 << 20, (1 << 20) + 1, 2 << 20, (3 << 20) + 1024 + 5, (5 << 20) + 7, 13 * 4 ** 10]
    files = [rand(n, n % 997) for n in sizes]
    bufs = [ctx.upload(f) for f in files]
    for f, b in zip(files, bufs):
        say("file_id", len(f))
        st, _, _ = dm.prove_verify_device(b.ptr, len(f))
        assert st[:32] == O.blake3(f), len(f)
    many = dm.prove_verify_many([(b.ptr, len(f)) for b, f in zip(bufs, files)])
    for f, r in zip(files, many):
        assert r[0][:32] == O.blake3(f), len(f)


def test_c1_golden_fixture_on_gpu():
    """the committed C1 fixture (oracle output) reproduced by the host-bytes prover and the fused
    job, and accepted by the GPU verifier"""
    data, st, recs, op, _ = G.load()
    assert dm.prove(data) == (st, recs, op)
    buf = dm.Context.default().upload(data)
    assert dm.prove_verify_device(buf.ptr, len(data)) == (st, recs, op)
    dm.verify(st, recs, op)


# ---------------- batches and the server at depth ----------------
def _file_params(i):
    return (O.derive_subseed(bytes([7] * 32), b"file-%d" % i), bytes([11] * 32))


def test_batch_with_deep_files_and_own_seeds():
    """a C4-shaped ragged batch (d = 8, 9 and 10, each file with its own generator seed): every job
    equals its solo run, through device and host files; a d = 9 file over 1 MiB and a d = 10 file
    equal the oracle"""
    sizes = [900_000, 2_100_000, 13 * 4 ** 9, 3_500_000, 6_000_000, 200_000]  # d = 9, 9, 9, 10, 10, 8
    files = [rand(n, 100 + i) for i, n in enumerate(sizes)]
    seeds = [_file_params(i) for i in range(len(files))]
    params = [params_of(s) for s in seeds]
    ctx = dm.Context.default()
    bufs = [ctx.upload(f) for f in files]
    say("deep batch, device files", sizes)
    dev = dm.prove_verify_many([(b.ptr, len(f)) for b, f in zip(bufs, files)], params)
    say("deep batch, host files")
    host = dm.prove_verify_many(files, params)
    assert host == dev
    for b, f, p, r in zip(bufs, files, params, dev):
        assert r == dm.prove_verify_device(b.ptr, len(f), p), len(f)
    assert dev[1] == oracle(files[1], seeds[1])
    assert dev[4] == oracle(files[4], seeds[4])


def test_depth_boundary_files_in_one_batch_match_oracle():
    """the largest file of depths 7, 8 and 9 (13 * 4^d bytes: a full witness, no padding) and one
    byte more (the next depth, a last coefficient of one byte; io.rs:72-76, 100-117), each with its own
    generator seed, proved in one mixed-depth batch: every proof equals the oracle's"""
    sizes = [n for d in (7, 8, 9) for n in (13 * 4 ** d, 13 * 4 ** d + 1)]
    files = [rand(n, 300 + i) for i, n in enumerate(sizes)]
    seeds = [_file_params(10 + i) for i in range(len(files))]
    ctx = dm.Context.default()
    bufs = [ctx.upload(f) for f in files]
    say("depth-boundary batch", sizes)
    got = dm.prove_verify_many([(b.ptr, len(f)) for b, f in zip(bufs, files)], [params_of(x) for x in seeds])
    for f, x, r in zip(files, seeds, got):
        st = W.decode_statement(r[0])
        assert st["original_len_bytes"] == len(f)
        assert r == oracle(f, x), len(f)
    depths = [W.decode_statement(r[0])["d"] for r in got]
    assert depths == [7, 8, 8, 9, 9, 10]


def test_server_mix_with_deep_requests():
    """requests of d = 7, 9, 10 and 11 submitted while earlier ones run: every served proof equals
    its solo run; the d = 11 and a d = 9 request equal the oracle"""
    files = [rand(2_000_000, 201), rand(16 << 20, 13), rand(5 << 20, 203), rand(1_500_000, 204),
             rand(100_000, 205), rand(3_000_000, 206)]
    seeds = [DEFAULT_SEEDS, DEFAULT_SEEDS, _file_params(2), DEFAULT_SEEDS, _file_params(4), DEFAULT_SEEDS]
    ctx = dm.Context.default()
    bufs = [ctx.upload(f) for f in files]
    srv = dm.Server(ctx, max_batch=4, max_inflight=3)
    try:
        tickets = []
        for i, (b, f) in enumerate(zip(bufs, files)):
            say("deep server submit", i, len(f))
            tickets.append(srv.submit((b.ptr, len(f)), params_of(seeds[i])))
            time.sleep(0.001)
        res = [srv.wait(t, timeout_ms=120_000) for t in tickets]
    finally:
        srv.close()
    assert all(r is not None for r in res), "a served request did not finish"
    for b, f, s, (r, lat) in zip(bufs, files, seeds, res):
        assert r == dm.prove_verify_device(b.ptr, len(f), params_of(s)), len(f)
    assert res[1][0] == oracle(files[1])
    assert res[0][0] == oracle(files[0])


def test_graph_shapes_alternate_on_one_workspace():
    """one cached (depth, 1) workspace holds the plain path's three-stream graph and the server's
    one-stream graph; used alternately, with the resident-generator graph of a batch between them,
    each result equals the oracle"""
    data = rand(1_500_000, 301)  # d = 9, more than 1024 chunks
    want = oracle(data)
    ctx = dm.Context.default()
    buf = ctx.upload(data)
    for rep in range(2):
        say("three-stream graph", rep)
        assert dm.prove_verify_device(buf.ptr, len(data)) == want
        srv = dm.Server(ctx, max_batch=1)
        try:
            say("one-stream graph", rep)
            r, _ = srv.wait(srv.submit((buf.ptr, len(data))), timeout_ms=60_000)
        finally:
            srv.close()
        assert r == want
        say("resident-generator graph", rep)
        assert dm.prove_verify_many([(buf.ptr, len(data))]) == [want]


# ---------------- resampling of rejected challenge candidates ----------------
class _Reject:
    """DAMASCUS_TEST_REJECT on the device and the same rejection count in the oracle"""

    def __init__(self, spec, n_oracle):
        self.spec, self.n = spec, n_oracle

    def __enter__(self):
        os.environ["DAMASCUS_TEST_REJECT"] = self.spec
        O.set_challenge_reject(self.n)

    def __exit__(self, *a):
        os.environ.pop("DAMASCUS_TEST_REJECT", None)
        O.set_challenge_reject(0)


def test_rejected_candidates_resample_like_the_oracle():
    """Every challenge rejects its first candidates (counter 0, or 0..2) and takes the next: at
    d = 6 the round-0 challenges run through the chunk and root kernels (chunk level) and the later
    ones through the single-workgroup kernel. The device proofs equal the oracle's under the same
    knob, differ from the normal proof, and the GPU verifier accepts them."""
    data = G.c1_input()
    normal = oracle(data)
    buf = dm.Context.default().upload(data)
    for n in (1, 3):
        with _Reject(str(n), n):
            say("reject", n)
            rc, ost, orecs, oop = O.full_flow(data)
            assert rc == 0
            got = dm.prove(data)
            assert got == (ost, orecs, oop)
            assert got[1] != normal[1]
            assert dm.prove_verify_device(buf.ptr, len(data)) == got
            dm.verify(*got)
    assert dm.prove(data) == normal


def test_group_level_resampling_at_d12_replays_on_the_oracle_verifier():
    """At d = 12 the round-0 transcript messages exceed 1024 chunks, so a rejected candidate
    re-hashes the counter's chunk, its group value and the root. The fused job (its GPU verifier
    accepts) is replayed record by record by the oracle's verifier under the same knob, which
    recomputes every challenge itself and accepts the final opening."""
    n = 13 * 4 ** 11 + 1
    data = rand(n, 12)
    buf = dm.Context.default().upload(data)
    with _Reject("2", 2):
        say("d = 12 fused job, reject 2")
        st, recs, op = dm.prove_verify_device(buf.ptr, n)
        del buf
        say("oracle verifier replay")
        v = O.Verifier(st)
        for j, r in enumerate(W.split_records(recs)):
            assert v.update_commitment(r) == 0, j
        assert v.verify_final_opening(op) == 0


# ---------------- server robustness ----------------
def test_server_admission_waits_within_a_small_budget():
    """a device-memory budget that holds one d = 7 workspace at a time: jobs queue for memory,
    every one finishes (no deadlock) and equals its solo run; a job whose workspace alone exceeds
    the budget is refused at submission (io.rs:46-49's memory guard, on the server's budget)"""
    sizes = [60_000, 150_000, 200_000, 100_000, 70_000, 180_000, 5_000, 40_000]  # d = 7 x 6, d = 6 x 2
    files = [rand(n, 400 + i) for i, n in enumerate(sizes)]
    ctx = dm.Context.default()
    bufs = [ctx.upload(f) for f in files]
    budget = int(1.5 * dm.job_workspace_size(200_000))
    srv = dm.Server(ctx, hbm_budget=budget, max_batch=8, max_inflight=4)
    try:
        tickets = [srv.submit((b.ptr, len(f))) for b, f in zip(bufs, files)]
        res = [srv.wait(t, timeout_ms=60_000) for t in tickets]
        big = ctx.upload(rand(1_000_000, 499))  # d = 9
        with pytest.raises(dm.DamascusError) as e:
            srv.submit((big.ptr, 1_000_000))
        assert e.value.name == "ERR_MEMORY_LIMIT"
        stats = srv.stats()
    finally:
        srv.close()
    assert all(r is not None for r in res), "admission did not finish every job"
    say("small-budget server stats", stats)
    assert stats["batches"] >= 4  # at most one d = 7 job per batch fits
    for b, f, (r, lat) in zip(bufs, files, res):
        assert r == dm.prove_verify_device(b.ptr, len(f)), len(f)


def test_failing_job_is_isolated_from_its_batch():
    """one job of a batch fails on the device (its challenges reject more candidates than the
    device's bounded resampling allows, DAMASCUS_TEST_REJECT=100@len): it alone reports ERR_HIP,
    the other jobs of the same launch finish with their solo proofs, and the server goes on to
    prove later requests"""
    sizes = [60_000, 150_001, 90_000, 120_000]  # one d = 7 batch
    files = [rand(n, 500 + i) for i, n in enumerate(sizes)]
    ctx = dm.Context.default()
    bufs = [ctx.upload(f) for f in files]
    jobs = [(b.ptr, len(f)) for b, f in zip(bufs, files)]
    os.environ["DAMASCUS_TEST_REJECT"] = "100@150001"
    try:
        many = dm.prove_verify_many(jobs, raise_on_error=False)
        srv = dm.Server(ctx, max_batch=4)
        try:
            tickets = [srv.submit(j) for j in jobs]
            served = [srv.wait(t, timeout_ms=60_000) for t in tickets]
            later = srv.wait(srv.submit(jobs[0]), timeout_ms=60_000)
        finally:
            srv.close()
    finally:
        os.environ.pop("DAMASCUS_TEST_REJECT", None)
    solo = [dm.prove_verify_device(p, n) for p, n in jobs[:1] + jobs[2:]]
    assert many[1] == dm.ERR_CODE["ERR_HIP"]
    assert [many[0]] + many[2:] == solo
    assert served[1][0] == dm.ERR_CODE["ERR_HIP"]
    assert [served[0][0]] + [r for r, _ in served[2:]] == solo
    assert later[0] == solo[0]


def test_file_pointers_of_the_wrong_kind_are_refused():
    """device files must be device memory inside one allocation, host files must not be device
    memory: checked before any launch (a mismatch would otherwise fault the GPU or the host)"""
    ctx = dm.Context.default()
    data = rand(70_000, 600)
    buf = ctx.upload(data)
    host = np.frombuffer(data, dtype=np.uint8).copy()
    srv = dm.Server(ctx)
    hsrv = dm.Server(ctx, files_on_host=True)
    try:
        with pytest.raises(TypeError):
            srv.submit(data)
        with pytest.raises(TypeError):
            dm.prove_verify_many([data, (buf.ptr, len(data))])
        for call in (lambda: srv.submit((host.ctypes.data, len(data))),
                     lambda: dm.prove_verify_device(host.ctypes.data, len(data)),
                     lambda: dm.prove_verify_many([(host.ctypes.data, len(data))]),
                     lambda: hsrv.submit((buf.ptr, len(data))),
                     lambda: dm.prove_verify_many([(buf.ptr, len(data))], files_on_host=True)):
            with pytest.raises(dm.DamascusError) as e:
                call()
            assert e.value.name == "ERR_ARG"
        whole = dm.DeviceBuffer(ctx, 1 << 20)  # a file running past the end of its allocation
        with pytest.raises(dm.DamascusError) as e:
            dm.prove_verify_device(whole.ptr + (1 << 20) - 1000, 2000)
        assert e.value.name == "ERR_ARG"
        # the right kinds still work on the same servers
        r, _ = srv.wait(srv.submit((buf.ptr, len(data))), timeout_ms=60_000)
        h, _ = hsrv.wait(hsrv.submit(data), timeout_ms=60_000)
    finally:
        srv.close()
        hsrv.close()
    assert r == h == dm.prove(data)
