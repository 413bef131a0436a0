#!/usr/bin/env python3
"""Bit-exact commit+fold throughput on one MI355X, on a mixed d=7..11 served load.

Headline (`value`): a served trace of open-loop Poisson arrivals of commit+fold requests, file
sizes log-uniform 64 KiB..16 MiB (d = 7..11, up to 4,194,304 coefficients), seeded uniform
bytes, K = 8, default SystemParams, files resident in HBM before the clock starts. Each request
is the whole job: file_id, witness expansion, the initial commitment, every fold round with both
invariant commitments and every Fiat-Shamir challenge, the verifier's replay (its folds and its
own challenges) and the final opening check, then the proof bytes copied to the host. The
served path is the continuous-batching server (dmx_server_*), which keeps the generator matrix
of each (seed, depth) resident in HBM (derived once, in the untimed warm-up pass); the plain
path runs the same trace one request at a time through dmx_prove_verify_device, which expands
the generators on every call (once: the prover's first transform also fills the verifier's
matrix). `value` is MiB of request files per
second from the first arrival to the last verified result; p50/p99 are arrival-to-verified-
result latencies. Every served proof is compared with the plain path's.

Also reported on the same JSON line:
  single_file   one 8 MiB file (645,278 coefficients, d = 10), the plain path, graph replays
  batched_files 256 files log-uniform 4 KiB..4 MiB (d = 6..10), each with its own generator seed,
                one ragged batch (dmx_prove_verify_many) against a loop of the plain path
  ragged_products 4,096 negacyclic products, n uniform in 2^6..2^12, device-resident: one
                persistent launch against per-degree launch sequences
  roofline      the degree-1024 forward transform (the dominant kernel), from HIP events around
                each of its launches in prover runs on the 8 MiB file
  cpu_baseline  the oracle (C restatement of the reference CPU algorithm) on a stratified sample of
                the served trace (one file per depth, weighted by the trace's bytes per depth), or on
                the 8 MiB file with --no-mixed
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FILE_BYTES = 8 << 20
SEED = 2
PEAK_LANE_CYCLES = 256 * 64 * 2.4e9  # CUs x lanes x clock: 39.32 T lane-cycles/s
WORK_WEIGHTS = "butterfly 19, mod-p reduction 42, Z_q mul-add 204, BLAKE3 compression 860 lane-cycles"
DEFAULT_REQUESTS = 400
DEFAULT_RATE = 2000.0  # requests per second (about 5.6 GiB/s offered)


def rand_bytes(rng, n):
    import numpy as np

    return rng.integers(0, 256, size=int(n), dtype=np.uint8).tobytes()


def single_file():
    import numpy as np

    return np.random.default_rng(SEED).integers(0, 256, size=FILE_BYTES, dtype=np.uint8).tobytes()


def served_trace(requests, rate, seed=5):
    """(arrival_seconds, file_bytes) with Poisson arrivals and log-uniform 64 KiB..16 MiB sizes"""
    import numpy as np

    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=requests)
    arrivals = np.cumsum(gaps) - gaps[0]
    sizes = np.floor(65536 * 256.0 ** rng.random(requests)).astype(np.int64)
    return [(float(a), rand_bytes(rng, n)) for a, n in zip(arrivals, sizes)]


def batch_files(count=256, seed=4):
    """log-uniform 4 KiB..4 MiB, per-file generator seeds derived from the index"""
    import numpy as np

    import damascus000_amd as dm

    rng = np.random.default_rng(seed)
    sizes = np.floor(4096 * 1024.0 ** rng.random(count)).astype(np.int64)
    files = [rand_bytes(rng, n) for n in sizes]
    params = [dm.SystemParams(dm.derive_subseed(bytes([7] * 32), b"file-%d" % i), bytes([11] * 32))
              for i in range(count)]
    return files, params


def pct(xs, q):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, int(q * len(xs)))]


INFLIGHT = 4


def run_served(dm, ctx, dev, arrivals, files_on_host=False):
    srv = dm.Server(ctx, max_batch=64, max_inflight=INFLIGHT, files_on_host=files_on_host)
    n = len(dev)
    tickets = [None] * n
    t_sub = [0.0] * n
    t0 = time.perf_counter()

    def submitter():
        for i, (d, a) in enumerate(zip(dev, arrivals)):
            now = time.perf_counter() - t0
            if now < a:
                time.sleep(a - now)
            t_sub[i] = time.perf_counter() - t0
            tickets[i] = srv.submit(d)

    th = threading.Thread(target=submitter)
    th.start()
    results, lat = [None] * n, [0.0] * n
    for i in range(n):
        while tickets[i] is None:
            time.sleep(0.0002)
        r, server_lat = srv.wait(tickets[i])
        results[i] = r
        lat[i] = (t_sub[i] - arrivals[i]) + server_lat
    th.join()
    global LAST_STATS
    LAST_STATS = srv.stats()
    srv.close()
    makespan = max(a + l for a, l in zip(arrivals, lat)) - arrivals[0]
    return results, lat, makespan


LAST_STATS = None
LAST_TRACE = None  # (files, served proofs) of the last mixed_served run, for the CPU baseline


def run_plain(dm, dev, arrivals):
    n = len(dev)
    results, lat = [None] * n, [0.0] * n
    t0 = time.perf_counter()
    for i, ((p, ln), a) in enumerate(zip(dev, arrivals)):
        now = time.perf_counter() - t0
        if now < a:
            time.sleep(a - now)
        results[i] = dm.prove_verify_device(p, ln)
        lat[i] = time.perf_counter() - t0 - a
    makespan = max(a + l for a, l in zip(arrivals, lat)) - arrivals[0]
    return results, lat, makespan


def mixed_served(dm, ctx, requests, rate):
    trace = served_trace(requests, rate)
    arrivals = [a for a, _ in trace]
    bufs = [ctx.upload(f) for _, f in trace]
    dev = [(b.ptr, len(f)) for b, (_, f) in zip(bufs, trace)]
    total = sum(len(f) for _, f in trace) / (1 << 20)
    run_served(dm, ctx, dev, arrivals)  # warm: workspaces and captured graphs per (depth, batch)
    served, s_lat, s_span = run_served(dm, ctx, dev, arrivals)
    # the same trace handed over in host memory (PCIe-inclusive: pinned staging + async upload)
    host = [f for _, f in trace]
    run_served(dm, ctx, host, arrivals, files_on_host=True)
    h_res, h_lat, h_span = run_served(dm, ctx, host, arrivals, files_on_host=True)
    run_plain(dm, dev[: min(len(dev), 32)], arrivals[: min(len(dev), 32)])  # warm
    plain, p_lat, p_span = run_plain(dm, dev, arrivals)
    if h_res != served:
        raise SystemExit("host-file and device-file served proofs differ")
    if any(not isinstance(r, tuple) for r in served):
        raise SystemExit("a served request failed")
    if served != plain:
        raise SystemExit("served and plain proofs differ")
    global LAST_TRACE
    LAST_TRACE = (host, served)
    return {
        "mib": total, "served_mib_s": total / s_span, "plain_mib_s": total / p_span,
        "served_p50_ms": pct(s_lat, 0.5) * 1e3, "served_p99_ms": pct(s_lat, 0.99) * 1e3,
        "plain_p50_ms": pct(p_lat, 0.5) * 1e3, "plain_p99_ms": pct(p_lat, 0.99) * 1e3,
        "offered_mib_s": total / max(arrivals[-1] - arrivals[0], 1e-9),
        "host_files_mib_s": total / h_span, "host_files_p50_ms": pct(h_lat, 0.5) * 1e3,
        "host_files_p99_ms": pct(h_lat, 0.99) * 1e3,
    }


def batched_files(dm, ctx):
    files, params = batch_files()
    bufs = [ctx.upload(f) for f in files]
    dev = [(b.ptr, len(f)) for b, f in zip(bufs, files)]
    total = sum(len(f) for f in files) / (1 << 20)
    dm.prove_verify_many(dev, params)  # warm
    t0 = time.perf_counter()
    batched = dm.prove_verify_many(dev, params)
    tb = time.perf_counter() - t0
    plain = [dm.prove_verify_device(p, n, pr) for (p, n), pr in zip(dev[:16], params[:16])]  # warm
    t0 = time.perf_counter()
    plain = [dm.prove_verify_device(p, n, pr) for (p, n), pr in zip(dev, params)]
    tp = time.perf_counter() - t0
    if batched != plain:
        raise SystemExit("batched and plain proofs differ")
    return {"files": len(files), "mib": round(total, 2), "batched_mib_s": round(total / tb, 1),
            "plain_mib_s": round(total / tp, 1), "speedup": round(tp / tb, 2), "bit_exact_vs_plain": True}


def ragged_products(dm, ctx, count=4096, reps=5, seed=3):
    """C3: negacyclic products with n uniform in 2^6..2^12 and coefficients uniform in [0, q),
    device-resident; one persistent launch against per-degree launch sequences"""
    import numpy as np

    rng = np.random.default_rng(seed)
    logs = rng.integers(6, 13, size=count).astype(np.uint32)
    total = int(sum(1 << int(l) for l in logs))

    def uniform_zq(n):  # 112-bit rejection sampling: values >= q are redrawn
        lo = rng.integers(0, 1 << 63, size=n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, size=n, dtype=np.uint64)
        hi = rng.integers(0, 1 << 48, size=n, dtype=np.uint64)
        bad = (hi == (1 << 48) - 1) & (lo > np.uint64((1 << 64) - 76))
        while bad.any():
            lo[bad] = rng.integers(0, 1 << 63, size=int(bad.sum()), dtype=np.uint64)
            bad = (hi == (1 << 48) - 1) & (lo > np.uint64((1 << 64) - 76))
        return np.stack([lo, hi], axis=1)

    A, B = uniform_zq(total), uniform_zq(total)
    da, db = ctx.upload(A.tobytes()), ctx.upload(B.tobytes())
    outs, times = [], []
    for mode in (0, 1):
        do = dm.DeviceBuffer(ctx, A.nbytes)
        dm.ring_mul_batch_device(da.ptr, db.ptr, do.ptr, logs, mode)  # warm
        t0 = time.perf_counter()
        for _ in range(reps):
            dm.ring_mul_batch_device(da.ptr, db.ptr, do.ptr, logs, mode)
        times.append((time.perf_counter() - t0) / reps)
        outs.append(do.read())
    if outs[0] != outs[1]:
        raise SystemExit("persistent and per-degree ring products differ")
    return {"pairs": count, "coefficients": total, "persistent_ms": round(times[0] * 1e3, 3),
            "per_degree_ms": round(times[1] * 1e3, 3), "speedup": round(times[1] / times[0], 2),
            "persistent_products_per_s": round(count / times[0], 1), "bit_exact_between_modes": True}


def drop_in_latency(dm, reps=50):
    """Per-call latency of the reference-compatible entry points, host buffers in and out as the
    reference's FFI has them: dmx_ntt_batch at ring length 1024 (cyclic size 2048, 8 primes) and
    dmx_fold_batch over 4096 values"""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O

    S = 2048
    rng = np.random.default_rng(8)
    a = np.stack([rng.integers(0, p, size=S, dtype=np.uint64) for p in O.PRIMES])
    b = np.stack([rng.integers(0, p, size=S, dtype=np.uint64) for p in O.PRIMES])
    roots, iroots, inv = [], [], []
    for p in O.PRIMES:
        g = O.lib().orc_multiplicative_generator(p)
        w = pow(g, (p - 1) // S, p)
        wi, ln = pow(w, p - 2, p), 2
        while ln <= S:
            roots.append(pow(w, S // ln, p))
            iroots.append(pow(wi, S // ln, p))
            ln *= 2
        inv.append(pow(S, p - 2, p))
    dm.ntt_batch(a, b, S, O.PRIMES, roots, iroots, inv)  # plan tables built on the first call
    t0 = time.perf_counter()
    for _ in range(reps):
        rc, _ = dm.ntt_batch(a, b, S, O.PRIMES, roots, iroots, inv)
    ntt_us = (time.perf_counter() - t0) / reps * 1e6
    if rc != 0:
        raise SystemExit("ntt_batch failed")
    n = 4096
    left = [int(x) for x in rng.integers(0, 1 << 62, size=n)]
    right = [int(x) for x in rng.integers(0, 1 << 62, size=n)]
    dm.fold_batch(left, right, 12345)
    t0 = time.perf_counter()
    for _ in range(reps):
        rc, _ = dm.fold_batch(left, right, 12345)
    fold_us = (time.perf_counter() - t0) / reps * 1e6
    return {"ntt_batch_ring1024_us": round(ntt_us, 1), "fold_batch_4096_us": round(fold_us, 1),
            "note": "per call through the Python mirror (ctypes, list conversion for fold_batch included)"}


def depth_of(n):
    """the witness depth of an n-byte file (io.rs:72-120: the smallest d >= 6 with 4^d >= ceil(n/13))"""
    c, d = max(1, -(-n // 13)), 6
    while (1 << (2 * d)) < c:
        d += 1
    return d


def cpu_baseline(data, gpu_proof, threads):
    """the oracle (C restatement of the reference CPU algorithm) on the same job, all threads"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O

    O.set_threads(threads)
    t0 = time.perf_counter()
    rc, st, recs, op = O.full_flow(data)
    dt = time.perf_counter() - t0
    same = rc == 0 and (st, recs, op) == gpu_proof
    return {"value": round(FILE_BYTES / (1 << 20) / dt, 4), "unit": "MiB/s", "cores": threads, "kind": "port",
            "sample": f"the full job (prove + verify) on the 8 MiB single file, {dt:.1f} s",
            "bit_exact_with_gpu": bool(same)}


def cpu_baseline_trace(files, gpu_proofs, threads):
    """The oracle on a stratified sample of the served trace: per depth, the file nearest that
    depth's median size, proved and verified on all threads and compared with its served GPU
    proof; the per-depth rates, weighted by the trace's bytes at each depth, give the trace's rate."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O

    O.set_threads(threads)
    by_d = {}
    for i, f in enumerate(files):
        by_d.setdefault(depth_of(len(f)), []).append(i)
    total_s, cpu_s, same, parts = 0.0, 0.0, True, []
    for d, idx in sorted(by_d.items()):
        sizes = sorted(len(files[i]) for i in idx)
        med = sizes[len(sizes) // 2]
        i = min(idx, key=lambda j: abs(len(files[j]) - med))
        t0 = time.perf_counter()
        rc, st, recs, op = O.full_flow(files[i])
        dt = time.perf_counter() - t0
        same = same and rc == 0 and (st, recs, op) == gpu_proofs[i]
        d_bytes = sum(len(files[j]) for j in idx)
        total_s += dt * d_bytes / len(files[i])  # the trace's depth-d bytes at the sample's rate
        cpu_s += dt
        parts.append(f"d={d}: {len(files[i]) / 1024:.0f} KiB in {dt:.2f} s")
    total_mib = sum(len(f) for f in files) / (1 << 20)
    return {"value": round(total_mib / total_s, 4), "unit": "MiB/s", "cores": threads, "kind": "port",
            "sample": f"the served trace, one file per depth (the one nearest the depth's median size), prove + "
                      f"verify each ({'; '.join(parts)}; {cpu_s:.1f} s), rates weighted by the trace's bytes "
                      f"per depth",
            "bit_exact_with_gpu": bool(same)}


def load_valu_issue():
    path = os.path.join(ROOT, "profiles", "ntt_fwd_c2_valu.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("valu_issue_frac")


def load_traffic():
    path = os.path.join(ROOT, "profiles", "ntt_fwd_c2_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def calibrate(dm, ctx, mib):
    """three device copies of `mib` MiB: the reference launch for HBM byte counters"""
    if not mib:
        return False
    n = mib << 20
    a, b = dm.DeviceBuffer(ctx, n), dm.DeviceBuffer(ctx, n)
    for _ in range(3):
        b.copy_from(a)
    return True


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed graph replays of the 8 MiB single file")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--requests", type=int, default=DEFAULT_REQUESTS, help="served trace length")
    ap.add_argument("--rate", type=float, default=DEFAULT_RATE, help="served trace arrival rate (requests/s)")
    ap.add_argument("--inflight", type=int, default=4, help="server batch slots")
    ap.add_argument("--no-mixed", action="store_true", help="skip the served trace and the 256-file batch")
    ap.add_argument("--only", choices=["roofline", "served", "ragged"], default=None,
                    help="profiling runs: only the prover runs of the roofline, the served trace, or the "
                         "ragged products")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--calibrate-copy", type=int, default=0, help="run a device copy of N MiB and exit")
    ap.add_argument("--calibrate-after", type=int, default=0,
                    help="after the job, run a device copy of N MiB (HBM-counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    global INFLIGHT
    INFLIGHT = args.inflight
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (jobs shard across GPUs "
              "with no exchange; see the design notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import damascus000_amd as dm

    if not dm.backend_ready():
        print("no HIP device: " + dm.device_summary(), file=sys.stderr)
        sys.exit(1)
    ctx = dm.Context.default()
    if args.calibrate_copy:
        n = args.calibrate_copy << 20
        a, b = dm.DeviceBuffer(ctx, n), dm.DeviceBuffer(ctx, n)
        for _ in range(3):
            b.copy_from(a)
        print(json.dumps({"copy_bytes": n, "launches": 3}))
        return

    if args.only == "served":
        trace = served_trace(args.requests, args.rate)
        arrivals = [a for a, _ in trace]
        bufs = [ctx.upload(f) for _, f in trace]
        dev = [(b.ptr, len(f)) for b, (_, f) in zip(bufs, trace)]
        total = sum(len(f) for _, f in trace) / (1 << 20)
        run_served(dm, ctx, dev, arrivals)
        _, lat, span = run_served(dm, ctx, dev, arrivals)
        print(json.dumps({"served_mib_s": round(total / span, 1), "p50_ms": round(pct(lat, 0.5) * 1e3, 2),
                          "p99_ms": round(pct(lat, 0.99) * 1e3, 2), "requests": args.requests, "rate": args.rate,
                          "server": LAST_STATS}))
        return
    if args.only == "ragged":
        print(json.dumps(ragged_products(dm, ctx)))
        return
    data = single_file()
    if args.only == "roofline":
        ctx.kernel_timing(True, log_m=10)
        ctx.kernel_timing()
        for _ in range(3):
            dm.prove(data)
        k_ms, k_launches, k_work = ctx.kernel_timing()
        ctx.kernel_timing(False)
        calibrate(dm, ctx, args.calibrate_after)
        print(json.dumps({"forward_transform_launches": k_launches, "ms": round(k_ms, 3),
                          "avg_us": round(k_ms * 1e3 / max(k_launches, 1), 2),
                          "achieved_T_lane_cycles_s": round(k_work / (k_ms * 1e-3) / 1e12, 3)}))
        return

    # ---- single 8 MiB file, plain path (graph replays) ----
    buf = ctx.upload(data)
    proof = None
    for _ in range(args.warmup):
        proof = dm.prove_verify_device(buf.ptr, FILE_BYTES)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = dm.prove_verify_device(buf.ptr, FILE_BYTES)
    t1 = time.perf_counter()
    if proof is not None and out != proof:
        raise SystemExit("proof changed between runs")
    single_ms = (t1 - t0) * 1e3 / args.steps
    single = {"mib_s": round(FILE_BYTES / (1 << 20) / (single_ms * 1e-3), 1), "ms_per_job": round(single_ms, 3),
              "file_bytes": FILE_BYTES, "depth": 10, "steps": args.steps}
    # the largest served depth: a 16 MiB file is d = 11 (2048 x 2048 witness)
    import numpy as np

    big = np.random.default_rng(SEED + 11).integers(0, 256, size=16 << 20, dtype=np.uint8).tobytes()
    bbuf = ctx.upload(big)
    dm.prove_verify_device(bbuf.ptr, len(big))
    t0 = time.perf_counter()
    for _ in range(3):
        dm.prove_verify_device(bbuf.ptr, len(big), want_proof=False)
    d11_ms = (time.perf_counter() - t0) * 1e3 / 3
    W_PROVER_D11 = 225.25e9  # lane-cycles of the d = 11 prover at the fixed weights
    single["d11"] = {"file_bytes": len(big), "ms_per_job": round(d11_ms, 3),
                     "mib_s": round(len(big) / (1 << 20) / (d11_ms * 1e-3), 1),
                     "prover_work_frac_of_peak": round(W_PROVER_D11 / (d11_ms * 1e-3) / PEAK_LANE_CYCLES, 3),
                     "note": "prove + verifier replay timed together; the prover's fixed-weight work over the "
                             "whole job's time is a lower bound on the prover's own fraction"}
    del bbuf

    # ---- roofline of the dominant kernel, the degree-1024 forward transform (k_ntt_fwd_reg<32,32>):
    # a prover run alone, HIP events around each of its launches ----
    ctx.kernel_timing(True, log_m=10)
    ctx.kernel_timing()  # clear
    timing_runs = 3
    for _ in range(timing_runs):
        if dm.prove(data) != out:
            raise SystemExit("prover API and fused job disagree")
    k_ms, k_launches, k_work = ctx.kernel_timing()
    ctx.kernel_timing(False)
    achieved = (k_work / max(k_launches, 1)) / ((k_ms / max(k_launches, 1)) * 1e-3) / 1e12 if k_ms > 0 else 0.0

    line = {
        "metric": "bit-exact commit+fold MiB/s on mixed d=6–11 load; p50/p99 latency vs plain GPU",
        "value": None,
        "unit": "MiB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (mod-p residues), u128 (Z_q)",
        "data": "synthetic: seeded uniform random bytes",
        "config": {"workload": f"served trace: {args.requests} commit+fold requests (prove + verifier replay), "
                               f"Poisson arrivals at {args.rate:g}/s, sizes log-uniform 64 KiB-16 MiB (d=7..11), "
                               "K=8, files resident in HBM", "requests": args.requests, "rate_per_s": args.rate,
                   "module_rank": 8, "parallelism": "one GPU; jobs of one depth share launches"},
        "single_file": single,
        "transform_tuning": {f"2^{lg}": dict(zip(("variant", "variants", "tuned_us"), ctx.transform_variant(lg)))
                             for lg in range(6, 12)},
        "roofline": {"bound": "valu", "kernel": "k_ntt_fwd_reg<32,32> (forward transform, degree 1024, 8 primes, "
                                                "fused row producer)",
                     "achieved": round(achieved, 3), "peak": round(PEAK_LANE_CYCLES / 1e12, 3),
                     "unit": "T lane-cycles/s", "frac": round(achieved / (PEAK_LANE_CYCLES / 1e12), 4),
                     "traffic": load_traffic(), "launches_per_job": k_launches // timing_runs,
                     "valu_issue_frac_counters": load_valu_issue(),
                     "note": "the fixed weights price 32-bit multiplies at quarter rate; gfx950 issues them at "
                             "full rate, so frac overstates VALU use. valu_issue_frac_counters is SQ_INSTS_VALU x 64 "
                             "/ duration / 39.32T for this kernel (profiles/valu_rates_gfx950.md)",
                     "kernel_ms_per_job": round(k_ms / timing_runs, 3), "work_weights": WORK_WEIGHTS,
                     "timing": f"HIP events around each degree-1024 forward-transform launch of {timing_runs} "
                               "prover runs on the 8 MiB file, after the timed region"},
    }
    if not args.no_mixed:
        mx = mixed_served(dm, ctx, args.requests, args.rate)
        line["value"] = round(mx["served_mib_s"], 3)
        line["ms_per_step"] = round(mx["mib"] / mx["served_mib_s"] * 1e3 / args.requests, 4)
        line["served"] = {k: round(v, 2) for k, v in mx.items()}
        line["served"]["speedup_vs_plain"] = round(mx["served_mib_s"] / mx["plain_mib_s"], 2)
        line["served"]["bit_exact_vs_plain"] = True
        line["batched_files"] = batched_files(dm, ctx)
        line["ragged_products"] = ragged_products(dm, ctx)
        line["drop_in"] = drop_in_latency(dm)
    else:
        line["value"] = single["mib_s"]
        line["ms_per_step"] = single["ms_per_job"]
        line["config"]["workload"] = "single 8 MiB file, plain path (mixed load skipped)"
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        if LAST_TRACE is not None:  # the workload `value` is quoted on
            line["cpu_baseline"] = cpu_baseline_trace(LAST_TRACE[0], LAST_TRACE[1], threads)
        else:
            line["cpu_baseline"] = cpu_baseline(data, out, threads)
    if calibrate(dm, ctx, args.calibrate_after):
        line["calibration_copy_bytes"] = args.calibrate_after << 20
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
