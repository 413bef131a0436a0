#!/usr/bin/env python3
"""Bit-exact commit+fold throughput on one MI355X, on a mixed d=7..11 served load.

Headline (`value`): a served trace of open-loop Poisson arrivals of commit+fold requests, file
sizes log-uniform 64 KiB..16 MiB (d = 7..11, up to 4,194,304 coefficients), seeded uniform
bytes, K = 8, default SystemParams, files resident in HBM before the clock starts. Each request
is a whole job (file_id, witness expansion, the initial commitment, every fold round with both
invariant commitments and every Fiat-Shamir challenge, the verifier's replay and the final
opening check, then the proof bytes copied to the host). A step is a segment of 100 requests:
`--steps` segments form one continuous Poisson stream, timed in one served pass, after `--warmup`
segments of another trace of the same mix (seed 6) in an untimed pass (workspaces, captured
graphs, the resident generator matrix of the trace's seed). The served path is the
continuous-batching server (dmx_server_*); the plain path runs the same trace one request at a
time through dmx_prove_verify_device, which expands the generators on every call. `value` is MiB
of request files per second from the first arrival to the last verified result; p50/p99 are
arrival-to-verified-result latencies. Every served proof is compared with the plain path's.

Also reported on the same JSON line:
  served        the headline pass and the plain path on the same trace; on its first 400 requests:
                the same files from host memory, a sweep of arrival rates, and requests that each
                carry their own generator seed (so no resident generator matrix applies)
  single_file   one 8 MiB file (645,278 coefficients, d = 10), the plain path, graph replays, and a
                16 MiB file (d = 11)
  batched_files 256 files log-uniform 4 KiB..4 MiB (d = 6..10), each with its own generator seed,
                one ragged batch (dmx_prove_verify_many) against a loop of the plain path
  ragged_products 4,096 negacyclic products, n uniform in 2^6..2^12, device-resident: one
                persistent launch against per-degree launch sequences
  drop_in       per-call latency of the reference-compatible dmx_ntt_batch / dmx_fold_batch
  roofline      the degree-1024 forward transform (the dominant kernel), from HIP events around
                each of its launches in prover runs on the 8 MiB file; algorithmic work at the
                fixed weights and at full-rate 32-bit multiplies; counters from the committed
                rocprofv3 runs
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
SEGMENT = 100  # requests per step: one step is a 100-request segment of the open-loop served trace
DEFAULT_STEPS = 4  # 400 requests, 0.2 s of arrivals at the default rate
COMPARE_REQUESTS = 400  # the rate sweep and the own-seed runs use the first 400 requests of the trace
DEFAULT_RATE = 2000.0  # requests per second (about 5.6 GiB/s offered)
SWEEP_RATES = (600.0, 1000.0, 3000.0, 4000.0)  # besides the headline rate; 600/s is near the plain path's capacity
K = 8

# ---- algorithmic work of a prover (fused negacyclic CRT-NTT formulation, 8 primes, K = 8) ----
# Op counts summed over the 6d + 1 commitments of a depth-d proof (the initial one with N rows at
# degree N, then per round j with m = N >> j: three with m/2 rows at degree m and three with m/2
# rows at degree m/2): B butterflies, P pointwise MACs, R mod-p reductions, C CRT lifts, Z Z_q
# multiply-adds of the folds, H BLAKE3 compressions (generators and transcript).
FIXED_WEIGHTS = {"B": 19, "P": 16, "R": 42, "C": 768, "Z": 204, "H": 860}
# the 32-bit multiplies inside each fixed weight (which prices one at 4 lane-cycles)
MULS_PER_OP = {"B": 3, "P": 2, "R": 9, "C": 84, "Z": 32, "H": 0}
# 32-bit multiplies at the full issue rate (as the dominant kernel's counters show on gfx950)
FULL_RATE_WEIGHTS = {k: FIXED_WEIGHTS[k] - 3 * MULS_PER_OP[k] for k in FIXED_WEIGHTS}


def b3_compressions(nbytes):
    """BLAKE3 compressions of an nbytes message: one per 64-byte block of each 1 KiB chunk (at
    least one), plus one per parent node"""
    chunks = max(1, -(-nbytes // 1024))
    blocks = sum(max(1, -(-min(1024, nbytes - 1024 * c) // 64)) for c in range(chunks)) if nbytes else 1
    return blocks + chunks - 1


def transcript_len(label_len, lens):
    return 96 + label_len + 4 + sum(8 + K * (8 + 16 * n) for n in lens)


def prover_ops(d):
    ops = {k: 0 for k in FIXED_WEIGHTS}
    N = 1 << d

    def commit(t, m):
        lg = m.bit_length() - 1
        ops["B"] += 8 * (t * (1 + K) + K) * (m // 2) * lg
        ops["P"] += 8 * t * K * m
        ops["R"] += 8 * t * (1 + K) * m
        ops["C"] += K * m

    commit(N, N)
    ops["H"] = K * 4 ** d
    for j in range(d):
        m = N >> j
        for _ in range(3):
            commit(m // 2, m)
        for _ in range(3):
            commit(m // 2, m // 2)
        ops["Z"] += (1 + K) * (m * m // 2 + m * m // 4)
        ops["H"] += b3_compressions(transcript_len(3, [m, m, m])) + b3_compressions(transcript_len(4, [m, m // 2, m // 2]))
    return ops


def work(ops, weights):
    return sum(ops[k] * weights[k] for k in ops)


def rand_bytes(rng, n):
    import numpy as np

    return rng.integers(0, 256, size=int(n), dtype=np.uint8).tobytes()


def single_file():
    import numpy as np

    return np.random.default_rng(SEED).integers(0, 256, size=FILE_BYTES, dtype=np.uint8).tobytes()


def served_trace(requests, rate, seed=5):
    """(arrival_seconds, file_bytes) with Poisson arrivals and log-uniform 64 KiB..16 MiB sizes;
    the sizes and bytes do not depend on the rate (it only scales the gaps)"""
    import numpy as np

    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=requests)
    arrivals = np.cumsum(gaps) - gaps[0]
    sizes = np.floor(65536 * 256.0 ** rng.random(requests)).astype(np.int64)

    def file_bytes(i):  # file i from its own stream, so files can be drawn in parallel
        return rand_bytes(np.random.default_rng((seed, i)), sizes[i])

    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        files = list(ex.map(file_bytes, range(requests)))
    return [(float(a), f) for a, f in zip(arrivals, files)]


def own_seed_params(dm, n, tag=b"request"):
    """per-request SystemParams: generator seeds derived from a tag and the request index"""
    return [dm.SystemParams(dm.derive_subseed(bytes([7] * 32), tag + b"-%d" % i), bytes([11] * 32)) for i in range(n)]


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


def run_served(dm, ctx, dev, arrivals, files_on_host=False, params=None):
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
            tickets[i] = srv.submit(d, params[i] if params else None)

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
LAST_TRACE = None  # (files, served proofs) of the headline pass, for the CPU baseline


def run_plain(dm, dev, arrivals, params=None):
    n = len(dev)
    results, lat = [None] * n, [0.0] * n
    t0 = time.perf_counter()
    for i, ((p, ln), a) in enumerate(zip(dev, arrivals)):
        now = time.perf_counter() - t0
        if now < a:
            time.sleep(a - now)
        results[i] = dm.prove_verify_device(p, ln, params[i] if params else None)
        lat[i] = time.perf_counter() - t0 - a
    makespan = max(a + l for a, l in zip(arrivals, lat)) - arrivals[0]
    return results, lat, makespan


def summary(mib, lat, span, arrivals):
    return {"mib_s": round(mib / span, 1), "p50_ms": round(pct(lat, 0.5) * 1e3, 2),
            "p99_ms": round(pct(lat, 0.99) * 1e3, 2),
            "offered_mib_s": round(mib / max(arrivals[-1] - arrivals[0], 1e-9), 1)}


def mixed_served(dm, ctx, requests, rate, warmup):
    """the headline pass (timed) and the comparisons on the same trace"""
    trace = served_trace(requests, rate)
    arrivals = [a for a, _ in trace]
    bufs = [ctx.upload(f) for _, f in trace]
    dev = [(b.ptr, len(f)) for b, (_, f) in zip(bufs, trace)]
    mib = sum(len(f) for _, f in trace) / (1 << 20)
    # untimed: another trace of the same size and depth mix (workspaces, graphs, resident generators)
    warm = served_trace(max(1, warmup), rate, seed=6)
    warm_bufs = [ctx.upload(f) for _, f in warm]
    run_served(dm, ctx, [(b.ptr, len(f)) for b, (_, f) in zip(warm_bufs, warm)], [a for a, _ in warm])
    del warm_bufs, warm
    served, s_lat, s_span = run_served(dm, ctx, dev, arrivals)  # the timed region
    head = summary(mib, s_lat, s_span, arrivals)
    head.update({"requests": requests, "rate_per_s": rate, "makespan_s": round(s_span, 4),
                 "mib": round(mib, 2), "server": LAST_STATS})
    out = {"headline": head}
    # the plain path on the same trace
    run_plain(dm, dev[: min(len(dev), 32)], arrivals[: min(len(dev), 32)])
    plain, p_lat, p_span = run_plain(dm, dev, arrivals)
    out["plain"] = summary(mib, p_lat, p_span, arrivals)
    if served != plain:
        raise SystemExit("served and plain proofs differ")
    if any(not isinstance(r, tuple) for r in served):
        raise SystemExit("a served request failed")
    out["speedup_vs_plain"] = round(out["plain"]["mib_s"] and head["mib_s"] / out["plain"]["mib_s"], 2)
    # The comparisons below use the first COMPARE_REQUESTS requests of the trace, so their cost
    # does not grow with the number of timed steps.
    c = min(len(dev), COMPARE_REQUESTS)
    cdev, carr, cserved = dev[:c], arrivals[:c], served[:c]
    cmib = sum(n for _, n in cdev) / (1 << 20)
    w = min(c, 32)
    out["comparison_requests"] = c
    # the same files handed over in host memory (PCIe-inclusive: pinned staging + async upload)
    host = [f for _, f in trace[:c]]
    run_served(dm, ctx, host[:w], carr[:w], files_on_host=True)
    h_res, h_lat, h_span = run_served(dm, ctx, host, carr, files_on_host=True)
    if h_res != cserved:
        raise SystemExit("host-file and device-file served proofs differ")
    out["host_files"] = summary(cmib, h_lat, h_span, carr)
    # arrival-rate sweep over the same files (the gaps scale with the rate)
    sweep = {}
    for r in SWEEP_RATES + (rate,):
        arr = [a * rate / r for a in carr]
        res, lat, span = run_served(dm, ctx, cdev, arr)
        if res != cserved:
            raise SystemExit(f"served proofs at {r}/s differ")
        sweep[f"{r:g}"] = summary(cmib, lat, span, arr)
        if r == SWEEP_RATES[0]:  # near the plain path's capacity: the plain path at the same rate
            _, pl_lat, pl_span = run_plain(dm, cdev, arr)
            sweep[f"{r:g}"]["plain"] = summary(cmib, pl_lat, pl_span, arr)
    out["rate_sweep"] = sweep
    # every request with its own generator seed, none of them seen before the timed pass (the
    # warm-up uses other seeds): no generator matrix is reused across requests, so this separates
    # batching and overlap from the generator cache
    params = own_seed_params(dm, c)
    run_served(dm, ctx, cdev[:w], carr[:w], params=own_seed_params(dm, w, b"warm-up"))
    o_res, o_lat, o_span = run_served(dm, ctx, cdev, carr, params=params)
    run_plain(dm, cdev[:16], carr[:16], params[:16])
    op_res, op_lat, op_span = run_plain(dm, cdev, carr, params)
    if o_res != op_res:
        raise SystemExit("own-seed served and plain proofs differ")
    out["own_seeds"] = {"served": summary(cmib, o_lat, o_span, carr), "plain": summary(cmib, op_lat, op_span, carr)}
    out["own_seeds"]["speedup_vs_plain"] = round((cmib / o_span) / (cmib / op_span), 2)
    out["bit_exact_vs_plain"] = True
    global LAST_TRACE
    LAST_TRACE = ([f for _, f in trace], served)
    return out


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
    # per product of degree n: 3 transforms x 8 primes x (n/2) log2 n butterflies, 2 x 8 n mod-p
    # reductions, 8 n pointwise products, n CRT lifts (full-rate weights)
    w = FULL_RATE_WEIGHTS
    lane_cycles = sum((1 << int(l)) * (12 * int(l) * w["B"] + 16 * w["R"] + 8 * w["P"] + w["C"]) for l in logs)
    return {"pairs": count, "coefficients": total, "persistent_ms": round(times[0] * 1e3, 3),
            "per_degree_ms": round(times[1] * 1e3, 3), "speedup": round(times[1] / times[0], 2),
            "persistent_products_per_s": round(count / times[0], 1), "bit_exact_between_modes": True,
            "work_G_lane_cycles_full_rate": round(lane_cycles / 1e9, 3),
            "frac_of_peak_per_call": round(lane_cycles / times[0] / PEAK_LANE_CYCLES, 3)}


def primitive_root(p):
    """the smallest generator of the multiplicative group mod the prime p (trial division of p - 1)"""
    m, fs, f = p - 1, [], 2
    while f * f <= m:
        if m % f == 0:
            fs.append(f)
            while m % f == 0:
                m //= f
        f += 1
    if m > 1:
        fs.append(m)
    g = 2
    while any(pow(g, (p - 1) // f, p) == 1 for f in fs):
        g += 1
    return g


def drop_in_latency(dm, reps=50):
    """Per-call latency of the reference-compatible entry points, host buffers in and out as the
    reference's FFI has them: dmx_ntt_batch at ring length 1024 (cyclic size 2048, 8 primes) and
    dmx_fold_batch over 4096 values"""
    import numpy as np

    S = 2048
    rng = np.random.default_rng(8)
    a = np.stack([rng.integers(0, p, size=S, dtype=np.uint64) for p in dm.PRIMES])
    b = np.stack([rng.integers(0, p, size=S, dtype=np.uint64) for p in dm.PRIMES])
    roots, iroots, inv = [], [], []
    for p in dm.PRIMES:
        g = primitive_root(p)
        w = pow(g, (p - 1) // S, p)
        wi, ln = pow(w, p - 2, p), 2
        while ln <= S:
            roots.append(pow(w, S // ln, p))
            iroots.append(pow(wi, S // ln, p))
            ln *= 2
        inv.append(pow(S, p - 2, p))
    dm.ntt_batch(a, b, S, dm.PRIMES, roots, iroots, inv)  # plan tables built on the first call
    t0 = time.perf_counter()
    for _ in range(reps):
        rc, _ = dm.ntt_batch(a, b, S, dm.PRIMES, roots, iroots, inv)
    ntt_us = (time.perf_counter() - t0) / reps * 1e6
    if rc != 0:
        raise SystemExit("ntt_batch failed")
    n = 4096
    arrs = [np.ascontiguousarray(rng.integers(0, 1 << 48, size=n, dtype=np.uint64)) for _ in range(6)]
    ptrs = [a.ctypes.data for a in arrs]
    fold = dm.lib().dmx_fold_batch
    fold(*ptrs, n, 12345, 678)
    t0 = time.perf_counter()
    for _ in range(reps):
        rc = fold(*ptrs, n, 12345, 678)
    fold_us = (time.perf_counter() - t0) / reps * 1e6
    if rc != 0:
        raise SystemExit("fold_batch failed")
    return {"ntt_batch_ring1024_us": round(ntt_us, 1), "fold_batch_4096_us": round(fold_us, 1),
            "note": "per call, host buffers in and out: ntt_batch through the Python mirror, fold_batch on "
                    "numpy arrays through the C ABI"}


def depth_of(n):
    """the witness depth of an n-byte file (io.rs:72-120: the smallest d >= 6 with 4^d >= ceil(n/13))"""
    c, d = max(1, -(-n // 13)), 6
    while (1 << (2 * d)) < c:
        d += 1
    return d


CPU_KIND_NOTE = ("oracle: the C restatement of the reference's CPU algorithm (2n-point cyclic NTTs per "
                 "product, %-based modular arithmetic, exact per-coefficient CRT lifts); its commitment loop "
                 "runs over rows on all `cores` threads where the reference's is sequential (sis.rs:163-165)")


def host_cores():
    """(threads used, nproc, cpu_count): the baseline uses the box's CPU share, at most 16 threads
    per GPU, whatever the host reports"""
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return min(16, nproc), nproc, os.cpu_count() or 1


def cpu_baseline(data, gpu_proof, threads):
    """the oracle on the same job (the 8 MiB file), all threads"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O

    O.set_threads(threads)
    t0 = time.perf_counter()
    rc, st, recs, op = O.full_flow(data)
    dt = time.perf_counter() - t0
    same = rc == 0 and (st, recs, op) == gpu_proof
    return {"value": round(FILE_BYTES / (1 << 20) / dt, 4), "unit": "MiB/s", "cores": threads, "kind": "port",
            "sample": f"the full job (prove + verify) on the 8 MiB single file, {dt:.1f} s",
            "bit_exact_with_gpu": bool(same), "note": CPU_KIND_NOTE}


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
            "bit_exact_with_gpu": bool(same), "note": CPU_KIND_NOTE}


def source_digest():
    """SHA-256 over the kernel sources with comments and blank lines removed, to tell whether a
    committed counter profile was taken on the kernels of this build (a comment edit keeps it)"""
    import hashlib
    import re

    h = hashlib.sha256()
    for f in ("dmx_ntt.hip", "dmx_hash.hip", "dmx_arith.hpp", "dmx_kernels.hpp", "dmx_transcript.hpp", "dmx_blake3.hpp",
              "dmx_b3quad.hpp"):
        with open(os.path.join(ROOT, "damascus000_amd", "csrc", f)) as fh:
            text = re.sub(r"/\*.*?\*/", "", fh.read(), flags=re.S)
        for ln in text.splitlines():
            ln = re.sub(r"//.*", "", ln).rstrip()
            if ln:
                h.update(ln.encode() + b"\n")
    return h.hexdigest()[:16]


def load_profile(name, kernel):
    """a committed counter summary (profiles/<name>.json) and whether it was taken on these kernel
    sources and on `kernel`"""
    path = os.path.join(ROOT, "profiles", name + ".json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        prof = json.load(f)
    same_kernel = prof.get("kernel", "").replace(" ", "").startswith(kernel.replace(" ", "") + "(")
    prof["current"] = prof.get("source_digest") == source_digest() and same_kernel
    return prof


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
    ap.add_argument("--steps", type=int, default=DEFAULT_STEPS,
                    help="segments of %d requests in the timed served trace (a step is one segment; the "
                         "segments form one continuous Poisson stream)" % SEGMENT)
    ap.add_argument("--warmup", type=int, default=DEFAULT_STEPS,
                    help="segments of %d requests of another trace, served untimed before it" % SEGMENT)
    ap.add_argument("--rate", type=float, default=DEFAULT_RATE, help="served trace arrival rate (requests/s)")
    ap.add_argument("--c2-reps", type=int, default=10, help="timed graph replays of the 8 MiB single file")
    ap.add_argument("--inflight", type=int, default=4, help="server batch slots")
    ap.add_argument("--no-mixed", action="store_true", help="skip the served trace and the 256-file batch")
    ap.add_argument("--only", choices=["roofline", "served", "ragged", "prover11", "fused"], default=None,
                    help="profiling runs: the prover runs of the roofline, the served trace, the ragged products, "
                         "d = 11 prover runs, or fused prove + verify jobs on the 8 MiB file")
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
        trace = served_trace(args.steps * SEGMENT, args.rate)
        arrivals = [a for a, _ in trace]
        bufs = [ctx.upload(f) for _, f in trace]
        dev = [(b.ptr, len(f)) for b, (_, f) in zip(bufs, trace)]
        mib = sum(len(f) for _, f in trace) / (1 << 20)
        run_served(dm, ctx, dev, arrivals)
        _, lat, span = run_served(dm, ctx, dev, arrivals)
        out = summary(mib, lat, span, arrivals)
        out.update({"requests": args.steps * SEGMENT, "rate": args.rate, "server": LAST_STATS})
        print(json.dumps(out))
        return
    if args.only == "ragged":
        print(json.dumps(ragged_products(dm, ctx)))
        return
    if args.only == "prover11":  # two d = 11 prover runs (dmx_prove), for a kernel trace
        import numpy as np

        big = np.random.default_rng(SEED + 11).integers(0, 256, size=16 << 20, dtype=np.uint8).tobytes()
        for _ in range(2):
            dm.prove(big)
        print(json.dumps({"prover_runs": 2, "depth": 11,
                          "work_fixed_lane_cycles": work(prover_ops(11), FIXED_WEIGHTS),
                          "work_full_rate_lane_cycles": work(prover_ops(11), FULL_RATE_WEIGHTS)}))
        return
    data = single_file()
    if args.only == "fused":  # four prove + verify jobs of the 8 MiB file (graph replays), for a kernel trace
        buf = ctx.upload(data)
        for _ in range(4):
            dm.prove_verify_device(buf.ptr, FILE_BYTES)
        print(json.dumps({"fused_jobs": 4, "depth": 10}))
        return
    if args.only == "roofline":
        ctx.kernel_timing(True, log_m=10)
        ctx.kernel_timing()
        for _ in range(3):
            dm.prove(data)
        k_ms, k_launches, k_work, k_full = ctx.kernel_timing(full_rate=True)
        ctx.kernel_timing(False)
        calibrate(dm, ctx, args.calibrate_after)
        print(json.dumps({"forward_transform_launches": k_launches, "ms": round(k_ms, 3),
                          "avg_us": round(k_ms * 1e3 / max(k_launches, 1), 2),
                          "achieved_T_lane_cycles_s": round(k_work / (k_ms * 1e-3) / 1e12, 3),
                          "achieved_full_rate_T_lane_cycles_s": round(k_full / (k_ms * 1e-3) / 1e12, 3)}))
        return

    # ---- single 8 MiB file, plain path (graph replays) ----
    buf = ctx.upload(data)
    proof = None
    for _ in range(2):
        proof = dm.prove_verify_device(buf.ptr, FILE_BYTES)
    t0 = time.perf_counter()
    for _ in range(args.c2_reps):
        out = dm.prove_verify_device(buf.ptr, FILE_BYTES)
    t1 = time.perf_counter()
    if proof is not None and out != proof:
        raise SystemExit("proof changed between runs")
    single_ms = (t1 - t0) * 1e3 / args.c2_reps
    single = {"mib_s": round(FILE_BYTES / (1 << 20) / (single_ms * 1e-3), 1), "ms_per_job": round(single_ms, 3),
              "file_bytes": FILE_BYTES, "depth": 10, "reps": args.c2_reps}
    # the largest served depth: a 16 MiB file is d = 11 (2048 x 2048 witness)
    import numpy as np

    big = np.random.default_rng(SEED + 11).integers(0, 256, size=16 << 20, dtype=np.uint8).tobytes()
    bbuf = ctx.upload(big)
    dm.prove_verify_device(bbuf.ptr, len(big))
    t0 = time.perf_counter()
    for _ in range(3):
        dm.prove_verify_device(bbuf.ptr, len(big), want_proof=False)
    d11_ms = (time.perf_counter() - t0) * 1e3 / 3
    ops11 = prover_ops(11)
    w11_fixed, w11_full = work(ops11, FIXED_WEIGHTS), work(ops11, FULL_RATE_WEIGHTS)
    single["d11"] = {"file_bytes": len(big), "ms_per_job": round(d11_ms, 3),
                     "mib_s": round(len(big) / (1 << 20) / (d11_ms * 1e-3), 1),
                     "prover_work_G_lane_cycles": {"fixed_weights": round(w11_fixed / 1e9, 2),
                                                   "full_rate_multiplies": round(w11_full / 1e9, 2)},
                     "prover_frac_full_rate_lower_bound": round(w11_full / (d11_ms * 1e-3) / PEAK_LANE_CYCLES, 3),
                     "note": "prove + verifier replay timed together, so the prover's work over the whole job's "
                             "time is a lower bound on the prover's own fraction; the prover-only fraction from "
                             "a kernel trace is in profiles/kernel_trace_prover_16MiB_d11.md"}
    del bbuf

    # ---- roofline of the dominant kernel, the degree-1024 forward transform (the variant the context
    # chose): prover runs on the 8 MiB file, HIP events around each of its launches on its own stream ----
    dom_kernel = ctx.transform_kernel(10)
    ctx.kernel_timing(True, log_m=10)
    ctx.kernel_timing()  # clear
    timing_runs = 3
    for _ in range(timing_runs):
        if dm.prove(data) != out:
            raise SystemExit("prover API and fused job disagree")
    k_ms, k_launches, k_work, k_full = ctx.kernel_timing(full_rate=True)
    ctx.kernel_timing(False)
    per = max(k_launches, 1)
    achieved = (k_work / per) / ((k_ms / per) * 1e-3) / 1e12 if k_ms > 0 else 0.0
    achieved_full = (k_full / per) / ((k_ms / per) * 1e-3) / 1e12 if k_ms > 0 else 0.0
    peak_t = PEAK_LANE_CYCLES / 1e12
    counters = load_profile("ntt_fwd_c2_counters", dom_kernel)
    traffic = load_profile("ntt_fwd_c2_traffic", dom_kernel)
    roof = {"bound": "valu", "kernel": f"{dom_kernel} (forward transform, degree 1024, 8 primes, fused row producer)",
            "achieved": round(achieved_full, 3), "peak": round(peak_t, 3), "unit": "T lane-cycles/s",
            "frac": round(achieved_full / peak_t, 4), "frac_reweighted": round(achieved_full / peak_t, 4),
            "traffic": traffic.get("hbm_bytes_per_launch") if traffic else None,
            "weights": "full-rate 32-bit multiplies: butterfly 10, mod-p reduction 15, Z_q mul-add 108, BLAKE3 "
                       "compression 860 lane-cycles",
            "achieved_fixed_weights": round(achieved, 3), "frac_fixed_weights": round(achieved / peak_t, 4),
            "fixed_weights": "butterfly 19, mod-p reduction 42, Z_q mul-add 204, BLAKE3 compression 860 (a 32-bit "
                             "multiply priced at 4 lane-cycles; gfx950 issues it at 1, so this fraction can pass 1)",
            "launches_per_job": k_launches // timing_runs, "kernel_ms_per_job": round(k_ms / timing_runs, 3),
            "timing": f"HIP events around each degree-1024 forward-transform launch of {timing_runs} prover runs "
                      "on the 8 MiB file, after the timed region"}
    if counters:
        roof["counters"] = {k: counters.get(k) for k in ("valu_issue_frac", "quad_cycles_per_valu_instr",
                                                          "int32_share_of_valu", "lds_bank_conflicts_per_launch",
                                                          "launches", "avg_us")}
        roof["counters"]["current_build"] = counters["current"]
    if traffic:
        roof["traffic_current_build"] = traffic["current"]

    threads, nproc, ncpu = host_cores()
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
        "config": {"workload": f"served trace: {args.steps * SEGMENT} commit+fold requests (prove + verifier replay), "
                               f"Poisson arrivals at {args.rate:g}/s, sizes log-uniform 64 KiB-16 MiB (d=7..11), "
                               f"K=8, files resident in HBM; a step is one segment of {SEGMENT} requests",
                   "requests": args.steps * SEGMENT,
                   "rate_per_s": args.rate, "module_rank": 8, "parallelism": "one GPU; jobs of one depth share launches"},
        "single_file": single,
        "transform_tuning": {f"2^{lg}": dict(zip(("variant", "variants", "tuned_us"), ctx.transform_variant(lg)))
                             for lg in range(6, 12)},
        "mac_tuning": {f"2^{lg}": dict(zip(("shape", "shapes", "tuned_us"), ctx.mac_shape(lg)))
                       for lg in range(6, 11)},
        "tuning_source": ctx.tuning_source()[0],
        "roofline": roof,
        "host": {"nproc": nproc, "cpu_count": ncpu},
    }
    if not args.no_mixed:
        mx = mixed_served(dm, ctx, args.steps * SEGMENT, args.rate, args.warmup * SEGMENT)
        head = mx["headline"]
        line["value"] = head["mib_s"]
        line["ms_per_step"] = round(head["makespan_s"] * 1e3 / args.steps, 4)
        line["timed_region_s"] = head["makespan_s"]
        line["served"] = mx
        line["batched_files"] = batched_files(dm, ctx)
        line["ragged_products"] = ragged_products(dm, ctx)
        line["drop_in"] = drop_in_latency(dm)
    else:
        line["value"] = single["mib_s"]
        line["ms_per_step"] = single["ms_per_job"]
        line["steps"], line["warmup"] = args.c2_reps, 2
        line["config"]["workload"] = "single 8 MiB file, plain path (mixed load skipped); a step is one job"
    if not args.no_cpu_baseline:
        t = args.cpu_threads or threads
        if LAST_TRACE is not None:  # the workload `value` is quoted on
            line["cpu_baseline"] = cpu_baseline_trace(LAST_TRACE[0], LAST_TRACE[1], t)
        else:
            line["cpu_baseline"] = cpu_baseline(data, out, t)
    if calibrate(dm, ctx, args.calibrate_after):
        line["calibration_copy_bytes"] = args.calibrate_after << 20
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
