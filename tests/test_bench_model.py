"""The algorithmic work model bench.py prices its roofline with: op counts per depth from the
fused-formulation formulas, checked against a per-depth table computed independently of bench.py
(butterflies, pointwise MACs, reductions, Z_q multiply-adds, BLAKE3 compressions, and the total at
the fixed weights), and BLAKE3 compression counts against the closed form for whole-chunk
messages."""
import pytest

import bench

APPENDIX_B = {  # d: (B, P, R, Z, H, W_prover) in G (Z, H in M)
    6: (0.003, 0.001, 0.001, 0.04, 0.034, 0.17),
    7: (0.016, 0.004, 0.005, 0.15, 0.134, 0.71),
    8: (0.071, 0.017, 0.019, 0.59, 0.530, 3.01),
    9: (0.322, 0.067, 0.075, 2.36, 2.108, 12.69),
    10: (1.437, 0.268, 0.302, 9.44, 8.410, 53.51),
    11: (6.348, 1.074, 1.208, 37.75, 33.598, 225.25),
}


@pytest.mark.parametrize("d", sorted(APPENDIX_B))
def test_prover_work_matches_the_independent_table(d):
    B, P, R, Z, H, W = APPENDIX_B[d]
    ops = bench.prover_ops(d)
    assert round(ops["B"] / 1e9, 3) == B
    assert round(ops["P"] / 1e9, 3) == P
    assert round(ops["R"] / 1e9, 3) == R
    assert round(ops["Z"] / 1e6, 2) == Z
    assert round(ops["H"] / 1e6, 3) == H
    assert round(bench.work(ops, bench.FIXED_WEIGHTS) / 1e9, 2) == W


def test_full_rate_weights_remove_three_lane_cycles_per_multiply():
    assert bench.FULL_RATE_WEIGHTS == {"B": 10, "P": 10, "R": 15, "C": 516, "Z": 108, "H": 860}
    ops = bench.prover_ops(11)
    assert bench.work(ops, bench.FULL_RATE_WEIGHTS) < bench.work(ops, bench.FIXED_WEIGHTS)


def test_blake3_compression_counts():
    # whole-chunk messages: one compression per 64-byte block plus one per parent node
    for n in (1024, 8 << 20, 16 << 20):
        assert bench.b3_compressions(n) == n // 64 + n // 1024 - 1
    assert bench.b3_compressions(0) == 1 and bench.b3_compressions(1) == 1 and bench.b3_compressions(1025) == 16 + 1 + 1


def test_primitive_root_is_the_smallest_generator():
    # the drop-in latency leg builds its root tables from these: brute-force orders on small primes,
    # the oracle's generator search on the eight CRT primes
    def order(g, p):
        k, x = 1, g % p
        while x != 1:
            x, k = x * g % p, k + 1
        return k

    for p in (17, 97, 7681, 12289):
        g = bench.primitive_root(p)
        assert order(g, p) == p - 1
        assert all(order(h, p) < p - 1 for h in range(2, g))
    import _oracle as O

    for p in O.PRIMES:
        assert bench.primitive_root(p) == O.lib().orc_multiplicative_generator(p)


def test_served_trace_is_a_seeded_open_loop_stream():
    """The served workload: a step is a 100-request segment of one Poisson stream. The trace is a
    function of (requests, rate, seed) alone, arrivals start at 0 and never decrease, sizes stay
    in 64 KiB..16 MiB (d = 7..11), and the rate only scales the gaps (the same files at any
    rate, which the rate sweep relies on)."""
    n = 3 * bench.SEGMENT
    a = bench.served_trace(n, 2000.0)
    b = bench.served_trace(n, 2000.0)
    c = bench.served_trace(n, 500.0)
    assert len(a) == n and a == b
    arr = [t for t, _ in a]
    assert arr[0] == 0.0 and all(x <= y for x, y in zip(arr, arr[1:]))
    assert all(65536 <= len(f) < 16 << 20 for _, f in a)
    assert [f for _, f in c] == [f for _, f in a]
    assert all(abs(tc - 4 * ta) < 1e-9 for (tc, _), (ta, _) in zip(c, a))
    # mean gap near 1/rate over 300 requests (exponential gaps: within 25 %)
    assert abs(arr[-1] / (n - 1) - 1 / 2000.0) < 0.25 / 2000.0
    # the warm-up trace (another seed) is another set of files of the same mix
    w = bench.served_trace(bench.SEGMENT, 2000.0, seed=6)
    assert [f for _, f in w] != [f for _, f in a[: bench.SEGMENT]]
