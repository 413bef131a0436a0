"""The register-blocked forward transforms' index algebra, checked on the CPU for every register
blocking the library launches (read from damascus000_amd/csrc/dmx_ntt.hip): the two passes (and
the three passes of the 2^8..2^11 kernel) equal the per-layer negacyclic transform on random
input, the padded LDS layouts keep each prime slice's elements on distinct words of its own slice,
and no 32-lane LDS access of any pass repeats a bank."""
import os
import random
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = 3892314113  # an NTT prime of the reference's list, 2^13 | P - 1 suffices here


def _shapes():
    src = open(os.path.join(ROOT, "damascus000_amd", "csrc", "dmx_ntt.hip")).read()
    shapes = sorted({(int(e), int(t)) for e, t in re.findall(r"launch_reg<(\d+), (\d+)>\(s1, s2", src)})
    assert shapes, "no register blockings found"
    return shapes


def _brv(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2)


def _zeta():
    g = 2
    while pow(g, (P - 1) // 2, P) == 1:
        g += 1
    psi = pow(g, (P - 1) >> 14, P)  # primitive 2^14-th root
    assert pow(psi, 1 << 13, P) == P - 1
    return [pow(psi, _brv(k, 13), P) for k in range(1 << 13)]


ZETA = _zeta()


def _butterfly(a, i0, i1, w):
    u, v = a[i0], a[i1] * w % P
    a[i0], a[i1] = (u + v) % P, (u - v) % P


def per_layer(x):
    """Cooley-Tukey, natural order in, bit-reversed out, twiddle zeta[m/(2 len) + block]"""
    s, m = list(x), len(x)
    half, ln = m // 2, m // 2
    while ln >= 1:
        lg = ln.bit_length() - 1
        for b in range(half):
            blk = b >> lg
            j = (blk << (lg + 1)) | (b & (ln - 1))
            _butterfly(s, j, j + ln, ZETA[(half >> lg) + blk])
        ln //= 2
    return s


def register_blocked(x, E, TP):
    """pass 1: thread l holds l + TP*i and runs the layers len >= TP in registers; pass 2: after the
    LDS transpose thread l holds E*l .. E*l+E-1 and runs the layers len < TP"""
    M = E * TP
    S = list(x)
    for l in range(TP):
        a = [S[l + TP * i] for i in range(E)]
        s = E // 2
        while s >= 1:
            for b in range(E // (2 * s)):
                w = ZETA[E // (2 * s) + b]
                for k in range(s):
                    _butterfly(a, 2 * s * b + k, 2 * s * b + k + s, w)
            s //= 2
        for i in range(E):
            S[l + TP * i] = a[i]
    out = [None] * M
    for l in range(TP):
        a = [S[E * l + e] for e in range(E)]
        ln = TP // 2
        while ln >= 1:
            base = M // (2 * ln) + (E * l) // (2 * ln)
            for b in range(E // (2 * ln)):
                for k in range(ln):
                    _butterfly(a, 2 * ln * b + k, 2 * ln * b + k + ln, ZETA[base + b])
            ln //= 2
        out[E * l:E * l + E] = a
    return out


def _slot(E, TP, q, x):
    """RegNtt::slot: one pad word after every E elements, prime slices TP * (E + 1) words apart"""
    return q * TP * (E + 1) + x + (x >> (E.bit_length() - 1))


@pytest.mark.parametrize("E,TP", _shapes())
def test_register_blocking_equals_per_layer_transform(E, TP):
    assert TP <= E and TP <= 32 and 8 * TP <= 256  # the kernel's shape rules
    rng = random.Random(E * 100 + TP)
    x = [rng.randrange(P) for _ in range(E * TP)]
    assert register_blocked(x, E, TP) == per_layer(x)


def test_register_slot_formula_is_the_kernels():
    src = open(os.path.join(ROOT, "damascus000_amd", "csrc", "dmx_ntt.hip")).read()
    assert "static constexpr int SL = TP * (E + 1);" in src
    assert "return q * SL + x + (x >> LOG_E);" in src


@pytest.mark.parametrize("E,TP", _shapes())
def test_lds_layout_is_injective_without_bank_repeats(E, TP):
    M, SL = E * TP, TP * (E + 1)
    for q in range(8):  # each prime slice's M elements land on distinct words of its own SL
        slots = [_slot(E, TP, q, x) for x in range(M)]
        assert len(set(slots)) == M and min(slots) >= q * SL and max(slots) < (q + 1) * SL
    for base in range(0, 8 * TP, 32):  # the 32 lanes of one LDS pass: prime q = t // TP, lane l = t % TP
        lanes = [(t // TP, t % TP) for t in range(base, min(base + 32, 8 * TP))]
        for i in range(E):
            for banks in ([_slot(E, TP, q, l + TP * i) % 32 for q, l in lanes],
                          [_slot(E, TP, q, E * l + i) % 32 for q, l in lanes]):
                assert len(banks) == len(set(banks)), (E, TP, base, i)


# ---------------- the three-pass register transform (k_ntt_fwd_reg3, degrees 2^8..2^11) ----------------
def _reg3_shapes():
    src = open(os.path.join(ROOT, "damascus000_amd", "csrc", "dmx_ntt.hip")).read()
    shapes = sorted({(int(d), int(t or 64)) for d, t in re.findall(r"launch_reg3<(\d+)(?:, (\d+))?>\(s1, s2", src)})
    assert shapes, "no three-pass launches found"
    return shapes


def _r3_slot(x, log_m, tp):
    """Reg3::slot: padded rows for 64 threads per slice at 2^10 and 2^11, swizzled inside 32-word
    rows at 2^9 and for 128 threads per slice at 2^11"""
    r = x >> 5
    if log_m == 9:
        return x ^ ((r & 7) | (((r >> 1) & 3) << 3))
    if tp == 128:
        return x ^ ((r & 15) | (((r >> 2) & 1) << 4))
    if tp == 32:
        return x ^ (r ^ (r << 2))
    return x + r


def _r3_words(log_m, tp):
    """Reg3::SL"""
    m = 1 << log_m
    return m if log_m == 9 or tp != 64 else m + m // 32


def _r3_p2_lane(G, E, l):
    """Reg3::p2_h / p2_b: the pass-2 (h, b) of lane l"""
    return {2: (l >> 1, l & 1), 4: (l % E, l // E), 8: (l >> 3, l & 7)}[G]


def _r3_passes(log_m, tp):
    """the element x that register i of lane l holds, per pass"""
    M = 1 << log_m
    E = M // tp
    G = tp // E

    def p1(l, i):
        return l + tp * i

    def p2(l, e):
        h, b = _r3_p2_lane(G, E, l)
        return b + G * e + tp * h

    def p3(l, e):
        return E * l + e

    return M, E, G, (p1, p2, p3)


def three_pass(x, log_m, tp):
    """k_ntt_fwd_reg3 through a simulated LDS slice: pass 1 layers len = TP s (s = E/2 .. 1) with
    zeta[E/(2s) + i/(2s)], pass 2 layers len = G s with zeta[m/(2Gs) + (TP/2)h/(Gs) + e/(2s)],
    pass 3 layers len = t (t = G/2 .. 1) with zeta[m/(2t) + E l/(2t) + e/(2t)]"""
    M, E, G, (p1, p2, p3) = _r3_passes(log_m, tp)
    lds = [None] * _r3_words(log_m, tp)
    for c in range(M):
        lds[_r3_slot(c, log_m, tp)] = x[c]

    def run(xof, layers):
        for l in range(tp):
            a = [lds[_r3_slot(xof(l, i), log_m, tp)] for i in range(E)]
            for s, zeta_of in layers(l):
                for b in range(E // (2 * s)):
                    for k in range(s):
                        _butterfly(a, 2 * s * b + k, 2 * s * b + k + s, ZETA[zeta_of(b)])
            for i in range(E):
                lds[_r3_slot(xof(l, i), log_m, tp)] = a[i]

    def halving(n):
        while n >= 1:
            yield n
            n //= 2

    run(p1, lambda l: [(s, lambda b, s=s: E // (2 * s) + b) for s in halving(E // 2)])
    run(p2, lambda l: [(s, lambda b, s=s: M // (2 * G * s) + (tp // 2 // (G * s)) * _r3_p2_lane(G, E, l)[0] + b)
                       for s in halving(E // 2)])
    run(p3, lambda l: [(t, lambda b, t=t: M // (2 * t) + (E // (2 * t)) * l + b) for t in halving(G // 2)])
    return [lds[_r3_slot(c, log_m, tp)] for c in range(M)]


def test_three_pass_lane_split_is_the_kernels():
    src = open(os.path.join(ROOT, "damascus000_amd", "csrc", "dmx_ntt.hip")).read()
    assert "p2_h(int l) { return G == 2 ? l >> 1 : G == 4 ? l & (E - 1) : l >> 3; }" in src
    assert "p2_b(int l) { return G == 2 ? l & 1 : G == 4 ? l >> LOG_E : l & 7; }" in src
    assert "const int base = M / (2 * G * s) + (TP / 2 / (G * s)) * h;" in src
    assert "const int base = M / (2 * t) + (E / (2 * t)) * l;" in src
    assert "static constexpr int SWIZZLE = LOG_M == 9 ? 1 : TP == 128 ? 2 : TP == 32 ? 3 : 0;" in src
    assert "if (SWIZZLE == 3) return x ^ (r ^ (r << 2));" in src
    assert "static constexpr int SL = SWIZZLE ? M : M + M / 32;" in src
    assert "if (SWIZZLE == 1) return x ^ ((r & 7) | (((r >> 1) & 3) << 3));" in src
    assert "if (SWIZZLE == 2) return x ^ ((r & 15) | (((r >> 2) & 1) << 4));" in src
    assert "return x + r;" in src


@pytest.mark.parametrize("log_m,tp", _reg3_shapes())
def test_three_pass_transform_equals_per_layer_transform(log_m, tp):
    rng = random.Random(log_m * 1000 + tp)
    x = [rng.randrange(P) for _ in range(1 << log_m)]
    assert three_pass(x, log_m, tp) == per_layer(x)


@pytest.mark.parametrize("log_m,tp", _reg3_shapes())
def test_three_pass_lds_accesses_have_no_bank_repeats(log_m, tp):
    M, E, G, passes = _r3_passes(log_m, tp)
    slots = [_r3_slot(x, log_m, tp) for x in range(M)]
    assert len(set(slots)) == M and max(slots) < _r3_words(log_m, tp)
    for xof in passes:  # each lane's registers are a permutation of the slice's elements
        assert sorted(xof(l, i) for l in range(tp) for i in range(E)) == list(range(M))
        for half in range(0, tp, 32):  # the 32 lanes of one LDS access of a wave
            for i in range(E):
                banks = [_r3_slot(xof(l, i), log_m, tp) % 32 for l in range(half, half + 32)]
                assert len(set(banks)) == 32, (log_m, tp, half, i)
    for base in range(0, M, 32):  # the producer: 32 lanes write consecutive coefficients c
        banks = [_r3_slot(c, log_m, tp) % 32 for c in range(base, base + 32)]
        assert len(set(banks)) == 32


# ---------------- the LDS-resident ragged product kernel (k_ring_mul_lds) ----------------
def _ring_shapes():
    src = open(os.path.join(ROOT, "damascus000_amd", "csrc", "dmx_ntt.hip")).read()
    shapes = sorted({(int(c), int(e), int(t)) for c, e, t in
                     re.findall(r"log_n == (\d+) \? RingCfg<(\d+), (\d+)>::RPB", src)})
    launched = {(int(e), int(t)) for e, t in re.findall(r"ring_item<(\d+), (\d+)>\(sh,", src)}
    assert launched == {(e, t) for _, e, t in shapes}, "launched blockings differ from the item table"
    assert shapes, "no ring blockings found"
    threads = int(re.search(r"constexpr int RING_LDS_THREADS = (\d+);", src).group(1))
    words = 8 * ((1 << int(re.search(r"RING_LDS_MAX_LOG = (\d+);", src).group(1))) + 64)
    return shapes, threads, words


RING_SHAPES, RING_THREADS, RING_WORDS = _ring_shapes()
IZETA = [pow(z, P - 2, P) for z in ZETA]


def _ring_slot(E, TP, q, x):
    return q * (E * TP + TP) + x + (x >> (E.bit_length() - 1))


def _ring_strided(E, TP, q, l, i):
    return q * (E * TP + TP) + l + TP * i + (i >> ((E.bit_length() - 1) - (TP.bit_length() - 1)))


def _ring_contiguous(E, TP, q, l, e):
    return q * (E * TP + TP) + (E + 1) * l + e


def _gs(a, i0, i1, w):
    u, v = a[i0], a[i1]
    a[i0], a[i1] = (u + v) % P, (u - v) * w % P


def ring_forward(x, E, TP, lds):
    """ring_fwd through a simulated LDS (prime slice 0): registers hold l + TP*i, then E*l + e"""
    M = E * TP
    regs = [[x[l + TP * i] for i in range(E)] for l in range(TP)]
    for l in range(TP):
        a = regs[l]
        s = E // 2
        while s >= 1:
            for b in range(E // (2 * s)):
                for k in range(s):
                    _butterfly(a, 2 * s * b + k, 2 * s * b + k + s, ZETA[E // (2 * s) + b])
            s //= 2
        for i in range(E):
            lds[_ring_strided(E, TP, 0, l, i)] = a[i]
    for l in range(TP):
        regs[l] = [lds[_ring_contiguous(E, TP, 0, l, e)] for e in range(E)]
    for l in range(TP):
        a, ln = regs[l], TP // 2
        while ln >= 1:
            base = M // (2 * ln) + (E * l) // (2 * ln)
            for b in range(E // (2 * ln)):
                for k in range(ln):
                    _butterfly(a, 2 * ln * b + k, 2 * ln * b + k + ln, ZETA[base + b])
            ln //= 2
    return regs


def ring_inverse(regs, E, TP, lds):
    """ring_inv: Gentleman-Sande back through the same two layouts (no 1/m)"""
    M = E * TP
    for l in range(TP):
        a, ln = regs[l], 1
        while ln < TP:
            base = M // (2 * ln) + (E * l) // (2 * ln)
            for b in range(E // (2 * ln)):
                for k in range(ln):
                    _gs(a, 2 * ln * b + k, 2 * ln * b + k + ln, IZETA[base + b])
            ln *= 2
        for e in range(E):
            lds[_ring_contiguous(E, TP, 0, l, e)] = a[e]
    out = [None] * M
    for l in range(TP):
        a = [lds[_ring_strided(E, TP, 0, l, i)] for i in range(E)]
        s = 1
        while s < E:
            for b in range(E // (2 * s)):
                for k in range(s):
                    _gs(a, 2 * s * b + k, 2 * s * b + k + s, IZETA[E // (2 * s) + b])
            s *= 2
        for i in range(E):
            out[l + TP * i] = a[i]
    return out


@pytest.mark.parametrize("log_n,E,TP", RING_SHAPES)
def test_ring_blocking_shape_rules(log_n, E, TP):
    assert E * TP == 1 << log_n and TP <= E and TP <= 64
    tpr = 8 * TP
    assert tpr >= 64 and RING_THREADS % tpr == 0  # a slice never leaves its wave; items fill the workgroup
    assert (RING_THREADS // tpr) * 8 * (E * TP + TP) <= RING_WORDS


@pytest.mark.parametrize("log_n,E,TP", RING_SHAPES)
def test_ring_product_algebra_equals_schoolbook(log_n, E, TP):
    """forward both operands, pointwise product times 1/m, inverse: the negacyclic product mod p"""
    M = E * TP
    rng = random.Random(log_n)
    x = [rng.randrange(P) for _ in range(M)]
    y = [rng.randrange(P) for _ in range(M)]
    lds = [None] * (E * TP + TP)
    fx = ring_forward(x, E, TP, lds)
    fy = ring_forward(y, E, TP, lds)
    assert [v for r in fx for v in r] == per_layer(x)  # contiguous layout = bit-reversed order
    inv_m = pow(M, P - 2, P)
    prod = [[a * b * inv_m % P for a, b in zip(ra, rb)] for ra, rb in zip(fx, fy)]
    got = ring_inverse(prod, E, TP, lds)
    if M <= 256:
        want = [0] * M
        for i in range(M):
            for j in range(M):
                k, v = i + j, x[i] * y[j]
                if k >= M:
                    want[k - M] -= v
                else:
                    want[k] += v
        assert got == [w % P for w in want]
    else:  # the inverse undoes the forward transform exactly
        assert ring_inverse(ring_forward(x, E, TP, lds), E, TP, lds) == [v * M % P for v in x]


@pytest.mark.parametrize("log_n,E,TP", RING_SHAPES)
def test_ring_lds_layout_is_a_bijection_without_bank_repeats(log_n, E, TP):
    M, tpr = E * TP, 8 * TP
    for q in range(8):
        slots = [_ring_slot(E, TP, q, x) for x in range(M)]
        assert len(set(slots)) == M and min(slots) >= q * (M + TP) and max(slots) < (q + 1) * (M + TP)
        for l in range(TP):
            for i in range(E):
                assert _ring_strided(E, TP, q, l, i) == _ring_slot(E, TP, q, l + TP * i)
                assert _ring_contiguous(E, TP, q, l, i) == _ring_slot(E, TP, q, E * l + i)
    for base in range(0, tpr, 32):  # 32 lanes of one LDS pass: prime t // TP, lane t % TP
        lanes = [(t // TP, t % TP) for t in range(base, base + 32)]
        for i in range(E):
            for banks in ([_ring_strided(E, TP, q, l, i) % 32 for q, l in lanes],
                          [_ring_contiguous(E, TP, q, l, i) % 32 for q, l in lanes]):
                assert len(banks) == len(set(banks)), (E, TP, base, i)


def test_kernel_names_per_variant_follow_the_launch_switch():
    """Context.transform_kernel's table names the kernel each (degree, variant) case launches"""
    import sys
    sys.path.insert(0, ROOT)
    from damascus000_amd import FWD_KERNELS
    src = open(os.path.join(ROOT, "damascus000_amd", "csrc", "dmx_ntt.hip")).read()
    launched = {}
    for case, e, t, d, tp in re.findall(r"case (\d+): return launch_reg(?:<(\d+), (\d+)>|3<(\d+)(?:, (\d+))?>)\(s1, s2",
                                        src):
        lg, v = divmod(int(case), 2)
        launched[(lg, v)] = f"k_ntt_fwd_reg<{e}, {t}>" if e else f"k_ntt_fwd_reg3<{d}, {tp or 64}>"
    assert launched == {(lg, v): FWD_KERNELS[lg][v] for lg in FWD_KERNELS for v in (0, 1)}
