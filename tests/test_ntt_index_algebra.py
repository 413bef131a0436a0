"""The register-blocked forward transform's index algebra, checked on the CPU for every register
blocking the library launches (read from damascus000_amd/csrc/dmx_ntt.hip): the two passes equal
the per-layer negacyclic transform on random input, the LDS swizzle is a bijection per prime slice,
and no 32-lane LDS access of either pass repeats a bank."""
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
    log_e, swm = E.bit_length() - 1, min(E, 32) - 1
    return q * E * TP + (x ^ ((x >> log_e) & swm) ^ ((q * TP) & 31))


@pytest.mark.parametrize("E,TP", _shapes())
def test_register_blocking_equals_per_layer_transform(E, TP):
    assert TP <= E and TP <= 32 and 8 * TP <= 256  # the kernel's shape rules
    rng = random.Random(E * 100 + TP)
    x = [rng.randrange(P) for _ in range(E * TP)]
    assert register_blocked(x, E, TP) == per_layer(x)


@pytest.mark.parametrize("E,TP", _shapes())
def test_lds_swizzle_is_a_bijection_without_bank_repeats(E, TP):
    M = E * TP
    for q in range(8):
        assert sorted(_slot(E, TP, q, x) for x in range(M)) == list(range(q * M, (q + 1) * M))
    for base in range(0, 8 * TP, 32):  # the 32 lanes of one LDS pass: prime q = t // TP, lane l = t % TP
        lanes = [(t // TP, t % TP) for t in range(base, min(base + 32, 8 * TP))]
        for i in range(E):
            for banks in ([_slot(E, TP, q, l + TP * i) % 32 for q, l in lanes],
                          [_slot(E, TP, q, E * l + i) % 32 for q, l in lanes]):
                assert len(banks) == len(set(banks)), (E, TP, base, i)
