"""CPU checks of libgrns_hip.so (no GPU needed):

* the library loads and exports every function the public headers declare;
* host-only contexts (device = -1) hold the same constants as the CPU restatement;
* the hot-path routines, which are __host__ __device__, run on the CPU through the debug hooks and
  agree bit for bit with the CPU restatement;
* the fused bn_err (one fma) returns the reference's unfused RN(RN(PHI1*|c|) + ETA) across all
  exponent ranges the kernels can see, and the glibc-equivalent ceil(log2(.)) agrees with libm.
"""
import math
import os
import random
import re
import struct

import numpy as np
import pytest

import grns_amd as G
from oracle.oracle import Oracle, lib as olib
from tests.helpers import SETS, eval_inputs, to_res

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PHI1 = 2.0 ** -53 * (1 + 2 * 2.0 ** -53)
ETA = 2.0 ** -1074


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(grns_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    L = G.lib()
    names = _declared("grns.h") + _declared("grns_debug.h")
    assert len(names) >= 20
    missing = [nm for nm in names if not hasattr(L, nm)]
    assert not missing, missing
    assert G.version().startswith("grns_amd")


def test_bad_arguments_rejected():
    L = G.lib()
    import ctypes as C
    h = C.c_void_p()
    arr = np.array([3, 5], np.int32)
    assert L.grns_ctx_create(arr.ctypes.data, 0, -1, C.byref(h)) == 1
    arr = np.array([3, 1], np.int32)
    assert L.grns_ctx_create(arr.ctypes.data, 2, -1, C.byref(h)) == 1
    c = G.Context(SETS[8], device=-1)
    # device entry points refuse a host-only context instead of touching the GPU
    assert L.grns_eval(c.handle, None, 1, 1, 0, None, 0, None) in (0, 1)
    assert L.grns_eval(c.handle, C.c_void_p(8), 1, 1, 1, C.c_void_p(8), 0, None) == 1


@pytest.mark.parametrize("n", [4, 8, 32, 64])
def test_host_context_constants_match_oracle(n):
    c = G.Context(SETS[n], device=-1)
    o = Oracle(SETS[n])
    oc, cc = o.constants(), c.constants()
    for k in cc:
        assert oc[k] == cc[k], k
    p2, mi = c.tables()
    assert (p2 == o.pow2_table()).all() and (mi == o.mrc_inv()).all()


@pytest.mark.parametrize("n", [4, 8, 32, 64])
def test_host_hot_path_matches_oracle(n):
    mod = SETS[n]
    c = G.Context(mod, device=-1)
    o = Oracle(mod)
    vals = eval_inputs(mod, 2500, seed=100 + n)
    hot = 0
    for fast in (False, True):
        ref = o.eval_array(to_res(vals, mod), fast=fast)
        for k, v in enumerate(vals):
            ok, iv = c.debug_eval_hot([v % m for m in mod], fast)
            if ok == 1:
                hot += 1
                assert iv.tobytes() == ref[k].tobytes(), v
    # most numbers take the streaming path (in {3,5,7,11} many have a zero residue, which does not)
    assert hot > (len(vals) if n > 4 else len(vals) // 2)


@pytest.mark.parametrize("n", [8, 32, 64])
def test_filter_bounds(n):
    """The comparison/arg-max filter: |F - x/M mod 1| within the stated bound, and `regular`
    only when the CPU-path evaluation takes its common path with canonical reductions."""
    from fractions import Fraction
    mod = SETS[n]
    M = G.product(mod)
    c = G.Context(mod, device=-1)
    o = Oracle(mod)
    err = G.lib().grns_debug_filter_err(c.handle)
    vals = eval_inputs(mod, 1500, seed=200 + n)
    for v in vals:
        r = [v % m for m in mod]
        Fv, reg, can = c.debug_filter(r)
        ex = Fraction(v, M)
        d = abs(Fraction(Fv) - ex)
        d = min(d, abs(d - 1))
        assert d <= Fraction(err), v
        if reg:
            ok, _ = c.debug_eval_hot(r, False)
            assert ok == 1
            iv = o.eval1(r)
            lo = Fraction(float(iv["low"]["frac"])) * Fraction(2) ** int(iv["low"]["exp"])
            up = Fraction(float(iv["upp"]["frac"])) * Fraction(2) ** int(iv["upp"]["exp"])
            assert lo <= ex <= up


@pytest.mark.parametrize("n", [4, 8, 32, 64])
def test_max_key_bound(n):
    """Arg-max keys: |F - x/M mod 1| < 2^-31 (inside err) for canonical residues and for any int32
    residues, which stand for their value mod m_i; checked with exact rationals."""
    from fractions import Fraction
    mod = SETS[n]
    M = G.product(mod)
    c = G.Context(mod, device=-1)
    err = Fraction(G.lib().grns_debug_filter_err(c.handle))
    rnd = random.Random(400 + n)
    vals = eval_inputs(mod, 600, seed=500 + n) + [0, 1, M - 1, M // 2, (M + 1) // 2]
    worst = Fraction(0)
    for k, v in enumerate(vals):
        r = [v % m for m in mod]
        if k % 3 == 1:  # non-canonical spellings of the same value: r_i + t m_i within int32
            r = [ri + rnd.choice([-1, 0, 1]) * m if -2 ** 31 <= ri + m * 1 < 2 ** 31 and ri - m >= -2 ** 31 else ri
                 for ri, m in zip(r, mod)]
        F = c.debug_max_key(r)
        assert 0.0 <= F <= 1.0
        d = abs(Fraction(F) - Fraction(v % M, M))
        d = min(d, 1 - d)
        worst = max(worst, d)
    assert worst < Fraction(1, 2 ** 31) < err, float(worst)


def _rn_unfused(c):
    return PHI1 * abs(c) + ETA  # Python floats: IEEE binary64, round to nearest, no contraction


def test_fused_bound_equals_reference():
    L = G.lib()
    rnd = random.Random(7)
    cases = [0.0, -0.0, ETA, 2 * ETA, 3 * ETA, 1.5, 3.0, 2.0 ** -52, 1.0 - 2.0 ** -53]
    cases += [1.5 * 2.0 ** k for k in range(-900, 1000, 7)]
    cases += [k * ETA for k in range(1, 2000, 13)]
    for _ in range(40000):
        e = rnd.randrange(-917, 1020)
        cases.append(rnd.choice([-1, 1]) * rnd.uniform(1, 2) * 2.0 ** e)
    for _ in range(4000):  # significands right at rounding midpoints of PHI1*c
        m = rnd.randrange(1 << 52, 1 << 53)
        cases.append(m * 2.0 ** rnd.randrange(-960, 900))
    for c in cases:
        assert L.grns_debug_bn_err(c) == _rn_unfused(c), c


def test_directed_bound_is_neighbour_double():
    """For positive c >= 2^-960 the bounded forms c - bn_err(c) and c + bn_err(c) of the reference
    are the doubles next to c (what the evaluation's streaming path uses, on values >= 2^-32),
    including at binade ends; just above 2^-1022 the ETA term makes them differ."""
    L = G.lib()
    rnd = random.Random(9)
    cases = [2.0 ** k for k in range(-960, 1000, 3)] + [(2 - 2.0 ** -52) * 2.0 ** k for k in range(-960, 1000, 3)]
    cases += [rnd.uniform(1, 2) * 2.0 ** rnd.randrange(-960, 1020) for _ in range(40000)]
    cases += [m * 2.0 ** -52 * 2.0 ** rnd.randrange(-60, 8) for m in range(1 << 52, (1 << 52) + 300)]
    cases += [((1 << 53) - m) * 2.0 ** -52 * 2.0 ** rnd.randrange(-60, 8) for m in range(1, 300)]
    for c in cases:
        e = L.grns_debug_bn_err(c)
        assert c - e == math.nextafter(c, -math.inf), c
        assert c + e == math.nextafter(c, math.inf), c


def test_ceil_log2_matches_libm():
    L = G.lib()
    rnd = random.Random(8)
    cases = []
    for e in range(-1074, 60):
        p = 2.0 ** e
        cases.append(p)
        for k in range(1, 40):
            cases.append(p * (1 + k * 2.0 ** -52) if e > -1022 else p + k * ETA)
        cases.append(p * (1 - 2.0 ** -53))
        cases.append(p * rnd.uniform(1, 2))
    for s in cases:
        if s <= 0 or math.isinf(s):
            continue
        assert L.grns_debug_ceil_log2(s) == math.ceil(math.log2(s)), s


def test_scalbn_and_reductions_match_oracle():
    L = G.lib()
    O = olib()
    rnd = random.Random(9)
    for _ in range(20000):
        x = rnd.choice([0.0, 1.0, -1.5, rnd.uniform(1, 2), -rnd.uniform(1, 2)])
        n = rnd.choice([rnd.randrange(-2100, 1100), -1024, -1023, -1025, -2047, -2048, 1023])
        a, b = L.grns_debug_fast_scalbn(x, n), O.orc_fast_scalbn(x, n)
        assert struct.pack("<d", a) == struct.pack("<d", b), (x, n)
    mods = SETS[32][:6] + [3, 5, 7, 11, 2147483647]
    for m in mods:
        d = 1.0 / m
        for _ in range(4000):
            x = rnd.randrange(-m, m + 2) if rnd.random() < 0.3 else rnd.randrange(m)
            y = rnd.randrange(m)
            want = O.orc_mod_mul(x, y, m, d)
            assert L.grns_debug_mulmod_recip(x, y, m, d) == want
            assert L.grns_debug_mulmod_recip_d(float(x), float(y), float(m), d) == float(want)
            assert L.grns_debug_addmod_recip(x, y, m, d) == O.orc_mod_add(x, y, m, d)


@pytest.mark.parametrize("n", [8, 32, 64])
def test_cmp_lane_decisions_exact(n):
    """The in-lane comparison (filter, equality, near-tie certificate) is the exact sign whenever it
    decides, agrees with the CPU restatement, and settles the +-1 near-ties itself."""
    mod = SETS[n]
    M = G.product(mod)
    c = G.Context(mod, device=-1)
    o = Oracle(mod)
    rnd = random.Random(300 + n)
    decided_near = 0
    for k in range(3000):
        x = rnd.randrange(M)
        kind = k % 4
        y = x + (1, -1, 0, None)[kind] if kind < 3 else rnd.randrange(M)
        if not 0 <= y < M:
            y = x
        if k % 97 == 0:
            y = (x + rnd.randrange(1 << 70)) % M
        xr, yr = [x % m for m in mod], [y % m for m in mod]
        r = c.debug_cmp_lane(xr, yr)
        if r == 2:
            continue
        assert r == (x > y) - (x < y), (x, y)
        assert r == o.cmp1(xr, yr)
        decided_near += kind < 2
    assert decided_near > 1400
    # outside the regular case: identical residues and zero against a canonical nonzero number are
    # settled in-lane (their rns_cmp result does not depend on the evaluations), the rest is not
    edge = [0, 1, 2, M - 1, M - 2, 1 << 40, M - (1 << 40)]
    for a in edge:
        for b in edge:
            xr, yr = [a % m for m in mod], [b % m for m in mod]
            r = c.debug_cmp_lane(xr, yr)
            if a == b or 0 in (a, b):
                assert r == (a > b) - (a < b) == o.cmp1(xr, yr), (a, b)
            elif r != 2:
                assert r == o.cmp1(xr, yr), (a, b)
    # a non-canonical zero is not the zero of the reference's zero test
    nz = list(mod)
    assert c.debug_cmp_lane([0] * n, nz) == 2


def _rare_values(M, rnd):
    L = M.bit_length()
    vals = [0, 1, 2, 3, M - 1, M - 2, M - 3, M // 2, M // 2 + 1, M // 3]
    vals += [1 << j for j in range(0, L - 1, 13)] + [M - (1 << j) for j in range(0, L - 1, 17)]
    vals += [rnd.randrange(1 << rnd.randrange(1, 70)) for _ in range(40)]
    vals += [M - 1 - rnd.randrange(1 << rnd.randrange(1, 70)) for _ in range(40)]
    vals += [rnd.randrange(M >> 30) for _ in range(10)] + [M - 1 - rnd.randrange(M >> 30) for _ in range(10)]
    vals += [rnd.randrange(M) for _ in range(20)]
    return vals


@pytest.mark.parametrize("n", [8, 32, 64])
def test_rare_lane_routines_match_oracle(n):
    """The rare passes' in-lane evaluation (ambiguity settled by position certificates, refinement
    loop in-lane) and comparison agree bit for bit with the CPU restatement whenever they decide."""
    mod = SETS[n]
    M = G.product(mod)
    c = G.Context(mod, device=-1)
    o = Oracle(mod)
    rnd = random.Random(700 + n)
    vals = _rare_values(M, rnd)
    for v in vals:
        r = [v % m for m in mod]
        # positions neither the key nor the 2^62 certificate can place: at least 2^62 from 0 and M
        # but within 2^-28 M of one of them
        zone = (1 << 62) <= v <= (M >> 28) or M - (M >> 28) <= v <= M - (1 << 62)
        for fast in (False, True):
            ok, iv, st = c.debug_eval_full(r, fast)
            assert ok == 1 or (zone and not fast), v
            if ok == 1:
                # status: the refinement left the reference's table (eval_fast of a number near M),
                # reported as a NaN interval exactly as the CPU restatement reports it
                assert st == 0 or math.isnan(float(iv["low"]["frac"]))
                assert iv.tobytes() == o.eval1(r, fast=fast).tobytes(), (v, fast)
    def certified(v):  # within 2^62 of 0 or of M: the position certificate applies
        return v < (1 << 62) or v > M - (1 << 62)

    dc = 0
    for _ in range(1500):
        a = rnd.choice(vals)
        b = rnd.choice([rnd.choice(vals), a + rnd.choice([-2, -1, 1, 2]) if 2 < a < M - 3 else a])
        xr, yr = [a % m for m in mod], [b % m for m in mod]
        s = c.debug_cmp_full(xr, yr)
        assert s != 2 or not (certified(a) and certified(b)), (a, b)
        if s != 2:
            dc += 1
            assert s == o.cmp1(xr, yr), (a, b)
    assert dc > 250


def test_word_division_estimate_is_exact():
    """The quotient estimate of the word division (three words by a normalised two-word divisor,
    with the two-word reciprocal floor((2^96 - 1) / d) - 2^32) returns divmod(u, d) for every
    <u2, u1> < <d1, d0>: the extremes of d1, d0, the top of u and u0 and random values, checked with
    Python's integers; the reciprocal is checked against its definition."""
    import ctypes as C
    L = G.lib()
    rnd = random.Random(32)
    B = 1 << 32
    d1s = [1 << 31, (1 << 31) + 1, B - 1, B - 2, 3 << 30] + [rnd.randrange(1 << 31, B) for _ in range(120)]
    d0s = [0, 1, B - 1, B - 2, 1 << 31] + [rnd.randrange(B) for _ in range(3)]
    r1, r0, v = C.c_uint(), C.c_uint(), C.c_uint()
    checked = 0
    for d1 in d1s:
        for d0 in d0s:
            d = d1 * B + d0
            tops = [0, 1, d - 1, d - 2, d // 2, d1 * B, (d1 - 1) * B + B - 1] + [rnd.randrange(d) for _ in range(12)]
            tops = [t for t in tops if t < d]
            u0s = [0, 1, B - 1, d0, (d0 + 1) % B] + [rnd.randrange(B) for _ in range(3)]
            for top in tops:
                for u0 in u0s:
                    q = L.grns_debug_div32(top >> 32, top & (B - 1), u0, d1, d0, C.byref(r1), C.byref(r0), C.byref(v))
                    assert (q, r1.value * B + r0.value) == divmod(top * B + u0, d), (top, u0, d1, d0)
                    checked += 1
            assert v.value == (B ** 3 - 1) // d - B, (d1, d0)
    assert checked > 100000
