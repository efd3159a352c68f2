"""Pins the CPU restatement (oracle/) against what mathematics and the reference fix, so that it
can serve as the checker of the HIP library:

* constants of rns_const_init (src/rnsbase.cuh:324-451) against closed forms computed with exact
  Python integers and rationals, and the default params.h set (src/params.h:16,32: 8 moduli,
  RNS_MODULI_PRODUCT_LOG2 = 242);
* BoundNear1 operations and the pairwise sums as rigorous enclosures (exact rationals);
* extended-range arithmetic against exact rational values;
* reciprocal modular reduction: congruence always, canonical away from the rounding corner, and
  the analytic "returns m instead of 0" case;
* mixed-radix digits reconstructing x exactly;
* interval evaluation enclosing x/M for every x of the {3,5,7,11} set (src/params/moduli-10-69-bit/
  params.4m.10b.h) and for random / edge / 2^j inputs (the sweep of tests/test_perf_rnseval.cu:655-687);
* comparison, division, power-of-two scaling and arg-max against exact integer answers, division
  exhaustively over the small set;
* mpint add/sub/mul against the known answers of tests/test_verify_mpint.cu:92-93 (x = -55,
  y = 13) and random signed integers.
"""
import math
import random
from fractions import Fraction as F

import numpy as np
import pytest

from grns_amd.moduli import SMALL_SET, largest_primes_below, product
from oracle.oracle import ER_DTYPE, ErFloat, Oracle, lib
from tests.helpers import SETS, cmp_pairs, div_inputs, eval_inputs, from_res, interval_value, to_res

DEFAULT_PARAMS_SET = [1283742825, 1283742827, 1283742829, 1283742833, 1283742839, 1283742841, 1283742847, 1283742851]
ETA = 2.0 ** -1074
PHI1 = 2.0 ** -53 * (1 + 2 * 2.0 ** -53)


def er_val(frac, exp):
    return F(frac) * F(2) ** int(exp)


@pytest.fixture(scope="module", params=[4, 8, 32, 64])
def orc(request):
    o = Oracle(SETS[request.param])
    yield o
    o.close()


# ----------------------------------------------------------------------------- constants
def _half_conv_bounds(value: F):
    """(RD, RU) of a positive rational to 53 bits as (frac in [0.5,1), exp)."""
    e = 0
    while value >= 1:
        value /= 2
        e += 1
    while value < F(1, 2):
        value *= 2
        e -= 1
    scaled = value * (1 << 53)
    lo = math.floor(scaled)
    hi = math.ceil(scaled)
    out = []
    for v in (lo, hi):
        if v == 1 << 53:
            out.append((0.5, e + 1))
        else:
            out.append((v / 2.0 ** 53, e))
    return out


def test_default_params_set_log2():
    o = Oracle(DEFAULT_PARAMS_SET)
    assert o.log2M == 242


@pytest.mark.parametrize("n", [4, 8, 32, 64])
def test_constants_closed_forms(n):
    mod = SETS[n]
    o = Oracle(mod)
    c = o.constants()
    M = product(mod)
    assert c["log2M"] == M.bit_length() - 1
    for i, m in enumerate(mod):
        assert c["part_inv"][i] == pow(M // m, -1, m)
        rd, ru = c["recip_rd"][i], c["recip_ru"][i]
        assert F(rd) <= F(1, m) <= F(ru)
        assert F(ru) - F(rd) <= 4 * F(1, m) * F(1, 2 ** 52)
    # accuracy constant and refinement factor (src/rnsbase.cuh:389-391)
    acc = 4 * 2.0 ** -52 * n * math.log2(n) * (1 + 1e-7 / 2) / 1e-7
    assert c["accuracy"] == acc
    assert c["ref_factor"] == math.floor(math.log2(1 / (2 * acc)))
    # M rounded down to 53 bits, then directed 53-bit roundings of 1/Md and 1 - 1/Md
    sh = max(M.bit_length() - 53, 0)
    Md = (M >> sh) << sh
    ul, uu = _half_conv_bounds(F(1, Md))
    il, iu = _half_conv_bounds(1 - F(1, Md))
    assert c["unit"] == (ul, uu)
    assert c["inv_unit"] == (il, iu)
    # tables
    pow2 = o.pow2_table()
    for i in range(0, o.log2M + 1, max(1, o.log2M // 37)):
        assert pow2[i].tolist() == [pow(2, i, m) for m in mod]
    inv = o.mrc_inv()
    for i in range(n):
        for j in range(i + 1, n):
            assert (mod[i] * int(inv[i, j])) % mod[j] == 1
    for t in range(30):
        assert c["prod_pow2_res"][t] == M % (1 << (t + 1))
        for j, m in enumerate(mod):
            assert c["part_pow2_res"][t][j] == (M // m) % (1 << (t + 1))
            assert (c["pow2_inv"][t][j] * pow(2, t + 1, m)) % m == 1


def test_accuracy_table_values():
    # values quoted for these sets: 2.13163e-07/21, 1.42109e-06/18, 3.41061e-06/17
    for n, acc, rf in ((8, 2.13163e-07, 21), (32, 1.42109e-06, 18), (64, 3.41061e-06, 17)):
        c = Oracle(SETS[n]).constants()
        assert abs(c["accuracy"] - acc) / acc < 1e-5 and c["ref_factor"] == rf


# ----------------------------------------------------------------------------- BoundNear1
def _rand_doubles(rnd, k):
    out = [0.0, -0.0, 1.0, 1.5, 2.0 ** -1074, 2.0 ** -1022, 1e300, -3.25]
    for _ in range(k):
        out.append(rnd.choice([-1, 1]) * rnd.random() * 2.0 ** rnd.randrange(-1060, 1000))
    return out


def test_boundnear1_enclosures():
    L = lib()
    rnd = random.Random(1)
    xs = _rand_doubles(rnd, 300)
    for a in xs[:120]:
        for b in xs[:60]:
            ex = F(a) + F(b)
            assert F(L.orc_dadd_rd(a, b)) <= ex <= F(L.orc_dadd_ru(a, b))
            ex = F(a) - F(b)
            assert F(L.orc_dsub_rd(a, b)) <= ex <= F(L.orc_dsub_ru(a, b))
            p = F(a) * F(b)
            if abs(p) < F(2) ** 1000:
                assert F(L.orc_dmul_rd(a, b)) <= p <= F(L.orc_dmul_ru(a, b))
            if b != 0 and abs(F(a) / F(b)) < F(2) ** 1000:
                q = F(a) / F(b)
                assert F(L.orc_ddiv_rd(a, b)) <= q <= F(L.orc_ddiv_ru(a, b))
    assert L.orc_dadd_rd(0.0, 0.0) == -ETA and L.orc_dadd_ru(0.0, 0.0) == ETA


def _iter_tree(vals, up):
    """Level-by-level formulation of the padded tree (an absent second half adds 0, an absent
    first half is absent) -- written independently of the recursive restatement."""
    L = lib()
    add = L.orc_dadd_ru if up else L.orc_dadd_rd
    n = len(vals)
    if n == 1:
        return vals[0]
    if n == 3:
        return add(add(vals[0], vals[1]), vals[2])
    K = 2 if n == 2 else 1 << (n.bit_length())
    level = [(v, True) for v in vals] + [(0.0, False)] * (K - n)
    while len(level) > 1:
        nxt = []
        for k in range(0, len(level), 2):
            (a, pa), (b, pb) = level[k], level[k + 1]
            nxt.append((add(a, b if pb else 0.0), True) if pa else (0.0, False))
        level = nxt
    return level[0][0]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 13, 16, 31, 32, 33, 64, 100])
def test_pairwise_sums(n):
    L = lib()
    rnd = random.Random(n)
    for trial in range(40):
        vals = [rnd.random() * 2.0 ** rnd.randrange(-40, 3) for _ in range(n)]
        if trial == 0:
            vals = [1.0] * n  # tests/test_verify_reduction.cu:16-23 pattern
        arr = np.array(vals, np.float64)
        lo = L.orc_psum_rd(arr.ctypes.data, n)
        up = L.orc_psum_ru(arr.ctypes.data, n)
        ex = sum(F(v) for v in vals)
        assert F(lo) <= ex <= F(up)
        assert lo == _iter_tree(vals, False) and up == _iter_tree(vals, True)


# ----------------------------------------------------------------------------- extended range
def _er(frac, exp):
    e = ErFloat()
    e.frac, e.exp = frac, exp
    return e


def test_extended_range_ops():
    L = lib()
    rnd = random.Random(3)
    for _ in range(400):
        x = rnd.choice([-1, 1]) * rnd.uniform(1e-300, 1e300)
        r = ErFloat()
        L.orc_er_set_d(r, x)
        assert er_val(r.frac, r.exp) == F(x) and 1 <= abs(r.frac) < 2
    for _ in range(400):
        a = _er(rnd.uniform(1, 2), rnd.randrange(-1200, 10))
        b = _er(rnd.uniform(1, 2), a.exp - rnd.randrange(-60, 60))
        va, vb = er_val(a.frac, a.exp), er_val(b.frac, b.exp)
        lo, up = ErFloat(), ErFloat()
        L.orc_er_add_rd(lo, _er(a.frac, a.exp), _er(b.frac, b.exp))
        L.orc_er_add_ru(up, _er(a.frac, a.exp), _er(b.frac, b.exp))
        assert er_val(lo.frac, lo.exp) <= va + vb <= er_val(up.frac, up.exp)
        L.orc_er_sub_rd(lo, _er(a.frac, a.exp), _er(b.frac, b.exp))
        L.orc_er_sub_ru(up, _er(a.frac, a.exp), _er(b.frac, b.exp))
        assert er_val(lo.frac, lo.exp) <= va - vb <= er_val(up.frac, up.exp)
        z = _er(rnd.uniform(1, 2), rnd.randrange(-300, 300))
        vz = er_val(z.frac, z.exp)
        L.orc_er_md_rd(lo, a, b, z)
        L.orc_er_md_ru(up, a, b, z)
        assert er_val(lo.frac, lo.exp) <= va * vb / vz <= er_val(up.frac, up.exp)
        want = (va > vb) - (va < vb)
        assert L.orc_er_ucmp(a, b) == want


# ----------------------------------------------------------------------------- modular
def test_reciprocal_reduction():
    L = lib()
    rnd = random.Random(4)
    mod = SETS[32]
    for m in mod[:8] + [3, 5, 7, 11]:
        d = 1.0 / m
        for _ in range(3000):
            x, y = rnd.randrange(m), rnd.randrange(m)
            r = L.orc_mod_mul(x, y, m, d)
            # congruent always (the int32 wrap of a result >= 2^31 is a multiple of 2^32 away)
            assert (r - x * y) % m == 0 or (r + (1 << 32) - x * y) % m == 0
            s = (x * y) % m
            if 1 << 11 <= s < m - (1 << 11):
                assert r == s
            a = L.orc_mod_add(x, y, m, d)
            assert a == (x + y) % m or (a == m and x + y == m)
        # x + y == m returns m exactly when RN(m * RN(1/m)) < 1
        a = L.orc_mod_add(1, m - 1, m, d)
        assert a == (m if m * d < 1 else 0)


def test_mrc_reconstructs(orc):
    mod = orc.moduli
    M = product(mod)
    rnd = random.Random(5)
    vals = eval_inputs(mod, 300, seed=6) if M > 2000 else list(range(M))
    for x in vals:
        mr = orc.mrc([x % m for m in mod])
        b, tot = 1, 0
        for i, m in enumerate(mod):
            assert 0 <= mr[i] < m
            tot += int(mr[i]) * b
            b *= m
        assert tot == x
    del rnd


# ----------------------------------------------------------------------------- evaluation
@pytest.mark.parametrize("fast", [False, True])
def test_eval_encloses(orc, fast):
    mod = orc.moduli
    M = product(mod)
    vals = eval_inputs(mod, 1500, seed=7)
    if fast:
        vals = [v for v in vals if v < M // 2]  # "guaranteed not too large"
    out = orc.eval_array(to_res(vals, mod), fast=fast)
    for x, iv in zip(vals, out):
        lo, up = interval_value(iv)
        if x == 0:
            assert lo == up == 0
            continue
        assert lo <= F(x, M) <= up, x
        assert up - lo <= F(1, 10 ** 6) * F(x, M) + F(2, 10 ** 15)


def test_eval_powers_of_two(orc):
    """All 2^j, j <= log2 M, built by chained reciprocal products as the reference's sweep does."""
    mod = orc.moduli
    M = product(mod)
    L = lib()
    entry = [1] * len(mod)
    cols = []
    for j in range(orc.log2M + 1):
        cols.append(list(entry))
        entry = [L.orc_mod_mul(e, 2, m, 1.0 / m) for e, m in zip(entry, mod)]
    x = np.array(cols, np.int64).T.astype(np.int32)
    out = orc.eval_array(x)
    for j, iv in enumerate(out):
        lo, up = interval_value(iv)
        assert lo <= F(2 ** j, M) <= up


# ----------------------------------------------------------------------------- cmp / div / scale / max
def test_cmp_exact(orc):
    mod = orc.moduli
    M = product(mod)
    if M < 2000:
        xs, ys = zip(*[(a, b) for a in range(M) for b in range(0, M, 7)])
    else:
        xs, ys = cmp_pairs(mod, 6000, seed=8)
    got = orc.cmp_array(to_res(xs, mod), to_res(ys, mod))
    truth = np.array([(a > b) - (a < b) for a, b in zip(xs, ys)])
    assert (got == truth).all()


def test_div_exact(orc):
    mod = orc.moduli
    M = product(mod)
    if M < 2000:
        xs, ds = zip(*[(a, b) for a in range(0, M, 3) for b in range(1, M, 5)])
        cnt_fast = True
    else:
        xs, ds = div_inputs(mod, {8: 300, 32: 40, 64: 16}[len(mod)], seed=9)
        cnt_fast = False
    for fast in (True, False):
        q, r = orc.div_array(to_res(xs, mod), to_res(ds, mod), fast=fast)
        qv, rv = from_res(q, mod), from_res(r, mod)
        for x, d, a, b in zip(xs, ds, qv, rv):
            assert (a, b) == divmod(x, d), (x, d)
    del cnt_fast


def test_scale2pow_exact(orc):
    mod = orc.moduli
    M = product(mod)
    rnd = random.Random(10)
    vals = eval_inputs(mod, 400, seed=11)
    Ds = [rnd.choice([0, 1, 29, 30, 31, 59, 60, 61, rnd.randrange(1, orc.log2M + 5)]) for _ in vals]
    y = orc.scale2pow_array(to_res(vals, mod), np.array(Ds, np.uint32))
    for x, D, v in zip(vals, Ds, from_res(y, mod)):
        assert v == x >> D, (x, D)
    del M


def test_max_exact(orc):
    mod = orc.moduli
    M = product(mod)
    rnd = random.Random(12)
    for case in range(4):
        vals = [rnd.randrange(M) for _ in range(2000)]
        if case == 1:
            vals = [vals[0]] * 50
        if case == 2:
            k = rnd.randrange(2000)
            vals[k] = M - 1
        idx, res = orc.max_array(to_res(vals, mod))
        truth = max(range(len(vals)), key=lambda j: (vals[j], -j))
        assert idx == truth and res.tolist() == [vals[truth] % m for m in mod]
        assert orc.max_eval_array(to_res(vals, mod)) in [j for j, v in enumerate(vals) if v == vals[truth]]


# ----------------------------------------------------------------------------- mpint
def _mpint_in(orc, values):
    mod = orc.moduli
    mags = [abs(v) for v in values]
    signs = np.array([1 if v < 0 else 0 for v in values], np.int32)
    d = to_res(mags, mod)
    ev = orc.eval_array(d)
    e = np.zeros((len(values), 2), ER_DTYPE)
    e[:, 0] = ev["low"]
    e[:, 1] = ev["upp"]
    return d, signs, e


def _signed(orc, rd, rs):
    M = product(orc.moduli)
    return [(-v if s else v) for v, s in zip(from_res(rd, orc.moduli), rs)], M


@pytest.mark.parametrize("n", [8, 32])
def test_mpint_known_answers(n):
    o = Oracle(SETS[n])
    X = _mpint_in(o, [-55])
    Y = _mpint_in(o, [13])
    for op, want in (("add", -42), ("sub", -68), ("mul", -715)):
        rd, rs, re, ovf = o.mpint_array(op, *X, *Y)
        vals, _ = _signed(o, rd, rs)
        assert vals == [want] and ovf[0] == 0
    # mpint_div is Euclidean (divE in the same test): -55 = 13 * (-5) + 10
    rd, rs, re = o.mpint_div_array(X[0], X[1], Y[0], Y[1])
    vals, _ = _signed(o, rd, rs)
    assert vals == [-5]


def test_mpint_div_euclidean(orc):
    """mpint_div over signed values: the Euclidean quotient (remainder in [0, |y|)), sign 0 for 0."""
    mod = orc.moduli
    M = product(mod)
    rnd = random.Random(17)
    bits = max(2, M.bit_length() - 2)
    xs = [rnd.choice([-1, 1]) * rnd.randrange(1 << rnd.randrange(1, bits)) for _ in range(400)] + [0, -1, 1]
    ys = [rnd.choice([-1, 1]) * max(1, rnd.randrange(1 << rnd.randrange(1, bits))) for _ in range(400)] + [5, 7, -7]
    X, Y = _mpint_in(orc, xs), _mpint_in(orc, ys)
    rd, rs, re = orc.mpint_div_array(X[0], X[1], Y[0], Y[1])
    vals, _ = _signed(orc, rd, rs)
    for a, b, v, s in zip(xs, ys, vals, rs):
        q = abs(a) // abs(b) + (1 if a < 0 and abs(a) % abs(b) else 0)
        want = -q if (a < 0) != (b < 0) else q
        assert v == want, (a, b)
        assert s == (1 if want < 0 else 0), (a, b)


def test_mpint_random(orc):
    mod = orc.moduli
    M = product(mod)
    if M < 2000:
        pytest.skip("range too small for products")
    rnd = random.Random(13)
    bits = (M.bit_length() - 4) // 2
    xs = [rnd.choice([-1, 1]) * rnd.randrange(1 << rnd.randrange(1, bits)) for _ in range(500)]
    ys = [rnd.choice([-1, 1]) * rnd.randrange(1 << rnd.randrange(1, bits)) for _ in range(500)]
    X, Y = _mpint_in(orc, xs), _mpint_in(orc, ys)
    for op, f in (("add", lambda a, b: a + b), ("sub", lambda a, b: a - b), ("mul", lambda a, b: a * b)):
        rd, rs, re, ovf = orc.mpint_array(op, *X, *Y)
        vals, _ = _signed(orc, rd, rs)
        for a, b, v, e in zip(xs, ys, vals, re):
            want = f(a, b)
            if want == 0:
                assert v == 0
                continue
            assert v == want, (op, a, b)
            lo, up = er_val(e[0]["frac"], e[0]["exp"]), er_val(e[1]["frac"], e[1]["exp"])
            assert lo <= F(abs(want), M) <= up
        assert (ovf == 0).all()
