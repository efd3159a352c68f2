"""Pins the bit patterns of the oracle's floating-point results (interval evaluation, the pairwise
sums, the extended-range operations and the mpint evaluation fields) against tests/bitmodel.py, a
model built from pred/succ (math.nextafter), Python's round-to-nearest arithmetic and exact
integers instead of the BoundNear1 formulas and the recursive trees of the restatement.

Closed forms used on top of the model (D = log2 of the tree width that src/pairwise.cuh:384-458
gives for n moduli; every zero leaf and every empty subtree is smaller than half an ulp of a
nonzero partial sum, so each tree level applies exactly one pred or succ to it):

* one nonzero leaf, no refinement (x = t M/m_i, s_i = t, ACCURACY <= t/m_i < 1 - 2^-19):
      low = er(pred^(D+1)(RN(t RD_i))),  upp = er(succ^(D+1)(RN(t RU_i)))
  with RD_i, RU_i = pred, succ of RN(1/m_i) (src/rnseval.cuh:45-47);
* one nonzero leaf, refinement (t/m_i < ACCURACY): the refinement multiplies s_i by 2^K without
  reaching m_i, and divides (ddiv, src/rnseval.cuh:85,93) instead of multiplying by a reciprocal:
      low = er(pred^(D+1)(RN(t/m_i))), upp = er(succ^(D+1)(RN(t/m_i)))   (the 2^K scaling is exact);
* two nonzero leaves i < j, no refinement: they meet at level h = bitlength(i xor j), so
      low = er(pred^(D-h+1)(RN(pred^h(RN(t_i RD_i)) + pred^h(RN(t_j RD_j)))))   (upp likewise);
* all-zero sums: psum_rd of n zeros (n = 2^j, width 2n, so n nodes) is -n*ETA, and of n values
  -ETA is -2n*ETA; psum_ru likewise with +ETA;
* er_add_rd/ru of operands whose aligned sum is exact: pred/succ of that exact sum (value and
  normalisation frac in [1, 2)); er_md_rd/ru likewise from two exactly representable steps.

Which mistakes these catch is shown by tests/test_oracle_mutants.py.
"""
from __future__ import annotations

import json
import math
import os
import random

import numpy as np
import pytest

from grns_amd.moduli import largest_primes_below
from oracle.oracle import ER_DTYPE, ErFloat, Oracle, lib
from tests.bitmodel import ETA, Model, NotCovered, er_of, er_value, pred, succ, tree_width
from tests.helpers import SETS, eval_inputs, to_res

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "params_sets.json")) as _f:
    SHIPPED = {e["file"]: e["moduli"] for e in json.load(_f)}

PIN_SETS = {
    "small4": SETS[4],
    "p8": SETS[8],
    "p12": largest_primes_below(12),
    "p20": largest_primes_below(20),
    "p32": SETS[32],
    "p64": SETS[64],
    "params.h": SHIPPED["src/params.h"],
    "16m.256b": SHIPPED["src/params/moduli-64-4096-bit/params.16m.256b.h"],
    "32m.800b": SHIPPED["src/params/moduli-100-6400-bit/params.32m.800b.h"],
    "10m.36b": SHIPPED["src/params/moduli-10-69-bit/params.10m.36b.h"],
}


def _iter(f, v, k):
    for _ in range(k):
        v = f(v)
    return v


def _bits(er):
    return (float(er[0]), int(er[1]))


def _got(iv):
    return _bits((iv["low"]["frac"], iv["low"]["exp"])), _bits((iv["upp"]["frac"], iv["upp"]["exp"]))


# ----------------------------------------------------------------------------- pin checks
# Each returns a list of mismatches (empty = pinned); tests/test_oracle_mutants.py runs the same
# checks on restatements with planted mistakes and expects mismatches.

def single_leaf_cases(model: Model, per_leaf=3, seed=0):
    """(x, expected (low, upp), refined?) for one-nonzero-leaf inputs."""
    rnd = random.Random(seed)
    n, M = model.n, model.M
    if n < 4:
        return []
    D = int(math.log2(tree_width(n)))
    pos = list(range(n)) if n <= 16 else sorted(set([0, 1, n // 2, n - 2, n - 1] + rnd.sample(range(n), 6)))
    out = []
    for i in pos:
        m = model.mod[i]
        lo_t = max(1 << 12, math.ceil(model.acc * m * 1.01) + 1)
        hi_t = m - (1 << 12)
        ts = []
        if lo_t < hi_t:
            ts += [rnd.randrange(lo_t, hi_t) for _ in range(per_leaf)]
        elif m < 1 << 12:
            ts += [t for t in range(1, m) if t / m >= model.acc * 1.01][:per_leaf]
        for t in ts:
            low = er_of(_iter(pred, t * model.rd[i], D + 1))
            upp = er_of(_iter(succ, t * model.ru[i], D + 1))
            out.append((t * (M // m), (low, upp), False))
        # refinement: t / m below the accuracy constant, t * 2^K never reaches m
        small = [t for t in (1, 2, 3, 5, 77, 1000) if t / m < model.acc * 0.9 and t < m]
        for t in small[:per_leaf]:
            q = t / m
            out.append((t * (M // m), (er_of(_iter(pred, q, D + 1)), er_of(_iter(succ, q, D + 1))), True))
    return out


def two_leaf_cases(model: Model, count=40, seed=1):
    rnd = random.Random(seed)
    n, M = model.n, model.M
    if n < 4:
        return []
    D = int(math.log2(tree_width(n)))
    out = []
    for _ in range(count):
        i, j = sorted(rnd.sample(range(n), 2))
        mi, mj = model.mod[i], model.mod[j]
        h = (i ^ j).bit_length()
        for _try in range(50):
            ti = rnd.randrange(1, mi)
            tj = rnd.randrange(1, mj)
            v = ti / mi + tj / mj
            if min(ti / mi, tj / mj) > max(model.acc * 1.01, 2.0 ** -19) and v < 1 - 2.0 ** -19 \
                    and max(ti / mi, tj / mj) < 1 - 2.0 ** -19:
                break
        else:
            continue
        a = _iter(pred, ti * model.rd[i], h)
        b = _iter(pred, tj * model.rd[j], h)
        low = er_of(_iter(pred, a + b, D - h + 1))
        a = _iter(succ, ti * model.ru[i], h)
        b = _iter(succ, tj * model.ru[j], h)
        upp = er_of(_iter(succ, a + b, D - h + 1))
        out.append(((ti * (M // mi) + tj * (M // mj)) % M, (low, upp)))
    return out


def check_closed_forms(orc: Oracle, model: Model):
    bad = []
    for x, want, _ in single_leaf_cases(model):
        for fast in (False, True):
            got = _got(orc.eval1(model.residues(x), fast=fast))
            if got != want:
                bad.append(("single", fast, x, got, want))
    for x, want in two_leaf_cases(model):
        for fast in (False, True):
            got = _got(orc.eval1(model.residues(x), fast=fast))
            if got != want:
                bad.append(("two", fast, x, got, want))
    return bad


def wrap_refinement_inputs(model: Model, count=30, seed=2):
    """Two leaves whose fractions sum to 1 + eps with 0 < eps < 1/m_j: the integer part is dropped
    and the refinement multiplies both residues past their moduli."""
    rnd = random.Random(seed)
    n, M = model.n, model.M
    out = []
    for _ in range(count):
        i, j = sorted(rnd.sample(range(n), 2))
        mi, mj = model.mod[i], model.mod[j]
        ti = rnd.randrange(mi // 4, 3 * mi // 4)
        num = mj * (mi - ti)
        tj = -(-num // mi)
        if tj >= mj:
            continue
        out.append((ti * (M // mi) + tj * (M // mj)) % M)
    return out


def model_inputs(model: Model, count, seed):
    vals = eval_inputs(model.mod, count, seed)
    # the 2^j sweep of tests/test_perf_rnseval.cu:682-687 (every j for the small sets)
    step = max(1, model.log2M // 160)
    vals += [1 << j for j in sorted(set(range(0, model.log2M + 1, step)) | {model.log2M})]
    vals += list(range(1, 10)) + [model.M - k for k in range(1, 11)]  # tests/test_verify_rnseval.cu:87-149
    vals += wrap_refinement_inputs(model)
    return [v % model.M for v in vals]


def check_model(orc: Oracle, model: Model, count=600, seed=3):
    """Oracle == model, bit for bit, on every covered input; returns (mismatches, covered, total)."""
    bad, covered, total = [], 0, 0
    vals = model_inputs(model, count, seed)
    x = to_res(vals, model.mod)
    got = {f: orc.eval_array(x, fast=f) for f in (False, True)}
    for k, v in enumerate(vals):
        for fast in (False, True):
            if fast and v >= model.M // 2:
                continue
            total += 1
            try:
                want = model.eval(v, fast=fast)
            except NotCovered:
                continue
            covered += 1
            g = _got(got[fast][k])
            if g != (_bits(want[0]), _bits(want[1])):
                bad.append((fast, v, g, want))
    return bad, covered, total


def check_zero_sums(L):
    bad = []
    for n in (4, 8, 16, 32, 64, 128, 256):
        z = np.zeros(n)
        me = np.full(n, -ETA)
        pe = np.full(n, ETA)
        for got, want in ((L.orc_psum_rd(z.ctypes.data, n), -n * ETA), (L.orc_psum_ru(z.ctypes.data, n), n * ETA),
                          (L.orc_psum_rd(me.ctypes.data, n), -2 * n * ETA),
                          (L.orc_psum_ru(pe.ctypes.data, n), 2 * n * ETA)):
            if got != want:
                bad.append((n, got, want))
    return bad


def _er(fe):
    e = ErFloat()
    e.frac, e.exp = fe
    return e


def check_er_exact(L, count=300, seed=4):
    """er_add_rd/ru and er_md_rd/ru on operands whose intermediate results are exact."""
    rnd = random.Random(seed)
    bad = []
    for _ in range(count):
        a = (1 + rnd.randrange(1 << 20) / (1 << 20)) * rnd.choice([1, -1])
        b = (1 + rnd.randrange(1 << 20) / (1 << 20)) * rnd.choice([1, -1])
        ea = rnd.randrange(-400, 400)
        eb = ea - rnd.randrange(0, 25)
        exact = (a + math.ldexp(b, eb - ea))  # exact: both have <= 21 significant bits, shift <= 24
        for down, fn in ((True, L.orc_er_add_rd), (False, L.orc_er_add_ru)):
            r = ErFloat()
            fn(r, _er((a, ea)), _er((b, eb)))
            want_v = (pred if down else succ)(exact)
            want = (0.0, 0) if want_v == 0 else er_of(want_v)
            want = (want[0], want[1] + ea)
            if (r.frac, r.exp) != want:
                bad.append(("add", down, a, ea, b, eb, (r.frac, r.exp), want))
        c = 1 + rnd.randrange(1 << 20) / (1 << 20)
        z = (1.0, rnd.randrange(-50, 50)) if rnd.random() < 0.5 else (2.0 ** -1 * 2, rnd.randrange(-50, 50))
        ez = z[1]
        for down, fn in ((True, L.orc_er_md_rd), (False, L.orc_er_md_ru)):
            r = ErFloat()
            fn(r, _er((abs(a), ea)), _er((c, eb)), _er(z))
            p = abs(a) * c  # exact (<= 42 significant bits)
            step = pred if down else succ
            v = step(step(p) / z[0])
            want = er_of(v)
            want = (want[0], want[1] + ea + eb - ez)
            if (r.frac, r.exp) != want:
                bad.append(("md", down, (r.frac, r.exp), want))
    return bad


def check_mpint_eval(orc: Oracle, model: Model, count=200, seed=5):
    """The eval fields of mpint add/sub/mul against the model's interval sums and products."""
    rnd = random.Random(seed)
    M = model.M
    bits = max(4, (M.bit_length() - 6) // 2)
    xs = [rnd.choice([-1, 1]) * rnd.randrange(1, 1 << rnd.randrange(2, bits)) for _ in range(count)]
    ys = [rnd.choice([-1, 1]) * rnd.randrange(1, 1 << rnd.randrange(2, bits)) for _ in range(count)]
    ys[:10] = [-v for v in xs[:10]]  # cancellation to zero
    ys[10:20] = [-v + rnd.choice([-1, 1]) for v in xs[10:20]]

    def side(vals):
        d = to_res([abs(v) for v in vals], model.mod)
        s = np.array([1 if v < 0 else 0 for v in vals], np.int32)
        ev = orc.eval_array(d)
        e = np.zeros((len(vals), 2), ER_DTYPE)
        e[:, 0] = ev["low"]
        e[:, 1] = ev["upp"]
        return d, s, e

    X, Y = side(xs), side(ys)
    bad = []
    for op in ("add", "sub", "mul"):
        rd, rs, re, _ = orc.mpint_array(op, *X, *Y)
        for k in range(count):
            ex = (_bits(X[2][k][0]), _bits(X[2][k][1]))
            ey = (_bits(Y[2][k][0]), _bits(Y[2][k][1]))
            try:
                if op == "mul":
                    lo, up = model.mpint_mul_eval(ex, ey)
                    sign = None
                else:
                    sy = int(Y[1][k]) ^ (1 if op == "sub" else 0)
                    sign, lo, up = model.mpint_addsub_eval(int(X[1][k]), ex, sy, ey)
            except NotCovered:
                continue
            got = (_bits(re[k][0]), _bits(re[k][1]))
            if got != (_bits(lo), _bits(up)) or (sign is not None and int(rs[k]) != sign):
                bad.append((op, xs[k], ys[k], got, (lo, up), sign, int(rs[k])))
    return bad


# ----------------------------------------------------------------------------- tests
@pytest.fixture(scope="module", params=sorted(PIN_SETS))
def pair(request):
    mod = PIN_SETS[request.param]
    o = Oracle(mod)
    yield o, Model(mod)
    o.close()


def test_model_constants_match_oracle(pair):
    """The constants the model derives from closed forms are the oracle's."""
    orc, model = pair
    c = orc.constants()
    assert c["recip_rd"] == model.rd and c["recip_ru"] == model.ru
    assert c["accuracy"] == model.acc and c["ref_factor"] == model.ref
    assert c["unit"] == model.unit and c["inv_unit"] == model.inv_unit


def test_eval_closed_forms(pair):
    orc, model = pair
    cases = single_leaf_cases(model)
    if model.n >= 4:
        assert cases and any(r for _, _, r in cases) or model.acc * min(model.mod) < 1
    bad = check_closed_forms(orc, model)
    assert not bad, bad[:3]


def test_eval_bits_equal_model(pair):
    orc, model = pair
    bad, covered, total = check_model(orc, model)
    assert not bad, bad[:3]
    assert covered >= 0.95 * total, (covered, total)


def test_model_covers_every_branch():
    """Across the pinned sets, the model's comparison runs every branch of rns_eval_compute:
    plain, refinement, and both outcomes of the ambiguity test (mr > 0, mr == 0)."""
    seen = {k: 0 for k in ("plain", "mrc_upper", "mrc_zero", "refine")}
    for name in ("p8", "p32", "params.h", "16m.256b"):
        mod = PIN_SETS[name]
        o, m = Oracle(mod), Model(mod)
        check_model(o, m, count=300)
        for k, v in m.branches.items():
            seen[k] += v
    assert all(v > 0 for v in seen.values()), seen


def test_zero_sums():
    assert not check_zero_sums(lib())


def test_er_exact_operations():
    assert not check_er_exact(lib())


def test_mpint_eval_fields(pair):
    orc, model = pair
    if model.M < 1 << 40:
        pytest.skip("range too small for the mpint operand sizes")
    bad = check_mpint_eval(orc, model)
    assert not bad, bad[:3]
