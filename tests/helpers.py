"""Seeded inputs built from exact Python integers (shared by the CPU and GPU tests)."""
from __future__ import annotations

import random
from fractions import Fraction

import numpy as np

from grns_amd.moduli import SMALL_SET, largest_primes_below, product

SETS = {4: SMALL_SET, 8: largest_primes_below(8), 32: largest_primes_below(32), 64: largest_primes_below(64)}


def to_res(values, moduli) -> np.ndarray:
    """(n, count) int32 modulus-major residues of Python integers (negative values wrap mod m)."""
    out = np.zeros((len(moduli), len(values)), np.int64)
    for j, v in enumerate(values):
        for i, m in enumerate(moduli):
            out[i, j] = v % m
    return out.astype(np.int32)


def from_res(res, moduli) -> list:
    """CRT back to integers in [0, M)."""
    M = product(moduli)
    basis = []
    for m in moduli:
        Mi = M // m
        basis.append(Mi * pow(Mi, -1, m))
    res = np.asarray(res, dtype=np.int64)
    vals = []
    for j in range(res.shape[1]):
        vals.append(sum(int(res[i, j]) * basis[i] for i in range(len(moduli))) % M)
    return vals


def edge_values(M):
    return [0, 1, 2, 3, M - 1, M - 2, M // 2, M // 2 - 1, M // 2 + 1, M // 3, 2 * M // 3]


def eval_inputs(moduli, count, seed):
    M = product(moduli)
    rnd = random.Random(seed)
    if M <= count:
        return list(range(M))
    vals = edge_values(M)
    vals += [1 << j for j in range(0, M.bit_length() - 1, 7)]
    vals += [rnd.randrange(1 << rnd.randrange(1, 64)) % M for _ in range(count // 10)]
    vals += [(M - 1 - rnd.randrange(1 << 40)) % M for _ in range(count // 20)]
    while len(vals) < count:
        vals.append(rnd.randrange(M))
    return vals[:count]


def cmp_pairs(moduli, count, seed):
    M = product(moduli)
    rnd = random.Random(seed)
    xs, ys = [], []
    for v in edge_values(M):
        for d in (-1, 0, 1):
            xs.append(v); ys.append(min(max(v + d, 0), M - 1))
    while len(xs) < count:
        x = rnd.randrange(M)
        k = rnd.randrange(5)
        if k == 0:
            y = x
        elif k in (1, 2):
            y = x + (1 if k == 1 else -1)
            if not 0 <= y < M:
                y = x - (y - x)
        elif k == 3:
            y = min(max(x + rnd.randrange(-(1 << 80), 1 << 80), 0), M - 1)
        else:
            y = rnd.randrange(M)
        xs.append(x); ys.append(y)
    return xs[:count], ys[:count]


def div_inputs(moduli, count, seed):
    M = product(moduli)
    L = M.bit_length() - 1
    rnd = random.Random(seed)
    xs, ds = [], []
    for _ in range(count):
        x = rnd.randrange(M)
        b = rnd.randrange(1, L)
        d = rnd.randrange(1 << (b - 1), 1 << b)
        xs.append(x); ds.append(d)
    # small cases: d > x, d == x, d == 1
    xs += [5, 7, 7, M - 1]; ds += [7, 7, 1, 1]
    return xs, ds


def interval_value(iv):
    """Exact rational values of an interval record."""
    lo = Fraction(float(iv['low']['frac'])) * Fraction(2) ** int(iv['low']['exp'])
    up = Fraction(float(iv['upp']['frac'])) * Fraction(2) ** int(iv['upp']['exp'])
    return lo, up
