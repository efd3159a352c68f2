"""A second model of the bit patterns the reference's CPU path produces, written from what the
mathematics of its operations fixes rather than from its code.  It is test infrastructure: the
oracle (oracle/) is checked against it, bit for bit.

The facts it is built on:

* BoundNear1 (src/dinterval.cuh:37-147, the method of Rump, Zimmermann, Boldo and Melquiond,
  "Computing predecessor and successor in rounding to nearest", BIT 49 (2009)) returns, for a
  round-to-nearest result c = RN(a op b), pred(c) for the downward and succ(c) for the upward
  operation, whenever |c| lies outside [2^-1022, 2^-1020] (the only range where the method may
  step two floating-point numbers).  pred and succ are math.nextafter here; RN is Python's own
  binary64 arithmetic.  The model refuses any value inside that range (NotCovered).
* The pairwise sums (src/pairwise.cuh:384-458) add adjacent elements level by level in a tree whose
  width the reference's dispatch table fixes (WIDTH below); a subtree with no elements is the
  literal 0.  For two elements i < j the first common node is log2 of the smallest power of two
  above i xor j levels up.
* s_i = x_i w_i mod m_i and every residue product are exact integers (the reciprocal reduction of
  src/modular.cuh:74-79 returns the canonical residue when the exact quotient is more than 2^-20
  from an integer; the model refuses the other cases).
* The top mixed-radix digit of x is floor(x / (m_0 ... m_{n-2})).
* The constants: RECIP_RD/RU = pred/succ(RN(1/m)), UNIT and INV_UNIT from M rounded down to 53 bits
  (src/rnsbase.cuh:326,393-416), the accuracy and refinement factor by their formula (:389-391).

What it does with them is the sequence of steps of rns_eval_compute(_fast) (src/rnseval.cuh:29-167)
and of the mpint evaluation fields (src/mpint.cuh:160-278).
"""
from __future__ import annotations

import math
from fractions import Fraction as F

ETA = 2.0 ** -1074
INF = math.inf


class NotCovered(Exception):
    """The input leaves the range where the facts above hold."""


def pred(c: float) -> float:
    if 2.0 ** -1022 <= abs(c) <= 2.0 ** -1020:
        raise NotCovered("value near the underflow threshold")
    return math.nextafter(c, -INF)


def succ(c: float) -> float:
    if 2.0 ** -1022 <= abs(c) <= 2.0 ** -1020:
        raise NotCovered("value near the underflow threshold")
    return math.nextafter(c, INF)


# src/pairwise.cuh:384-458: length -> tree width ("For 4 ... 7" -> psum8, ..., "For 4096 ... 8192" -> psum8192)
WIDTH = [(4096, 8192), (2048, 4096), (1024, 2048), (512, 1024), (256, 512), (128, 256), (64, 128), (32, 64),
         (16, 32), (8, 16), (4, 8)]


def tree_width(length: int) -> int:
    for lo, w in WIDTH:
        if length >= lo:
            return w
    raise ValueError("lengths below 4 have their own forms")


def tree_sum(leaves, down: bool) -> float:
    """Directed pairwise sum of the leaves (pred/succ of the RN sum at every node)."""
    step = pred if down else succ
    n = len(leaves)
    if n == 1:
        return leaves[0]
    if n == 2:
        return step(leaves[0] + leaves[1])
    if n == 3:
        return step(step(leaves[0] + leaves[1]) + leaves[2])
    level = list(leaves) + [None] * (tree_width(n) - n)
    while len(level) > 1:
        nxt = []
        for k in range(0, len(level), 2):
            a, b = level[k], level[k + 1]
            nxt.append(None if a is None else step(a + (0.0 if b is None else b)))
        level = nxt
    return level[0]


def levels_to_join(i: int, j: int) -> int:
    """Tree level (1 = the pairs of adjacent leaves) of the first node above leaves i and j."""
    return (i ^ j).bit_length()


def er_of(v: float):
    """er_set_d (src/extrange.cuh:29-42) of a normal double or zero: frac in [1, 2), exact."""
    if v == 0:
        return (0.0, 0)
    if abs(v) < 2.0 ** -1022:
        raise NotCovered("subnormal bound")
    m, e = math.frexp(v)
    return (2 * m, e - 1)


def er_norm(frac: float, exp: int):
    """er_adjust (src/extrange.cuh:116-125): the same value with frac in [1, 2)."""
    if frac == 0:
        return (0.0, 0)
    f, e = er_of(frac)
    return (f, exp + e)


def er_value(er) -> F:
    return F(er[0]) * F(2) ** int(er[1])


def _half_conv(value: F, up: bool):
    e = 0
    while value >= 1:
        value /= 2
        e += 1
    while value < F(1, 2):
        value *= 2
        e -= 1
    v = math.ceil(value * (1 << 53)) if up else math.floor(value * (1 << 53))
    return (0.5, e + 1) if v == 1 << 53 else (v / 2.0 ** 53, e)


class Model:
    """One moduli set."""

    def __init__(self, moduli):
        self.mod = [int(m) for m in moduli]
        self.n = n = len(self.mod)
        M = 1
        for m in self.mod:
            M *= m
        self.M = M
        self.log2M = M.bit_length() - 1
        self.w = [pow(M // m, -1, m) for m in self.mod]
        self.top_radix = M // self.mod[-1]
        self.rd = [pred(1.0 / m) for m in self.mod]
        self.ru = [succ(1.0 / m) for m in self.mod]
        self.acc = 4 * 2.0 ** -52 * n * math.log2(n) * (1 + 1e-7 / 2) / 1e-7 if n > 1 else 0.0
        self.ref = math.floor(math.log2(1 / (2 * self.acc))) if n > 1 else 0
        sh = max(M.bit_length() - 53, 0)
        Md = (M >> sh) << sh
        self.unit = (_half_conv(F(1, Md), False), _half_conv(F(1, Md), True))
        self.inv_unit = (_half_conv(1 - F(1, Md), False), _half_conv(1 - F(1, Md), True))
        self.branches = {"plain": 0, "mrc_upper": 0, "mrc_zero": 0, "refine": 0}

    # ------------------------------------------------------------------ residue products
    def _product(self, a: int, b: int, m: int) -> int:
        r = a * b
        f = F(r % m, m)
        if r and not (F(1, 1 << 20) < f < 1 - F(1, 1 << 20)):
            raise NotCovered("reciprocal quotient next to an integer")
        return r % m

    def residues(self, x: int):
        return [x % m for m in self.mod]

    # ------------------------------------------------------------------ evaluation
    def eval(self, x: int, fast: bool = False):
        """(low, upp) as er pairs for 0 <= x < M given by canonical residues."""
        res = self.residues(x)
        if all(r == 0 for r in res):
            return (0.0, 0), (0.0, 0)
        s = [self._product(r, w, m) for r, w, m in zip(res, self.w, self.mod)]
        suml = tree_sum([pred(si * rd) for si, rd in zip(s, self.rd)], True)
        sumu = tree_sum([succ(si * ru) for si, ru in zip(s, self.ru)], False)
        if not (-1 < suml < 2 ** 32 and -1 < sumu < 2 ** 32):
            raise NotCovered("sum outside the unsigned range")
        whl, whu = int(suml), int(sumu)
        suml -= whl
        sumu -= whu
        low, upp = er_of(suml), er_of(sumu)
        if not fast:
            mr = x // self.top_radix if whl != whu else -1
            if mr > 0:
                self.branches["mrc_upper"] += 1
                return low, self.inv_unit[1]
            if mr == 0:
                self.branches["mrc_zero"] += 1
                low = self.unit[0]
        if sumu >= self.acc:
            self.branches["plain"] += 1
            return low, upp
        self.branches["refine"] += 1
        K = 0
        while sumu < self.acc:
            k = max(-(math.ceil(math.log2(sumu)) + 1), self.ref)
            if not 0 <= k <= self.log2M:
                raise NotCovered("refinement factor outside RNS_POW2")
            s = [self._product(si, pow(2, k, m), m) for si, m in zip(s, self.mod)]
            sumu = tree_sum([succ(si / m) for si, m in zip(s, self.mod)], False)
            sumu -= int(sumu)
            K += k
        suml = tree_sum([pred(si / m) for si, m in zip(s, self.mod)], True)
        suml -= int(suml)
        low, upp = er_of(suml), er_of(sumu)
        return (low[0], low[1] - K), (upp[0], upp[1] - K)

    # ------------------------------------------------------------------ extended range
    @staticmethod
    def _aligned(x, y):
        """Both operands' fractions at the larger exponent (exact), and that exponent."""
        (xf, xe), (yf, ye) = x, y
        if xf == 0:
            xe = 0
        if yf == 0:
            ye = 0
        d = (xe - ye) if (xf != 0 and yf != 0) else 0
        if d > 0:
            b = math.ldexp(yf, -d)
            if b != 0 and abs(b) < 2.0 ** -1022:
                raise NotCovered("aligned operand below the normal range")
            return xf, b, xe
        if d < 0:
            a = math.ldexp(xf, d)
            if a != 0 and abs(a) < 2.0 ** -1022:
                raise NotCovered("aligned operand below the normal range")
            return a, yf, ye
        return xf, yf, (ye if xe == 0 else xe)

    def er_add(self, x, y, down: bool):
        a, b, e = self._aligned(x, y)
        c = a + b
        return er_norm(pred(c) if down else succ(c), e)

    def er_md(self, x, y, z, down: bool):
        step = pred if down else succ
        return er_norm(step(step(x[0] * y[0]) / z[0]), x[1] + y[1] - z[1])

    # ------------------------------------------------------------------ mpint evaluation fields
    def mpint_addsub_eval(self, sx, ex, sy, ey):
        """Signed interval sum of x and (+/-)y (sy already includes the subtraction's flip):
        lower = round-down sum of the lower ends, upper = round-up sum of the upper ends; the
        result is negated when both ends are negative.  Returns (sign, low, upp)."""
        def signed(s, e):
            lo, up = e
            return ((-up[0], up[1]), (-lo[0], lo[1])) if s else (lo, up)
        xl, xu = signed(sx, ex)
        yl, yu = signed(sy, ey)
        lo = self.er_add(xl, yl, True)
        up = self.er_add(xu, yu, False)
        if lo[0] < 0 and up[0] < 0:
            return 1, (-up[0], up[1]), (-lo[0], lo[1])
        return 0, lo, up

    def mpint_mul_eval(self, ex, ey):
        return self.er_md(ex[0], ey[0], self.unit[1], True), self.er_md(ex[1], ey[1], self.unit[0], False)
