"""Seeded synthetic workloads on the device (libgrns_workload.so) for the benchmark and the
full-size tests.  Values are functions of (seed, stream, element) only.

Workloads (sizes as named by the benchmark configurations):
  * eval_cmp(n, N): x uniform in [0, M); 1% of the entries forced round-robin to 0, M-1,
    floor(M/2)-1, floor(M/2), floor(M/2)+1; pairs (x, y) with y = x + delta, delta drawn from
    {-1, 0, +1, uniform} (delta = uniform draws y independently; x + delta outside [0, M) is
    reflected to x - delta);
  * max_planted(n, N): x uniform in [0, h * M/m_n) with h = (m_n - 1) // 2 (mixed-radix digits
    with a bounded top digit, so every x < M/2); V = floor(M/2) - 1 planted at one seeded index
    and V - 1 at 1000 other seeded indices;
  * div(n, N): x uniform in [0, M); divisor bit length uniform in [1, L], divisor uniform with that
    exact bit length; scale exponents uniform in [1, K].
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .moduli import product

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgrns_workload.so")
_wl = None


def wlib():
    global _wl
    if _wl is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build with `python -m grns_amd.build`")
        L = C.CDLL(LIB_PATH)
        P, I, LL, U = C.c_void_p, C.c_int, C.c_longlong, C.c_ulonglong
        L.gw_uniform.argtypes = [P, I, LL, U, U, P, LL, P]
        L.gw_mixed_radix.argtypes = [P, I, P, I, LL, U, U, P, LL, P]
        L.gw_bits.argtypes = [P, I, P, I, P, I, LL, U, U, P, LL, P]
        L.gw_place.argtypes = [I, I, P, P, P, LL, P]
        L.gw_uniform_u32.argtypes = [LL, C.c_uint, C.c_uint, U, U, P, P]
        _wl = L
    return _wl


def residues_of(value: int, moduli) -> list:
    return [value % m for m in moduli]


def _p(t):
    return C.c_void_p(t.data_ptr())


def _s():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ok(rc, where):
    if rc != 0:
        raise RuntimeError(f"{where} failed ({rc})")


class Synth:
    """Generators for one moduli set on one device."""

    def __init__(self, moduli, device="cuda"):
        import torch
        self.moduli = [int(v) for v in moduli]
        self.n = len(self.moduli)
        self.M = product(self.moduli)
        self.device = device
        self.mod_t = torch.tensor(self.moduli, dtype=torch.int32, device=device)
        n = self.n
        mrb = np.zeros((n, n), np.int64)  # B_k mod m_i
        for i, m in enumerate(self.moduli):
            b = 1
            for k in range(n):
                mrb[k, i] = b % m
                b = b * self.moduli[k]
        self.mrb_t = torch.tensor(mrb.astype(np.int32), device=device)
        words = (self.M.bit_length() + 31) // 32 + 1
        p32 = np.zeros((words, n), np.int64)
        for i, m in enumerate(self.moduli):
            for w in range(words):
                p32[w, i] = pow(2, 32 * w, m)
        self.words = words
        self.p32_t = torch.tensor(p32.astype(np.int32), device=device)

    def empty(self, count):
        import torch
        return torch.empty((self.n, count), dtype=torch.int32, device=self.device)

    def uniform(self, count, seed, stream=0, out=None):
        out = self.empty(count) if out is None else out
        _ok(wlib().gw_uniform(_p(self.mod_t), self.n, count, seed, stream, _p(out), out.stride(0), _s()), "gw_uniform")
        return out

    def half_range(self, count, seed, stream=0, out=None):
        out = self.empty(count) if out is None else out
        top = (self.moduli[-1] - 1) // 2
        _ok(wlib().gw_mixed_radix(_p(self.mod_t), self.n, _p(self.mrb_t), top, count, seed, stream, _p(out),
                                  out.stride(0), _s()), "gw_mixed_radix")
        return out

    def uniform_u32(self, count, lo, hi, seed, stream=0):
        import torch
        out = torch.empty(count, dtype=torch.int32, device=self.device)
        _ok(wlib().gw_uniform_u32(count, lo, hi, seed, stream, _p(out), _s()), "gw_uniform_u32")
        return out

    def bits(self, bits_t, seed, stream=0, top_set=True, out=None):
        count = bits_t.numel()
        out = self.empty(count) if out is None else out
        _ok(wlib().gw_bits(_p(self.mod_t), self.n, _p(self.p32_t), self.words, _p(bits_t), int(top_set), count, seed,
                           stream, _p(out), out.stride(0), _s()), "gw_bits")
        return out

    def place(self, out, values, idx):
        import torch
        if len(values) == 0:
            return out
        vals = torch.tensor(np.array([residues_of(v, self.moduli) for v in values], dtype=np.int64).astype(np.int32),
                            device=self.device)
        ix = torch.tensor(np.asarray(idx, dtype=np.int64), device=self.device)
        _ok(wlib().gw_place(self.n, len(values), _p(vals), _p(ix), _p(out), out.stride(0), _s()), "gw_place")
        return out

    # ------------------------------------------------------------------ benchmark workloads
    def eval_cmp(self, count, seed=1):
        """(x, y, forced) for the interval-evaluation / comparison workload."""
        import torch
        M = self.M
        x = self.uniform(count, seed, 1)
        rng = np.random.default_rng(seed)
        nforce = count // 100
        pos = rng.choice(count, size=nforce, replace=False) if nforce else np.zeros(0, np.int64)
        edges = [0, M - 1, M // 2 - 1, M // 2, M // 2 + 1]
        forced = {int(p): edges[k % len(edges)] for k, p in enumerate(pos)}
        # y = x + delta residue-wise; delta code 0:-1 1:0 2:+1 3:uniform
        code = torch.tensor(rng.integers(0, 4, size=count).astype(np.int32), device=self.device)
        delta = (code - 1).clamp(-1, 1)
        m = self.mod_t.view(-1, 1)
        y = torch.remainder(x + delta.view(1, -1), m).to(torch.int32)
        yu = self.uniform(count, seed, 2)
        y = torch.where((code == 3).view(1, -1), yu, y).contiguous()
        if forced:
            idx = list(forced.keys())
            xv = [forced[p] for p in idx]
            cd = code[torch.tensor(idx, device=self.device)].cpu().numpy()
            yv = []
            for v, c in zip(xv, cd):
                d = int(c) - 1
                if c == 3:
                    yv.append(None)
                    continue
                t = v + d
                if t < 0 or t >= M:
                    t = v - d
                yv.append(t)
            self.place(x, xv, idx)
            keep = [(p, v) for p, v in zip(idx, yv) if v is not None]
            self.place(y, [v for _, v in keep], [p for p, _ in keep])
        return x, y, forced

    def max_planted(self, count, seed=2, ties=1000):
        """(x, planted index, V) with V = floor(M/2) - 1 unique maximum and `ties` copies of V - 1."""
        M = self.M
        x = self.half_range(count, seed, 3)
        rng = np.random.default_rng(seed)
        k = min(ties + 1, count)
        pos = rng.choice(count, size=k, replace=False)
        V = M // 2 - 1
        p = int(pos[0])
        self.place(x, [V] + [V - 1] * (k - 1), [int(v) for v in pos])
        return x, p, V

    def div(self, count, seed=3, max_bits=1980, max_shift=1900):
        """(x, d, D): dividends uniform in [0, M), divisors of uniform bit length, scale exponents."""
        import torch
        x = self.uniform(count, seed, 4)
        nb = self.uniform_u32(count, 1, max_bits, seed, 5)
        d = self.bits(nb, seed, 6, top_set=True)
        D = self.uniform_u32(count, 1, max_shift, seed, 7)
        return x, d, D, nb


# ---------------------------------------------------------------------- multi-request workloads
ALIGN = 64  # elements: 256 bytes of int32 digits


def mpint_batch(ctx, syn, vectors=4096, max_len=65536, seed=4, bits=480):
    """The batched mpint workload: `vectors` vectors with lengths floor(e^U[0, ln max_len]) (log-uniform
    in [1, max_len]), an operation (0 add, 1 sub, 2 mul) drawn per vector, operands X and Y of uniform
    `bits`-bit magnitudes with random signs and their evaluations (ctx.eval) as the eval fields.

    The vectors share one digit-major array per operand, one after another, each starting on a
    256-byte boundary (a multiple of ALIGN elements), and the row stride `span` is a multiple of
    ALIGN too, so that every digit row of every vector starts on a 256-byte boundary: with rows at
    unaligned offsets each wave's 256-byte row segment spans three 128-byte lines instead of two, and
    an elementwise pass over 2.5e7 elements runs at 3.95 instead of 5.33 TB/s on MI355X.  The padding
    elements between vectors hold valid numbers and belong to no vector.

    Returns a dict: X, Y ((digits (n, span), sign (span,), eval (span, 2, 2))), lens, offs, ops
    (numpy arrays), total (elements in vectors) and span."""
    import torch
    rng = np.random.default_rng(seed)
    lens = np.floor(np.exp(rng.uniform(0, np.log(max_len), size=vectors))).astype(np.int64).clip(1, max_len)
    ops = rng.integers(0, 3, size=vectors)
    total = int(lens.sum())
    slots = (lens + ALIGN - 1) // ALIGN * ALIGN
    offs = np.concatenate([[0], np.cumsum(slots)[:-1]]).astype(np.int64)
    span = int(slots.sum())

    def operand(stream):
        b = torch.full((span,), bits, dtype=torch.int32, device=syn.device)
        d = syn.bits(b, seed, stream, top_set=False)
        s = syn.uniform_u32(span, 0, 1, seed, stream + 100)
        ev = ctx.eval(d)  # interval records have the mp_int_t eval[2] layout
        return d, s, ev.view(span, 2, 2)

    return {"X": operand(10), "Y": operand(20), "lens": lens, "offs": offs, "ops": ops, "total": total,
            "span": span}


MPINT_OPS = ["mpint_add", "mpint_sub", "mpint_mul"]


def mpint_view(T, off, length):
    """The (digits, sign, eval) slices of one vector of an mpint operand triple."""
    return T[0][:, off:off + length], T[1][off:off + length], T[2][off:off + length]


SERVING_MIX = (("evalcmp", 0.3), ("max", 0.2), ("div", 0.2), ("scale", 0.1), ("mpint", 0.2))


def serving_requests(n_req, seed, lo=10, hi=22, sets=(8, 32, 64)):
    """The serving workload's requests: a mix of 30% eval+cmp, 20% arg-max, 20% div_fast, 10%
    scale2pow and 20% mpint mul, sizes floor(e^U[ln 2^lo, ln 2^hi]), each on one of the resident
    moduli sets of 8, 32 and 64 largest primes below 2^31, operands drawn like the single-operation
    workloads (edge values in eval+cmp, a planted maximum with near-ties in arg-max, divisors of
    uniform bit length, exponents in [1, 1900], mpint magnitudes of up to 480 bits within the range).

    Each request is a dict: ctx, tasks (grns_batch_task array), plain (a closure running it through
    the per-call entry points into out_p), out_s (the served outputs), out_p, keep (its inputs, alive
    until the run ends), kind, n, count, and truth (for arg-max: planted index and value; for the
    others: the input tensors the checker samples)."""
    import torch

    import grns_amd as G

    rng = np.random.default_rng(seed)
    setmod = {k: G.largest_primes_below(k) for k in sets}
    ctx = {k: G.Context(m, 0) for k, m in setmod.items()}
    syn = {k: Synth(m) for k, m in setmod.items()}
    names = [k for k, _ in SERVING_MIX]
    kinds = rng.choice(names, size=n_req, p=[p for _, p in SERVING_MIX])
    sizes = np.floor(np.exp(rng.uniform(np.log(2 ** lo), np.log(2 ** hi), size=n_req))).astype(np.int64)
    which = rng.choice(list(sets), size=n_req)
    reqs = []
    for r, (kind, cnt, k) in enumerate(zip(kinds, sizes, which)):
        cnt = int(cnt)
        c, s, n = ctx[k], syn[k], int(k)
        s_seed = seed * 1000 + r
        truth = {}
        if kind == "evalcmp":
            x, y, _ = s.eval_cmp(cnt, s_seed)

            def outs():
                return (torch.empty((cnt, 4), dtype=torch.float64, device="cuda"),
                        torch.empty(cnt, dtype=torch.int32, device="cuda"))
            o_s, o_p = outs(), outs()
            tasks = [G.make_task("eval", cnt, a=x, out0=o_s[0]), G.make_task("cmp", cnt, a=x, b=y, out0=o_s[1])]

            def plain(c=c, x=x, y=y, o=o_p):
                c.eval(x, out=o[0])
                c.cmp(x, y, out=o[1])
            keep = (x, y)
        elif kind == "max":
            x, p, V = s.max_planted(cnt, s_seed)
            truth = {"index": p, "value": V}
            o_s = (torch.empty(n, dtype=torch.int32, device="cuda"), torch.empty(1, dtype=torch.int64, device="cuda"))
            o_p = (torch.empty(n, dtype=torch.int32, device="cuda"), torch.empty(1, dtype=torch.int64, device="cuda"))
            tasks = [G.make_task("max", cnt, a=x, out0=o_s[0], out1=o_s[1])]
            ws = torch.empty(c.max_workspace_size(cnt), dtype=torch.uint8, device="cuda")

            def plain(c=c, x=x, o=o_p, ws=ws):
                c.max(x, workspace=ws, out=o[0], out_idx=o[1])
            keep = (x, ws)
        elif kind == "div":
            L = product(setmod[k]).bit_length() - 1
            x, d, _, _ = s.div(cnt, seed=s_seed, max_bits=L - 1, max_shift=L)
            o_s = (torch.empty_like(x), torch.empty_like(x))
            o_p = (torch.empty_like(x), torch.empty_like(x))
            tasks = [G.make_task("div_fast", cnt, a=x, b=d, out0=o_s[0], out1=o_s[1])]

            def plain(c=c, x=x, d=d, o=o_p):
                c.div(x, d, fast=True, q=o[0], r=o[1])
            keep = (x, d)
        elif kind == "scale":
            x = s.uniform(cnt, s_seed, 8)
            D = s.uniform_u32(cnt, 1, 1900, s_seed, 9)
            o_s, o_p = (torch.empty_like(x),), (torch.empty_like(x),)
            tasks = [G.make_task("scale2pow", cnt, a=x, out0=o_s[0], D=D)]

            def plain(c=c, x=x, D=D, o=o_p):
                c.scale2pow(x, D, y=o[0])
            keep = (x, D)
        else:
            L = product(setmod[k]).bit_length() - 1
            bits = torch.full((cnt,), min(480, (L - 4) // 2), dtype=torch.int32, device="cuda")

            def operand(st, c=c, s=s, bits=bits, s_seed=s_seed, cnt=cnt):
                dg = s.bits(bits, s_seed, st, top_set=False)
                sg = s.uniform_u32(cnt, 0, 1, s_seed, st + 100)
                return dg, sg, c.eval(dg).view(cnt, 2, 2)
            X, Y = operand(10), operand(20)
            o_s = (torch.empty_like(X[0]), torch.empty_like(X[1]), torch.empty_like(X[2]))
            o_p = (torch.empty_like(X[0]), torch.empty_like(X[1]), torch.empty_like(X[2]))
            tasks = [G.make_task("mpint_mul", cnt, mx=X, my=Y, mz=o_s)]

            def plain(c=c, X=X, Y=Y, o=o_p, cnt=cnt):
                c.mpint("mul", X, Y, o, cnt)
            keep = (X, Y)
        reqs.append({"ctx": c, "tasks": G.Server.task_array(tasks), "task_list": tasks, "plain": plain,
                     "out_s": o_s, "out_p": o_p, "keep": keep, "kind": str(kind), "n": n, "count": cnt,
                     "truth": truth, "moduli": setmod[k]})
    torch.cuda.synchronize()
    return reqs, ctx
