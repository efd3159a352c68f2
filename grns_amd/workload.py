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
