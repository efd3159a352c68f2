"""ctypes binding of the CPU restatement (oracle/liboracle.so).  TEST INFRASTRUCTURE ONLY:
imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the checker,
never by the product package.

Residue arrays are numpy int32 in modulus-major layout, shape (n, count): element j,
modulus i at x[i, j].
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

MAX_MODULI = 256
THRESH = 30


class ErFloat(C.Structure):
    _fields_ = [("frac", C.c_double), ("exp", C.c_long)]


class Interval(C.Structure):
    _fields_ = [("low", ErFloat), ("upp", ErFloat)]


class OrcCtx(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("log2M", C.c_int),
        ("moduli", C.c_int * MAX_MODULI),
        ("part_inv", C.c_int * MAX_MODULI),
        ("recip_rd", C.c_double * MAX_MODULI),
        ("recip_ru", C.c_double * MAX_MODULI),
        ("pow2", C.POINTER(C.c_int)),
        ("prod_pow2_res", C.c_int * THRESH),
        ("part_pow2_res", (C.c_int * MAX_MODULI) * THRESH),
        ("pow2_inv", (C.c_int * MAX_MODULI) * THRESH),
        ("mrc_inv", C.POINTER(C.c_int)),
        ("accuracy", C.c_double),
        ("ref_factor", C.c_int),
        ("unit", Interval),
        ("inv_unit", Interval),
        ("one", C.c_int * MAX_MODULI),
    ]


# numpy view of orc_er_t / interval arrays: (frac f8, exp i8)
ER_DTYPE = np.dtype([("frac", "<f8"), ("exp", "<i8")])
INTERVAL_DTYPE = np.dtype([("low", ER_DTYPE), ("upp", ER_DTYPE)])


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


_libs = {}


def build_mutant(k: int, out_dir: str) -> str:
    """Compile the restatement with planted mistake k (-DORC_MUTANT=k, listed in grns_oracle.c) into
    out_dir; used only by the tests that show the pins fail on a wrong restatement."""
    out = os.path.join(out_dir, "liboracle_mutant%d.so" % k)
    subprocess.run(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-std=gnu11", "-shared",
                    "-DORC_MUTANT=%d" % k, "-o", out, os.path.join(_HERE, "grns_oracle.c"), "-lm"], check=True)
    return out


def lib(path: str | None = None):
    """The restatement's shared library (the default build, or another build of it at `path`)."""
    key = path or _LIB_PATH
    if key not in _libs:
        if path is None and not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(key)
        P = C.c_void_p
        L.orc_ctx_create.restype = C.POINTER(OrcCtx)
        L.orc_ctx_create.argtypes = [P, C.c_int]
        L.orc_ctx_destroy.argtypes = [C.POINTER(OrcCtx)]
        for nm in ("dadd_rd", "dadd_ru", "dsub_rd", "dsub_ru", "dmul_rd", "dmul_ru", "ddiv_rd", "ddiv_ru"):
            f = getattr(L, "orc_" + nm)
            f.restype = C.c_double
            f.argtypes = [C.c_double, C.c_double]
        L.orc_fast_scalbn.restype = C.c_double
        L.orc_fast_scalbn.argtypes = [C.c_double, C.c_int]
        for nm in ("psum_rd", "psum_ru"):
            f = getattr(L, "orc_" + nm)
            f.restype = C.c_double
            f.argtypes = [P, C.c_int]
        for nm in ("mod_add", "mod_mul"):
            f = getattr(L, "orc_" + nm)
            f.restype = C.c_int
            f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double]
        for nm in ("mod_mul3", "mod_psub"):
            f = getattr(L, "orc_" + nm)
            f.restype = C.c_int
            f.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_er_set_d.argtypes = [C.POINTER(ErFloat), C.c_double]
        L.orc_er_adjust.argtypes = [C.POINTER(ErFloat)]
        for nm in ("add_rd", "add_ru", "sub_rd", "sub_ru", "mul_rd", "mul_ru", "div_rd", "div_ru"):
            getattr(L, "orc_er_" + nm).argtypes = [C.POINTER(ErFloat)] * 3
        for nm in ("md_rd", "md_ru"):
            getattr(L, "orc_er_" + nm).argtypes = [C.POINTER(ErFloat)] * 4
        for nm in ("ucmp", "cmp"):
            f = getattr(L, "orc_er_" + nm)
            f.restype = C.c_int
            f.argtypes = [C.POINTER(ErFloat)] * 2
        CP = C.POINTER(OrcCtx)
        L.orc_mrc.argtypes = [CP, P, P]
        L.orc_mrc_compare.restype = C.c_int
        L.orc_mrc_compare.argtypes = [CP, P, P]
        L.orc_eval.argtypes = [CP, P, P]
        L.orc_eval_fast.argtypes = [CP, P, P]
        L.orc_cmp.restype = C.c_int
        L.orc_cmp.argtypes = [CP, P, P]
        for nm in ("div", "div_fast"):
            f = getattr(L, "orc_" + nm)
            f.restype = C.c_int
            f.argtypes = [CP, P, P, P, P]
        L.orc_scale2pow.argtypes = [CP, P, P, C.c_uint]
        L.orc_check_overflow.restype = C.c_int
        L.orc_check_overflow.argtypes = [CP, P, P]
        for nm in ("mpint_add", "mpint_sub", "mpint_mul"):
            f = getattr(L, "orc_" + nm)
            f.restype = C.c_int
            f.argtypes = [CP, P, P, P, P, C.c_int, P, P, C.c_int, P]
        L.orc_eval_array.argtypes = [CP, P, C.c_size_t, C.c_size_t, P, C.c_int, C.c_int]
        L.orc_cmp_array.argtypes = [CP, P, P, C.c_size_t, C.c_size_t, P, C.c_int]
        L.orc_div_array.argtypes = [CP, P, P, C.c_size_t, C.c_size_t, P, P, C.c_int, C.c_int]
        L.orc_scale2pow_array.argtypes = [CP, P, P, C.c_size_t, C.c_size_t, P, C.c_int]
        L.orc_max_array.restype = C.c_long
        L.orc_max_array.argtypes = [CP, P, C.c_size_t, C.c_size_t, P, C.c_int]
        L.orc_max_eval_array.restype = C.c_long
        L.orc_max_eval_array.argtypes = [CP, P, C.c_size_t, C.c_size_t, C.c_int]
        L.orc_mpint_array.argtypes = [CP, C.c_int, C.c_size_t, C.c_size_t, P, P, P, P, P, P, P, P, P, P, C.c_int]
        L.orc_mpint_div_array.argtypes = [CP, C.c_size_t, C.c_size_t, P, P, P, P, P, P, P, C.c_int]
        _libs[key] = L
    return _libs[key]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


class Oracle:
    """One moduli set of the CPU restatement."""

    def __init__(self, moduli, lib_path: str | None = None):
        self.moduli = [int(v) for v in moduli]
        self.n = len(self.moduli)
        self._L = lib(lib_path)
        arr = _i32(self.moduli)
        self._p = self._L.orc_ctx_create(_ptr(arr), self.n)
        if not self._p:
            raise ValueError("bad moduli set")
        self.c = self._p.contents

    def close(self):
        if self._p:
            self._L.orc_ctx_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- constants -------------------------------------------------------------------
    @property
    def log2M(self) -> int:
        return self.c.log2M

    def pow2_table(self) -> np.ndarray:
        n = self.n
        return np.ctypeslib.as_array(self.c.pow2, shape=((self.c.log2M + 1) * n,)).reshape(-1, n).copy()

    def mrc_inv(self) -> np.ndarray:
        n = self.n
        return np.ctypeslib.as_array(self.c.mrc_inv, shape=(n * n,)).reshape(n, n).copy()

    def constants(self) -> dict:
        c, n = self.c, self.n
        return {
            "log2M": c.log2M,
            "part_inv": list(c.part_inv[:n]),
            "recip_rd": list(c.recip_rd[:n]),
            "recip_ru": list(c.recip_ru[:n]),
            "prod_pow2_res": list(c.prod_pow2_res),
            "part_pow2_res": [list(c.part_pow2_res[t][:n]) for t in range(THRESH)],
            "pow2_inv": [list(c.pow2_inv[t][:n]) for t in range(THRESH)],
            "accuracy": c.accuracy,
            "ref_factor": c.ref_factor,
            "unit": ((c.unit.low.frac, c.unit.low.exp), (c.unit.upp.frac, c.unit.upp.exp)),
            "inv_unit": ((c.inv_unit.low.frac, c.inv_unit.low.exp), (c.inv_unit.upp.frac, c.inv_unit.upp.exp)),
        }

    # --- single numbers (python lists of residues) -----------------------------------
    def mrc(self, x):
        a = _i32(x)
        out = np.zeros(self.n, np.int32)
        self._L.orc_mrc(self._p, _ptr(out), _ptr(a))
        return out

    def eval1(self, x, fast=False):
        a = _i32(x)
        out = np.zeros(1, INTERVAL_DTYPE)
        (self._L.orc_eval_fast if fast else self._L.orc_eval)(self._p, _ptr(out), _ptr(a))
        return out[0]

    def cmp1(self, x, y):
        return self._L.orc_cmp(self._p, _ptr(_i32(x)), _ptr(_i32(y)))

    def div1(self, x, d, fast=True):
        q = np.zeros(self.n, np.int32)
        r = np.zeros(self.n, np.int32)
        it = (self._L.orc_div_fast if fast else self._L.orc_div)(self._p, _ptr(q), _ptr(r), _ptr(_i32(x)), _ptr(_i32(d)))
        return q, r, it

    def scale1(self, x, D):
        out = np.zeros(self.n, np.int32)
        self._L.orc_scale2pow(self._p, _ptr(out), _ptr(_i32(x)), int(D))
        return out

    # --- arrays (modulus-major (n, count)) -------------------------------------------
    def eval_array(self, x, fast=False, nthreads=0):
        x = _i32(x)
        cnt = x.shape[1]
        out = np.zeros(cnt, INTERVAL_DTYPE)
        self._L.orc_eval_array(self._p, _ptr(x), cnt, cnt, _ptr(out), int(fast), nthreads)
        return out

    def cmp_array(self, x, y, nthreads=0):
        x, y = _i32(x), _i32(y)
        cnt = x.shape[1]
        out = np.zeros(cnt, np.int32)
        self._L.orc_cmp_array(self._p, _ptr(x), _ptr(y), cnt, cnt, _ptr(out), nthreads)
        return out

    def div_array(self, x, d, fast=True, nthreads=0):
        x, d = _i32(x), _i32(d)
        cnt = x.shape[1]
        q = np.zeros_like(x)
        r = np.zeros_like(x)
        self._L.orc_div_array(self._p, _ptr(x), _ptr(d), cnt, cnt, _ptr(q), _ptr(r), int(fast), nthreads)
        return q, r

    def scale2pow_array(self, x, D, nthreads=0):
        x = _i32(x)
        D = np.ascontiguousarray(D, dtype=np.uint32)
        cnt = x.shape[1]
        y = np.zeros_like(x)
        self._L.orc_scale2pow_array(self._p, _ptr(x), _ptr(D), cnt, cnt, _ptr(y), nthreads)
        return y

    def max_array(self, x, nthreads=0):
        x = _i32(x)
        cnt = x.shape[1]
        res = np.zeros(self.n, np.int32)
        idx = self._L.orc_max_array(self._p, _ptr(x), cnt, cnt, _ptr(res), nthreads)
        return idx, res

    def max_eval_array(self, x, nthreads=0):
        """Arg-max by the reference's evaluate-then-reduce pipeline (CPU path)."""
        x = _i32(x)
        cnt = x.shape[1]
        return self._L.orc_max_eval_array(self._p, _ptr(x), cnt, cnt, nthreads)

    def mpint_div_array(self, xd, xs, yd, ys, nthreads=0):
        """mpint_div (Euclidean) of the magnitudes xd / yd with signs: (digits, sign, eval)."""
        xd, yd = _i32(xd), _i32(yd)
        xs, ys = _i32(xs), _i32(ys)
        cnt = xd.shape[1]
        rd = np.zeros_like(xd)
        rs = np.zeros(cnt, np.int32)
        re = np.zeros((cnt, 2), ER_DTYPE)
        self._L.orc_mpint_div_array(self._p, cnt, cnt, _ptr(xd), _ptr(xs), _ptr(yd), _ptr(ys), _ptr(rd), _ptr(rs),
                                  _ptr(re), nthreads)
        return rd, rs, re

    def mpint_array(self, op, xd, xs, xe, yd, ys, ye, nthreads=0):
        """op: 'add' | 'sub' | 'mul'.  xe/ye: ER_DTYPE arrays of shape (count, 2)."""
        opc = {"add": 0, "sub": 1, "mul": 2}[op]
        xd, yd = _i32(xd), _i32(yd)
        xs, ys = _i32(xs), _i32(ys)
        xe = np.ascontiguousarray(xe, dtype=ER_DTYPE)
        ye = np.ascontiguousarray(ye, dtype=ER_DTYPE)
        cnt = xd.shape[1]
        rd = np.zeros_like(xd)
        rs = np.zeros(cnt, np.int32)
        re = np.zeros((cnt, 2), ER_DTYPE)
        ovf = np.zeros(cnt, np.uint8)
        self._L.orc_mpint_array(self._p, opc, cnt, cnt, _ptr(xd), _ptr(xs), _ptr(xe), _ptr(yd), _ptr(ys), _ptr(ye),
                              _ptr(rd), _ptr(rs), _ptr(re), _ptr(ovf), nthreads)
        return rd, rs, re, ovf
