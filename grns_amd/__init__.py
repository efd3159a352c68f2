"""grns_amd -- Python binding of libgrns_hip.so, the MI355X (gfx950) implementation of the GRNS
data-parallel RNS hot path.

The binding calls the C ABI of include/grns.h through ctypes.  PyTorch is used only to hold
device memory and streams (torch.cuda on ROCm is HIP).  Residue arrays are int32 tensors of
shape (n, count) in modulus-major layout unless strides are given.

There is no fallback: if the shared library is missing the import fails, and every operation
runs the HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .moduli import SMALL_SET, largest_primes_below, params_header, product, workload_set  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgrns_hip.so")

GRNS_OK = 0
GRNS_ERR_ARG = 1
GRNS_ERR_UNSUPPORTED = 4
STATUS_REFINE_RANGE = 1  # GRNS_STATUS_REFINE_RANGE: a refinement factor outside RNS_POW2
STATUS_DIV_LOOP = 2      # GRNS_STATUS_DIV_LOOP: a division loop that would not terminate
STATUS_NAMES = {0: "OK", 1: "ERR_ARG", 2: "ERR_HIP", 3: "ERR_OOM", 4: "ERR_UNSUPPORTED", 5: "ERR_ADMISSION",
                6: "ERR_BUSY"}

ER_DTYPE = np.dtype([("frac", "<f8"), ("exp", "<i8")])
INTERVAL_DTYPE = np.dtype([("low", ER_DTYPE), ("upp", ER_DTYPE)])


class GrnsError(RuntimeError):
    def __init__(self, code, where):
        super().__init__(f"{where}: {STATUS_NAMES.get(code, code)}")
        self.code = code


class Er(C.Structure):
    _fields_ = [("frac", C.c_double), ("exp", C.c_longlong)]


class Interval(C.Structure):
    _fields_ = [("low", Er), ("upp", Er)]


class MpintArray(C.Structure):
    _fields_ = [("digits", C.c_void_p), ("sign", C.c_void_p), ("eval", C.c_void_p), ("ld", C.c_longlong)]


class BatchTask(C.Structure):
    _fields_ = [("op", C.c_int), ("count", C.c_size_t),
                ("a", C.c_void_p), ("a_mod_stride", C.c_longlong), ("a_num_stride", C.c_longlong),
                ("b", C.c_void_p), ("b_mod_stride", C.c_longlong), ("b_num_stride", C.c_longlong),
                ("out0", C.c_void_p), ("out1", C.c_void_p), ("D", C.c_void_p),
                ("mx", MpintArray), ("my", MpintArray), ("mz", MpintArray), ("overflow", C.c_void_p)]


class BatchConfig(C.Structure):
    _fields_ = [("tile", C.c_int), ("blocks_per_cu", C.c_int), ("threads_per_number", C.c_int),
                ("use_graph", C.c_int)]


OPS = {"eval": 0, "eval_fast": 1, "cmp": 2, "mpint_add": 3, "mpint_sub": 4, "mpint_mul": 5, "div": 6,
       "div_fast": 7, "scale2pow": 8}

_lib = None

EXPORTS = [
    "grns_version", "grns_ctx_create", "grns_ctx_destroy", "grns_ctx_n", "grns_ctx_log2m", "grns_ctx_constants",
    "grns_ctx_tables", "grns_ctx_status", "grns_eval", "grns_cmp", "grns_div", "grns_scale2pow",
    "grns_max_workspace_size", "grns_max", "grns_max_noalloc", "grns_max_mrc", "grns_mpint_add", "grns_mpint_sub", "grns_mpint_mul", "grns_mpint_div",
    "grns_mpint_aos", "grns_mpint_aos_bytes",
    "grns_batch_create", "grns_batch_set_config", "grns_batch_get_config", "grns_batch_tune", "grns_batch_launch",
    "grns_batch_destroy", "grns_serve_create", "grns_serve_submit", "grns_serve_poll", "grns_serve_wait",
    "grns_serve_drain", "grns_serve_stats_get", "grns_serve_destroy", "grns_ctx_words", "grns_from_binary",
    "grns_to_binary", "grns_mrc",
]
DEBUG_EXPORTS = [
    "grns_debug_eval_hot", "grns_debug_filter", "grns_debug_cmp_lane", "grns_debug_max_key", "grns_debug_ceil_log2",
    "grns_debug_eval_full", "grns_debug_cmp_full",
    "grns_debug_mulmod_recip_d",
    "grns_debug_mulmod_recip", "grns_debug_addmod_recip", "grns_debug_bn_err", "grns_debug_fast_scalbn",
    "grns_debug_div32",
    "grns_debug_filter_tau", "grns_debug_filter_err", "grns_debug_stage_timing", "grns_debug_stage_read",
    "grns_debug_reference_loops", "grns_debug_grid_cap", "grns_debug_max_blocks", "grns_debug_batch_per_call",
    "grns_debug_rare_staging", "grns_debug_rare_spread",
]


def lib():
    """The loaded libgrns_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m grns_amd.build` "
                          "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P, I, LL, SZ, D = C.c_void_p, C.c_int, C.c_longlong, C.c_size_t, C.c_double
    L.grns_version.restype = C.c_char_p
    L.grns_ctx_create.argtypes = [P, I, I, C.POINTER(P)]
    L.grns_ctx_destroy.argtypes = [P]
    L.grns_ctx_n.argtypes = [P]
    L.grns_ctx_log2m.argtypes = [P]
    L.grns_ctx_constants.argtypes = [P, P, P, P, P, P, P, P]
    L.grns_ctx_tables.argtypes = [P, P, P]
    L.grns_ctx_status.argtypes = [P, P]
    L.grns_eval.argtypes = [P, P, LL, LL, SZ, P, I, P]
    L.grns_cmp.argtypes = [P, P, LL, LL, P, LL, LL, SZ, P, P]
    L.grns_div.argtypes = [P, P, P, LL, LL, SZ, P, P, I, P]
    L.grns_scale2pow.argtypes = [P, P, LL, LL, P, SZ, P, P]
    L.grns_max_workspace_size.restype = SZ
    L.grns_ctx_words.argtypes = [P]
    L.grns_mrc.argtypes = [P, P, LL, LL, SZ, P, LL, LL, P]
    L.grns_from_binary.argtypes = [P, P, LL, LL, I, P, SZ, P, LL, LL, P]
    L.grns_to_binary.argtypes = [P, P, LL, LL, SZ, P, LL, LL, P]
    L.grns_max_workspace_size.argtypes = [P, SZ]
    L.grns_max.argtypes = [P, P, LL, LL, SZ, P, P, P, SZ, P]
    L.grns_max_noalloc.argtypes = [P, P, LL, LL, SZ, P, P, P]
    L.grns_mpint_aos_bytes.restype = SZ
    L.grns_mpint_aos_bytes.argtypes = [P]
    L.grns_mpint_aos.argtypes = [P, I, P, P, P, SZ, SZ, P, P]
    L.grns_max_mrc.argtypes = [P, P, LL, LL, SZ, P, P, P, SZ, P]
    for nm in ("grns_mpint_add", "grns_mpint_sub", "grns_mpint_mul", "grns_mpint_div"):
        getattr(L, nm).argtypes = [P, C.POINTER(MpintArray), C.POINTER(MpintArray), C.POINTER(MpintArray), SZ, P, P]
    L.grns_batch_create.argtypes = [P, C.POINTER(BatchTask), I, C.POINTER(P)]
    L.grns_batch_set_config.argtypes = [P, C.POINTER(BatchConfig)]
    L.grns_batch_get_config.argtypes = [P, C.POINTER(BatchConfig)]
    L.grns_batch_tune.argtypes = [P, P, I, C.POINTER(BatchConfig), C.POINTER(C.c_double)]
    L.grns_batch_launch.argtypes = [P, P]
    L.grns_batch_destroy.argtypes = [P]
    L.grns_debug_eval_hot.argtypes = [P, P, P, I]
    L.grns_debug_filter.restype = D
    L.grns_debug_filter.argtypes = [P, P, P, P]
    L.grns_debug_ceil_log2.argtypes = [D]
    L.grns_debug_cmp_lane.argtypes = [P, P, P]
    L.grns_debug_mulmod_recip_d.restype = D
    L.grns_debug_mulmod_recip_d.argtypes = [D, D, D, D]
    L.grns_debug_mulmod_recip.argtypes = [I, I, I, D]
    L.grns_debug_addmod_recip.argtypes = [I, I, I, D]
    L.grns_debug_bn_err.restype = D
    L.grns_debug_bn_err.argtypes = [D]
    L.grns_debug_fast_scalbn.restype = D
    L.grns_debug_fast_scalbn.argtypes = [D, I]
    L.grns_debug_div32.restype = C.c_uint
    L.grns_debug_div32.argtypes = [C.c_uint] * 5 + [C.POINTER(C.c_uint)] * 3
    L.grns_debug_eval_full.restype = C.c_int
    L.grns_debug_eval_full.argtypes = [P, P, P, C.c_int, P]
    L.grns_debug_cmp_full.restype = C.c_int
    L.grns_debug_cmp_full.argtypes = [P, P, P]
    L.grns_debug_max_key.restype = D
    L.grns_debug_max_key.argtypes = [P, P]
    L.grns_debug_filter_tau.restype = D
    L.grns_debug_filter_tau.argtypes = [P]
    L.grns_debug_filter_err.restype = D
    L.grns_debug_filter_err.argtypes = [P]
    L.grns_debug_stage_timing.argtypes = [P, I]
    L.grns_debug_reference_loops.argtypes = [P, I]
    L.grns_debug_grid_cap.argtypes = [P, I]
    L.grns_debug_max_blocks.argtypes = [P, I]
    L.grns_debug_rare_staging.argtypes = [P, I]
    L.grns_debug_rare_spread.argtypes = [P, I]
    L.grns_debug_batch_per_call.argtypes = [P, P]
    L.grns_debug_stage_read.argtypes = [P, I, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]
    _lib = L
    return L


def _check(code, where):
    if code != GRNS_OK:
        raise GrnsError(code, where)


def _np_ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _tptr(t):
    return C.c_void_p(t.data_ptr())


def _stream_ptr(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


def mpint_array(digits, sign, ev):
    """grns_mpint_array view of device tensors: digits (n, ld) int32, sign (count,) int32,
    eval (count, 2, 2) float64 storage holding (frac, exp-bits) pairs."""
    return MpintArray(digits.data_ptr(), sign.data_ptr(), ev.data_ptr(), digits.stride(0))


class Context:
    """One moduli set resident on a device (grns_ctx_create).  device=-1 makes a host-only
    context (constants and host test hooks only)."""

    def __init__(self, moduli, device: int = 0):
        self.moduli = [int(v) for v in moduli]
        self.n = len(self.moduli)
        self.device = device
        arr = np.ascontiguousarray(self.moduli, dtype=np.int32)
        h = C.c_void_p()
        _check(lib().grns_ctx_create(_np_ptr(arr), self.n, device, C.byref(h)), "grns_ctx_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().grns_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ----------------------------------------------------------------- constants
    @property
    def log2M(self) -> int:
        return lib().grns_ctx_log2m(self._h)

    def constants(self) -> dict:
        n = self.n
        pi = np.zeros(n, np.int32)
        rd = np.zeros(n, np.float64)
        ru = np.zeros(n, np.float64)
        acc = C.c_double()
        rf = C.c_int()
        unit = Interval()
        inv = Interval()
        _check(lib().grns_ctx_constants(self._h, _np_ptr(pi), _np_ptr(rd), _np_ptr(ru), C.byref(acc), C.byref(rf),
                                        C.byref(unit), C.byref(inv)), "grns_ctx_constants")
        return {
            "log2M": self.log2M, "part_inv": pi.tolist(), "recip_rd": rd.tolist(), "recip_ru": ru.tolist(),
            "accuracy": acc.value, "ref_factor": rf.value,
            "unit": ((unit.low.frac, unit.low.exp), (unit.upp.frac, unit.upp.exp)),
            "inv_unit": ((inv.low.frac, inv.low.exp), (inv.upp.frac, inv.upp.exp)),
        }

    def tables(self):
        n = self.n
        pow2 = np.zeros((self.log2M + 1) * n, np.int32)
        mrc = np.zeros(n * n, np.int32)
        _check(lib().grns_ctx_tables(self._h, _np_ptr(pow2), _np_ptr(mrc)), "grns_ctx_tables")
        return pow2.reshape(-1, n), mrc.reshape(n, n)

    def status(self) -> int:
        b = C.c_int()
        _check(lib().grns_ctx_status(self._h, C.byref(b)), "grns_ctx_status")
        return b.value

    # ----------------------------------------------------------------- device operations
    # Residue tensors: int32, shape (n, count) (modulus-major) or any 2-D view whose strides the
    # kernels can follow.  Outputs are allocated with torch on the tensor's device.
    @staticmethod
    def _strides(x):
        return x.stride(0), x.stride(1)

    def eval(self, x, fast=False, out=None, stream=None):
        import torch
        cnt = x.shape[1]
        if out is None:
            out = torch.empty((cnt, 4), dtype=torch.float64, device=x.device)
        sm, sn = self._strides(x)
        _check(lib().grns_eval(self._h, _tptr(x), sm, sn, cnt, _tptr(out), int(fast), _stream_ptr(stream)),
               "grns_eval")
        return out

    def cmp(self, x, y, out=None, stream=None):
        import torch
        cnt = x.shape[1]
        if out is None:
            out = torch.empty(cnt, dtype=torch.int32, device=x.device)
        xs, ys = self._strides(x), self._strides(y)
        _check(lib().grns_cmp(self._h, _tptr(x), xs[0], xs[1], _tptr(y), ys[0], ys[1], cnt, _tptr(out),
                              _stream_ptr(stream)), "grns_cmp")
        return out

    def div(self, x, d, fast=True, q=None, r=None, stream=None, reference=False):
        """Quotients and remainders (grns_div); reference=True runs the reference's loop for every
        pair (GRNS_DIV_REFERENCE: the CPU path's results bit for bit, wrapped ones included)."""
        import torch
        if q is None:
            q = torch.empty_like(x)
        if r is None:
            r = torch.empty_like(x)
        sm, sn = self._strides(x)
        assert self._strides(d) == (sm, sn) and self._strides(q) == (sm, sn) and self._strides(r) == (sm, sn)
        flags = (1 if fast else 0) | (2 if reference else 0)
        _check(lib().grns_div(self._h, _tptr(x), _tptr(d), sm, sn, x.shape[1], _tptr(q), _tptr(r), flags,
                              _stream_ptr(stream)), "grns_div")
        return q, r

    def scale2pow(self, x, D, y=None, stream=None):
        import torch
        if y is None:
            y = torch.empty_like(x)
        sm, sn = self._strides(x)
        assert self._strides(y) == (sm, sn)
        _check(lib().grns_scale2pow(self._h, _tptr(x), sm, sn, _tptr(D), x.shape[1], _tptr(y),
                                    _stream_ptr(stream)), "grns_scale2pow")
        return y

    def mrc(self, x, out=None, stream=None):
        """Mixed-radix digits (n, count) of x (mrc, src/mrc.cuh:31-44)."""
        import torch
        out = torch.empty_like(x) if out is None else out
        sm, sn = self._strides(x)
        dm, dn = self._strides(out)
        _check(lib().grns_mrc(self._h, _tptr(x), sm, sn, x.shape[1], _tptr(out), dm, dn, _stream_ptr(stream)),
               "grns_mrc")
        return out

    @property
    def words(self) -> int:
        """32-bit words of M (the length grns_to_binary writes)."""
        return lib().grns_ctx_words(self._h)

    def from_binary(self, w, sign=None, out=None, stream=None):
        """Residues (n, count) of the integers whose little-endian 32-bit words are w[j, k]
        (an int32 (count, nwords) tensor holding the bit patterns); sign[j] != 0 marks negatives."""
        import torch
        count, nw = w.shape
        out = torch.empty((self.n, count), dtype=torch.int32, device=w.device) if out is None else out
        sm, sn = self._strides(out)
        _check(lib().grns_from_binary(self._h, _tptr(w), w.stride(1), w.stride(0), nw,
                                      _tptr(sign) if sign is not None else None, count, _tptr(out), sm, sn,
                                      _stream_ptr(stream)), "grns_from_binary")
        return out

    def to_binary(self, x, out=None, stream=None):
        """Words (count, words) of the CRT values of x (int32 bit patterns of the uint32 words)."""
        import torch
        count = x.shape[1]
        out = torch.empty((count, self.words), dtype=torch.int32, device=x.device) if out is None else out
        sm, sn = self._strides(x)
        _check(lib().grns_to_binary(self._h, _tptr(x), sm, sn, count, _tptr(out), out.stride(1), out.stride(0),
                                    _stream_ptr(stream)), "grns_to_binary")
        return out

    def max_workspace_size(self, count) -> int:
        return lib().grns_max_workspace_size(self._h, count)

    def max(self, x, workspace=None, out=None, out_idx=None, stream=None, mrc=False):
        """Arg-max (residues, index) of the numbers x (n, count); mrc=True runs the mixed-radix
        pipeline (grns_max_mrc), which returns the same result."""
        import torch
        cnt = x.shape[1]
        launch = stream if stream is not None else torch.cuda.current_stream()
        # temporaries belong to the launch stream: the caching allocator hands a freed block out
        # again only to work ordered after that stream's kernels
        with torch.cuda.stream(launch):
            if workspace is None:
                workspace = torch.empty(self.max_workspace_size(cnt), dtype=torch.uint8, device=x.device)
            if out is None:
                out = torch.empty(self.n, dtype=torch.int32, device=x.device)
            if out_idx is None:
                out_idx = torch.empty(1, dtype=torch.int64, device=x.device)
        sm, sn = self._strides(x)
        fn, name = (lib().grns_max_mrc, "grns_max_mrc") if mrc else (lib().grns_max, "grns_max")
        _check(fn(self._h, _tptr(x), sm, sn, cnt, _tptr(out), _tptr(out_idx), _tptr(workspace), workspace.numel(),
                  _stream_ptr(stream)), name)
        return out, out_idx

    def mpint(self, op, x, y, z, count, overflow=None, stream=None):
        """op in {'add','sub','mul','div'}; x, y, z: (digits, sign, eval) device tensor triples."""
        fn = {"add": lib().grns_mpint_add, "sub": lib().grns_mpint_sub, "mul": lib().grns_mpint_mul,
              "div": lib().grns_mpint_div}[op]
        xa, ya, za = mpint_array(*x), mpint_array(*y), mpint_array(*z)
        _check(fn(self._h, C.byref(xa), C.byref(ya), C.byref(za), count,
                  _tptr(overflow) if overflow is not None else None, _stream_ptr(stream)), f"grns_mpint_{op}")

    def grid_cap(self, blocks_per_cu=0):
        """Cap the streaming kernels' grids at blocks_per_cu blocks per CU, 0 for none (test hook)."""
        _check(lib().grns_debug_grid_cap(self._h, int(blocks_per_cu)), "grns_debug_grid_cap")

    def max_blocks(self, nblocks=0):
        """Stage-1 grid of the arg-max: nblocks blocks, 0 for the default (test hook)."""
        _check(lib().grns_debug_max_blocks(self._h, int(nblocks)), "grns_debug_max_blocks")

    def rare_staging(self, on=True):
        """Stage queued numbers' residues for the rare pass (on, the default) or not (test hook)."""
        _check(lib().grns_debug_rare_staging(self._h, int(on)), "grns_debug_rare_staging")

    def rare_spread(self, on=True):
        """Spread rare items over the sub-queues (on, the default) or use one queue (test hook)."""
        _check(lib().grns_debug_rare_spread(self._h, int(on)), "grns_debug_rare_spread")

    def reference_loops(self, on=True):
        """Division and scaling through the reference's loops for every number (test hook)."""
        _check(lib().grns_debug_reference_loops(self._h, int(on)), "grns_debug_reference_loops")

    def stage_timing(self, enable=True):
        """Time the streaming kernel of eval / cmp / max with HIP events on the launch stream."""
        _check(lib().grns_debug_stage_timing(self._h, int(enable)), "grns_debug_stage_timing")

    def stage_read(self, op):
        """(summed kernel ms, launches) of op 'eval' / 'cmp' / 'max' since the last read."""
        ms, k = C.c_double(), C.c_longlong()
        _check(lib().grns_debug_stage_read(self._h, {"eval": 0, "cmp": 1, "max": 2}[op], C.byref(ms), C.byref(k)),
               "grns_debug_stage_read")
        return ms.value, k.value

    # ----------------------------------------------------------------- host test hooks
    def debug_eval_hot(self, residues, fast=False):
        a = np.ascontiguousarray(residues, dtype=np.int32)
        out = np.zeros(1, INTERVAL_DTYPE)
        ok = lib().grns_debug_eval_hot(self._h, _np_ptr(a), _np_ptr(out), int(fast))
        return ok, out[0]

    def debug_cmp_lane(self, x, y):
        a = np.ascontiguousarray(x, dtype=np.int32)
        b = np.ascontiguousarray(y, dtype=np.int32)
        return lib().grns_debug_cmp_lane(self._h, _np_ptr(a), _np_ptr(b))

    def debug_eval_full(self, residues, fast=False):
        a = np.ascontiguousarray(residues, dtype=np.int32)
        out = np.zeros(1, INTERVAL_DTYPE)
        st = C.c_int()
        ok = lib().grns_debug_eval_full(self._h, _np_ptr(a), _np_ptr(out), int(fast), C.byref(st))
        return ok, out[0], st.value

    def debug_cmp_full(self, x, y):
        a = np.ascontiguousarray(x, dtype=np.int32)
        b = np.ascontiguousarray(y, dtype=np.int32)
        return lib().grns_debug_cmp_full(self._h, _np_ptr(a), _np_ptr(b))

    def debug_max_key(self, residues):
        a = np.ascontiguousarray(residues, dtype=np.int32)
        return lib().grns_debug_max_key(self._h, _np_ptr(a))

    def debug_filter(self, residues):
        a = np.ascontiguousarray(residues, dtype=np.int32)
        rg, cn = C.c_int(), C.c_int()
        F = lib().grns_debug_filter(self._h, _np_ptr(a), C.byref(rg), C.byref(cn))
        return F, bool(rg.value), bool(cn.value)


def make_task(op, count, a=None, b=None, out0=None, out1=None, D=None, mx=None, my=None, mz=None, overflow=None):
    """grns_batch_task from device tensors.  a/b/out0/out1: (n, count)-shaped residue views
    (or the interval / sign outputs); mx/my/mz: (digits, sign, eval) triples for mpint ops.
    The task holds references to its tensors (`_keep`): Batch keeps its tasks for its lifetime and
    Server keeps a request's tasks until the request completes, so the device memory they point to
    cannot be freed while a launch may still read or write it."""
    t = BatchTask()
    t._keep = tuple(v for v in (a, b, out0, out1, D, overflow) if v is not None) + \
        tuple(v for tri in (mx, my, mz) if tri is not None for v in tri)
    t.op = OPS[op] if isinstance(op, str) else int(op)
    t.count = int(count)
    if a is not None:
        t.a, t.a_mod_stride, t.a_num_stride = a.data_ptr(), a.stride(0), a.stride(1)
    if b is not None:
        t.b, t.b_mod_stride, t.b_num_stride = b.data_ptr(), b.stride(0), b.stride(1)
    t.out0 = out0.data_ptr() if out0 is not None else None
    t.out1 = out1.data_ptr() if out1 is not None else None
    t.D = D.data_ptr() if D is not None else None
    if mx is not None:
        t.mx, t.my, t.mz = mpint_array(*mx), mpint_array(*my), mpint_array(*mz)
    t.overflow = overflow.data_ptr() if overflow is not None else None
    return t


class Batch:
    """A fixed set of tasks on fixed buffers (grns_batch_*): one persistent launch per call."""

    def __init__(self, ctx: Context, tasks):
        arr = (BatchTask * max(1, len(tasks)))(*tasks)
        h = C.c_void_p()
        _check(lib().grns_batch_create(ctx.handle, arr, len(tasks), C.byref(h)), "grns_batch_create")
        self._h = h
        self.ctx = ctx
        self.tasks = tasks

    def config(self):
        c = BatchConfig()
        _check(lib().grns_batch_get_config(self._h, C.byref(c)), "grns_batch_get_config")
        return {"tile": c.tile, "blocks_per_cu": c.blocks_per_cu, "threads_per_number": c.threads_per_number,
                "use_graph": c.use_graph}

    def set_config(self, tile=None, blocks_per_cu=None, threads_per_number=None, use_graph=None):
        cur = self.config()
        c = BatchConfig(tile if tile is not None else cur["tile"],
                        blocks_per_cu if blocks_per_cu is not None else cur["blocks_per_cu"],
                        threads_per_number if threads_per_number is not None else cur["threads_per_number"],
                        int(use_graph) if use_graph is not None else cur["use_graph"])
        _check(lib().grns_batch_set_config(self._h, C.byref(c)), "grns_batch_set_config")

    def tune(self, reps=3, stream=None):
        c = BatchConfig()
        ms = C.c_double()
        _check(lib().grns_batch_tune(self._h, _stream_ptr(stream), reps, C.byref(c), C.byref(ms)), "grns_batch_tune")
        return self.config(), ms.value

    def launch(self, stream=None):
        _check(lib().grns_batch_launch(self._h, _stream_ptr(stream)), "grns_batch_launch")

    def run_per_call(self, stream=None):
        """Every task through its per-call entry point, one call per task, looped in C++ (the
        baseline the batched launch replaces; test and benchmark hook)."""
        _check(lib().grns_debug_batch_per_call(self._h, _stream_ptr(stream)), "grns_debug_batch_per_call")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().grns_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServeConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("slots", C.c_int), ("max_batch", C.c_int), ("tile", C.c_int),
                ("blocks_per_cu", C.c_int)]


class ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong), ("rejected", C.c_longlong),
                ("launches", C.c_longlong), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("max_ms", C.c_double)]


OPS["max"] = 9


class Server:
    """Request loop on one device (grns_serve_*): continuous batching with HBM-byte admission."""

    def __init__(self, device=0, hbm_budget=8 << 30, slots=4, max_batch=64, tile=1024, blocks_per_cu=4):
        L = lib()
        P, I = C.c_void_p, C.c_int
        L.grns_serve_create.argtypes = [I, C.POINTER(ServeConfig), C.POINTER(P)]
        L.grns_serve_submit.argtypes = [P, P, C.POINTER(BatchTask), I, C.POINTER(C.c_longlong)]
        L.grns_serve_poll.argtypes = [P, C.c_longlong, C.POINTER(I), C.POINTER(C.c_double)]
        L.grns_serve_wait.argtypes = [P, C.c_longlong, C.POINTER(C.c_double)]
        L.grns_serve_drain.argtypes = [P]
        L.grns_serve_stats_get.argtypes = [P, C.POINTER(ServeStats)]
        L.grns_serve_destroy.argtypes = [P]
        cfg = ServeConfig(hbm_budget, slots, max_batch, tile, blocks_per_cu)
        h = C.c_void_p()
        self._inflight = {}  # ticket -> (task array holding the request's tensors, context)
        _check(L.grns_serve_create(device, C.byref(cfg), C.byref(h)), "grns_serve_create")
        self._h = h

    @staticmethod
    def task_array(tasks):
        """The ctypes task table of a request (build it once to resubmit a request cheaply); it holds
        the tasks' tensor references."""
        arr = (BatchTask * len(tasks))(*tasks)
        arr._keep = tuple(getattr(t, "_keep", ()) for t in tasks)
        return arr

    def submit(self, ctx: Context, tasks):
        """Queue a request (a list of make_task() tasks or a task_array()); returns its ticket.  The
        request's tensors (and its context) stay referenced until wait() or poll() sees it complete,
        or drain() or close() returns."""
        arr = tasks if isinstance(tasks, C.Array) else self.task_array(tasks)
        t = C.c_longlong()
        rc = lib().grns_serve_submit(self._h, ctx.handle, arr, len(arr), C.byref(t))
        if rc != GRNS_OK:
            raise GrnsError(rc, "grns_serve_submit")
        self._inflight[t.value] = (arr, ctx)
        return t.value

    def poll(self, ticket):
        st, lat = C.c_int(), C.c_double()
        rc = lib().grns_serve_poll(self._h, ticket, C.byref(st), C.byref(lat))
        if rc == GRNS_OK and st.value in (2, 3):  # done or failed: nothing of it runs any more
            self._inflight.pop(ticket, None)
        return st.value, lat.value, rc

    def wait(self, ticket):
        lat = C.c_double()
        rc = lib().grns_serve_wait(self._h, ticket, C.byref(lat))
        self._inflight.pop(ticket, None)  # wait returns once the request has finished, either way
        _check(rc, "grns_serve_wait")
        return lat.value

    def drain(self):
        _check(lib().grns_serve_drain(self._h), "grns_serve_drain")
        self._inflight.clear()

    def stats(self):
        s = ServeStats()
        _check(lib().grns_serve_stats_get(self._h, C.byref(s)), "grns_serve_stats_get")
        return {f: getattr(s, f) for f, _ in ServeStats._fields_}

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().grns_serve_destroy(self._h)  # drains the requests still queued or running
            self._h = C.c_void_p()
            self._inflight.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def version() -> str:
    return lib().grns_version().decode()
