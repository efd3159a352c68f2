"""The C++ drop-in boundary (include/grns_dropin.hpp): a program written against the GRNS API
(tests/dropin/dropin_demo.cpp: a params.h, the reference's types, rns_const_init, cuda::rns_max with
an xinterval_t buffer, the array forms of eval / cmp / div / scale2pow and of mpint add / sub / mul /
div on mp_int_t arrays) compiles with hipcc against the drop-in header (CPU), and on the GPU its
outputs equal the oracle's for the reference's default moduli set (src/params.h) and a shipped
small-moduli set (src/params/moduli-64-4096-bit/params.16m.256b.h)."""
import os
import random
import subprocess

import numpy as np
import pytest

from grns_amd.moduli import product
from tests.dropin import build_demos as B
from tests.helpers import cmp_pairs, div_inputs, eval_inputs, to_res

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_dropin_header_compiles(tmp_path):
    """The drop-in header and the reference-shaped program compile (host code, no GPU needed)."""
    d = tmp_path / "inc"
    d.mkdir()
    from grns_amd.moduli import params_header
    (d / "params.h").write_text(params_header(B.moduli_of("params_h"), note="GRNS src/params.h"))
    obj = tmp_path / "demo.o"
    cmd = [B.HIPCC, "--offload-arch=gfx950", "-O1", "-std=c++17", "-Wall", "-Werror", "-I", str(d), "-I",
           os.path.join(ROOT, "include"), "-c", os.path.join(ROOT, "tests", "dropin", "dropin_demo.cpp"), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def _mp_records(orc, values):
    """mp_int_t records (digits, sign, pad, eval[2]) of signed integers, evaluations by the oracle."""
    n = orc.n
    mags = [abs(v) for v in values]
    d = to_res(mags, orc.moduli)
    ev = orc.eval_array(d)
    eoff = (4 * n + 4 + 7) // 8 * 8
    size = eoff + 32
    rec = np.zeros((len(values), size), np.uint8)
    rec[:, :4 * n] = np.ascontiguousarray(d.T).view(np.uint8).reshape(len(values), 4 * n)
    rec[:, 4 * n:4 * n + 4] = np.array([1 if v < 0 else 0 for v in values], np.int32).view(np.uint8).reshape(-1, 4)
    e = np.zeros((len(values), 2), ev["low"].dtype)
    e[:, 0], e[:, 1] = ev["low"], ev["upp"]
    rec[:, eoff:] = e.view(np.uint8).reshape(len(values), 32)
    return rec, d, np.array([1 if v < 0 else 0 for v in values], np.int32), e


def _split_mp(raw, n, count):
    """(digits (n, count), sign, eval (count, 2) ER) from mp_int_t records."""
    from oracle.oracle import ER_DTYPE
    eoff = (4 * n + 4 + 7) // 8 * 8
    rec = raw.reshape(count, eoff + 32)
    digits = np.ascontiguousarray(rec[:, :4 * n]).view(np.int32).reshape(count, n).T
    sign = np.ascontiguousarray(rec[:, 4 * n:4 * n + 4]).view(np.int32).reshape(count)
    ev = np.ascontiguousarray(rec[:, eoff:]).view(ER_DTYPE).reshape(count, 2)
    return digits, sign, ev


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(B.SETS))
def test_dropin_program_matches_oracle(name, tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from oracle.oracle import Oracle
    exe = B.binary(name)
    assert os.path.exists(exe), f"{exe} missing: run __graft_entry__.build()"
    mod = B.moduli_of(name)
    n, M = len(mod), product(mod)
    orc = Oracle(mod)
    rnd = random.Random(11 + sorted(B.SETS).index(name))
    N = 3000
    xs, ys = (v[:N] for v in cmp_pairs(mod, N, seed=5))
    vals = eval_inputs(mod, N, seed=6)[:N]
    dx, dd = (v[:N] for v in div_inputs(mod, N, seed=7))
    x, y = to_res(vals, mod), to_res(ys, mod)
    y[:, :N // 2] = to_res(xs[:N // 2], mod)  # half the comparisons against near neighbours
    xd, d = to_res(dx, mod), to_res(dd, mod)
    # eval/cmp/max/scale on x; the division on its own dividends (the program divides x by d):
    # use the division workload's dividends as x for half of the numbers
    x[:, N // 2:] = xd[:, N // 2:]
    D = np.array([rnd.randrange(0, orc.log2M + 3) for _ in range(N)], np.uint32)
    Nm = 1000
    bits = (M.bit_length() - 6) // 2
    mxv = [rnd.choice([-1, 1]) * rnd.randrange(1, 1 << rnd.randrange(2, bits)) for _ in range(Nm)]
    myv = [rnd.choice([-1, 1]) * rnd.randrange(1, 1 << rnd.randrange(2, bits)) for _ in range(Nm)]
    MX, mxd, mxs, mxe = _mp_records(orc, mxv)
    MY, myd, mys, mye = _mp_records(orc, myv)
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(np.array([N, n, Nm], np.int32).tobytes())
        for a in (x, y, d):
            f.write(np.ascontiguousarray(a.T).tobytes())  # number-major
        f.write(D.tobytes())
        f.write(MX.tobytes())
        f.write(MY.tobytes())
    out = tmp_path / "out.bin"
    env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(ROOT, "grns_amd"))
    r = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    raw = open(out, "rb").read()
    pos = 0

    def take(nbytes):
        nonlocal pos
        b = np.frombuffer(raw[pos:pos + nbytes], np.uint8)
        pos += nbytes
        return b

    rec = (4 * n + 4 + 7) // 8 * 8 + 32
    got_max = take(4 * n).view(np.int32)
    ev = take(32 * N).tobytes()
    evf = take(32 * N).tobytes()
    sg = take(4 * N).view(np.int32)
    div = [take(4 * N * n).view(np.int32).reshape(N, n).T for _ in range(4)]
    sc = take(4 * N * n).view(np.int32).reshape(N, n).T
    mp = [take(rec * Nm) for _ in range(5)]
    assert pos == len(raw)
    # arg-max: the reference returns the residues of the largest value (first of equal ones)
    vx = [v for v in vals[:N // 2]] + list(dx[N // 2:])
    k = max(range(N), key=lambda j: (vx[j], -j))
    assert got_max.tolist() == x[:, k].tolist()
    assert ev == orc.eval_array(x).tobytes()
    assert evf == orc.eval_array(x, fast=True).tobytes()
    assert (sg == orc.cmp_array(x, y)).all()
    for (q, rr), fast in ((div[0:2], False), (div[2:4], True)):
        oq, orr = orc.div_array(x, d, fast=fast)
        assert (q == oq).all() and (rr == orr).all(), fast
    assert (sc == orc.scale2pow_array(x, D)).all()
    for got, op in zip(mp[:3], ("add", "sub", "mul")):
        rd, rs, re, _ = orc.mpint_array(op, mxd, mxs, mxe, myd, mys, mye)
        gd, gs, ge = _split_mp(got, n, Nm)
        assert (gd == rd).all() and (gs == rs).all() and ge.tobytes() == re.tobytes(), op
    rd, rs, re = orc.mpint_div_array(mxd, mxs, myd, mys)
    gd, gs, ge = _split_mp(mp[3], n, Nm)
    assert (gd == rd).all() and (gs == rs).all() and ge.tobytes() == re.tobytes(), "div"
    assert mp[4].tobytes() == mp[0].tobytes(), "in-place add"
