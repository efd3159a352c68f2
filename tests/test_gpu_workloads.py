"""The batched and serving workloads at their configured shapes, checked against the oracle (the CPU
restatement) as well as against the per-call entry points:

* batched mpint (32 moduli, 4096 vectors of log-uniform length in [1, 65536], about 2.5e7 elements,
  magnitudes below 2^480 with random signs, add/sub/mul drawn per vector) in one persistent launch,
  tuned and replayed from a hipGraph: every element equal to the per-call launches, and a seeded
  sample of 2^16 elements spread over every vector equal to the oracle's mpint_add/sub/mul (digits,
  sign, evaluation bytes, overflow flags);
* serving (Poisson arrivals; 30% eval+cmp, 20% arg-max, 20% div_fast, 10% scale2pow, 20% mpint
  mul; sizes log-uniform in [2^10, 2^22]; moduli sets of 8, 32, 64 primes): every served output
  equal to the same request run alone through the per-call path, and a sample of every request's
  elements equal to the oracle (arg-max: the planted maximum's index and residues).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return "cuda"


def _np(t):
    return t.cpu().numpy()


def _bits_equal(a, b):
    """Byte equality of two tensors (float64 eval records compared as their bit patterns)."""
    return torch.equal(a.contiguous().view(torch.uint8), b.contiguous().view(torch.uint8))


def _sample_positions(lens, offs, target, seed):
    """About `target` element positions, at least one in every vector, the rest in proportion to
    the vectors' lengths."""
    rng = np.random.default_rng(seed)
    total = int(lens.sum())
    pos = []
    for ln, off in zip(lens, offs):
        k = int(min(ln, max(1, round(target * ln / total))))
        pos.append(off + rng.choice(ln, size=k, replace=False))
    return np.concatenate(pos)


def _orc_mpint(orc, op, X, Y, pos):
    """The oracle's result of op on the elements at `pos` of the device operands X, Y."""
    from oracle.oracle import ER_DTYPE
    ix = torch.from_numpy(pos).to(X[0].device)

    def side(T):
        d = _np(T[0].index_select(1, ix))
        s = _np(T[1].index_select(0, ix))
        e = _np(T[2].index_select(0, ix).contiguous()).view(ER_DTYPE).reshape(len(pos), 2)
        return d, s, e
    return orc.mpint_array(op, *side(X), *side(Y))


def _compare_mpint(got, want, pos, what):
    from oracle.oracle import ER_DTYPE
    Zd, Zs, Ze, Zo = got
    ix = torch.from_numpy(pos).to(Zd.device)
    rd, rs, re, ro = want
    assert (_np(Zd.index_select(1, ix)) == rd).all(), what + " digits"
    assert (_np(Zs.index_select(0, ix)) == rs).all(), what + " sign"
    ge = _np(Zs.new_empty(0)) if Ze is None else _np(Ze.index_select(0, ix).contiguous()).view(ER_DTYPE)
    assert ge.reshape(-1).tobytes() == re.reshape(-1).tobytes(), what + " eval"
    if Zo is not None:
        assert (_np(Zo.index_select(0, ix)) == ro).all(), what + " overflow"


def test_batched_mpint_full_workload(dev):
    import grns_amd as G
    from grns_amd.workload import MPINT_OPS, Synth, mpint_batch, mpint_view
    from oracle.oracle import Oracle
    mod = G.largest_primes_below(32)
    ctx = G.Context(mod, 0)
    orc = Oracle(mod)
    W = mpint_batch(ctx, Synth(mod), vectors=4096, max_len=65536, seed=4)
    X, Y, lens, offs, ops, total = W["X"], W["Y"], W["lens"], W["offs"], W["ops"], W["total"]
    print(f"batched mpint workload: {len(lens)} vectors, {total} elements")
    Zb = (torch.empty_like(X[0]), torch.empty_like(X[1]), torch.empty_like(X[2]))
    # zero-filled outputs: the padding between vectors belongs to no task and stays 0 in both forms
    Zp = (torch.zeros_like(X[0]), torch.zeros_like(X[1]), torch.zeros_like(X[2]))
    ob = torch.full((W["span"],), 7, dtype=torch.uint8, device=dev)
    op_ = torch.zeros((W["span"],), dtype=torch.uint8, device=dev)
    tasks = [G.make_task(MPINT_OPS[o], int(ln), mx=mpint_view(X, f, ln), my=mpint_view(Y, f, ln),
                         mz=mpint_view(Zb, f, ln), overflow=ob[f:f + ln])
             for o, f, ln in zip(ops, offs, lens)]
    batch = G.Batch(ctx, tasks)
    cfg, ms = batch.tune(reps=1)
    batch.set_config(use_graph=True)
    for o, f, ln in zip(ops, offs, lens):
        ctx.mpint(MPINT_OPS[o][6:], mpint_view(X, f, ln), mpint_view(Y, f, ln), mpint_view(Zp, f, ln), int(ln),
                  overflow=op_[f:f + ln])
    for rep in range(2):  # the graph replayed twice, outputs cleared in between
        for t in (*Zb, ob):
            t.zero_()
        batch.launch()
        torch.cuda.synchronize()
        for a, b in zip((*Zb, ob), (*Zp, op_)):
            assert _bits_equal(a, b), f"replay {rep}: batched differs from per-call"
    # the oracle on a sample of 2^16 elements spread over every vector, per operation
    pos = _sample_positions(lens, offs, 1 << 16, seed=5)
    vec_of = np.searchsorted(offs, pos, side="right") - 1
    for o in range(3):
        p = pos[ops[vec_of] == o]
        if len(p) == 0:
            continue
        want = _orc_mpint(orc, MPINT_OPS[o][6:], X, Y, p)
        _compare_mpint((Zb[0], Zb[1], Zb[2], ob), want, p, MPINT_OPS[o])
    # the operands' evaluation fields (ctx.eval) are the oracle's too
    ix = torch.from_numpy(pos[:4096]).to(dev)
    ev = orc.eval_array(_np(X[0].index_select(1, ix)))
    assert _np(X[2].index_select(0, ix).contiguous()).tobytes() == ev.tobytes()
    assert ctx.status() == 0
    batch.close()
    ctx.close()
    orc.close()


def _check_request_against_oracle(q, orcs, rng):
    """A sample of the request's served outputs against the oracle (arg-max: the planted truth)."""
    from grns_amd.moduli import product
    from tests.helpers import from_res
    orc = orcs[q["n"]]
    mod = q["moduli"]
    cnt = q["count"]
    kind = q["kind"]
    inputs = q["keep"]
    k = min(cnt, {"evalcmp": 2048, "scale": 1024, "mpint": 1024, "div": 256 if q["n"] < 64 else 96}.get(kind, 0))
    pos = np.sort(rng.choice(cnt, size=k, replace=False)) if k else None
    ix = torch.from_numpy(pos).to("cuda") if k else None
    if kind == "max":
        r, i = q["out_s"]
        assert int(i.item()) == q["truth"]["index"]
        assert _np(r).tolist() == [q["truth"]["value"] % m for m in mod]
    elif kind == "evalcmp":
        x, y = inputs
        xs, ys = _np(x.index_select(1, ix)), _np(y.index_select(1, ix))
        ev, sg = q["out_s"]
        assert _np(ev.index_select(0, ix)).tobytes() == orc.eval_array(xs).tobytes(), "eval"
        assert (_np(sg.index_select(0, ix)) == orc.cmp_array(xs, ys)).all(), "cmp"
    elif kind == "div":
        x, d = inputs
        xs, ds = _np(x.index_select(1, ix)), _np(d.index_select(1, ix))
        qq, rr = (_np(t.index_select(1, ix)) for t in q["out_s"])
        oq, orr = orc.div_array(xs, ds, fast=True)
        same = (qq == oq).all(axis=0) & (rr == orr).all(axis=0)
        # where the reference's loop wraps int32 its result is not the quotient; the exact one is returned
        for j in np.nonzero(~same)[0]:
            X, D = from_res(xs[:, [j]], mod)[0], from_res(ds[:, [j]], mod)[0]
            Q, R = divmod(X, D)
            assert from_res(qq[:, [j]], mod)[0] == Q and from_res(rr[:, [j]], mod)[0] == R, "div"
            assert (from_res(oq[:, [j]], mod)[0], from_res(orr[:, [j]], mod)[0]) != (Q, R)
    elif kind == "scale":
        x, D = inputs
        xs, Ds = _np(x.index_select(1, ix)), _np(D.index_select(0, ix)).astype(np.uint32)
        (y,) = q["out_s"]
        assert (_np(y.index_select(1, ix)) == orc.scale2pow_array(xs, Ds)).all(), "scale2pow"
    else:
        X, Y = inputs
        want = _orc_mpint(orc, "mul", X, Y, pos)
        Z = q["out_s"]
        _compare_mpint((Z[0], Z[1], Z[2], None), want, pos, "mpint_mul")
    del product


def test_serving_mixed_stream(dev):
    import time

    import grns_amd as G
    from grns_amd.workload import serving_requests
    from oracle.oracle import Oracle
    reqs, ctxs = serving_requests(24, seed=17, lo=10, hi=22)
    orcs = {n: Oracle(c.moduli) for n, c in ctxs.items()}
    print("serving requests:", [(q["kind"], q["n"], q["count"]) for q in reqs])
    for q in reqs:  # each request alone, one after another
        q["plain"]()
    torch.cuda.synchronize()
    S = G.Server(0, hbm_budget=8 << 30, slots=4, max_batch=16)
    rng = np.random.default_rng(18)
    gaps = rng.exponential(0.002, size=len(reqs))  # Poisson arrivals, 500 requests/s offered
    tickets = []
    for q, g in zip(reqs, gaps):
        time.sleep(g)
        tickets.append(S.submit(q["ctx"], q["tasks"]))
    lats = [S.wait(t) for t in tickets]
    st = S.stats()
    S.close()
    assert st["completed"] == len(reqs) and st["rejected"] == 0 and all(v > 0 for v in lats)
    print(f"served p50 {st['p50_ms']:.3f} ms, p99 {st['p99_ms']:.3f} ms, {st['launches']} launches")
    for q in reqs:
        for a, b in zip(q["out_s"], q["out_p"]):
            assert _bits_equal(a, b), f"{q['kind']} n={q['n']} count={q['count']}: served differs from alone"
    for q in reqs:
        _check_request_against_oracle(q, orcs, rng)
    for c in ctxs.values():
        assert c.status() == 0
