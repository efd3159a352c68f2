"""Serving form: requests of mixed operations and moduli sets, submitted while others run, give
the same results as the same requests run alone through the per-call entry points; admission
by HBM budget queues what does not fit and rejects what never can."""
import random
import time

import numpy as np
import pytest

from tests.helpers import SETS, cmp_pairs, div_inputs, eval_inputs, to_res

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return "cuda"


def _d(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


class Req:
    """One request: inputs, its own outputs, the task list, and the expected bytes."""

    def __init__(self, kind, ctx, o, mod, rnd, seed):
        import grns_amd as G
        self.kind, self.ctx = kind, ctx
        self.keep = []  # the request's device inputs: the server reads them until the request completes
        cnt = rnd.choice([1, 17, 256, 1000, 3000])
        if kind == "evalcmp":
            xs, ys = cmp_pairs(mod, cnt, seed)
            xn, yn = to_res(xs, mod), to_res(ys, mod)
            x, y = _d(xn), _d(yn)
            self.oracle = [o.eval_array(xn).tobytes(), o.cmp_array(xn, yn).astype(np.int32).tobytes()]
            self.keep += [x, y]
            ev = torch.zeros((cnt, 4), dtype=torch.float64, device="cuda")
            sg = torch.zeros(cnt, dtype=torch.int32, device="cuda")
            self.tasks = [G.make_task("eval", cnt, a=x, out0=ev), G.make_task("cmp", cnt, a=x, b=y, out0=sg)]
            self.outs = [ev, sg]
            self.expect = [ctx.eval(x), ctx.cmp(x, y)]
        elif kind == "max":
            vals = eval_inputs(mod, cnt, seed)
            x = _d(to_res(vals, mod))
            best = max(range(cnt), key=lambda j: (vals[j], -j))  # the lowest index among equal maxima
            self.oracle = [np.array([vals[best] % m for m in mod], np.int32).tobytes(),
                           np.array([best], np.int64).tobytes()]
            self.keep.append(x)
            r = torch.zeros(len(mod), dtype=torch.int32, device="cuda")
            i = torch.zeros(1, dtype=torch.int64, device="cuda")
            self.tasks = [G.make_task("max", cnt, a=x, out0=r, out1=i)]
            self.outs = [r, i]
            self.expect = list(ctx.max(x))
        elif kind == "div":
            k = max(1, cnt // {4: 1, 8: 20, 32: 100, 64: 300}[len(mod)])
            xs, ds = div_inputs(mod, k, seed)
            xn, dn = to_res(xs, mod), to_res(ds, mod)
            x, d = _d(xn), _d(dn)
            self.oracle = [a.tobytes() for a in o.div_array(xn, dn, fast=True)]
            self.keep += [x, d]
            q, r = torch.zeros_like(x), torch.zeros_like(x)
            self.tasks = [G.make_task("div_fast", x.shape[1], a=x, b=d, out0=q, out1=r)]
            self.outs = [q, r]
            self.expect = list(ctx.div(x, d, fast=True))
        elif kind == "scale":
            vals = eval_inputs(mod, cnt, seed)
            xn = to_res(vals, mod)
            Dn = np.array([rnd.randrange(0, 80) for _ in vals], np.int32)
            x, D = _d(xn), _d(Dn)
            self.oracle = [o.scale2pow_array(xn, Dn.astype(np.uint32)).tobytes()]
            self.keep += [x, D]
            y = torch.zeros_like(x)
            self.tasks = [G.make_task("scale2pow", cnt, a=x, out0=y, D=D)]
            self.outs = [y]
            self.expect = [ctx.scale2pow(x, D)]
        else:  # mpint mul
            mags = [rnd.randrange(1 << rnd.randrange(1, (o.log2M - 4) // 2)) for _ in range(cnt)]
            sg = np.array([rnd.randrange(2) if m else 0 for m in mags], np.int32)
            dg = to_res(mags, mod)
            ev = o.eval_array(dg)
            e = np.zeros((cnt, 2), dtype=ev["low"].dtype)
            e[:, 0], e[:, 1] = ev["low"], ev["upp"]
            X = (_d(dg), _d(sg), _d(e.view(np.float64).reshape(cnt, 2, 2)))
            rd, rs, re, _ = o.mpint_array("mul", dg, sg, e, dg, sg, e)
            self.oracle = [rd.tobytes(), rs.tobytes(), re.tobytes()]
            self.keep += list(X)
            Z = (torch.zeros_like(X[0]), torch.zeros_like(X[1]), torch.zeros_like(X[2]))
            self.tasks = [G.make_task("mpint_mul", cnt, mx=X, my=X, mz=Z)]
            self.outs = list(Z)
            E = (torch.empty_like(X[0]), torch.empty_like(X[1]), torch.empty_like(X[2]))
            ctx.mpint("mul", X, X, E, cnt)
            self.expect = list(E)
        torch.cuda.synchronize()

    def check(self):
        """Served outputs == the request run alone through the per-call path == the oracle."""
        for a, b, w in zip(self.outs, self.expect, self.oracle):
            got = a.cpu().numpy().tobytes()
            assert got == b.cpu().numpy().tobytes(), self.kind
            assert got == w, self.kind + " (oracle)"


def _requests(n_req, seed):
    import grns_amd as G
    from oracle.oracle import Oracle
    rnd = random.Random(seed)
    ctxs = {n: (G.Context(SETS[n], 0), Oracle(SETS[n]), SETS[n]) for n in (8, 32, 64)}
    reqs = []
    kinds = ["evalcmp"] * 3 + ["max"] * 2 + ["div"] * 2 + ["scale"] + ["mpint"] * 2
    for k in range(n_req):
        n = rnd.choice([8, 32, 64])
        c, o, mod = ctxs[n]
        reqs.append(Req(rnd.choice(kinds), c, o, mod, rnd, seed + k))
    return reqs, ctxs


def test_serve_alone_equals_mixed(dev):
    import grns_amd as G
    reqs, ctxs = _requests(60, seed=11)
    S = G.Server(0, hbm_budget=64 << 20, slots=4, max_batch=16)
    rnd = random.Random(3)
    tickets = []
    for r in reqs:
        tickets.append(S.submit(r.ctx, r.tasks))
        if rnd.random() < 0.3:
            time.sleep(rnd.random() * 0.002)  # arrivals while earlier requests run
    lats = [S.wait(t) for t in tickets]
    S.drain()
    for r in reqs:
        r.check()
    st = S.stats()
    assert st["completed"] == len(reqs) and st["rejected"] == 0
    assert st["launches"] < len(reqs)  # requests were fused
    assert all(v > 0 for v in lats) and st["p99_ms"] >= st["p50_ms"] > 0
    for c, o, _ in ctxs.values():
        assert c.status() == 0
    S.close()


def test_serve_admission(dev):
    import grns_amd as G
    reqs, ctxs = _requests(16, seed=21)
    S = G.Server(0, hbm_budget=1 << 20, slots=2, max_batch=4)
    # a request that can never fit the budget is refused at submission (nothing is launched)
    from grns_amd import GrnsError
    dummy_x = torch.zeros((8, 1), dtype=torch.int32, device="cuda")
    dummy_o = torch.zeros((1, 4), dtype=torch.float64, device="cuda")
    huge = G.make_task("eval", 1 << 22, a=dummy_x, out0=dummy_o)
    with pytest.raises(GrnsError) as ei:
        S.submit(ctxs[8][0], [huge])
    assert ei.value.code == 5
    # requests below the budget queue behind it and all complete (<= 1040 bytes per number)
    small = [r for r in reqs if sum(int(t.count) for t in r.tasks) * 1040 <= (1 << 20)]
    assert len(small) >= 4
    tickets = [S.submit(r.ctx, r.tasks) for r in small]
    for t in tickets:
        S.wait(t)
    for r in small:
        r.check()
    st = S.stats()
    assert st["rejected"] == 1 and st["completed"] == len(small)
    S.close()
