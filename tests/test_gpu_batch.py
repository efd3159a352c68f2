"""Batched form: many uneven arrays and mixed operations in one persistent launch must give exactly
the per-call results, for every mapping and tile shape, also when replayed from a hipGraph."""
import random

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


def _mp_operands(o, mod, cnt, seed, bits):
    """Device (digits, sign, eval) of random signed magnitudes, and the same as numpy arrays."""
    rnd = random.Random(seed)
    mags = [rnd.randrange(1 << rnd.randrange(1, bits)) for _ in range(cnt)]
    signs = np.array([rnd.randrange(2) if m else 0 for m in mags], np.int32)
    d = to_res(mags, mod)
    ev = o.eval_array(d)
    e = np.zeros((cnt, 2), dtype=ev["low"].dtype)
    e[:, 0], e[:, 1] = ev["low"], ev["upp"]
    return (_d(d), _d(signs), _d(e.view(np.float64).reshape(cnt, 2, 2))), (d, signs, e)


def _build(n, seed):
    """Tasks with uneven lengths plus the buffers of their per-call twins."""
    import grns_amd as G
    from oracle.oracle import Oracle
    mod = SETS[n]
    ctx = G.Context(mod, 0)
    o = Oracle(mod)
    rnd = random.Random(seed)
    jobs = []  # (kind, inputs, batch outputs, the oracle's output bytes)
    for k in range(12):
        cnt = rnd.choice([1, 3, 63, 64, 65, 255, 700, 2000])
        kind = ["eval", "eval_fast", "cmp", "mpint_add", "mpint_sub", "mpint_mul"][k % 6]
        if kind.startswith("eval"):
            xn = to_res(eval_inputs(mod, cnt, seed + k), mod)
            want = o.eval_array(xn, fast=(kind == "eval_fast")).tobytes()
            jobs.append((kind, (_d(xn),), torch.zeros((cnt, 4), dtype=torch.float64, device="cuda"), want))
        elif kind == "cmp":
            xs, ys = cmp_pairs(mod, cnt, seed + k)
            a, b = to_res(xs, mod), to_res(ys, mod)
            jobs.append((kind, (_d(a), _d(b)), torch.zeros(cnt, dtype=torch.int32, device="cuda"),
                         o.cmp_array(a, b).astype(np.int32).tobytes()))
        else:
            half = (o.log2M - 2) // 2
            X, Xn = _mp_operands(o, mod, cnt, seed + 3 * k, half)
            Y, Yn = _mp_operands(o, mod, cnt, seed + 3 * k + 1, half)
            Z = (torch.zeros_like(X[0]), torch.zeros_like(X[1]), torch.zeros_like(X[2]))
            of = torch.zeros(cnt, dtype=torch.uint8, device="cuda")
            rd, rs, re, ro = o.mpint_array(kind.split("_")[1], *Xn, *Yn)
            jobs.append((kind, (X, Y), (Z, of), b"".join(a.tobytes() for a in (rd, rs, re, ro))))
    dcnt = {8: 40, 32: 12, 64: 6}[n]
    xs, ds = div_inputs(mod, dcnt, seed)
    xn, dn = to_res(xs, mod), to_res(ds, mod)
    xd, dd = _d(xn), _d(dn)
    q, r = o.div_array(xn, dn, fast=True)
    jobs.append(("div_fast", (xd, dd), (torch.zeros_like(xd), torch.zeros_like(xd)), q.tobytes() + r.tobytes()))
    vals = eval_inputs(mod, 300, seed + 99)
    xn = to_res(vals, mod)
    Dn = np.array([rnd.randrange(0, o.log2M + 3) for _ in vals], np.int32)
    jobs.append(("scale2pow", (_d(xn), _d(Dn)), torch.zeros_like(_d(xn)),
                 o.scale2pow_array(xn, Dn.astype(np.uint32)).tobytes()))
    return ctx, o, jobs


def _tasks(jobs):
    import grns_amd as G
    tasks = []
    for kind, ins, outs, _ in jobs:
        if kind.startswith("eval"):
            tasks.append(G.make_task(kind, ins[0].shape[1], a=ins[0], out0=outs))
        elif kind == "cmp":
            tasks.append(G.make_task(kind, ins[0].shape[1], a=ins[0], b=ins[1], out0=outs))
        elif kind.startswith("mpint"):
            (X, Y), (Z, of) = ins, outs
            tasks.append(G.make_task(kind, X[0].shape[1], mx=X, my=Y, mz=Z, overflow=of))
        elif kind == "div_fast":
            tasks.append(G.make_task(kind, ins[0].shape[1], a=ins[0], b=ins[1], out0=outs[0], out1=outs[1]))
        else:
            tasks.append(G.make_task(kind, ins[0].shape[1], a=ins[0], out0=outs, D=ins[1]))
    return tasks


def _plain(ctx, jobs):
    ref = []
    for kind, ins, outs, _ in jobs:
        if kind.startswith("eval"):
            ref.append(ctx.eval(ins[0], fast=(kind == "eval_fast")).cpu().numpy().tobytes())
        elif kind == "cmp":
            ref.append(ctx.cmp(ins[0], ins[1]).cpu().numpy().tobytes())
        elif kind.startswith("mpint"):
            (X, Y) = ins
            Z = (torch.empty_like(X[0]), torch.empty_like(X[1]), torch.empty_like(X[2]))
            of = torch.empty(X[0].shape[1], dtype=torch.uint8, device="cuda")
            ctx.mpint(kind.split("_")[1], X, Y, Z, X[0].shape[1], overflow=of)
            ref.append(b"".join(t.cpu().numpy().tobytes() for t in (*Z, of)))
        elif kind == "div_fast":
            q, r = ctx.div(ins[0], ins[1], fast=True)
            ref.append(q.cpu().numpy().tobytes() + r.cpu().numpy().tobytes())
        else:
            ref.append(ctx.scale2pow(ins[0], ins[1]).cpu().numpy().tobytes())
    return ref


def _got(jobs):
    out = []
    for kind, ins, outs, _ in jobs:
        if kind.startswith("mpint"):
            Z, of = outs
            out.append(b"".join(t.cpu().numpy().tobytes() for t in (*Z, of)))
        elif kind == "div_fast":
            out.append(outs[0].cpu().numpy().tobytes() + outs[1].cpu().numpy().tobytes())
        else:
            out.append(outs.cpu().numpy().tobytes())
    return out


def _zero(jobs):
    for kind, ins, outs, _ in jobs:
        if kind.startswith("mpint"):
            for t in (*outs[0], outs[1]):
                t.zero_()
        elif kind == "div_fast":
            outs[0].zero_()
            outs[1].zero_()
        else:
            outs.zero_()


@pytest.mark.parametrize("n", [8, 32, 64])
@pytest.mark.parametrize("tpn,tile,bpc", [(1, 256, 2), (1, 4096, 4), (0, 256, 4), (0, 1024, 8)])
def test_batch_matches_plain(dev, n, tpn, tile, bpc):
    import grns_amd as G
    ctx, o, jobs = _build(n, seed=1000 + n)
    ref = _plain(ctx, jobs)
    b = G.Batch(ctx, _tasks(jobs))
    b.set_config(tile=tile, blocks_per_cu=bpc, threads_per_number=tpn, use_graph=False)
    b.launch()
    torch.cuda.synchronize()
    got = _got(jobs)
    assert got == ref
    for (kind, _, _, want), g in zip(jobs, got):
        assert g == want, kind  # and the oracle's results
    assert ctx.status() == 0
    b.close()


@pytest.mark.parametrize("n", [8, 32])
def test_batch_graph_replays(dev, n):
    """Graph replays with eager work and synchronisations between them; every replay checked."""
    import grns_amd as G
    ctx, o, jobs = _build(n, seed=2000 + n)
    ref = _plain(ctx, jobs)
    b = G.Batch(ctx, _tasks(jobs))
    b.set_config(tile=512, blocks_per_cu=4, threads_per_number=1, use_graph=True)
    for rep in range(4):
        _zero(jobs)
        junk = torch.randn(1 << 16, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        b.launch()
        torch.cuda.synchronize()
        assert _got(jobs) == ref, f"replay {rep}"
        del junk
    b.close()


def test_batch_tuner_keeps_results(dev):
    import grns_amd as G
    ctx, o, jobs = _build(32, seed=3000)
    ref = _plain(ctx, jobs)
    b = G.Batch(ctx, _tasks(jobs))
    cfg, ms = b.tune(reps=1)
    assert ms > 0 and cfg["tile"] in (256, 1024, 4096, 16384)
    _zero(jobs)
    b.launch()
    torch.cuda.synchronize()
    assert _got(jobs) == ref
    b.close()


@pytest.mark.parametrize("n", [8, 64])
def test_batch_per_call_loop_matches(dev, n):
    """The C++ per-call loop over a batch's tasks (the batched form's baseline) gives the per-call
    and oracle results too."""
    import grns_amd as G
    ctx, o, jobs = _build(n, seed=4000 + n)
    ref = _plain(ctx, jobs)
    b = G.Batch(ctx, _tasks(jobs))
    _zero(jobs)
    b.run_per_call()
    torch.cuda.synchronize()
    got = _got(jobs)
    assert got == ref
    for (kind, _, _, want), g in zip(jobs, got):
        assert g == want, kind
    assert ctx.status() == 0
    b.close()
