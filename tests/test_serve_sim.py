"""Serving dispatcher on the CPU (grns_serve_create_sim: launches complete after a fixed delay and
run nothing): admission by HBM budget, fusion of pending requests, waits, drain and shutdown.
These are the failure-handling and scheduling checks that need no GPU."""
import ctypes as C
import random
import threading
import time

import pytest

import grns_amd as G
from tests.helpers import SETS


def _lib():
    L = G.lib()
    L.grns_serve_create_sim.argtypes = [C.POINTER(G.ServeConfig), C.c_int, C.POINTER(C.c_void_p)]
    L.grns_serve_submit.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(G.BatchTask), C.c_int,
                                    C.POINTER(C.c_longlong)]
    L.grns_serve_wait.argtypes = [C.c_void_p, C.c_longlong, C.POINTER(C.c_double)]
    L.grns_serve_poll.argtypes = [C.c_void_p, C.c_longlong, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.grns_serve_drain.argtypes = [C.c_void_p]
    L.grns_serve_stats_get.argtypes = [C.c_void_p, C.POINTER(G.ServeStats)]
    L.grns_serve_destroy.argtypes = [C.c_void_p]
    return L


@pytest.fixture(scope="module")
def ctxs():
    return {n: G.Context(SETS[n], device=-1) for n in (8, 32, 64)}


def _task(rnd):
    t = G.BatchTask()
    t.op = rnd.choice([0, 1, 2, 3, 5, 7, 8, 9])
    t.count = rnd.choice([1, 17, 256, 1000, 3000])
    t.a = t.b = t.out0 = t.out1 = t.D = 8  # never dereferenced in simulation
    for m in (t.mx, t.my, t.mz):
        m.digits = m.sign = m.eval = 8
        m.ld = t.count
    return t


def _server(L, budget, slots, max_batch, delay_us):
    cfg = G.ServeConfig(budget, slots, max_batch, 1024, 4)
    h = C.c_void_p()
    assert L.grns_serve_create_sim(C.byref(cfg), delay_us, C.byref(h)) == 0
    return h


def _submit(L, h, ctx, tasks):
    arr = (G.BatchTask * len(tasks))(*tasks)
    tk = C.c_longlong()
    rc = L.grns_serve_submit(h, ctx.handle, arr, len(tasks), C.byref(tk))
    return rc, tk.value


@pytest.mark.parametrize("seed", range(8))
def test_sim_all_requests_complete(ctxs, seed):
    L = _lib()
    rnd = random.Random(seed)
    h = _server(L, 64 << 20, rnd.choice([1, 2, 4]), 16, rnd.choice([0, 50, 500]))
    tickets = []
    for _ in range(60):
        n = rnd.choice([8, 32, 64])
        rc, tk = _submit(L, h, ctxs[n], [_task(rnd) for _ in range(rnd.choice([1, 2]))])
        assert rc == 0
        tickets.append(tk)
        if rnd.random() < 0.3:
            time.sleep(rnd.random() * 0.002)
    lat = C.c_double()
    for tk in tickets:
        assert L.grns_serve_wait(h, tk, C.byref(lat)) == 0
        assert lat.value >= 0
    assert L.grns_serve_drain(h) == 0
    st = G.ServeStats()
    L.grns_serve_stats_get(h, C.byref(st))
    assert st.completed == 60 and st.rejected == 0 and 1 <= st.launches <= 60
    assert st.p99_ms >= st.p50_ms >= 0
    assert L.grns_serve_destroy(h) == 0


def test_sim_admission_and_budget(ctxs):
    L = _lib()
    h = _server(L, 1 << 20, 2, 4, 200)
    big = G.BatchTask()
    big.op, big.count, big.a, big.out0 = 0, 1 << 22, 8, 8
    rc, tk = _submit(L, h, ctxs[8], [big])
    assert rc == 5  # can never fit: refused at submission
    st_, lat = C.c_int(), C.c_double()
    assert L.grns_serve_poll(h, tk, C.byref(st_), C.byref(lat)) == 5 and st_.value == 3
    # many requests that each fit, but not all at once: they queue behind the budget
    small = G.BatchTask()
    small.op, small.count, small.a, small.out0 = 0, 4000, 8, 8  # 4000 * 64 B = 256 KB at n = 8
    ts = [_submit(L, h, ctxs[8], [small])[1] for _ in range(20)]
    for t in ts:
        assert L.grns_serve_wait(h, t, C.byref(lat)) == 0
    s = G.ServeStats()
    L.grns_serve_stats_get(h, C.byref(s))
    assert s.completed == 20 and s.rejected == 1 and s.launches >= 5  # <= 4 fit the budget per launch
    assert L.grns_serve_destroy(h) == 0


def test_sim_concurrent_submitters(ctxs):
    L = _lib()
    h = _server(L, 32 << 20, 4, 8, 100)
    out, errs = [], []

    def worker(seed):
        rnd = random.Random(seed)
        try:
            for _ in range(40):
                rc, tk = _submit(L, h, ctxs[rnd.choice([8, 32, 64])], [_task(rnd)])
                assert rc == 0
                out.append(tk)
        except Exception as e:  # pragma: no cover
            errs.append(e)

    th = [threading.Thread(target=worker, args=(s,)) for s in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs
    lat = C.c_double()
    for tk in out:
        assert L.grns_serve_wait(h, tk, C.byref(lat)) == 0
    s = G.ServeStats()
    L.grns_serve_stats_get(h, C.byref(s))
    assert s.completed == 160
    assert L.grns_serve_destroy(h) == 0


def test_sim_bad_requests(ctxs):
    L = _lib()
    h = _server(L, 1 << 30, 1, 2, 0)
    bad = G.BatchTask()
    bad.op, bad.count = 42, 5
    assert _submit(L, h, ctxs[8], [bad])[0] == 1
    mx = G.BatchTask()
    mx.op, mx.count = 9, 5  # arg-max without its output pointers
    assert _submit(L, h, ctxs[8], [mx])[0] == 1
    # fused operations with a missing operand or output, and mpint arrays shorter than the task
    for op in range(0, 9):
        t = _task(random.Random(op))
        t.op, t.count = op, 100
        for m in (t.mx, t.my, t.mz):
            m.ld = 100
        assert _submit(L, h, ctxs[8], [t])[0] == 0
        for field in ("a", "b", "out0", "out1", "D"):
            u = _task(random.Random(op))
            u.op, u.count = op, 100
            for m in (u.mx, u.my, u.mz):
                m.ld = 100
            setattr(u, field, None)
            needed = {0: "a out0", 1: "a out0", 2: "a b out0", 6: "a b out0 out1", 7: "a b out0 out1",
                      8: "a D out0"}.get(op, "")
            assert _submit(L, h, ctxs[8], [u])[0] == (1 if field in needed.split() else 0), (op, field)
        if 3 <= op <= 5:
            for arr in ("mx", "my", "mz"):
                for field in ("digits", "sign", "eval"):
                    u = _task(random.Random(op))
                    u.op, u.count = op, 100
                    for m in (u.mx, u.my, u.mz):
                        m.ld = 100
                    setattr(getattr(u, arr), field, None)
                    assert _submit(L, h, ctxs[8], [u])[0] == 1, (op, arr, field)
            u = _task(random.Random(op))
            u.op, u.count = op, 100
            for m in (u.mx, u.my, u.mz):
                m.ld = 99  # shorter than the task
            assert _submit(L, h, ctxs[8], [u])[0] == 1
    L.grns_serve_drain(h)
    assert L.grns_serve_destroy(h) == 0
