"""grns_max_noalloc (the arg-max that allocates nothing: its candidates live in the library's
file-scope device array) against grns_max with a caller workspace, through the C ABI: two calls
in a row on each moduli set, the second on a smaller input, so that the second call reuses the
array the first one left behind.  Both must return the same index and residues, which are also
checked against the integers (a planted unique maximum; a seeded random input's largest value,
lowest index among equal ones).
"""
import ctypes as C
import random

import numpy as np
import pytest

from tests.helpers import SETS, to_res

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return "cuda"


def _p(t):
    return C.c_void_p(t.data_ptr())


def _run(G, c, x, noalloc, stream):
    L = G.lib()
    n, cnt = x.shape
    out = torch.full((n,), -7, dtype=torch.int32, device=x.device)
    oi = torch.full((1,), -7, dtype=torch.int64, device=x.device)
    sm, sn = x.stride(0), x.stride(1)
    s = C.c_void_p(stream.cuda_stream)
    if noalloc:
        rc = L.grns_max_noalloc(c._h, _p(x), sm, sn, cnt, _p(out), _p(oi), s)
    else:
        ws = torch.empty(L.grns_max_workspace_size(c._h, cnt), dtype=torch.uint8, device=x.device)
        rc = L.grns_max(c._h, _p(x), sm, sn, cnt, _p(out), _p(oi), _p(ws), ws.numel(), s)
    assert rc == 0, rc
    stream.synchronize()
    return out.cpu().numpy(), int(oi.item())


@pytest.mark.parametrize("n", [8, 32, 64])
def test_noalloc_max_equals_workspace_max(dev, n):
    import grns_amd as G
    from grns_amd.moduli import product
    from grns_amd.workload import Synth
    mod = SETS[n]
    M = product(mod)
    c = G.Context(mod, 0)
    stream = torch.cuda.current_stream()
    # call 1: a planted maximum over 2^22 numbers (the full stage-1 grid)
    syn = Synth(mod)
    x, p, V = syn.max_planted(1 << 22, seed=11 + n)
    want_res = np.array([V % m for m in mod], np.int32)
    r0, i0 = _run(G, c, x, False, stream)
    r1, i1 = _run(G, c, x, True, stream)
    assert i1 == i0 == p
    assert (r1 == r0).all() and (r1 == want_res).all()
    del x
    # call 2: a smaller input (fewer stage-1 blocks than call 1 left in the array)
    rnd = random.Random(n)
    vals = [rnd.randrange(M) for _ in range(3000)]
    vals[17] = vals[2900] = max(vals)  # two equal maxima: the lower index wins
    xs = torch.from_numpy(to_res(vals, mod)).to(dev)
    r0, i0 = _run(G, c, xs, False, stream)
    r1, i1 = _run(G, c, xs, True, stream)
    truth = max(range(len(vals)), key=lambda k: (vals[k], -k))
    assert i1 == i0 == truth
    assert (r1 == r0).all() and (r1 == np.array([vals[truth] % m for m in mod], np.int32)).all()
    c.close()


def test_noalloc_max_number_major_and_other_stream(dev):
    """The reference's number-major layout, on a stream other than torch's current one."""
    import grns_amd as G
    from grns_amd.moduli import product
    mod = SETS[32]
    M = product(mod)
    c = G.Context(mod, 0)
    rnd = random.Random(5)
    vals = [rnd.randrange(M // 2) for _ in range(50000)]
    k = rnd.randrange(len(vals))
    vals[k] = M // 2
    xm = torch.from_numpy(np.ascontiguousarray(to_res(vals, mod).T)).to(dev).t()  # strides (1, n)
    s = torch.cuda.Stream()
    r0, i0 = _run(G, c, xm, False, s)
    r1, i1 = _run(G, c, xm, True, s)
    assert i1 == i0 == k
    assert (r1 == r0).all() and (r1 == np.array([(M // 2) % m for m in mod], np.int32)).all()
    c.close()


def test_noalloc_rejects_unsupported_sets(dev):
    import grns_amd as G
    c = G.Context(SETS[4], 0)  # n = 4 is supported; n = 12 is not
    c12 = G.Context(G.largest_primes_below(12), 0)
    x = torch.zeros((12, 10), dtype=torch.int32, device=dev)
    out = torch.empty(12, dtype=torch.int32, device=dev)
    oi = torch.empty(1, dtype=torch.int64, device=dev)
    L = G.lib()
    rc = L.grns_max_noalloc(c12._h, _p(x), 10, 1, 10, _p(out), _p(oi), None)
    assert rc == G.GRNS_ERR_UNSUPPORTED
    assert L.grns_max_noalloc(c._h, None, 10, 1, 10, _p(out), _p(oi), None) == G.GRNS_ERR_ARG
    c.close()
    c12.close()
