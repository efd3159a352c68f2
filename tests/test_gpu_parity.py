"""GPU parity: every HIP entry point against the CPU restatement (oracle) on the same seeded
inputs, bit for bit (interval bounds compared as raw bytes), through the C ABI."""
import random

import numpy as np
import pytest

from tests.helpers import SETS, cmp_pairs, div_inputs, eval_inputs, from_res, to_res

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return "cuda"


@pytest.fixture(scope="module")
def ctxs(dev):
    import grns_amd as G
    from oracle.oracle import Oracle
    out = {}
    for n, mod in SETS.items():
        out[n] = (G.Context(mod, 0), Oracle(mod), mod)
    yield out
    for c, o, _ in out.values():
        c.close()
        o.close()


def _dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _intervals_bytes(t):
    return t.cpu().numpy().tobytes()


@pytest.mark.parametrize("n", [4, 8, 32, 64])
@pytest.mark.parametrize("fast", [False, True])
def test_eval_parity(ctxs, dev, n, fast):
    c, o, mod = ctxs[n]
    vals = eval_inputs(mod, 3000, seed=10 + n)
    x = to_res(vals, mod)
    got = c.eval(_dev(x, dev), fast=fast)
    ref = o.eval_array(x, fast=fast)
    g = got.cpu().numpy().view(np.uint8).reshape(len(vals), 32)
    r = ref.view(np.uint8).reshape(len(vals), 32)
    bad = np.nonzero((g != r).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} mismatches, first value {vals[bad[0]]}"
    assert c.status() == 0


@pytest.mark.parametrize("n", [8, 32])
def test_eval_number_major_layout(ctxs, dev, n):
    c, o, mod = ctxs[n]
    vals = eval_inputs(mod, 1000, seed=77)
    x = to_res(vals, mod)
    xt = _dev(x.T.copy(), dev).t()  # number-major storage viewed as (n, count)
    got = c.eval(xt)
    ref = o.eval_array(x)
    assert got.cpu().numpy().tobytes() == ref.tobytes()


@pytest.mark.parametrize("n", [4, 8, 32, 64])
def test_cmp_parity(ctxs, dev, n):
    c, o, mod = ctxs[n]
    xs, ys = cmp_pairs(mod, 4000, seed=20 + n)
    x, y = to_res(xs, mod), to_res(ys, mod)
    got = c.cmp(_dev(x, dev), _dev(y, dev)).cpu().numpy()
    ref = o.cmp_array(x, y)
    assert (got == ref).all(), f"{(got != ref).sum()} mismatches"
    truth = np.array([(a > b) - (a < b) for a, b in zip(xs, ys)], np.int32)
    # the CPU path is exact away from its reciprocal-reduction corner cases
    assert (ref == truth).mean() > 0.999


@pytest.mark.parametrize("n", [8, 32, 64])
@pytest.mark.parametrize("fast", [True, False])
def test_div_parity(ctxs, dev, n, fast):
    c, o, mod = ctxs[n]
    cnt = {8: 400, 32: 96, 64: 48}[n] if fast else {8: 200, 32: 48, 64: 24}[n]
    xs, ds = div_inputs(mod, cnt, seed=30 + n)
    x, d = to_res(xs, mod), to_res(ds, mod)
    q, r = c.div(_dev(x, dev), _dev(d, dev), fast=fast)
    rq, rr = o.div_array(x, d, fast=fast)
    assert (q.cpu().numpy() == rq).all() and (r.cpu().numpy() == rr).all()
    assert c.status() == 0


def _div_cases(mod, count, seed):
    """Dividends and divisors that exercise the word division: random sizes, divisors of one word,
    all-ones and power-of-2^32 words, quotients near word boundaries, capped quotient estimates,
    d > x, d == x, x == M - 1."""
    from grns_amd import product
    M = product(mod)
    L = M.bit_length()
    rnd = random.Random(seed)
    xs, ds = [], []
    for _ in range(count):
        x = rnd.randrange(M)
        d = rnd.getrandbits(rnd.randrange(1, L)) or 1
        xs.append(x); ds.append(min(d, M - 1))
    for _ in range(count // 4):  # x = q d + r with chosen r
        d = rnd.randrange(1, M)
        q = rnd.randrange(M // d + 1)
        xs.append(min(q * d + rnd.choice([0, d - 1, rnd.randrange(d)]), M - 1)); ds.append(d)
    for k in range(1, (L - 1) // 32 + 1):
        for v in ((1 << (32 * k)) - 1, 1 << (32 * k), (1 << (32 * k)) + 1):
            if v < M:
                xs += [v, v, M - 1, v, M - 1]
                ds += [1, (1 << 32) - 1, v, v, (1 << 31) + 1]
    xs += [0, 5, 7, 7, M - 1, M - 1, M - 1, M // 2]
    ds += [1, 7, 7, 1, 1, M - 1, 2, M // 2 + 1]
    # x = d 2^(32k) - 1 (quotient words 2^32 - 1, remainder d - 1): with three or more divisor
    # words the running remainder's top two words equal the divisor's, where the three-by-two
    # quotient estimate would reach 2^32 and is capped
    for _ in range(8):
        wd = rnd.randrange(3, max(4, (L - 1) // 32))
        d = rnd.getrandbits(32 * wd) | 1 | (1 << (32 * wd - 1 - rnd.randrange(31)))
        for k in (1, 2, 3):
            if d * (1 << (32 * k)) - 1 < M:
                xs.append(d * (1 << (32 * k)) - 1); ds.append(d)
    # quotients divisible by a modulus (the CPU path's reciprocal addition can return m_i there)
    for m in mod:
        for k in (1, 2, 12345):
            for dd in (1, 3, (1 << 40) + 7, rnd.randrange(1, max(2, M // (k * m)))):
                if k * m * dd < M:
                    xs.append(k * m * dd + rnd.randrange(dd)); ds.append(dd)
    return xs, ds


@pytest.mark.parametrize("n", [4, 8, 32, 64])
@pytest.mark.parametrize("fast", [True, False])
def test_div_exact(ctxs, dev, n, fast):
    """Quotient and remainder residues equal the CPU path's, and are congruent to those of exact
    integer division."""
    c, o, mod = ctxs[n]
    xs, ds = _div_cases(mod, 1500 if n < 64 else 600, seed=60 + n)
    x, d = to_res(xs, mod), to_res(ds, mod)
    q, r = c.div(_dev(x, dev), _dev(d, dev), fast=fast)
    qn, rn = q.cpu().numpy(), r.cpu().numpy()
    rq, rr = o.div_array(x, d, fast=fast)
    bad = np.nonzero((qn != rq).any(axis=0) | (rn != rr).any(axis=0))[0]
    assert bad.size == 0, f"{bad.size} mismatches with the CPU path, first x={xs[bad[0]]} d={ds[bad[0]]}"
    m = np.array(mod, np.int64)[:, None]
    eq = to_res([a // b for a, b in zip(xs, ds)], mod)
    er = to_res([a % b for a, b in zip(xs, ds)], mod)
    assert ((qn.astype(np.int64) % m) == eq).all() and ((rn.astype(np.int64) % m) == er).all()
    assert c.status() == 0


def _crt(col, mod, M):
    v = 0
    for r, m in zip(col, mod):
        Mi = M // m
        v += int(r) * Mi * pow(Mi, -1, m)
    return v % M


@pytest.mark.parametrize("n,count,fast", [(8, 1 << 20, True), (32, 1 << 19, True), (64, 1 << 20, True),
                                          (32, 1 << 16, False), (64, 1 << 14, False)])
def test_div_words_against_reference_loop(ctxs, dev, n, count, fast):
    """The word division against the reference's loop run on the device (same kernel code as the
    CPU restatement), on the division workload's distribution: every pair agrees, except pairs on
    which the loop itself is not the exact quotient (an int32 wrap in its reciprocal reduction)."""
    import grns_amd as G
    from grns_amd.workload import Synth
    c, o, mod = ctxs[n]
    M = G.product(mod)
    L = M.bit_length() - 1
    x, d, _, _ = Synth(mod).div(count, seed=500 + n, max_bits=min(1980, L - 1), max_shift=min(1900, L))
    q1, r1, q2, r2 = (torch.empty_like(x) for _ in range(4))
    c.div(x, d, fast=fast, q=q1, r=r1)
    c.reference_loops(True)
    try:
        c.div(x, d, fast=fast, q=q2, r=r2)
    finally:
        c.reference_loops(False)
    idx = torch.nonzero(((q1 != q2) | (r1 != r2)).any(dim=0)).flatten().cpu().numpy()
    assert idx.size <= 3
    for j in idx:
        xv, dv = _crt(x[:, j].cpu().numpy(), mod, M), _crt(d[:, j].cpu().numpy(), mod, M)
        exact = ([(xv // dv) % m for m in mod], [(xv % dv) % m for m in mod])
        assert (q1[:, j].cpu().numpy().tolist(), r1[:, j].cpu().numpy().tolist()) == exact
        assert (q2[:, j].cpu().numpy().tolist(), r2[:, j].cpu().numpy().tolist()) != exact
    assert c.status() == 0


def test_div_full_size_exact(ctxs, dev):
    """The division configuration at full size (64 moduli, 2^20 pairs, dividends uniform in [0, M),
    divisor bit lengths uniform in [1, 1980]): every pair satisfies x = q d + r modulo every m_i
    (int64 on the device); on a seeded sample of 4096 pairs, q and r are the exact quotient and
    remainder (0 <= r < d by CRT) and equal the oracle's, except pairs on which the reference's loop
    itself wraps int32 (counted and printed); with GRNS_DIV_REFERENCE the sample equals the oracle
    bit for bit, wrapped pairs included."""
    import grns_amd as G
    from grns_amd.workload import Synth
    c, o, mod = ctxs[64]
    M = G.product(mod)
    count = 1 << 20
    x, d, _, _ = Synth(mod).div(count, seed=808, max_bits=1980, max_shift=1900)
    q, r = c.div(x, d, fast=True)
    m = torch.tensor(mod, dtype=torch.int64, device=x.device).view(-1, 1)
    lhs = (q.long() * d.long() + r.long() - x.long()) % m
    assert int(lhs.abs().sum()) == 0, "x != q d + r modulo some m_i"
    rng = np.random.default_rng(809)
    pos = np.sort(rng.choice(count, size=4096, replace=False))
    ix = torch.from_numpy(pos).to(x.device)
    xs, ds = x.index_select(1, ix).cpu().numpy(), d.index_select(1, ix).cpu().numpy()
    qs, rs = q.index_select(1, ix).cpu().numpy(), r.index_select(1, ix).cpu().numpy()
    for j in range(len(pos)):
        xv, dv = _crt(xs[:, j], mod, M), _crt(ds[:, j], mod, M)
        qv, rv = _crt(qs[:, j], mod, M), _crt(rs[:, j], mod, M)
        assert (qv, rv) == divmod(xv, dv) and 0 <= rv < dv
    oq, orr = o.div_array(xs, ds, fast=True)
    differ = np.nonzero((oq != qs).any(axis=0) | (orr != rs).any(axis=0))[0]
    for j in differ:  # the oracle (the CPU path) is not the exact quotient there
        xv, dv = _crt(xs[:, j], mod, M), _crt(ds[:, j], mod, M)
        assert (_crt(oq[:, j], mod, M), _crt(orr[:, j], mod, M)) != divmod(xv, dv)
    print(f"division n=64, 2^20 pairs: sample of 4096, {len(differ)} differ from the CPU path (int32 wrap)")
    qr, rr = c.div(x.index_select(1, ix).contiguous(), d.index_select(1, ix).contiguous(), fast=True, reference=True)
    assert (qr.cpu().numpy() == oq).all() and (rr.cpu().numpy() == orr).all()
    assert c.status() == 0


def test_div_rare_queue_overflow(ctxs, dev):
    """More pairs for the reference loop than the rare queue holds (negative, non-canonical residues
    in every other pair): the overflow is marked and the rare pass still divides every pair."""
    import grns_amd as G
    from grns_amd.workload import Synth
    c, o, mod = ctxs[8]
    count = (1 << 21) + 5000
    L = G.product(mod).bit_length() - 1
    x, d, _, _ = Synth(mod).div(count, seed=77, max_bits=L - 1, max_shift=L)
    x[0, ::2] -= mod[0]
    q1, r1, q2, r2 = (torch.empty_like(x) for _ in range(4))
    c.div(x, d, fast=True, q=q1, r=r1)
    c.reference_loops(True)
    try:
        c.div(x, d, fast=True, q=q2, r=r2)
    finally:
        c.reference_loops(False)
    diff = ((q1 != q2) | (r1 != r2)).any(dim=0)
    assert not bool(diff[::2].any())  # both ran the reference loop on the non-canonical pairs
    M = G.product(mod)
    for j in torch.nonzero(diff).flatten().cpu().numpy()[:4]:  # canonical pairs: words exact
        xv, dv = _crt(x[:, j].cpu().numpy(), mod, M), _crt(d[:, j].cpu().numpy(), mod, M)
        assert q1[:, j].cpu().numpy().tolist() == [(xv // dv) % m for m in mod]
    assert int(diff.sum()) <= 4
    assert c.status() == 0


@pytest.mark.parametrize("n", [8, 32])
def test_div_noncanonical_residues(ctxs, dev, n):
    """Residues at or above their modulus take the reference loop; results equal the CPU path's."""
    c, o, mod = ctxs[n]
    xs, ds = div_inputs(mod, 40, seed=70 + n)
    x, d = to_res(xs, mod), to_res(ds, mod)
    rnd = random.Random(71)
    for j in range(0, x.shape[1], 3):
        i = rnd.randrange(len(mod))
        tgt = x if j % 2 else d
        if int(tgt[i, j]) + mod[i] < 2 ** 31:
            tgt[i, j] += mod[i]
    q, r = c.div(_dev(x, dev), _dev(d, dev), fast=True)
    rq, rr = o.div_array(x, d, fast=True)
    assert (q.cpu().numpy() == rq).all() and (r.cpu().numpy() == rr).all()
    assert c.status() == 0


@pytest.mark.parametrize("n,count", [(8, 1 << 18), (32, 1 << 16), (64, 1 << 20)])
def test_scale2pow_words_against_reference_steps(ctxs, dev, n, count):
    """The replayed word scaling equals the reference steps run on the device, for exponents from
    0 to past the range (small results make the reference's products wrap: those numbers take the
    reference steps)."""
    import grns_amd as G
    from grns_amd.workload import Synth
    c, o, mod = ctxs[n]
    L = G.product(mod).bit_length()
    syn = Synth(mod)
    x = syn.uniform(count, 900 + n, 1)
    D = syn.uniform_u32(count, 0, L + 70, 900 + n, 2)
    y1, y2 = torch.empty_like(x), torch.empty_like(x)
    c.scale2pow(x, D, y=y1)
    c.reference_loops(True)
    try:
        c.scale2pow(x, D, y=y2)
    finally:
        c.reference_loops(False)
    assert torch.equal(y1, y2)


@pytest.mark.parametrize("n", [4, 8, 32, 64])
def test_scale2pow_structured(ctxs, dev, n):
    """Scaling of values at fixed fractions of M, powers of two and word boundaries by exponents
    around the range, against the CPU path."""
    from grns_amd import product
    c, o, mod = ctxs[n]
    M = product(mod)
    L = M.bit_length()
    rnd = random.Random(95 + n)
    xs = [M * k // 100000 for k in (9104, 8450, 23150, 50000, 99900, 1, 77777)]
    xs += [(1 << k) - 1 for k in range(1, L - 1, 7)] + [1 << k for k in range(0, L - 1, 11)] + [M - 1, 0, 1]
    xs = [v for v in xs for _ in range(3)]
    Ds = [max(0, rnd.choice([0, 1, 29, 30, 31, 60, L - 30, L - 9, L - 7, L - 5, L - 2, L - 1, L, L + 3, 5000]))
          for _ in xs]
    x = to_res(xs, mod)
    D = np.array(Ds, np.uint32)
    y = c.scale2pow(_dev(x, dev), _dev(D.view(np.int32), dev)).cpu().numpy()
    ref = o.scale2pow_array(x, D)
    bad = np.nonzero((y != ref).any(axis=0))[0]
    assert bad.size == 0, f"{bad.size} mismatches, first x={xs[bad[0]]} D={Ds[bad[0]]}"


@pytest.mark.parametrize("n", [8, 32, 64])
def test_scale2pow_parity(ctxs, dev, n):
    c, o, mod = ctxs[n]
    rnd = random.Random(40 + n)
    vals = eval_inputs(mod, 600, seed=41 + n)
    L = o.log2M
    D = np.array([rnd.choice([0, 1, 29, 30, 31, 60, rnd.randrange(1, L + 40)]) for _ in vals], np.uint32)
    x = to_res(vals, mod)
    y = c.scale2pow(_dev(x, dev), _dev(D.view(np.int32), dev))
    ref = o.scale2pow_array(x, D)
    assert (y.cpu().numpy() == ref).all()


@pytest.mark.parametrize("mrc", [False, True])
@pytest.mark.parametrize("n", [4, 8, 32, 64])
@pytest.mark.parametrize("case", ["random", "planted", "equal", "small"])
def test_max_parity(ctxs, dev, n, case, mrc):
    """Arg-max by keys + exact order (grns_max) and by mixed-radix digits (grns_max_mrc): the
    index of the largest value (lowest among equal ones) and its residues, against the integers
    and the CPU path."""
    c, o, mod = ctxs[n]
    from grns_amd.moduli import product
    M = product(mod)
    rnd = random.Random(50 + n)
    if case == "small":
        vals = [rnd.randrange(M) for _ in range(3)]
    elif case == "equal":
        v = rnd.randrange(M)
        vals = [v] * 1500
    else:
        vals = [rnd.randrange(M // 2) for _ in range(20000)]
        if case == "planted":
            V = M // 2 - 1
            idx = rnd.sample(range(len(vals)), 40)
            vals[idx[0]] = V
            for k in idx[1:]:
                vals[k] = V - 1
    x = to_res(vals, mod)
    res, idx = c.max(_dev(x, dev), mrc=mrc)
    ridx, rres = o.max_array(x)
    truth = max(range(len(vals)), key=lambda k: (vals[k], -k))
    assert int(idx.item()) == truth == ridx
    assert (res.cpu().numpy() == rres).all()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [16, 32, 64])
def test_max_any_stage1_grid(ctxs, dev, n):
    """The arg-max of a planted workload (one maximum, 200 values one below it) is the same for
    every stage-1 grid: the default, small and odd grids, and one block per 256 numbers."""
    import grns_amd as G
    from oracle.oracle import Oracle
    own = n not in ctxs
    if own:
        mod = G.largest_primes_below(n)
        c, o = G.Context(mod, 0), Oracle(mod)
    else:
        c, o, mod = ctxs[n]
    M = G.product(mod)
    rnd = random.Random(90 + n)
    count = 1 << 18
    vals = [rnd.randrange(M // 2) for _ in range(count)]
    idx = rnd.sample(range(count), 201)
    vals[idx[0]] = M // 2 - 1
    for k in idx[1:]:
        vals[k] = M // 2 - 2
    x = to_res(vals, mod)
    xd = _dev(x, dev)
    ridx, rres = o.max_array(x)
    assert ridx == idx[0]
    try:
        for nb in (0, 1, 7, 1000, 2979, 1024):
            c.max_blocks(nb)
            for mrc in (False, True):
                res, got = c.max(xd, mrc=mrc)
                assert int(got.item()) == ridx, (nb, mrc)
                assert (res.cpu().numpy() == rres).all(), (nb, mrc)
    finally:
        c.max_blocks(0)
        if own:
            c.close()
            o.close()


@pytest.mark.parametrize("n", [8, 32, 64])
def test_max_noncanonical_residues(ctxs, dev, n):
    """Residues stand for their value mod m_i (some raised by m_i, some made negative): both
    arg-max pipelines return the index of the largest canonical value, lowest among equal ones,
    and that number's residues as given."""
    c, o, mod = ctxs[n]
    from grns_amd.moduli import product
    M = product(mod)
    rnd = random.Random(70 + n)
    vals = [rnd.randrange(M) for _ in range(6000)]
    top = max(vals)
    for k in rnd.sample(range(len(vals)), 5):  # ties at the maximum, the first of them wins
        vals[k] = top
    x = to_res(vals, mod).astype(np.int64)
    for j in range(0, x.shape[1], 3):
        i = rnd.randrange(len(mod))
        x[i, j] += mod[i] if x[i, j] + mod[i] < 2 ** 31 else -mod[i]
    for j in range(1, x.shape[1], 7):
        i = rnd.randrange(len(mod))
        x[i, j] -= mod[i]  # negative residue standing for the same value
    x = x.astype(np.int32)
    truth = max(range(len(vals)), key=lambda k: (vals[k], -k))
    for mrc in (False, True):
        res, idx = c.max(_dev(x, dev), mrc=mrc)
        assert int(idx.item()) == truth, mrc
        assert res.cpu().numpy().tolist() == x[:, truth].tolist(), mrc


@pytest.mark.parametrize("n", [8, 32, 64])
@pytest.mark.parametrize("case", ["ascending", "near_M", "random_small"])
def test_max_small_magnitudes(ctxs, dev, n, case):
    """2^20 numbers whose values lie within 2^-40 M of 0 or of M, where the arg-max keys carry no
    order: x_j = j (ascending), x_j = M - 1 - j (the maximum is the first), and random values of at
    most log2(M) - 41 bits with a planted maximum 2^(log2 M - 40) and 100 near-ties one below it.
    Index and residues against the integers, both pipelines; stage 2 of the key pipeline stays
    bounded (a few pivot rounds, then a tree of exact comparisons)."""
    import time
    c, o, mod = ctxs[n]
    from grns_amd.moduli import product
    from grns_amd.workload import Synth
    M = product(mod)
    count = 1 << 20
    j = np.arange(count, dtype=np.int64)
    if case == "ascending":
        x = _dev(np.stack([j % m for m in mod]).astype(np.int32), dev)
        truth, value = count - 1, count - 1
    elif case == "near_M":
        x = _dev(np.stack([(m - 1 - j) % m for m in mod]).astype(np.int32), dev)
        truth, value = 0, M - 1
    else:
        syn = Synth(mod)
        L = M.bit_length() - 1
        bits = torch.randint(1, L - 41, (count,), generator=torch.Generator().manual_seed(n), dtype=torch.int32)
        x = syn.bits(bits.to(dev), seed=7 + n, top_set=False)
        rnd = random.Random(n)
        pos = rnd.sample(range(count), 101)
        value = (1 << (L - 40)) + rnd.randrange(1 << 20)
        syn.place(x, [value] + [value - 1] * 100, pos)
        truth = pos[0]
    want = [value % m for m in mod]
    for mrc in (False, True):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res, idx = c.max(x, mrc=mrc)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        assert int(idx.item()) == truth, (case, mrc)
        assert res.cpu().numpy().tolist() == want, (case, mrc)
        print(f"max small magnitudes n={n} {case} mrc={mrc}: {dt * 1e3:.1f} ms")


@pytest.mark.parametrize("n", [8, 32])
def test_max_mrc_number_major(ctxs, dev, n):
    """Number-major input (residue i of number j at x[j * n + i]): the strided path of the mixed-
    radix arg-max equals the modulus-major one."""
    c, o, mod = ctxs[n]
    from grns_amd.moduli import product
    M = product(mod)
    rnd = random.Random(90 + n)
    vals = [rnd.randrange(M) for _ in range(5000)]
    x = to_res(vals, mod)
    xt = _dev(np.ascontiguousarray(x.T), dev).t()  # (n, count) view with strides (1, n)
    res, idx = c.max(xt, mrc=True)
    truth = max(range(len(vals)), key=lambda k: (vals[k], -k))
    assert int(idx.item()) == truth
    assert res.cpu().numpy().tolist() == x[:, truth].tolist()


def _mpint_operands(mod, o, count, seed, bits):
    rnd = random.Random(seed)
    mags = [rnd.randrange(1 << rnd.randrange(1, bits)) for _ in range(count)]
    mags[:4] = [0, 1, 55, 13]
    signs = [rnd.randrange(2) for _ in range(count)]
    signs[:4] = [0, 0, 1, 0]
    d = to_res(mags, mod)
    ev = o.eval_array(d)
    e = np.zeros((count, 2), dtype=ev["low"].dtype)
    e[:, 0] = ev["low"]
    e[:, 1] = ev["upp"]
    for j in range(count):  # mpint_set_mpz: sign 0 for zero
        if mags[j] == 0:
            signs[j] = 0
    return d, np.array(signs, np.int32), e


@pytest.mark.parametrize("n", [8, 32, 64])
@pytest.mark.parametrize("op", ["add", "sub", "mul"])
def test_mpint_parity(ctxs, dev, n, op):
    c, o, mod = ctxs[n]
    half = (o.log2M - 2) // 2
    cnt = 3000
    xd, xs, xe = _mpint_operands(mod, o, cnt, 60 + n, half)
    yd, ys, ye = _mpint_operands(mod, o, cnt, 70 + n, half)
    rd, rs, re, ovf = o.mpint_array(op, xd, xs, xe, yd, ys, ye)

    def dev_triple(d, s, e):
        return (_dev(d, dev), _dev(s, dev), _dev(e.view(np.float64).reshape(cnt, 2, 2), dev))

    X, Y = dev_triple(xd, xs, xe), dev_triple(yd, ys, ye)
    Z = (torch.empty_like(X[0]), torch.empty_like(X[1]), torch.empty_like(X[2]))
    of = torch.empty(cnt, dtype=torch.uint8, device=dev)
    c.mpint(op, X, Y, Z, cnt, overflow=of)
    assert (Z[0].cpu().numpy() == rd).all()
    assert (Z[1].cpu().numpy() == rs).all()
    assert Z[2].cpu().numpy().tobytes() == re.tobytes()
    assert (of.cpu().numpy() == ovf).all()


def _column(mod, v, count, dev):
    from grns_amd.workload import residues_of
    r = torch.tensor(np.array(residues_of(v, mod), dtype=np.int64).astype(np.int32), device=dev)
    return r.view(-1, 1).expand(-1, count).contiguous()


@pytest.mark.parametrize("n", [8])
def test_rare_queue_overflow(ctxs, dev, n):
    """More rare items than the per-stream queue holds (2^20): the overflow is marked in the
    outputs and resolved by the rare pass's scan; the next calls on the stream start clean."""
    c, o, mod = ctxs[n]
    from grns_amd.moduli import product
    M = product(mod)
    count = (1 << 20) + 4099
    x = _column(mod, M - 1, count, dev)  # x = M - 1 needs the mixed-radix path in every number
    y = _column(mod, M - 2, count, dev)
    ev = c.eval(x)
    want = o.eval_array(to_res([M - 1], mod)).tobytes()
    g = ev.cpu().numpy().view(np.uint8).reshape(count, 32)
    assert (g == np.frombuffer(want, np.uint8)).all()
    sg = c.cmp(x, y).cpu().numpy()
    assert (sg == o.cmp1(to_res([M - 1], mod)[:, 0], to_res([M - 2], mod)[:, 0])).all()
    # the counters were cleared: an ordinary call right after is exact
    xs, ys = cmp_pairs(mod, 3000, seed=61)
    xa, ya = to_res(xs, mod), to_res(ys, mod)
    assert (c.cmp(_dev(xa, dev), _dev(ya, dev)).cpu().numpy() == o.cmp_array(xa, ya)).all()
    assert _intervals_bytes(c.eval(_dev(xa, dev))) == o.eval_array(xa).tobytes()
    assert c.status() == 0


@pytest.mark.parametrize("n", [8, 32])
@pytest.mark.parametrize("staging,spread", [(True, True), (False, True), (True, False)])
def test_rare_items_distinct(ctxs, dev, n, staging, spread):
    """2^17 + 999 distinct numbers just below M in shuffled order (each needs the rare pass of
    evaluation and of comparison): every queued item is resolved for its own number (the first
    2^16 of the queue from residues staged by the streaming pass, the rest, or all of them with
    staging off, from the inputs), with the items spread over the sub-queues or all in one; every
    interval (as bytes) and sign equals the CPU path's."""
    c, o, mod = ctxs[n]
    c.rare_staging(staging)
    c.rare_spread(spread)
    from grns_amd.moduli import product
    M = product(mod)
    count = (1 << 17) + 999
    rnd = random.Random(66 + n)
    ks = list(range(count))
    rnd.shuffle(ks)
    xs = [M - 1 - k for k in ks]
    ys = [v - 1 if i % 2 else v + (1 if v + 1 < M else -2) for i, v in enumerate(xs)]
    xa, ya = to_res(xs, mod), to_res(ys, mod)
    xd, yd = _dev(xa, dev), _dev(ya, dev)
    assert _intervals_bytes(c.eval(xd)) == o.eval_array(xa).tobytes()
    assert (c.cmp(xd, yd).cpu().numpy() == o.cmp_array(xa, ya)).all()
    c.rare_staging(True)
    c.rare_spread(True)
    assert c.status() == 0


def test_first_call_on_nonblocking_stream(dev):
    """The first call of a fresh context on a fresh non-blocking stream (torch streams are) creates
    that stream's rare-item queue; its counters must be clear before the call's kernels push to
    it, or queued numbers are lost.  Scaling with non-canonical residues queues every such number."""
    import grns_amd as G
    from oracle.oracle import Oracle
    mod = SETS[8]
    o = Oracle(mod)
    vals = eval_inputs(mod, 20000, seed=64)
    x = to_res(vals, mod).astype(np.int64)
    for j in range(0, x.shape[1], 3):  # one residue raised by its modulus: the rare path
        if x[0, j] + mod[0] < 2 ** 31:
            x[0, j] += mod[0]
    x = x.astype(np.int32)
    rnd = random.Random(65)
    D = np.array([rnd.randrange(1, 200) for _ in vals], np.uint32)
    ref = o.scale2pow_array(x, D)
    xd, Dd = _dev(x, dev), _dev(D.view(np.int32), dev)
    torch.cuda.synchronize()  # the inputs are in place before another stream reads them
    for _ in range(4):
        c = G.Context(mod, 0)
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            y = c.scale2pow(xd, Dd)
        s.synchronize()
        assert (y.cpu().numpy() == ref).all()
        c.close()
    o.close()


def test_rare_queue_per_stream(ctxs, dev):
    """Calls on two streams use their own rare-item queues."""
    c, o, mod = ctxs[32]
    xs, ys = cmp_pairs(mod, 5000, seed=62)
    xa, ya = to_res(xs, mod), to_res(ys, mod)
    xd, yd = _dev(xa, dev), _dev(ya, dev)
    ref = o.cmp_array(xa, ya)
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    outs = []
    for s in (s1, s2):
        with torch.cuda.stream(s):
            outs.append(c.cmp(xd, yd))
    torch.cuda.synchronize()
    for t in outs:
        assert (t.cpu().numpy() == ref).all()


@pytest.mark.parametrize("n", [8, 32])
def test_scale2pow_large_exponents(ctxs, dev, n):
    """Exponents far above log2 M: the steps stop once the number is 0 (the reference's further
    steps map 0 to 0), so the result is 0 and the call takes O(log2 M) steps for any D."""
    c, o, mod = ctxs[n]
    rnd = random.Random(90 + n)
    vals = eval_inputs(mod, 200, seed=91 + n)
    L = o.log2M
    D_mid = np.array([rnd.randrange(L + 1, L + 3000) for _ in vals], np.uint32)
    x = to_res(vals, mod)
    y = c.scale2pow(_dev(x, dev), _dev(D_mid.view(np.int32), dev))
    assert (y.cpu().numpy() == o.scale2pow_array(x, D_mid)).all()
    D_huge = np.array([rnd.choice([1 << 31, 0xFFFFFFFF, rnd.randrange(1 << 20, 1 << 32)]) for _ in vals], np.uint32)
    y = c.scale2pow(_dev(x, dev), _dev(D_huge.view(np.int32), dev))
    assert (y.cpu().numpy() == 0).all()


def _words(v, nw):
    return [(v >> (32 * k)) & 0xFFFFFFFF for k in range(nw)]


@pytest.mark.parametrize("n", [4, 8, 32, 64])
def test_from_binary(ctxs, dev, n):
    """rns_from_binary: residues of integers of any length (beyond M and beyond a group's width),
    with signs (mpz_mod_ui's nonnegative remainder), in number-major and word-major layouts."""
    c, o, mod = ctxs[n]
    from grns_amd import product
    M = product(mod)
    rnd = random.Random(200 + n)
    nw = max(3, (M.bit_length() + 31) // 32 + 70)  # longer than M and than 64 words
    vals = [rnd.getrandbits(rnd.randrange(1, 32 * nw)) for _ in range(700)]
    vals += [0, 1, M - 1, M, M + 1, 2 ** (32 * nw) - 1, 2 ** 32, 2 ** 32 - 1] + [m * 12345 for m in mod[:3]]
    signs = [rnd.randrange(2) for _ in vals]
    W = np.array([_words(v, nw) for v in vals], dtype=np.uint64).astype(np.uint32).view(np.int32)
    S = np.array(signs, np.int32)
    got = c.from_binary(_dev(W, dev), sign=_dev(S, dev)).cpu().numpy()
    exp = to_res([-v if s else v for v, s in zip(vals, signs)], mod)
    assert (got == exp).all()
    got2 = c.from_binary(_dev(W.T.copy(), dev).t()).cpu().numpy()  # word-major storage, no signs
    assert (got2 == to_res(vals, mod)).all()


@pytest.mark.parametrize("n", [4, 8, 32, 64])
def test_to_binary(ctxs, dev, n):
    """rns_to_binary: CRT values as words, for canonical, non-canonical and negative residues
    (a negative residue counts as x_i + 2^64, the reference's conversion to unsigned long)."""
    c, o, mod = ctxs[n]
    from grns_amd import product
    M = product(mod)
    rnd = random.Random(300 + n)
    vals = [rnd.randrange(M) for _ in range(600)] + [0, 1, M - 1, M // 2]
    x = to_res(vals, mod)
    got = c.to_binary(_dev(x, dev)).cpu().numpy().view(np.uint32)
    W = c.words
    assert got.shape == (len(vals), W)
    assert [sum(int(w) << (32 * k) for k, w in enumerate(row)) for row in got] == vals
    xr = x.astype(np.int64)
    for j in range(0, x.shape[1], 3):  # raise some residues by m (same value) or make them negative
        i = rnd.randrange(len(mod))
        if xr[i, j] + mod[i] < 2 ** 31:
            xr[i, j] += mod[i]
        else:
            xr[i, j] -= mod[i]
    got = c.to_binary(_dev(xr.astype(np.int32), dev)).cpu().numpy().view(np.uint32)
    Mi = [M // m for m in mod]
    B = [Mi[i] * pow(Mi[i] % mod[i], -1, mod[i]) for i in range(len(mod))]
    exp = [sum(((int(v) + 2 ** 64) if v < 0 else int(v)) * B[i] for i, v in enumerate(xr[:, j])) % M
           for j in range(x.shape[1])]
    assert [sum(int(w) << (32 * k) for k, w in enumerate(row)) for row in got] == exp
    # round trip through the words
    back = c.from_binary(c.to_binary(_dev(x, dev))).cpu().numpy()
    assert (back == x).all()


@pytest.mark.parametrize("n", [4, 8, 32, 64])
def test_mpint_div(ctxs, dev, n):
    """mpint_div (Euclidean) against the CPU path: digits, sign and the eval bounds as bytes; the
    digits are also those of the exact Euclidean quotient (up to the CPU path's m_i for 0)."""
    c, o, mod = ctxs[n]
    from grns_amd import product
    M = product(mod)
    L = M.bit_length() - 1
    rnd = random.Random(400 + n)
    cnt = {4: 1500, 8: 800, 32: 300, 64: 150}[n]
    xs, ys = [], []
    for _ in range(cnt):
        xs.append(rnd.getrandbits(rnd.randrange(1, L)))
        ys.append(max(1, rnd.getrandbits(rnd.randrange(1, L))))
    xs += [0, 5, 7, mod[-1] * 3 + 1, mod[0] * mod[1] * 9 + 2, M - 1]
    ys += [3, 7, 5, 3, 9, 1]
    sx = [rnd.randrange(2) if v else 0 for v in xs]
    sy = [rnd.randrange(2) for _ in ys]
    xd, yd = to_res(xs, mod), to_res(ys, mod)
    cnt = len(xs)
    X = (_dev(xd, dev), _dev(np.array(sx, np.int32), dev), torch.zeros((cnt, 2, 2), dtype=torch.float64, device=dev))
    Y = (_dev(yd, dev), _dev(np.array(sy, np.int32), dev), torch.zeros((cnt, 2, 2), dtype=torch.float64, device=dev))
    Z = (torch.empty_like(X[0]), torch.empty_like(X[1]), torch.empty_like(X[2]))
    c.mpint("div", X, Y, Z, cnt)
    rd, rs, re = o.mpint_div_array(xd, np.array(sx, np.int32), yd, np.array(sy, np.int32))
    assert (Z[0].cpu().numpy() == rd).all()
    assert (Z[1].cpu().numpy() == rs).all()
    assert Z[2].cpu().numpy().tobytes() == re.tobytes()
    m = np.array(mod, np.int64)[:, None]
    q = [a // b + (1 if s and a % b else 0) for a, b, s in zip(xs, ys, sx)]
    assert ((Z[0].cpu().numpy().astype(np.int64) % m) == to_res(q, mod)).all()
    assert c.status() == 0


@pytest.mark.parametrize("n", [4, 8, 32, 64])
def test_mrc(ctxs, dev, n):
    """Mixed-radix digits against the CPU path (also for non-canonical residues) and against the
    integer they expand to."""
    from grns_amd import product
    c, o, mod = ctxs[n]
    M = product(mod)
    rnd = random.Random(500 + n)
    vals = [rnd.randrange(M) for _ in range(3000)] + [0, 1, M - 1, M // 2]
    x = to_res(vals, mod)
    for j in range(0, x.shape[1], 5):  # some residues at or above their modulus
        i = rnd.randrange(len(mod))
        if int(x[i, j]) + mod[i] < 2 ** 31:
            x[i, j] += mod[i]
    got = c.mrc(_dev(x, dev)).cpu().numpy()
    for j in range(x.shape[1]):
        assert got[:, j].tolist() == o.mrc(x[:, j]).tolist(), j
    xt = _dev(np.ascontiguousarray(x.T), dev).t()  # number-major input (and output)
    assert (c.mrc(xt).cpu().numpy() == got).all()
    canon = to_res(vals, mod)
    dg = c.mrc(_dev(canon, dev)).cpu().numpy()
    for j, v in enumerate(vals):
        acc, base = 0, 1
        for k, m in enumerate(mod):
            acc += int(dg[k, j]) * base
            base *= m
        assert acc == v


@pytest.mark.parametrize("fast", [False, True])
def test_eval_cmp_small_config_full_size(ctxs, dev, fast):
    """The CPU-runnable configuration at full size: 8 largest primes below 2^31, N = 65536 numbers
    uniform in [0, M) with 1% forced to 0, M-1, M/2 and M/2 +- 1, pairs (x, x + delta) with delta in
    {-1, 0, +1, uniform}: every interval (as bytes) and every sign equal the CPU path's."""
    from grns_amd.workload import Synth
    c, o, mod = ctxs[8]
    x, y, forced = Synth(mod).eval_cmp(65536, seed=11)
    assert len(forced) == 655
    xh, yh = x.cpu().numpy(), y.cpu().numpy()
    ev = c.eval(x, fast=fast).cpu().numpy()
    assert ev.tobytes() == o.eval_array(xh, fast=fast).tobytes()
    if not fast:
        assert (c.cmp(x, y).cpu().numpy() == o.cmp_array(xh, yh)).all()
    assert c.status() == 0


def test_max_headline_config_full_size(ctxs, dev):
    """The arg-max configuration at full size: 32 largest primes below 2^31, N = 2^24 numbers in
    [0, M/2), a planted unique maximum and 1000 near-ties one below it: index and residues of both
    arg-max pipelines (keys, mixed-radix digits) equal the planted ones and the CPU path's
    evaluate-then-reduce arg-max."""
    from grns_amd.workload import Synth
    c, o, mod = ctxs[32]
    x, p, V = Synth(mod).max_planted(1 << 24, seed=2)
    res, idx = c.max(x)
    assert int(idx.item()) == p
    assert res.cpu().numpy().tolist() == [V % m for m in mod]
    res, idx = c.max(x, mrc=True)
    assert int(idx.item()) == p
    assert res.cpu().numpy().tolist() == [V % m for m in mod]
    assert o.max_eval_array(x.cpu().numpy()) == p


@pytest.mark.parametrize("n", [12, 20])
def test_other_moduli_counts(dev, n):
    """A moduli count with no per-lane kernel (12: groups of 16 lanes, 20: groups of 32 with idle
    lanes) runs the group kernels: evaluation, comparison, mixed-radix digits, division and scaling
    against the CPU path, the binary round trip against the integers."""
    import grns_amd as G
    from grns_amd import product
    from oracle.oracle import Oracle
    mod = G.largest_primes_below(n)
    M = product(mod)
    c, o = G.Context(mod, 0), Oracle(mod)
    try:
        vals = eval_inputs(mod, 1500, seed=40 + n)
        x = to_res(vals, mod)
        for fast in (False, True):
            assert c.eval(_dev(x, dev), fast=fast).cpu().numpy().tobytes() == o.eval_array(x, fast=fast).tobytes()
        xs, ys = cmp_pairs(mod, 1500, seed=41 + n)
        xa, ya = to_res(xs, mod), to_res(ys, mod)
        assert (c.cmp(_dev(xa, dev), _dev(ya, dev)).cpu().numpy() == o.cmp_array(xa, ya)).all()
        dg = c.mrc(_dev(x, dev)).cpu().numpy()
        for j in range(0, x.shape[1], 7):
            assert dg[:, j].tolist() == o.mrc(x[:, j]).tolist(), j
        dx, dd = div_inputs(mod, 200, seed=42 + n)
        q, r = c.div(_dev(to_res(dx, mod), dev), _dev(to_res(dd, mod), dev), fast=True)
        rq, rr = o.div_array(to_res(dx, mod), to_res(dd, mod), fast=True)
        assert (q.cpu().numpy() == rq).all() and (r.cpu().numpy() == rr).all()
        rnd = random.Random(43 + n)
        D = np.array([rnd.randrange(0, M.bit_length() + 5) for _ in vals], np.uint32)
        y = c.scale2pow(_dev(x, dev), _dev(D.view(np.int32), dev)).cpu().numpy()
        assert (y == o.scale2pow_array(x, D)).all()
        w = c.to_binary(_dev(x, dev)).cpu().numpy().view(np.uint32)
        assert [sum(int(v) << (32 * k) for k, v in enumerate(row)) for row in w] == [v % M for v in vals]
        assert (c.from_binary(c.to_binary(_dev(x, dev))).cpu().numpy() == x).all()
        assert c.status() == 0
    finally:
        c.close()
        o.close()


SHIPPED_SETS = ["src/params.h", "src/params/moduli-64-4096-bit/params.16m.256b.h",
                "src/params/moduli-64-4096-bit/params.8m.128b.h", "src/params/moduli-10-69-bit/params.10m.36b.h",
                "src/params/moduli-10-69-bit/params.16m.69b.h", "src/params/moduli-100-6400-bit/params.32m.800b.h",
                "src/params/moduli-100-6400-bit/params.64m.1600b.h"]


@pytest.mark.parametrize("name", SHIPPED_SETS)
def test_shipped_moduli_sets(dev, name):
    """The moduli sets GRNS ships as params.h files (tests/golden/params_sets.json): pairwise coprime
    but not prime, from 3..31 up to ~2^31, so that every modulus-size-dependent path (reciprocal
    reductions of small moduli, the key and certificate bounds, the word division with few words)
    runs.  Every operation against the oracle bit for bit; arg-max against the integers."""
    import json
    import os

    import grns_amd as G
    from oracle.oracle import Oracle
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "params_sets.json")) as f:
        mod = {e["file"]: e["moduli"] for e in json.load(f)}[name]
    M = G.product(mod)
    c, o = G.Context(mod, 0), Oracle(mod)
    rnd = random.Random(len(mod) * 1000 + mod[0] % 1000)
    vals = eval_inputs(mod, 3000, seed=3)
    x = to_res(vals, mod)
    xd = _dev(x, dev)
    for fast in (False, True):
        assert c.eval(xd, fast=fast).cpu().numpy().tobytes() == o.eval_array(x, fast=fast).tobytes(), fast
    xs, ys = cmp_pairs(mod, 3000, seed=4)
    a, b = to_res(xs, mod), to_res(ys, mod)
    assert (c.cmp(_dev(a, dev), _dev(b, dev)).cpu().numpy() == o.cmp_array(a, b)).all()
    dx, dd = div_inputs(mod, 400, seed=5)
    p, q = to_res(dx, mod), to_res(dd, mod)
    for fast in (True, False):
        gq, gr = c.div(_dev(p, dev), _dev(q, dev), fast=fast)
        oq, orr = o.div_array(p, q, fast=fast)
        assert (gq.cpu().numpy() == oq).all() and (gr.cpu().numpy() == orr).all(), fast
    D = np.array([rnd.randrange(0, o.log2M + 3) for _ in vals], np.int32)
    assert (c.scale2pow(xd, _dev(D, dev)).cpu().numpy() == o.scale2pow_array(x, D.astype(np.uint32))).all()
    assert (c.mrc(xd).cpu().numpy() == np.stack([o.mrc(x[:, j]) for j in range(x.shape[1])], axis=1)).all()
    truth = max(range(len(vals)), key=lambda k: (vals[k], -k))
    for mrc in (False, True):
        res, idx = c.max(xd, mrc=mrc)
        assert int(idx.item()) == truth and res.cpu().numpy().tolist() == x[:, truth].tolist(), mrc
    if M > 1 << 40:
        bits = (M.bit_length() - 6) // 2
        mv = [rnd.choice([-1, 1]) * rnd.randrange(1 << rnd.randrange(1, bits)) for _ in range(600)]
        nv = [rnd.choice([-1, 1]) * rnd.randrange(1 << rnd.randrange(1, bits)) for _ in range(600)]

        def side(v):
            dg = to_res([abs(t) for t in v], mod)
            sg = np.array([1 if t < 0 else 0 for t in v], np.int32)
            ev = o.eval_array(dg)
            e = np.zeros((len(v), 2), ev["low"].dtype)
            e[:, 0], e[:, 1] = ev["low"], ev["upp"]
            return (dg, sg, e), (_dev(dg, dev), _dev(sg, dev), _dev(e.view(np.float64).reshape(len(v), 2, 2), dev))
        (Xn, X), (Yn, Y) = side(mv), side(nv)
        for op in ("add", "sub", "mul"):
            Z = (torch.empty_like(X[0]), torch.empty_like(X[1]), torch.empty_like(X[2]))
            c.mpint(op, X, Y, Z, len(mv))
            rd, rs, re, _ = o.mpint_array(op, *Xn, *Yn)
            assert (Z[0].cpu().numpy() == rd).all() and (Z[1].cpu().numpy() == rs).all(), op
            assert Z[2].cpu().numpy().tobytes() == re.tobytes(), op
    # the status bit for a refinement factor outside RNS_POW2 (the reference reads past its table)
    # is raised exactly when the oracle met such an input (it returns NaN bounds there)
    ev = [o.eval_array(x, fast=f) for f in (False, True)]
    past_table = any(np.isnan(e["low"]["frac"]).any() or np.isnan(e["upp"]["frac"]).any() for e in ev)
    assert c.status() == (G.STATUS_REFINE_RANGE if past_table else 0)
    c.close()
    o.close()


LARGE_SETS = ["src/params/moduli-64-4096-bit/params.128m.2048b.h", "src/params/moduli-64-4096-bit/params.256m.4096b.h",
              "src/params/moduli-100-6400-bit/params.128m.3200b.h", "src/params/moduli-100-6400-bit/params.256m.6400b.h"]


def _shipped(name):
    import json
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "params_sets.json")) as f:
        return {e["file"]: e["moduli"] for e in json.load(f)}[name]


@pytest.mark.parametrize("name", LARGE_SETS)
def test_large_moduli_sets(dev, name):
    """128 and 256 moduli (wider than a wave: one thread per number through the CPU path's steps):
    evaluation, comparison, division (both loops), scaling, mixed-radix digits, both arg-max
    pipelines, mpint add/sub/mul/div on arrays and on mp_int_t structs, against the oracle bit for bit
    and the integers."""
    import grns_amd as G
    from oracle.oracle import ER_DTYPE, Oracle
    mod = _shipped(name)
    n, M = len(mod), G.product(mod)
    c, o = G.Context(mod, 0), Oracle(mod)
    rnd = random.Random(n + mod[0] % 97)
    vals = eval_inputs(mod, 600, seed=3)[:600]
    x = to_res(vals, mod)
    xd = _dev(x, dev)
    for fast in (False, True):
        assert c.eval(xd, fast=fast).cpu().numpy().tobytes() == o.eval_array(x, fast=fast).tobytes(), fast
    xs, ys = cmp_pairs(mod, 600, seed=4)
    a, b = to_res(xs, mod), to_res(ys, mod)
    assert (c.cmp(_dev(a, dev), _dev(b, dev)).cpu().numpy() == o.cmp_array(a, b)).all()
    dx, dd = div_inputs(mod, 40, seed=5)
    p, q = to_res(dx, mod), to_res(dd, mod)
    for fast in (True, False):
        gq, gr = c.div(_dev(p, dev), _dev(q, dev), fast=fast)
        oq, orr = o.div_array(p, q, fast=fast)
        assert (gq.cpu().numpy() == oq).all() and (gr.cpu().numpy() == orr).all(), fast
        assert [(v1, v2) for v1, v2 in zip(from_res(gq.cpu().numpy(), mod), from_res(gr.cpu().numpy(), mod))] == \
            [divmod(u, w) for u, w in zip(dx, dd)], fast
    D = np.array([rnd.choice([0, 1, 29, 30, 31, rnd.randrange(o.log2M + 3)]) for _ in vals], np.int32)
    sc = c.scale2pow(xd, _dev(D, dev)).cpu().numpy()
    assert (sc == o.scale2pow_array(x, D.astype(np.uint32))).all()
    assert from_res(sc, mod) == [v >> int(k) for v, k in zip(vals, D)]
    assert (c.mrc(xd).cpu().numpy() == np.stack([o.mrc(x[:, j]) for j in range(x.shape[1])], axis=1)).all()
    truth = max(range(len(vals)), key=lambda k: (vals[k], -k))
    xn = x.astype(np.int64)
    for j in range(0, xn.shape[1], 5):  # non-canonical residues stand for their value mod m_i
        i = rnd.randrange(n)
        xn[i, j] += mod[i] if xn[i, j] + mod[i] < 2 ** 31 else -mod[i]
    for mrc in (False, True):
        for arr in (x, xn.astype(np.int32)):
            res, idx = c.max(_dev(arr, dev), mrc=mrc)
            assert int(idx.item()) == truth and res.cpu().numpy().tolist() == arr[:, truth].tolist(), mrc
    bits = (M.bit_length() - 6) // 2
    mv = [rnd.choice([-1, 1]) * rnd.randrange(1, 1 << rnd.randrange(2, bits)) for _ in range(200)]
    nv = [rnd.choice([-1, 1]) * rnd.randrange(1, 1 << rnd.randrange(2, bits)) for _ in range(200)]

    def side(v):
        dg = to_res([abs(t) for t in v], mod)
        sg = np.array([1 if t < 0 else 0 for t in v], np.int32)
        ev = o.eval_array(dg)
        e = np.zeros((len(v), 2), ev["low"].dtype)
        e[:, 0], e[:, 1] = ev["low"], ev["upp"]
        return (dg, sg, e), (_dev(dg, dev), _dev(sg, dev), _dev(e.view(np.float64).reshape(len(v), 2, 2), dev))
    (Xn, X), (Yn, Y) = side(mv), side(nv)
    for op in ("add", "sub", "mul", "div"):
        Z = (torch.empty_like(X[0]), torch.empty_like(X[1]), torch.empty_like(X[2]))
        c.mpint(op, X, Y, Z, len(mv))
        if op == "div":
            rd, rs, re = o.mpint_div_array(Xn[0], Xn[1], Yn[0], Yn[1])
        else:
            rd, rs, re, _ = o.mpint_array(op, *Xn, *Yn)
        assert (Z[0].cpu().numpy() == rd).all() and (Z[1].cpu().numpy() == rs).all(), op
        assert Z[2].cpu().numpy().tobytes() == re.tobytes(), op
    # the same operations on mp_int_t structs (the reference's layout)
    eoff = (4 * n + 4 + 7) // 8 * 8
    rec = eoff + 32
    L = G.lib()
    assert L.grns_mpint_aos_bytes(c.handle) == rec

    def records(T):
        r = np.zeros((len(mv), rec), np.uint8)
        r[:, :4 * n] = np.ascontiguousarray(T[0].T).view(np.uint8).reshape(len(mv), 4 * n)
        r[:, 4 * n:4 * n + 4] = T[1].view(np.uint8).reshape(-1, 4)
        r[:, eoff:] = T[2].view(np.uint8).reshape(len(mv), 32)
        return torch.from_numpy(r).to(dev)
    RX, RY = records(Xn), records(Yn)
    for code, op in ((3, "add"), (4, "sub"), (5, "mul"), (10, "div")):
        RZ = torch.zeros_like(RX)
        assert L.grns_mpint_aos(c.handle, code, RX.data_ptr(), RY.data_ptr(), RZ.data_ptr(), len(mv), rec, None,
                                None) == 0
        z = RZ.cpu().numpy()
        if op == "div":
            rd, rs, re = o.mpint_div_array(Xn[0], Xn[1], Yn[0], Yn[1])
        else:
            rd, rs, re, _ = o.mpint_array(op, *Xn, *Yn)
        assert (np.ascontiguousarray(z[:, :4 * n]).view(np.int32).T == rd).all(), op
        assert (np.ascontiguousarray(z[:, 4 * n:4 * n + 4]).view(np.int32).reshape(-1) == rs).all(), op
        assert np.ascontiguousarray(z[:, eoff:]).view(ER_DTYPE).reshape(-1, 2).tobytes() == re.tobytes(), op
    ev = [o.eval_array(x, fast=f) for f in (False, True)]
    past_table = any(np.isnan(e["low"]["frac"]).any() or np.isnan(e["upp"]["frac"]).any() for e in ev)
    assert c.status() == (G.STATUS_REFINE_RANGE if past_table else 0)
    c.close()
    o.close()


@pytest.mark.parametrize("n", [12, 20])
def test_mpint_and_max_other_sizes(dev, n):
    """Moduli counts without thread-per-number kernels (n = 12, 20): mpint and the arg-max through the
    serial forms, against the oracle and the integers."""
    import grns_amd as G
    from oracle.oracle import Oracle
    mod = G.largest_primes_below(n)
    M = G.product(mod)
    c, o = G.Context(mod, 0), Oracle(mod)
    rnd = random.Random(n)
    vals = [rnd.randrange(M) for _ in range(3000)]
    vals[1234] = max(vals)
    x = to_res(vals, mod)
    truth = max(range(len(vals)), key=lambda k: (vals[k], -k))
    for mrc in (False, True):
        res, idx = c.max(_dev(x, dev), mrc=mrc)
        assert int(idx.item()) == truth and res.cpu().numpy().tolist() == x[:, truth].tolist()
    bits = (M.bit_length() - 6) // 2
    mv = [rnd.choice([-1, 1]) * rnd.randrange(1 << rnd.randrange(1, bits)) for _ in range(500)]
    nv = [rnd.choice([-1, 1]) * rnd.randrange(1, 1 << rnd.randrange(2, bits)) for _ in range(500)]

    def side(v):
        dg = to_res([abs(t) for t in v], mod)
        sg = np.array([1 if t < 0 else 0 for t in v], np.int32)
        ev = o.eval_array(dg)
        e = np.zeros((len(v), 2), ev["low"].dtype)
        e[:, 0], e[:, 1] = ev["low"], ev["upp"]
        return (dg, sg, e), (_dev(dg, dev), _dev(sg, dev), _dev(e.view(np.float64).reshape(len(v), 2, 2), dev))
    (Xn, X), (Yn, Y) = side(mv), side(nv)
    for op in ("add", "sub", "mul", "div"):
        Z = (torch.empty_like(X[0]), torch.empty_like(X[1]), torch.empty_like(X[2]))
        c.mpint(op, X, Y, Z, len(mv))
        if op == "div":
            rd, rs, re = o.mpint_div_array(Xn[0], Xn[1], Yn[0], Yn[1])
        else:
            rd, rs, re, _ = o.mpint_array(op, *Xn, *Yn)
        assert (Z[0].cpu().numpy() == rd).all() and (Z[1].cpu().numpy() == rs).all(), op
        assert Z[2].cpu().numpy().tobytes() == re.tobytes(), op
    assert c.status() == 0


@pytest.mark.parametrize("n", [8, 64])
def test_streaming_kernels_persistent_grid(ctxs, dev, n):
    """Grid-stride loops that take several passes per block: eval (a persistent grid of one block
    per CU by default at n = 64) and cmp under a cap of one block per CU, on 2^17 + 77 numbers,
    against the oracle."""
    c, o, mod = ctxs[n]
    cnt = (1 << 17) + 77
    vals = eval_inputs(mod, cnt, seed=90 + n)[:cnt]
    xs, ys = (v[:cnt] for v in cmp_pairs(mod, cnt, seed=91 + n))
    x, xa, ya = to_res(vals, mod), to_res(xs, mod), to_res(ys, mod)
    try:
        c.grid_cap(1)
        assert _intervals_bytes(c.eval(_dev(x, dev))) == o.eval_array(x).tobytes()
        assert (c.cmp(_dev(xa, dev), _dev(ya, dev)).cpu().numpy() == o.cmp_array(xa, ya)).all()
        c.grid_cap(0)
        assert _intervals_bytes(c.eval(_dev(x, dev))) == o.eval_array(x).tobytes()
        assert c.status() == 0
    finally:
        c.grid_cap(0)
