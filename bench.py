"""Benchmark of the GRNS hot path on one MI355X.

Headline workload (the `value` of the JSON line): rns_max over N = 2^24 RNS numbers of the 32
largest primes below 2^31 (~990-bit range), modulus-major int32 residues resident in HBM, a
unique maximum planted at a seeded index plus 1000 near-ties one below it.  A step is one
arg-max over the whole array; value = numbers reduced per second.

Reported beside it, at the same n = 32 and N = 2^24: rns_eval_compute and rns_cmp (pairs
(x, x + delta), delta in {-1, 0, +1, uniform}, 1% forced edge cases), each as ops/s and as a
fraction of the HBM roofline.

Before timing, every workload's result is checked: the arg-max index and residues against the
planted maximum, and evaluation / comparison results on a seeded sample of 4096 elements against
the CPU restatement.

Timing: W untimed warm-up steps, then K steps bracketed by torch.cuda.synchronize() and HIP
events recorded on the stream the kernels are launched on (torch's current stream).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--only max,eval,cmp,div,batch,serve]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_MODULI = 32
COUNT = 1 << 24
# the list kernel of the batched mpint workload (n = 32, op class mpint) as the profilers name it
BATCH_KERNEL = "k_batch_thread<32, 4>"
METRIC = "RNS ops/sec + % HBM roofline (eval/cmp/max, n=32); batched p50/p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(steps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps  # ms per step


def kernel_ms(ctx, op, fn, steps):
    """Average duration of op's streaming kernel over `steps` calls: HIP events recorded by the
    library on the launch stream around that kernel alone."""
    ctx.stage_timing(True)
    for _ in range(steps):
        fn()
    ms, launches = ctx.stage_read(op)
    ctx.stage_timing(False)
    return ms / max(1, launches)


def roofline(bytes_per_launch, ms, kernel):
    """HBM roofline of one kernel: algorithmic bytes per launch over its average launch time."""
    gbs = bytes_per_launch / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": load_traffic(kernel), "kernel": kernel,
            "kernel_us": round(ms * 1e3, 2), "bytes_per_launch": bytes_per_launch}


def load_traffic(name):
    """HBM bytes per launch of kernel `name` measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    each scaled by the same counter's reading for a plain copy of known size), recorded in
    profiles/traffic.json by `bench.py --traffic`; null when absent."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            k = json.load(f)["kernels"].get(name)
        return None if k is None else k["traffic_bytes"]
    except (OSError, ValueError, KeyError):
        return None


COPY_BYTES = 1 << 31


def pmc_workload():
    """The program the counter runs profile: a plain 2 GiB device copy (calibration of FETCH_SIZE
    and WRITE_SIZE), then two launches each of arg-max (both pipelines), evaluation and comparison on
    the headline workload (n = 32, N = 2^24), of division and scaling on 2^18 pairs of the n = 64
    workload, and of the batched mpint workload (4096 vectors, n = 32)."""
    import torch

    import grns_amd as G
    from grns_amd.workload import Synth
    mod = G.largest_primes_below(N_MODULI)
    ctx, syn = G.Context(mod, 0), Synth(mod)
    src = torch.zeros(COPY_BYTES // 4, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    torch.cuda.empty_cache()
    x, _, _ = syn.max_planted(COUNT, seed=2)
    ws = torch.empty(ctx.max_workspace_size(COUNT), dtype=torch.uint8, device="cuda")
    out = torch.empty(N_MODULI, dtype=torch.int32, device="cuda")
    oi = torch.empty(1, dtype=torch.int64, device="cuda")
    for _ in range(2):
        ctx.max(x, workspace=ws, out=out, out_idx=oi)
    for _ in range(2):
        ctx.max(x, workspace=ws, out=out, out_idx=oi, mrc=True)
    torch.cuda.synchronize()
    del x
    torch.cuda.empty_cache()
    xe, ye, _ = syn.eval_cmp(COUNT, seed=1)
    ev = torch.empty((COUNT, 4), dtype=torch.float64, device="cuda")
    sg = torch.empty(COUNT, dtype=torch.int32, device="cuda")
    for _ in range(2):
        ctx.eval(xe, out=ev)
        ctx.cmp(xe, ye, out=sg)
    torch.cuda.synchronize()
    del xe, ye, ev, sg
    torch.cuda.empty_cache()
    mod64 = G.largest_primes_below(64)
    c64 = G.Context(mod64, 0)
    x, d, D, _ = Synth(mod64).div(1 << 18, seed=3, max_bits=1980, max_shift=1900)
    q, r = torch.empty_like(x), torch.empty_like(x)
    for _ in range(2):
        c64.div(x, d, fast=True, q=q, r=r)
        c64.scale2pow(x, D, y=q)
    torch.cuda.synchronize()
    del x, d, D, q, r
    torch.cuda.empty_cache()
    # the batched mpint workload: two launches of its list kernel at the configuration the tuner
    # picks on MI355X (tiles of 256 numbers, 2 persistent blocks per CU)
    from grns_amd.workload import MPINT_OPS, mpint_batch, mpint_view
    W = mpint_batch(ctx, syn)
    X, Y = W["X"], W["Y"]
    Z = tuple(torch.empty_like(a) for a in X)
    tasks = [G.make_task(MPINT_OPS[o], int(ln), mx=mpint_view(X, f, ln), my=mpint_view(Y, f, ln),
                         mz=mpint_view(Z, f, ln)) for o, f, ln in zip(W["ops"], W["offs"], W["lens"])]
    batch = G.Batch(ctx, tasks)
    batch.set_config(tile=256, blocks_per_cu=2, threads_per_number=1, use_graph=False)
    for _ in range(2):
        batch.launch()
    torch.cuda.synchronize()
    batch.close()
    log("[bench] pmc workload done")


def _short_kernel(name):
    name = name[5:] if name.startswith("void ") else name
    name = name.split("(")[0] if not name.startswith("(") else name
    return name.replace("grns::", "")


def _counter_rows(d):
    """Rows of a rocprofv3 counter_collection CSV in dispatch order."""
    import csv
    import glob
    (path,) = glob.glob(os.path.join(d, "*counter_collection.csv"))
    return sorted(csv.DictReader(open(path)), key=lambda r: int(r["Dispatch_Id"]))


def _first_dispatches(rows, counter, per_kernel=2):
    """{kernel: [value of each of its first `per_kernel` dispatches]} for one counter.  The counter
    program launches each measured kernel twice on its workload before anything else (input
    generation of a later workload) dispatches the same kernel again."""
    seen = {}
    for r in rows:
        if r["Counter_Name"] != counter:
            continue
        k = _short_kernel(r["Kernel_Name"])
        vals = seen.setdefault(k, {})
        did = int(r["Dispatch_Id"])
        if did in vals or len(vals) < per_kernel:
            vals[did] = vals.get(did, 0.0) + float(r["Counter_Value"])
    return {k: list(v.values()) for k, v in seen.items()}


def traffic_from_counters(fetch_dir, write_dir, out_path):
    """profiles/traffic.json from the counter_collection CSVs of a FETCH_SIZE run and a WRITE_SIZE
    run of --pmc-workload: per kernel, the average KiB over its first two dispatches times 1024
    times the factor that maps the counter's reading for the 2 GiB copy onto the bytes the copy
    moves."""

    def read(d, counter):
        rows = _counter_rows(d)
        copies = [(int(r["Grid_Size"]), float(r["Counter_Value"])) for r in rows
                  if r["Counter_Name"] == counter and r["Kernel_Name"].startswith("__amd_rocclr_copyBuffer")]
        big = max(g for g, _ in copies)
        copy_kib = sum(v for g, v in copies if g == big) / sum(1 for g, _ in copies if g == big)
        per = _first_dispatches(rows, counter)
        return {k: sum(v) / len(v) for k, v in per.items()}, copy_kib

    fetch, fcopy = read(fetch_dir, "FETCH_SIZE")
    write, wcopy = read(write_dir, "WRITE_SIZE")
    fs, ws_ = COPY_BYTES / (fcopy * 1024), COPY_BYTES / (wcopy * 1024)
    kernels = {}
    for k in sorted(fetch):
        if not k.startswith("k_") or k not in write:
            continue
        fb, wb = fetch[k] * 1024 * fs, write[k] * 1024 * ws_
        kernels[k] = {"fetch_kib": round(fetch[k], 3), "write_kib": round(write[k], 3), "fetch_bytes": round(fb),
                      "write_bytes": round(wb), "traffic_bytes": round(fb + wb)}
    res = {"calibration": {"copy_bytes": COPY_BYTES, "fetch_kib": round(fcopy, 3), "write_kib": round(wcopy, 3),
                           "fetch_scale": round(fs, 4), "write_scale": round(ws_, 4)},
           "kernels": kernels}
    with open(out_path, "w") as f:
        json.dump(res, f, indent=1)
    log(f"[bench] wrote {out_path}")


SQ_COUNTERS = ("SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_BUSY_CYCLES", "SQ_WAVE_CYCLES", "SQ_WAVES", "SQ_INSTS_SALU",
               "SQ_INSTS_VMEM_RD", "GRBM_GUI_ACTIVE")


def sq_summary(sq_dir):
    """Markdown table of the SQ counters of a --pmc-workload run (the SQ_COUNTERS in one pass), per
    kernel, averaged over its first two dispatches.  VALU busy = 4 x SQ_ACTIVE_INST_VALU / SIMDs /
    (GRBM_GUI_ACTIVE / XCDs): a wave64 VALU instruction occupies a 16-lane SIMD for 4 cycles; MI355X
    has 1024 SIMDs in 8 XCDs, and the GRBM counter sums the XCDs' clocks."""
    rows = _counter_rows(sq_dir)
    per = {c: _first_dispatches(rows, c) for c in SQ_COUNTERS}
    out = ["| Kernel | waves | VALU instr / wave | SALU instr / wave | VMEM loads / wave | GPU cycles | VALU busy |",
           "|---|---|---|---|---|---|---|"]
    for k in per["SQ_WAVES"]:
        if not k.startswith("k_"):
            continue
        avg = {c: sum(per[c][k]) / len(per[c][k]) for c in SQ_COUNTERS if k in per[c]}
        waves = max(avg["SQ_WAVES"], 1.0)
        cycles = avg["GRBM_GUI_ACTIVE"] / 8
        busy = 4 * avg["SQ_ACTIVE_INST_VALU"] / 1024 / max(cycles, 1.0)
        out.append(f"| `{k}` | {waves:.0f} | {avg['SQ_INSTS_VALU'] / waves:.0f} | {avg['SQ_INSTS_SALU'] / waves:.0f} | "
                   f"{avg['SQ_INSTS_VMEM_RD'] / waves:.1f} | {cycles:.0f} | {100 * busy:.0f}% |")
    print("\n".join(out))


CPU_RUNS = 7  # timed runs per CPU figure (after one untimed run); the median is reported


def median_rate(fn, count, runs=CPU_RUNS):
    """Median over `runs` timed calls of fn (after one untimed call) of count / seconds, and the runs."""
    import statistics
    fn()
    rates = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        rates.append(count / (time.perf_counter() - t0))
    return statistics.median(rates), rates


def cpu_threads():
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)


def cpu_headline(mod, syn, count):
    """The reference's arg-max pipeline on the CPU (the oracle: evaluate every number, then the
    rns_max_cmp reduction) over a sample of the headline workload: median of CPU_RUNS runs on all
    host threads (the baseline's value) and on one thread."""
    from oracle.oracle import Oracle
    orc = Oracle(mod)
    threads = cpu_threads()
    out = {}
    for nt, sample in ((threads, min(count, 1 << 22)), (1, min(count, 1 << 19))):
        xh = syn.max_planted(sample, seed=2)[0].cpu().numpy()
        med, rates = median_rate(lambda: orc.max_eval_array(xh, nthreads=nt), sample)
        out[nt] = {"value": round(med), "unit": "RNS ops/s", "cores": nt, "runs": [round(r) for r in rates],
                   "sample": f"rns_max pipeline (evaluate + rns_max_cmp reduction, CPU path) over 2^"
                             f"{sample.bit_length() - 1} numbers of the same workload, median of {CPU_RUNS} runs, "
                             f"{nt} OpenMP thread{'s' if nt > 1 else ''}"}
    base = dict(out[threads], kind="port")
    base["one_core"] = out[1]
    return base


def cpu_per_op(threads):
    """The oracle (the reference's CPU path restated) on bounded samples of the eval/cmp (n = 32),
    division/scaling (n = 64) and mpint mul (n = 32) workloads: operations per second, median of
    CPU_RUNS runs, on `threads` OpenMP threads and on one."""
    import numpy as np
    import torch

    import grns_amd as G
    from grns_amd.workload import Synth
    from oracle.oracle import ER_DTYPE, Oracle

    mod32, mod64 = G.largest_primes_below(32), G.largest_primes_below(64)
    o32, o64 = Oracle(mod32), Oracle(mod64)
    s32 = Synth(mod32)
    xe, ye, _ = s32.eval_cmp(1 << 18, seed=1)
    xh, yh = xe.cpu().numpy(), ye.cpu().numpy()
    x, d, D, _ = Synth(mod64).div(1 << 12, seed=3, max_bits=1980, max_shift=1900)
    dxh, ddh, Dh = x.cpu().numpy(), d.cpu().numpy(), D.cpu().numpy().view(np.uint32)
    c32 = G.Context(mod32, 0)
    bits = torch.full((1 << 18,), 480, dtype=torch.int32, device="cuda")
    X = [s32.bits(bits, 5, k, top_set=False) for k in (1, 2)]
    E = [np.ascontiguousarray(c32.eval(t).cpu().numpy()).view(ER_DTYPE).reshape(-1, 2) for t in X]
    sg = np.zeros(1 << 18, np.int32)
    xa, ya = X[0].cpu().numpy(), X[1].cpu().numpy()

    def cols(a, k):
        return np.ascontiguousarray(a[:, :k])

    # (name, sample at all threads, sample at one thread, function of (sample, threads), workload)
    ops = (("eval", 1 << 18, 1 << 15, lambda k, t: o32.eval_array(cols(xh, k), nthreads=t), "n=32"),
           ("cmp", 1 << 18, 1 << 14, lambda k, t: o32.cmp_array(cols(xh, k), cols(yh, k), nthreads=t), "n=32"),
           ("div_fast", 1 << 12, 1 << 8, lambda k, t: o64.div_array(cols(dxh, k), cols(ddh, k), fast=True, nthreads=t),
            "n=64"),
           ("scale2pow", 1 << 12, 1 << 9, lambda k, t: o64.scale2pow_array(cols(dxh, k), Dh[:k], nthreads=t), "n=64"),
           ("mpint_mul", 1 << 18, 1 << 16,
            lambda k, t: o32.mpint_array("mul", cols(xa, k), sg[:k], E[0][:k], cols(ya, k), sg[:k], E[1][:k],
                                         nthreads=t), "n=32"))
    out = {}
    for name, k_all, k_one, fn, wl in ops:
        rec = {}
        for nt, k in ((threads, k_all), (1, k_one)):
            med, _ = median_rate(lambda: fn(k, nt), k)
            rec[nt] = {"value": round(med), "unit": "ops/s", "cores": nt,
                       "sample": f"2^{k.bit_length() - 1} of the {wl} workload, median of {CPU_RUNS} runs"}
        out[name] = dict(rec[threads], one_core=rec[1])
    return out


def div_scale_leg(steps, warmup, n_mod=64, count=1 << 20, seed=3):
    """General division and power-of-two scaling: 64 largest primes below 2^31 (~1983-bit range),
    2^20 pairs, dividends uniform in [0, M), divisors of uniform bit length in [1, 1980] with
    uniform bits, exponents uniform in [1, 1900].  A sample of 256 results is checked against
    exact integer arithmetic before timing."""
    import numpy as np
    import torch

    import grns_amd as G
    from grns_amd.workload import Synth
    mod = G.largest_primes_below(n_mod)
    M = G.product(mod)
    ctx, syn = G.Context(mod, 0), Synth(mod)
    x, d, D, _ = syn.div(count, seed=seed, max_bits=1980, max_shift=1900)
    q, r, y = torch.empty_like(x), torch.empty_like(x), torch.empty_like(x)
    ctx.div(x, d, fast=True, q=q, r=r)
    ctx.scale2pow(x, D, y=y)
    torch.cuda.synchronize()
    idx = np.random.default_rng(seed).choice(count, 256, replace=False)
    xs, ds, qs, rs, ys = (t[:, idx].cpu().numpy() for t in (x, d, q, r, y))
    Ds = D[idx].cpu().numpy().view(np.uint32)
    inv = [pow(M // m, -1, m) for m in mod]

    def val(col):
        return sum(int(v) * (M // m) * w for v, m, w in zip(col, mod, inv)) % M
    for j in range(len(idx)):
        xv, dv = val(xs[:, j]), val(ds[:, j])
        if (val(qs[:, j]), val(rs[:, j])) != (xv // dv, xv % dv) or val(ys[:, j]) != xv >> int(Ds[j]):
            raise SystemExit("division / scaling sample differs from exact integer arithmetic")
    ms_df = timed(lambda: ctx.div(x, d, fast=True, q=q, r=r), steps, warmup)
    ms_d = timed(lambda: ctx.div(x, d, fast=False, q=q, r=r), max(1, steps // 4), 1)
    ms_s = timed(lambda: ctx.scale2pow(x, D, y=y), steps, warmup)
    wb = ctx.to_binary(x)
    if not torch.equal(ctx.from_binary(wb), x):
        raise SystemExit("binary round trip differs")
    ms_tb = timed(lambda: ctx.to_binary(x, out=wb), steps, warmup)
    ms_fb = timed(lambda: ctx.from_binary(wb, out=y), steps, warmup)
    return {"workload": f"n={n_mod} (largest primes below 2^31) N=2^{count.bit_length() - 1} pairs, divisor bits "
                        "uniform in [1, 1980], exponents uniform in [1, 1900]",
            "div_fast_ops_per_s": round(count / (ms_df * 1e-3)), "div_fast_ms": round(ms_df, 4),
            "div_ops_per_s": round(count / (ms_d * 1e-3)), "div_ms": round(ms_d, 4),
            "scale2pow_ops_per_s": round(count / (ms_s * 1e-3)), "scale2pow_ms": round(ms_s, 4),
            "to_binary_ms": round(ms_tb, 4), "from_binary_ms": round(ms_fb, 4),
            "sample_checked_exact": 256}


def batched_mpint_leg(steps, warmup, vectors=4096, max_len=65536, seed=4):
    """4096 mpint vectors (32 moduli) of log-uniform length in [1, 65536], random signs, |x| < 2^480,
    op add/sub/mul drawn per vector: one persistent launch (autotuned, hipGraph) against one
    per-call launch per vector, looped in C++ by the library (per_call_ms) and, for reference, from
    Python (per_call_python_ms).  Every vector's result is compared between the two forms."""
    import torch

    import grns_amd as G
    from grns_amd.workload import MPINT_OPS, Synth, mpint_batch, mpint_view

    mod = G.largest_primes_below(N_MODULI)
    n = len(mod)
    ctx = G.Context(mod, 0)
    W = mpint_batch(ctx, Synth(mod), vectors=vectors, max_len=max_len, seed=seed)
    X, Y, lens, offs, ops, total = W["X"], W["Y"], W["lens"], W["offs"], W["ops"], W["total"]
    # zero-filled outputs: the padding between vectors belongs to no task and stays 0 in both forms
    Zb = (torch.zeros_like(X[0]), torch.zeros_like(X[1]), torch.zeros_like(X[2]))
    Zp = (torch.zeros_like(X[0]), torch.zeros_like(X[1]), torch.zeros_like(X[2]))
    tasks = [G.make_task(MPINT_OPS[op], int(ln), mx=mpint_view(X, o, ln), my=mpint_view(Y, o, ln),
                         mz=mpint_view(Zb, o, ln)) for op, o, ln in zip(ops, offs, lens)]
    batch = G.Batch(ctx, tasks)
    cfg, tuned_ms = batch.tune(reps=2)
    batch.set_config(use_graph=True)
    # the same tasks writing Zp, run one per-call launch per task by the library's C++ loop
    per_call = G.Batch(ctx, [G.make_task(MPINT_OPS[op], int(ln), mx=mpint_view(X, o, ln), my=mpint_view(Y, o, ln),
                                         mz=mpint_view(Zp, o, ln)) for op, o, ln in zip(ops, offs, lens)])

    def plain():
        for op, o, ln in zip(ops, offs, lens):
            ctx.mpint(MPINT_OPS[op][6:], mpint_view(X, o, ln), mpint_view(Y, o, ln), mpint_view(Zp, o, ln), int(ln))

    per_call.run_per_call()
    batch.launch()
    torch.cuda.synchronize()
    # compare bit patterns: the eval records hold int64 exponents in float64 storage
    same = all(torch.equal(a.view(torch.int64) if a.dtype == torch.float64 else a,
                           b.view(torch.int64) if b.dtype == torch.float64 else b) for a, b in zip(Zb, Zp))
    if not same:
        raise SystemExit("batched mpint results differ from the per-call results")
    ms_b = timed(lambda: batch.launch(), steps, warmup)
    ms_p = timed(lambda: per_call.run_per_call(), max(2, steps // 4), 1)
    ms_py = timed(plain, max(2, steps // 4), 1)
    bytes_per = 3 * (4 * n + 36)
    out = {"workload": f"{vectors} mpint vectors n={n}, log-uniform lengths [1,{max_len}], {total} elements, "
                       "add/sub/mul per vector, digit rows 256-byte aligned",
           "elements": total, "batched_ms": round(ms_b, 4), "per_call_ms": round(ms_p, 4),
           "per_call_python_ms": round(ms_py, 4), "speedup": round(ms_p / ms_b, 2),
           "elem_ops_per_s": round(total / (ms_b * 1e-3)),
           "roofline": roofline(bytes_per * total, ms_b, BATCH_KERNEL), "tuned": cfg, "bit_exact_vs_per_call": same}
    batch.close()
    per_call.close()
    return out


def serving_leg(n_req=120, seed=7, budget=8 << 30, lo=10, hi=22, max_batch=16):
    """Serving stream: requests with Poisson arrivals over three resident moduli
    sets (8, 32, 64 largest primes below 2^31), op mix 30% eval+cmp, 20% arg-max, 20% div_fast,
    10% scale2pow, 20% mpint mul, sizes log-uniform in [2^10, 2^22], operands drawn like the
    single-operation workloads (edge values in eval+cmp, planted maximum and near-ties in arg-max,
    divisors of uniform bit length, exponents in [1, 1900], mpint magnitudes of up to 480 bits
    within the range).  The same requests also run one after another through the plain per-call
    entry points; every served output is compared with the per-call one byte for byte."""
    import numpy as np
    import torch

    import grns_amd as G

    from grns_amd.workload import serving_requests
    reqs, _ctx = serving_requests(n_req, seed, lo, hi)
    elements = sum(q["count"] for q in reqs)
    rng = np.random.default_rng(seed + 1)

    # plain per-call path: one request after another on the current stream
    for q in reqs[:4]:
        q["plain"]()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for q in reqs:
        q["plain"]()
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    mismatches = []  # served outputs that differ from the per-call ones (reported, never expected)

    def served(rate):
        """Submit every request at Poisson arrival times of the given rate; wall time and stats."""
        for q in reqs:
            for o in q["out_s"]:
                o.zero_()
        torch.cuda.synchronize()
        S = G.Server(0, hbm_budget=budget, slots=4, max_batch=max_batch)
        gaps = rng.exponential(1.0 / rate, size=n_req)
        tickets = []
        t0 = time.perf_counter()
        t_next = t0
        for q, g in zip(reqs, gaps):
            t_next += g
            while time.perf_counter() < t_next:
                pass
            tickets.append(S.submit(q["ctx"], q["tasks"]))
        for t in tickets:
            S.wait(t)
        wall = time.perf_counter() - t0
        st = S.stats()
        S.close()
        for q in reqs:
            for k, (a, b) in enumerate(zip(q["out_s"], q["out_p"])):
                if not torch.equal(a.view(torch.uint8), b.view(torch.uint8)):
                    nd = int((a.view(torch.uint8) != b.view(torch.uint8)).sum().item())
                    mismatches.append({"kind": q["kind"], "n": q["n"], "count": q["count"], "output": k,
                                       "bytes_differing": nd})
                    log(f"[bench] served result differs from the per-call one: {mismatches[-1]}")
        return wall, st

    cap = n_req / t_plain  # requests per second the per-call path sustains
    served(4 * cap)  # warm-up of the server's kernels and graphs
    t_sat, st_sat = served(16 * cap)  # offered load far above capacity: throughput
    lam = 0.5 * cap
    t_lat, st_lat = served(lam)  # offered load half the per-call capacity: latency
    return {"workload": f"{n_req} requests, Poisson arrivals, mix 30% eval+cmp / 20% max / 20% div_fast / "
                        f"10% scale2pow / 20% mpint mul, sizes log-uniform [2^{lo}, 2^{hi}], moduli sets of 8/32/64 "
                        f"primes, {elements} elements",
            "elements": elements, "per_call_s": round(t_plain, 4), "per_call_elem_per_s": round(elements / t_plain),
            "saturated": {"offered_req_per_s": round(16 * cap, 1), "served_s": round(t_sat, 4),
                          "served_elem_per_s": round(elements / t_sat), "speedup": round(t_plain / t_sat, 2),
                          "p50_ms": round(st_sat["p50_ms"], 3), "p99_ms": round(st_sat["p99_ms"], 3),
                          "launches": st_sat["launches"]},
            "half_load": {"offered_req_per_s": round(lam, 1), "served_s": round(t_lat, 4),
                          "p50_ms": round(st_lat["p50_ms"], 3), "p99_ms": round(st_lat["p99_ms"], 3),
                          "launches": st_lat["launches"]},
            "bit_exact_vs_per_call": not mismatches, "mismatches": mismatches[:20]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", default="")
    ap.add_argument("--count", type=int, default=COUNT)
    ap.add_argument("--pmc-workload", action="store_true", help="run the counter-profiling program and exit")
    ap.add_argument("--traffic", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "OUT"),
                    help="write the traffic summary from two rocprofv3 --pmc runs of --pmc-workload and exit")
    ap.add_argument("--sq-summary", metavar="SQ_DIR",
                    help="print the SQ counter table of a rocprofv3 --pmc run of --pmc-workload and exit")
    args = ap.parse_args()
    if args.sq_summary:
        sq_summary(args.sq_summary)
        return
    if args.traffic:
        traffic_from_counters(*args.traffic)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the RNS path would shard "
              "element ranges across GPUs, one process per GPU, exchanging only the arg-max winners).",
              file=sys.stderr)
        sys.exit(2)
    if args.pmc_workload:
        pmc_workload()
        return

    import numpy as np
    import torch

    import grns_amd as G
    from grns_amd.workload import Synth, residues_of

    N = args.count
    mod = G.largest_primes_below(N_MODULI)
    ctx = G.Context(mod, 0)
    syn = Synth(mod)
    n = len(mod)
    extras = {}
    only = set(args.only.split(",")) if args.only else {"max", "eval", "cmp"}

    # ------------------------------------------------------------------ arg-max (headline)
    t0 = time.time()
    x, p, V = syn.max_planted(N, seed=2)
    torch.cuda.synchronize()
    log(f"[bench] max workload generated in {time.time() - t0:.1f}s (n={n}, N={N})")
    ws = torch.empty(ctx.max_workspace_size(N), dtype=torch.uint8, device="cuda")
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    oidx = torch.empty(1, dtype=torch.int64, device="cuda")
    ctx.max(x, workspace=ws, out=out, out_idx=oidx)
    got_idx = int(oidx.item())
    ok_max = got_idx == p and out.cpu().numpy().tolist() == residues_of(V, mod)
    if not ok_max:
        raise SystemExit(f"arg-max check failed: index {got_idx}, planted {p}")
    ms_max = timed(lambda: ctx.max(x, workspace=ws, out=out, out_idx=oidx), args.steps, args.warmup)
    max_bytes = 4 * n * N
    kms_max = kernel_ms(ctx, "max", lambda: ctx.max(x, workspace=ws, out=out, out_idx=oidx), args.steps)
    log(f"[bench] max: {ms_max:.4f} ms/step, stage-1 kernel {kms_max:.4f} ms")
    if "max" in only:
        # the mixed-radix pipeline on the same numbers (VALU-bound: n(n-1)/2 exact FP64 modular
        # products per number); its result is checked like the headline's
        ctx.max(x, workspace=ws, out=out, out_idx=oidx, mrc=True)
        if int(oidx.item()) != p or out.cpu().numpy().tolist() != residues_of(V, mod):
            raise SystemExit("mixed-radix arg-max check failed")
        ms_mrc = timed(lambda: ctx.max(x, workspace=ws, out=out, out_idx=oidx, mrc=True), args.steps, args.warmup)
        extras["max_mrc"] = {"ms_per_step": round(ms_mrc, 4), "ops_per_s": round(N / (ms_mrc * 1e-3)),
                             "bound": "valu", "digit_products_per_number": n * (n - 1) // 2}
        log(f"[bench] max by mixed-radix digits: {ms_mrc:.4f} ms/step")

    # ------------------------------------------------------------------ eval / cmp
    if only & {"eval", "cmp"}:
        from oracle.oracle import Oracle  # checker only: verifies a sample, never timed
        orc = Oracle(mod)
        del x
        torch.cuda.empty_cache()
        xe, ye, _ = syn.eval_cmp(N, seed=1)
        torch.cuda.synchronize()
        rng = np.random.default_rng(5)
        samp = np.sort(rng.choice(N, size=4096, replace=False))
        si = torch.from_numpy(samp).cuda()
        if "eval" in only:
            ev = torch.empty((N, 4), dtype=torch.float64, device="cuda")
            ctx.eval(xe, out=ev)
            ref = orc.eval_array(xe[:, si].cpu().numpy())
            if ev[si].cpu().numpy().tobytes() != ref.tobytes():
                raise SystemExit("eval sample mismatch against the CPU restatement")
            ms_eval = timed(lambda: ctx.eval(xe, out=ev), args.steps, args.warmup)
            kms = kernel_ms(ctx, "eval", lambda: ctx.eval(xe, out=ev), args.steps)
            eb = (4 * n + 32) * N
            extras["eval"] = {"ops_per_s": round(N / (ms_eval * 1e-3)), "ms_per_step": round(ms_eval, 4),
                              "roofline": roofline(eb, kms, "k_eval<32, false, 1>"), "bytes_per_number": 4 * n + 32,
                              "op_roofline_frac": round(eb / (ms_eval * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            log(f"[bench] eval: {ms_eval:.4f} ms/step, streaming kernel {kms:.4f} ms")
            del ev
        if "cmp" in only:
            sg = torch.empty(N, dtype=torch.int32, device="cuda")
            ctx.cmp(xe, ye, out=sg)
            ref = orc.cmp_array(xe[:, si].cpu().numpy(), ye[:, si].cpu().numpy())
            if not (sg[si].cpu().numpy() == ref).all():
                raise SystemExit("cmp sample mismatch against the CPU restatement")
            ms_cmp = timed(lambda: ctx.cmp(xe, ye, out=sg), args.steps, args.warmup)
            kms = kernel_ms(ctx, "cmp", lambda: ctx.cmp(xe, ye, out=sg), args.steps)
            cb = (8 * n + 4) * N
            extras["cmp"] = {"ops_per_s": round(N / (ms_cmp * 1e-3)), "ms_per_step": round(ms_cmp, 4),
                             "roofline": roofline(cb, kms, "k_cmp<32, 1>"), "bytes_per_pair": 8 * n + 4,
                             "op_roofline_frac": round(cb / (ms_cmp * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            log(f"[bench] cmp: {ms_cmp:.4f} ms/step, streaming kernel {kms:.4f} ms")
        status = ctx.status()
        if status:
            log(f"[bench] device status bits {status}")
        del xe, ye
        torch.cuda.empty_cache()

    if "div" in only or not args.only:
        extras["div_scale"] = div_scale_leg(args.steps, args.warmup)
        log(f"[bench] div_fast {extras['div_scale']['div_fast_ms']} ms, div {extras['div_scale']['div_ms']} ms, "
            f"scale2pow {extras['div_scale']['scale2pow_ms']} ms per 2^20")
        torch.cuda.empty_cache()

    if "batch" in only or not args.only:
        extras["batched"] = batched_mpint_leg(args.steps, args.warmup)
        log(f"[bench] batched mpint: {extras['batched']['batched_ms']} ms vs per-call "
            f"{extras['batched']['per_call_ms']} ms (C++ loop), {extras['batched']['per_call_python_ms']} ms (Python loop)")
        torch.cuda.empty_cache()

    if "serve" in only or not args.only:
        extras["serving"] = serving_leg()
        extras["serving_small"] = serving_leg(n_req=1000, seed=8, lo=4, hi=14, max_batch=64)
        sv = extras["serving"]
        log(f"[bench] serving: saturated {sv['saturated']['served_s']} s vs per-call {sv['per_call_s']} s "
            f"({sv['saturated']['speedup']}x); half load p50 {sv['half_load']['p50_ms']} ms, "
            f"p99 {sv['half_load']['p99_ms']} ms")
        ss = extras["serving_small"]
        log(f"[bench] serving small requests: saturated {ss['saturated']['served_s']} s vs per-call {ss['per_call_s']} s "
            f"({ss['saturated']['speedup']}x); half load p50 {ss['half_load']['p50_ms']} ms, "
            f"p99 {ss['half_load']['p99_ms']} ms")
        torch.cuda.empty_cache()

    # ------------------------------------------------------------------ CPU baseline
    cpu = None
    if not args.no_cpu:
        threads = cpu_threads()
        cpu = cpu_headline(mod, syn, N)
        log(f"[bench] cpu baseline {cpu['value']} ops/s on {threads} threads, {cpu['one_core']['value']} on one")
        extras["cpu_per_op"] = cpu_per_op(threads)
        log("[bench] cpu per op: " + ", ".join(f"{k} {v['value']:.3g}/s ({v['one_core']['value']:.3g} on one)"
                                               for k, v in extras["cpu_per_op"].items()))

    value = N / (ms_max * 1e-3)
    rl = roofline(max_bytes, kms_max, "k_max_partial<32, 1>")
    rl["op_frac"] = round(max_bytes / (ms_max * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    line = {
        "metric": METRIC,
        "value": round(value),
        "unit": "RNS ops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_max, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+f64",
        "data": "synthetic",
        "config": {"workload": f"rns_max n={n} (32 largest primes below 2^31) N=2^{N.bit_length() - 1} "
                               "modulus-major int32, planted unique max + 1000 near-ties",
                   "n_moduli": n, "count": N, "layout": "modulus-major"},
        "roofline": rl,
        "cpu_baseline": cpu,
    }
    line.update(extras)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
