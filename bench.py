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

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--only max|eval|cmp]
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


def roofline(bytes_per_step, ms):
    gbs = bytes_per_step / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4)}


def load_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE), recorded in
    profiles/traffic.json by the profiling run; null when absent."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            return json.load(f).get(name)
    except (OSError, ValueError):
        return None


def batched_mpint_leg(steps, warmup, vectors=4096, max_len=65536, seed=4):
    """4096 mpint vectors (32 moduli) of log-uniform length in [1, 65536], random signs, |x| < 2^480,
    op add/sub/mul drawn per vector: one persistent launch (autotuned, hipGraph) against one
    per-call launch per vector.  Every vector's result is compared between the two forms."""
    import numpy as np
    import torch

    import grns_amd as G
    from grns_amd.workload import Synth

    mod = G.largest_primes_below(N_MODULI)
    n = len(mod)
    ctx = G.Context(mod, 0)
    syn = Synth(mod)
    rng = np.random.default_rng(seed)
    lens = np.floor(np.exp(rng.uniform(0, np.log(max_len), size=vectors))).astype(np.int64).clip(1, max_len)
    ops = rng.integers(0, 3, size=vectors)
    total = int(lens.sum())
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]])

    def operand(stream):
        bits = torch.full((total,), 480, dtype=torch.int32, device="cuda")
        d = syn.bits(bits, seed, stream, top_set=False)
        s = syn.uniform_u32(total, 0, 1, seed, stream + 100)
        ev = ctx.eval(d)  # interval records have the mp_int_t eval[2] layout
        return d, s, ev.view(total, 2, 2)

    X, Y = operand(10), operand(20)
    Zb = (torch.empty_like(X[0]), torch.empty_like(X[1]), torch.empty_like(X[2]))
    Zp = (torch.empty_like(X[0]), torch.empty_like(X[1]), torch.empty_like(X[2]))
    names = ["mpint_add", "mpint_sub", "mpint_mul"]

    def view(T, o, ln):
        return (T[0][:, o:o + ln], T[1][o:o + ln], T[2][o:o + ln])

    tasks = [G.make_task(names[op], int(ln), mx=view(X, o, ln), my=view(Y, o, ln), mz=view(Zb, o, ln))
             for op, o, ln in zip(ops, offs, lens)]
    batch = G.Batch(ctx, tasks)
    cfg, tuned_ms = batch.tune(reps=2)
    batch.set_config(use_graph=True)

    def plain():
        for op, o, ln in zip(ops, offs, lens):
            ctx.mpint(names[op][6:], view(X, o, ln), view(Y, o, ln), view(Zp, o, ln), int(ln))

    plain()
    batch.launch()
    torch.cuda.synchronize()
    # compare bit patterns: the eval records hold int64 exponents in float64 storage
    same = all(torch.equal(a.view(torch.int64) if a.dtype == torch.float64 else a,
                           b.view(torch.int64) if b.dtype == torch.float64 else b) for a, b in zip(Zb, Zp))
    if not same:
        raise SystemExit("batched mpint results differ from the per-call results")
    ms_b = timed(lambda: batch.launch(), steps, warmup)
    ms_p = timed(plain, max(2, steps // 4), 1)
    bytes_per = 3 * (4 * n + 36)
    out = {"workload": f"{vectors} mpint vectors n={n}, log-uniform lengths [1,{max_len}], {total} elements, "
                       "add/sub/mul per vector",
           "elements": total, "batched_ms": round(ms_b, 4), "per_call_ms": round(ms_p, 4),
           "speedup": round(ms_p / ms_b, 2), "elem_ops_per_s": round(total / (ms_b * 1e-3)),
           "roofline": roofline(bytes_per * total, ms_b), "tuned": cfg, "bit_exact_vs_per_call": same}
    batch.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", default="")
    ap.add_argument("--count", type=int, default=COUNT)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the RNS path shards "
              "element ranges across GPUs with no exchange except the arg-max winners; see DESIGN.md).",
              file=sys.stderr)
        sys.exit(2)

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
    log(f"[bench] max: {ms_max:.4f} ms/step")

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
            eb = (4 * n + 32) * N
            extras["eval"] = {"ops_per_s": round(N / (ms_eval * 1e-3)), "ms_per_step": round(ms_eval, 4),
                              "roofline": roofline(eb, ms_eval), "bytes_per_number": 4 * n + 32}
            log(f"[bench] eval: {ms_eval:.4f} ms/step")
            del ev
        if "cmp" in only:
            sg = torch.empty(N, dtype=torch.int32, device="cuda")
            ctx.cmp(xe, ye, out=sg)
            ref = orc.cmp_array(xe[:, si].cpu().numpy(), ye[:, si].cpu().numpy())
            if not (sg[si].cpu().numpy() == ref).all():
                raise SystemExit("cmp sample mismatch against the CPU restatement")
            ms_cmp = timed(lambda: ctx.cmp(xe, ye, out=sg), args.steps, args.warmup)
            cb = (8 * n + 4) * N
            extras["cmp"] = {"ops_per_s": round(N / (ms_cmp * 1e-3)), "ms_per_step": round(ms_cmp, 4),
                             "roofline": roofline(cb, ms_cmp), "bytes_per_pair": 8 * n + 4}
            log(f"[bench] cmp: {ms_cmp:.4f} ms/step")
        status = ctx.status()
        if status:
            log(f"[bench] device status bits {status}")
        del xe, ye
        torch.cuda.empty_cache()

    if "batch" in only or not args.only:
        extras["batched"] = batched_mpint_leg(args.steps, args.warmup)
        log(f"[bench] batched mpint: {extras['batched']['batched_ms']} ms vs per-call "
            f"{extras['batched']['per_call_ms']} ms")
        torch.cuda.empty_cache()

    # ------------------------------------------------------------------ CPU baseline
    cpu = None
    if not args.no_cpu:
        from oracle.oracle import Oracle
        orc = Oracle(mod)
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        sample = min(N, 1 << 22)
        xh = syn.max_planted(sample, seed=2)[0].cpu().numpy()
        # repeated passes over the same array until about 12 s of CPU work have been timed
        passes, t0 = 0, time.perf_counter()
        while True:
            orc.max_eval_array(xh, nthreads=threads)
            passes += 1
            dt = time.perf_counter() - t0
            if dt >= 12.0 or passes >= 400:
                break
        cpu = {"value": round(passes * sample / dt), "unit": "RNS ops/s", "cores": threads, "kind": "port",
               "sample": f"rns_max pipeline (evaluate + rns_max_cmp reduction, CPU path) over 2^"
                         f"{sample.bit_length() - 1} numbers of the same workload, {passes} passes, "
                         f"{threads} OpenMP threads, {dt:.2f}s"}
        log(f"[bench] cpu baseline {cpu['value']} ops/s on {threads} threads")

    value = N / (ms_max * 1e-3)
    rl = roofline(max_bytes, ms_max)
    rl["traffic"] = load_traffic("k_max_partial")
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
