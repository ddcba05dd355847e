#!/usr/bin/env python3
"""Benchmark: pi(x) queries/s, batched vs one-at-a-time plain, on one MI355X.

A step is one batched solve of the 512-query workload (x log-uniform in
[1e9, 1e15], splitmix64 seed 3, tests/golden/queries_c3.json) through the
C ABI: every hard-leaf and P2 segment of every query is a tile of one
persistent kernel launch over shared resident tables, replayed from captured
hipGraphs.  The steps are pipelined two deep (cpx_pi_batch_submit/collect:
a step is planned and uploaded while the previous one runs); value = 512 *
steps / timed seconds (tables already resident).  sync_qps is the same
steps made one synchronous call at a time.

Beside it, in the same run:
  * plain: the same 512 x solved one at a time, each with its own table build
    (the reference's per-term, per-query table rebuilds), and plain_warm with
    tables kept resident;  speedup = value / plain qps;
  * served: Poisson streams (tests/golden/queries_c4.json) through the server
    in both modes -- continuous batching into one streaming launch, and one
    batch launch per admitted FIFO prefix -- p50/p99 latency and queries/s,
    every served value checked against the same x solved alone, beside the
    plain one-at-a-time rate on the same x list;
  * single: pi(1e16) wall time;
  * roofline of the persistent segment kernel (HIP events on its stream);
  * cpu_baseline: the oracle's C/OpenMP restatement of the reference terms on
    a bounded, stratified sample of the same 512 x.
Every batched result is checked bit-exact against the plain solve.
Prints ONE JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SMALLP = [2, 3, 5, 7, 11, 13, 17, 19]
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_json(name):
    with open(os.path.join(ROOT, "tests", "golden", name)) as f:
        return json.load(f)


def iroot(x, k):
    r = int(round(x ** (1.0 / k)))
    while r ** k > x:
        r -= 1
    while (r + 1) ** k <= x:
        r += 1
    return r


def primes_upto(n):
    import numpy as np
    s = np.ones(n + 1, dtype=bool)
    s[:2] = False
    for i in range(2, int(n ** 0.5) + 1):
        if s[i]:
            s[i * i::i] = False
    return np.nonzero(s)[0]


GROUP = 8   # queries per shared segment sieve (the library's group capacity)

# Tuning of the batched runs (segment threads, words per thread, alpha factor,
# wheel primes): the fastest of the autotuner's choices on the 512-query
# workload, and the one the committed counter profiles were taken at.
BENCH_TUNING = "512,32,0.3,8"


def sieve_groups(queries, primes):
    """The library's query groups: increasing x, the same wheel c, at most
    GROUP members, every member's P2 start (isqrt(x) rounded down to a
    multiple of 64) at or above every member's largest stage prime (the
    largest prime <= x^(1/4)), and a query with y >= 2^28 alone."""
    import numpy as np
    order = sorted((t for t in queries if t["x"] >= 16384), key=lambda t: t["x"])
    groups, cur, smax, p2min = [], [], 0, None
    for t in order:
        qi = np.searchsorted(primes, iroot(t["x"], 4), side="right")
        si = int(primes[qi - 1]) if qi > int(t["c"]) else 0
        pi = iroot(t["x"], 2) & ~63
        alone = t["y"] >= 1 << 28
        if cur and (len(cur) == GROUP or int(cur[0]["c"]) != int(t["c"]) or alone
                    or cur[-1]["y"] >= 1 << 28 or max(smax, si) > min(p2min, pi)):
            groups.append(cur)
            cur, smax, p2min = [], 0, None
        cur.append(t)
        smax = max(smax, si)
        p2min = pi if p2min is None else min(p2min, pi)
    if cur:
        groups.append(cur)
    return groups


def algorithmic_lds_bytes(queries, primes):
    """Segment-sieve LDS work of the DR staged sieve, per group of queries
    that share it (a single query is a group of one): with z and p_h the
    largest of the members' z = x/y and floor(x^(1/4)),
    W = B_h * ceil(z/2)/8 + 8 * sum_{b=c+1}^{c+B_h} ceil((z/2)/p_b),
    one popcount/prefix pass over the odd-only bitmap per stage plus one
    32-bit read-modify-write per cross-off.  (Leaf evaluations are not in W:
    they read a few words each.)"""
    import numpy as np
    tot = 0
    for g in sieve_groups(queries, primes):
        c = int(g[0]["c"])
        z = max(t["x"] // t["y"] for t in g)
        ph = max(iroot(t["x"], 4) for t in g)
        ps = primes[c:np.searchsorted(primes, ph, side="right")]
        bh = len(ps)
        half = (z + 1) // 2
        tot += bh * half / 8.0 + 8.0 * float(np.sum(np.ceil(half / ps.astype(np.float64))))
    return tot


def profiled_counters(workload):
    """Per-launch k_segments figures from the committed rocprofv3 --pmc
    summary of the same workload and tuning: lines `key: value` (tuning,
    valu_instr_per_launch, salu_instr_per_launch, lds_bytes_per_launch,
    hbm_bytes_per_launch).  Empty if the summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_segments_%s_pmc.md" % workload)
    out = {}
    try:
        with open(path) as f:
            for line in f:
                k, sep, v = line.partition(":")
                if sep and k.strip() in ("tuning", "valu_instr_per_launch", "salu_instr_per_launch",
                                         "lds_bytes_per_launch", "hbm_bytes_per_launch"):
                    v = v.strip()
                    out[k.strip()] = v if k.strip() == "tuning" else float(v)
    except (OSError, ValueError):
        pass
    return out


# Documented MI355X rates the roofline divides by (256 CUs at 2.4 GHz peak
# engine clock):
#  * VALU issue: the published FP32 vector peak, 157.3 TFLOP/s = 256 CUs x
#    2.4 GHz x 128 FMA lanes per clock, i.e. two wave64 32-bit VALU
#    instructions per CU per clock: 1228.8 G wave64 instructions/s;
#  * LDS: 256 bytes per CU per clock on CDNA 4 (twice CDNA 3's 128, per AMD's
#    CDNA 4 white paper), 157.3 TB/s;
#  * HBM3E: 8 TB/s.
CU_COUNT = 256
CLOCK_GHZ = 2.4
VALU_PEAK_G = 2 * CU_COUNT * CLOCK_GHZ          # G wave64 instructions/s
LDS_PEAK_GBS = 256 * CU_COUNT * CLOCK_GHZ       # GB/s


def hbm_model_bytes(terms, primes, seg_bits):
    """W_HBM of one batch launch (a model, labelled as such): 4 B per hard-leaf
    candidate (the packed mu*lpf word of each odd m in (y/p, y], or the prime
    list entry when p^2 > y), 8 B per easy-leaf pair (one pi-table entry),
    8 B per P2 prime (one pi-table entry at x/p), and 8 B per wheel word
    filled into each segment of each query group (seg_bits odd numbers per
    segment; replicas of a segment not counted).  Tables are resident and
    not rebuilt inside the launch."""
    import numpy as np
    nseg = sum(-(-((max(t["x"] // t["y"] for t in g) + 1) // 2) // seg_bits)
               for g in sieve_groups(terms, primes))
    pi_at = lambda v: int(np.searchsorted(primes, v, side="right"))
    l_hard = l_easy = n_p2 = 0
    for t in terms:
        x, y, c = t["x"], t["y"], int(t["c"])
        q4, c3, sq = iroot(x, 4), iroot(x, 3), iroot(x, 2)
        ps = primes[c:pi_at(q4)].astype(np.int64)
        dense = ps[ps * ps <= y]
        l_hard += int(np.sum((y - y // dense + 1) // 2))
        for p in ps[ps * ps > y]:
            l_hard += max(0, pi_at(y) - pi_at(max(int(p), y // int(p))))
        for p in primes[pi_at(q4):pi_at(c3)]:
            hi = min(y, x // (int(p) * int(p)))
            if hi > p:
                l_easy += pi_at(hi) - pi_at(int(p))
        n_p2 += max(0, pi_at(sq) - pi_at(y))
    wheel = nseg * seg_bits // 32
    w = 4 * l_hard + 8 * l_easy + 8 * n_p2 + 8 * wheel
    return w, {"hard_candidates": l_hard, "easy_pairs": l_easy, "p2_primes": n_p2,
               "group_segments": nseg, "wheel_words": wheel}


def host_info():
    """lscpu model, logical CPUs, the CPUs this process may run on (nproc) and
    the OpenMP thread count the CPU baseline uses."""
    import subprocess
    info = {"nproc": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count()}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("Model name", "CPU(s)", "Thread(s) per core", "Core(s) per socket", "Socket(s)",
                     "CPU max MHz"):
                info["lscpu " + k] = v.strip()
    except (OSError, subprocess.SubprocessError):
        info["lscpu"] = "unavailable"
    env = os.environ.get("OMP_NUM_THREADS", "")
    info["threads"] = int(env) if env.isdigit() and int(env) > 0 else info["nproc"]
    return info


def median_time(fn, reps=3, warm=True):
    """1 warm-up call (optional), then the median wall time of `reps` calls
    (steady clock); returns (median seconds, value of the last call)."""
    if warm:
        fn()
    ts, v = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        v = fn()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2], v


def cpu_alpha(x):
    """y = alpha * cbrt(x) for the CPU oracle: the hard-leaf walk dominates on
    the CPU, so alpha grows with x (chosen by timing the sample strata)."""
    return max(1.0, min(12.0, math.log10(x) - 5.0))


def cpu_baseline(xs, full=False, budget_s=90.0):
    """The oracle (C + OpenMP, rebuilt here with -O3 -march=native) on the
    host: 1 warm-up and the median of 3 for every timing.
      * queries/s on a stratified sample of the 512-query list (the median x
        of each of 16 strata of 32 sorted queries; a stratum's time stands for
        its 32 queries);
      * C1 (x = 1e12, alpha 2, c = 6): P2, ordinary, trivial, easy and hard
        terms, and the brute segmented-sieve count of primes below 1e12 timed
        on the window [1e12 - 2^33, 1e12] and scaled by 1e12 / 2^33 (with
        full: the whole count);
      * C2 (x = 1e16, alpha 2): ordinary and trivial (with full: P2 too)."""
    from oracle import oracle as O
    flags = O.use_native()
    hi = host_info()
    th = hi["threads"]
    srt = sorted(xs)
    k = 16
    per = len(srt) // k
    sample = [srt[i * per + per // 2] for i in range(k)]
    O.pi_dr(sample[0], O.choose_y(sample[0], cpu_alpha(sample[0])), 6, th)      # warm-up
    est, spent, strata = 0.0, 0.0, []
    for x in sample:
        y = O.choose_y(x, cpu_alpha(x))
        reps = 3 if spent < budget_s else 1
        t, r = median_time(lambda: O.pi_dr(x, y, 6, th), reps=reps, warm=False)
        spent += t * reps
        est += t * per
        strata.append({"x": x, "y": y, "s": round(t, 4), "reps": reps})
    out = {"value": round(len(srt) / est, 4), "unit": "queries/s", "cores": th, "kind": "port",
           "sample": "median x of each of 16 strata of the sorted 512-query list, each timed as the "
                     "median of 3 (1 warm-up before the first), alpha = clamp(log10 x - 5, 1, 12); "
                     "%.1f s of CPU" % spent,
           "build": "gcc " + flags, "host": hi, "strata": strata}

    x1 = 10 ** 12
    y1 = O.choose_y(x1, 2.0)
    c1 = {"x": x1, "y": y1, "c": 6}
    for name, fn in (("p2", lambda: O.P2(x1, y1, th)), ("ordinary", lambda: O.Ordinary(x1, y1, 6, th)),
                     ("trivial", lambda: O.S2_trivial(x1, y1, 6)), ("easy", lambda: O.V(x1, y1, 6, th)),
                     ("hard", lambda: O.S3(x1, y1, 6, th)), ("pi_dr", lambda: O.pi_dr(x1, y1, 6, th))):
        t, v = median_time(fn)
        c1[name + "_ms"] = round(t * 1000, 3)
    c1["p2_ord_triv_ms"] = round(c1["p2_ms"] + c1["ordinary_ms"] + c1["trivial_ms"], 3)
    if v["pi"] != 37607912018:
        raise SystemExit("oracle pi(1e12) wrong: %d" % v["pi"])
    if full:
        t, v = median_time(lambda: O.count_primes(0, x1 + 1, th), reps=3)
        if v != 37607912018:
            raise SystemExit("brute count below 1e12 wrong: %d" % v)
        c1["brute_ms"] = round(t * 1000, 1)
        c1["brute_sample"] = "the whole count"
    else:
        w = 1 << 33
        t, _ = median_time(lambda: O.count_primes(x1 - w, x1 + 1, th))
        c1["brute_ms"] = round(t * 1000 * x1 / w, 1)
        c1["brute_sample"] = "window [1e12 - 2^33, 1e12] in %.3f s (median of 3), scaled by 1e12/2^33" % t
    out["c1"] = c1

    x2 = 10 ** 16
    y2 = O.choose_y(x2, 2.0)
    c2 = {"x": x2, "y": y2, "c": 6}
    for name, fn in (("ordinary", lambda: O.Ordinary(x2, y2, 6, th)), ("trivial", lambda: O.S2_trivial(x2, y2, 6))) \
            + ((("p2", lambda: O.P2(x2, y2, th)),) if full else ()):
        t, _ = median_time(fn)
        c2[name + "_ms"] = round(t * 1000, 3)
    out["c2"] = c2
    return out


def configs_run(args):
    """The single-query configurations as markdown rows on stdout: pi(1e12) at
    alpha 2; pi(1e16) over alpha in {1, 1.5, 2, 3, 4, 6, 8} and the segment
    sizes 2^18, 2^19, 2^20 bits (2^21 bits do not fit one block's LDS);
    pi(1e20) at alpha 2, 4 and the built-in model.  y = floor(alpha *
    icbrt(x)); every value is checked against the published pi(x); wall time
    of a call on a warm context (tables resident), best of --reps."""
    import cudapix_amd as C
    known = load_json("known.json")["pi10"]

    def icbrt(x):
        r = int(round(x ** (1.0 / 3.0)))
        while r ** 3 > x:
            r -= 1
        while (r + 1) ** 3 <= x:
            r += 1
        return r

    def timed(cx, x, y, reps):
        t = cx.pi(x, y=y, terms=True)               # warm-up: tables
        best = None
        for _ in range(reps):
            t0 = time.time()
            v = cx.pi(x, y=y)
            dt = (time.time() - t0) * 1000
            best = dt if best is None else min(best, dt)
        if v != known[str(round(math.log10(x)))]:
            raise SystemExit("pi(%d) wrong at y=%d: %d" % (x, y, v))
        return best, t

    print("| case | x | alpha | y | segment bits (threads x words) | ms | pi(x) |", flush=True)
    print("|---|---|---|---|---|---|---|", flush=True)
    with C.Context(device=0) as cx:
        x = 10 ** 12
        y = 2 * icbrt(x)
        ms, t = timed(cx, x, y, args.reps)
        print("| C1 | 1e12 | 2 | %d | 2^19 (512 x 32) | %.2f | %d |" % (y, ms, t["pi"]), flush=True)
    for shape in ((256, 32), (512, 32), (512, 64), (256, 128)):
        bits = 32 * shape[0] * shape[1]
        with C.Context(device=0, seg_threads=shape[0], seg_kwords=shape[1]) as cx:
            x = 10 ** 16
            for alpha in (1, 1.5, 2, 3, 4, 6, 8):
                y = int(alpha * icbrt(x))
                ms, t = timed(cx, x, y, args.reps)
                print("| C2 | 1e16 | %g | %d | 2^%d (%d x %d) | %.2f | %d |"
                      % (alpha, y, round(math.log2(bits)), shape[0], shape[1], ms, t["pi"]), flush=True)
    with C.Context(device=0) as cx:
        x = 10 ** 20
        for alpha in (2, 4, 0):
            y = int(alpha * icbrt(x)) if alpha else 0
            ms, t = timed(cx, x, y, 1)
            print("| C5 | 1e20 | %s | %d | 2^19 (512 x 32) | %.1f | %d |"
                  % (alpha if alpha else "model (%.1f)" % t["alpha"], t["y"], ms, t["pi"]), flush=True)


def profile_run(args):
    """One workload only, for profiler runs: `single` = pi(x) on a warm
    context (default shape), `batch` = the 512-query batched step (with
    --tuning as the timed run's autotune chose it), `copy` = a
    device-to-device copy of --copy-mib MiB (the byte-counter calibration),
    `probe` = the LDS and VALU peak microbenchmarks."""
    if args.profile == "copy":
        import torch
        n = args.copy_mib << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda:0")
        b = torch.empty_like(a)
        a.fill_(1)
        torch.cuda.synchronize()
        for _ in range(args.reps):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"profile": "copy", "bytes": n, "reps": args.reps}), flush=True)
        return
    import cudapix_amd as C
    ctx = C.Context(device=0)
    if args.profile == "probe":
        for _ in range(args.reps):
            lds, valu = ctx.probe_peaks()
            print(json.dumps({"profile": "probe", "lds_GBs": round(lds / 1e9, 1),
                              "valu_lane_ops": round(valu / 1e12, 3)}), flush=True)
        ctx.close()
        return
    if args.profile == "single":
        x = int(args.x)
        ctx.pi(x)
        ms = []
        for _ in range(args.reps):
            t0 = time.time()
            v = ctx.pi(x)
            ms.append((time.time() - t0) * 1000)
        print(json.dumps({"profile": "single", "x": x, "pi": v, "ms": [round(m, 3) for m in ms]}),
              flush=True)
    else:
        xs = load_json("queries_c3.json")["x"]
        if args.tuning != "auto":
            t, k, a, c = args.tuning.split(",")
            ctx.set_tuning(int(t), int(k), float(a), int(c))
        else:
            ctx.autotune(xs)
        ref = ctx.pi_batch(xs)
        ms = []
        for _ in range(args.reps):
            t0 = time.time()
            if ctx.pi_batch(xs) != ref:
                raise SystemExit("batched results changed between steps")
            ms.append((time.time() - t0) * 1000)
        print(json.dumps({"profile": "batch", "queries": len(xs), "ms": [round(m, 3) for m in ms]}),
              flush=True)
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU baseline: the whole brute count below 1e12 and P2 at 1e16 too")
    ap.add_argument("--cpu-only", action="store_true", help="print only the CPU baseline object and exit")
    ap.add_argument("--serve-n", type=int, default=400)
    ap.add_argument("--serve-lambdas", default="200,2000")
    ap.add_argument("--serve-saturate", type=int, default=1,
                    help="also offer the lambda=2000 stream's x list ten times as fast (1, default) or not (0)")
    ap.add_argument("--serve-xmax", type=float, default=1e16)
    ap.add_argument("--serve-full", action="store_true",
                    help="every stream of the served fixture: lambda 50..2000, 2000 requests, x < 1e17")
    ap.add_argument("--profile", choices=["single", "batch", "copy", "probe"],
                    help="run only one workload --reps times (for rocprofv3) and exit")
    ap.add_argument("--x", type=float, default=1e16, help="x of --profile single")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=1024, help="bytes of --profile copy")
    ap.add_argument("--configs", action="store_true",
                    help="time the single-query configurations (1e12, the 1e16 alpha x segment sweep, "
                         "1e20 at two alphas) as markdown rows and exit")
    ap.add_argument("--tuning", default=BENCH_TUNING,
                    help="threads,kwords,alpha_scale,c of the batched runs and --profile batch, or 'auto' "
                         "(autotune; the committed counters then may not match)")
    args = ap.parse_args()
    if args.serve_full:
        args.serve_lambdas, args.serve_n, args.serve_xmax = "50,200,500,1000,2000", 2000, 1e17
    if args.gpus != 1:
        print("bench.py measures one GPU only (the workload shards by query with no exchange); "
              "run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    if args.cpu_only:
        print(json.dumps(cpu_baseline(load_json("queries_c3.json")["x"], full=args.cpu_full)), flush=True)
        return
    if args.profile:
        profile_run(args)
        return
    if args.configs:
        configs_run(args)
        return

    import cudapix_amd as C
    xs = load_json("queries_c3.json")["x"]
    n = len(xs)
    ctx = C.Context(device=0)

    # ---- warmup (builds resident tables; the segment shape, alpha factor and
    # wheel are pinned to the tuning the committed counter profiles were
    # taken at, or autotuned with --tuning auto)
    log("warmup: tuning %s + %d batched steps" % (args.tuning, args.warmup))
    t0 = time.time()
    if args.tuning == "auto":
        ctx.autotune(xs)
    else:
        tt, tk, ta, tc = args.tuning.split(",")
        ctx.set_tuning(int(tt), int(tk), float(ta), int(tc))
    tuning = ctx.tuning()
    shape = (tuning["seg_threads"], tuning["seg_kwords"])
    tuning_key = "%d,%d,%g,%d" % (shape[0], shape[1], tuning["alpha_scale"], tuning["c"])
    log("tuning %s (%.1fs)" % (tuning_key, time.time() - t0))
    ref = ctx.pi_batch(xs, use_graph=True)
    # both batch slots get their arena and launch graphs before the timed run
    for r in ctx.pi_batches([xs] * max(2, args.warmup)):
        assert r == ref

    # ---- timed batched steps, pipelined two deep: a step's batch is planned
    #      and uploaded while the previous step's batch runs on the device
    seg_ms, dev_ms = [], []

    def collect(tk):
        r, st = ctx.pi_batch_collect(tk, stats=True)
        seg_ms.append(st["ms_segments"])
        dev_ms.append(st["ms_device"])
        if r != ref:
            raise SystemExit("batched results changed between steps")

    t0 = time.time()
    prev = None
    for i in range(args.steps):
        tk = ctx.pi_batch_submit(xs, use_graph=True)
        if prev is not None:
            collect(prev)
        prev = tk
    collect(prev)
    dt = time.time() - t0
    qps = n * args.steps / dt
    ms_step = 1000.0 * dt / args.steps
    log("batched (pipelined): %.1f q/s, %.2f ms/step (device %.2f ms, segments %.2f ms)"
        % (qps, ms_step, sum(dev_ms) / len(dev_ms), sum(seg_ms) / len(seg_ms)))
    # the same steps one call at a time (plan, upload, run, wait); their
    # segment-kernel times (no other batch beside it) feed the roofline
    seg_ms = []
    t0 = time.time()
    for i in range(args.steps):
        r, st = ctx.pi_batch(xs, use_graph=True, stats=True)
        seg_ms.append(st["ms_segments"])
        if r != ref:
            raise SystemExit("batched results changed between steps")
    dts = time.time() - t0
    log("batched (one call at a time): %.1f q/s, %.2f ms/step (segments %.2f ms)"
        % (n * args.steps / dts, 1000.0 * dts / args.steps, sum(seg_ms) / len(seg_ms)))

    # terms of the batch, for the roofline's algorithmic bytes
    terms = ctx.pi_batch(xs, terms=True)
    for x, t in zip(xs, terms):
        t["x"] = x
    primes = primes_upto(iroot(max(xs), 2) + 1)
    w_lds = algorithmic_lds_bytes(terms, primes)
    seg_avg_s = (sum(seg_ms) / len(seg_ms)) / 1000.0

    out = {"metric": "pi(x) queries/s batched+served vs plain (x<=1e16), p50/p99 ms, bit-exact",
           "value": round(qps, 2), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "u64/u128 integer",
           "data": "synthetic: 512 x log-uniform in [1e9,1e15] (splitmix64 seed 3)",
           "config": {"workload": "batched pi(x), 512 queries x in [1e9,1e15], one persistent "
                                  "segment launch per step",
                      "queries": n, "seg_threads": shape[0], "seg_kwords": shape[1],
                      "alpha_scale": tuning["alpha_scale"], "wheel_c": tuning["c"],
                      "parallelism": "single GPU"},
           "pipelined": "two batches in flight: each step's planning and upload overlap the previous step's kernels",
           "sync_qps": round(n * args.steps / dts, 2), "sync_ms_per_step": round(1000.0 * dts / args.steps, 3)}

    # ---- plain one-at-a-time solves (fresh tables per query) and warm-table variant
    if not args.no_plain:
        with C.Context(device=0, fresh_tables=True, use_graphs=False,
                       seg_threads=shape[0], seg_kwords=shape[1]) as pc:
            t0 = time.time()
            plain = [pc.pi(x) for x in xs]
            dtp = time.time() - t0
        with C.Context(device=0, use_graphs=False, seg_threads=shape[0],
                       seg_kwords=shape[1]) as wc:
            wc.pi(max(xs))
            t0 = time.time()
            warm = [wc.pi(x) for x in xs]
            dtw = time.time() - t0
        if plain != ref or warm != ref:
            bad = [i for i in range(n) if plain[i] != ref[i] or warm[i] != ref[i]]
            raise SystemExit("batched results differ from plain solves at %s" % bad[:5])
        out["plain_qps"] = round(n / dtp, 2)
        out["plain_warm_qps"] = round(n / dtw, 2)
        out["speedup_vs_plain"] = round(qps / (n / dtp), 2)
        out["speedup_vs_plain_warm"] = round(qps / (n / dtw), 2)
        out["bit_exact_vs_plain"] = True
        log("plain: %.1f q/s (fresh tables), %.1f q/s (warm tables)" % (n / dtp, n / dtw))

    # ---- single large query
    if not args.no_single:
        # a context with the default (single-query) tuning
        sc = C.Context(device=0)
        x16 = 10 ** 16
        sc.pi(x16)
        t0 = time.time()
        t = sc.pi(x16, terms=True)
        d16 = time.time() - t0
        if t["pi"] != 279238341033925:
            raise SystemExit("pi(1e16) wrong: %d" % t["pi"])
        _, st16 = sc.pi_batch([x16], stats=True)
        sc.close()
        t["x"] = x16
        w16 = algorithmic_lds_bytes([t], primes)
        out["single_1e16"] = {"ms": round(d16 * 1000, 2), "y": t["y"], "alpha": round(t["alpha"], 3),
                              "segments_ms": round(st16["ms_segments"], 3),
                              "lds_algorithmic_GB": round(w16 / 1e9, 2),
                              "pi": t["pi"]}
        log("pi(1e16) = %d in %.1f ms" % (t["pi"], d16 * 1000))
        # C1: pi(1e12) at alpha 2 (y = 20000), c = 6, warm context, median of 3
        c1 = C.Context(device=0)
        ms1, v1 = median_time(lambda: c1.pi(10 ** 12, y=20000, c=6))
        c1.close()
        if v1 != 37607912018:
            raise SystemExit("pi(1e12) wrong: %d" % v1)
        out["c1_gpu_ms"] = round(ms1 * 1000, 3)
        log("pi(1e12) at y=20000: %.2f ms" % (ms1 * 1000))

    # ---- served Poisson streams
    if not args.no_serve:
        c4 = load_json("queries_c4.json")["streams"]
        served = {}
        lams = [int(v) for v in args.serve_lambdas.split(",") if v]
        # (key rate, fixture stream, time divisor): the fixture's streams, and
        # with --serve-saturate the lambda = 2000 stream's x list offered ten
        # times as fast, so that the servers run at their capacity rather than
        # at the offered load
        # (every request of that stream up to serve_xmax, batches of up to 512)
        streams = [(lam, lam, 1) for lam in lams]
        if args.serve_saturate and 2000 in lams:
            streams.append((20000, 2000, 10))
        plain_cache = {}
        for lam, src, div in streams:
            s = c4[str(src)]
            pairs = [(tt, xx) for tt, xx in zip(s["t_ms"], s["x"]) if xx <= args.serve_xmax]
            if div == 1:
                pairs = pairs[:args.serve_n]
            t0 = pairs[0][0]
            sx = [p[1] for p in pairs]
            st_ms = [(p[0] - t0) / div for p in pairs]
            # every x of the stream solved alone (one batch), to check each served value
            uniq = sorted(set(sx))
            alone = dict(zip(uniq, ctx.pi_batch(uniq)))
            # the same x list one query at a time: warm tables, and fresh tables per
            # query (each solve builds its own, as the reference's runs do)
            pw, pf = plain_cache.get(tuple(sx), (None, None))
            if pw is None and not args.no_plain:
                with C.Context(device=0, use_graphs=False, seg_threads=shape[0],
                               seg_kwords=shape[1]) as wc:
                    wc.pi(max(sx))
                    t0p = time.time()
                    for x in sx:
                        if wc.pi(x) != alone[x]:
                            raise SystemExit("plain solve of %d differs" % x)
                    pw = len(sx) / (time.time() - t0p)
                with C.Context(device=0, fresh_tables=True, use_graphs=False, seg_threads=shape[0],
                               seg_kwords=shape[1]) as fc:
                    t0p = time.time()
                    for x in sx:
                        if fc.pi(x) != alone[x]:
                            raise SystemExit("plain solve of %d differs" % x)
                    pf = len(sx) / (time.time() - t0p)
                plain_cache[tuple(sx)] = (pw, pf)
            for mode, key in ((C.SERVER_STREAM, str(lam)), (C.SERVER_BATCHES, "%d_batches" % lam)):
                srv = C.Server(ctx, max_x=max(sx), mode=mode, max_batch=256 if div == 1 else 512)
                tickets = srv.replay(sx, st_ms)
                srv.drain()
                res = srv.poll(len(sx) + 16)
                stt = srv.stats()
                srv.close()
                got = {r["ticket"]: r for r in res}
                ok = all(r["status"] == 0 for r in res) and len(res) == len(sx)
                matches = sum(1 for tk, x in zip(tickets, sx)
                              if tk in got and got[tk]["status"] == 0 and got[tk]["pi"] == alone[x])
                if matches != len(sx):
                    raise SystemExit("served lambda=%d mode %d: %d of %d results differ from the same x "
                                     "solved alone" % (lam, mode, len(sx) - matches, len(sx)))
                waits = sorted(r["t_admit_ms"] - r["t_arrive_ms"] for r in res)
                qw = (lambda f: round(waits[min(len(waits) - 1, int(f * len(waits)))], 3)) if waits else \
                    (lambda f: None)
                served[key] = {"mode": "stream" if mode == C.SERVER_STREAM else "batches",
                               "lambda": lam, "n": len(sx), "qps": round(stt["qps"], 2),
                               "arrivals": ("the lambda=%d stream" % src) if div == 1 else
                                           ("the lambda=%d stream's times / %d, batches up to 512" % (src, div)),
                               "p50_ms": round(stt["p50_ms"], 3), "p99_ms": round(stt["p99_ms"], 3),
                               "mean_ms": round(stt["mean_ms"], 3),
                               "queue_p50_ms": qw(0.5), "queue_p99_ms": qw(0.99),
                               "submitted": stt["submitted"], "completed": stt["completed"],
                               "rejected": stt["rejected"], "batches": stt["batches"],
                               "max_inflight_bytes": stt["max_inflight_bytes"],
                               "launches": stt["launches"], "joined": stt["joined"],
                               "all_ok": ok, "matches_alone": matches,
                               "plain_warm_qps_same_x": round(pw, 2) if pw else None,
                               "speedup_vs_plain_warm": round(stt["qps"] / pw, 2) if pw else None,
                               "plain_qps_same_x": round(pf, 2) if pf else None,
                               "speedup_vs_plain": round(stt["qps"] / pf, 2) if pf else None}
                log("served lambda=%d: %s" % (lam, served[key]))
        out["served"] = served

    # ---- roofline of the persistent segment kernel.  The counters show VALU
    # issue as its busiest resource (LDS and HBM far below their peaks), so
    # the bound is VALU: instructions per launch (rocprofv3 SQ_INSTS_VALU on
    # the same workload and tuning, committed) over the live kernel time.
    # W_LDS and W_HBM are models, reported beside the counted bytes.
    lds_probe, int_probe = ctx.probe_peaks()
    valu_probe_g = int_probe / 64.0 / 1e9              # lane ops -> wave64 instructions
    pc = profiled_counters("batch")
    same = pc.get("tuning") == tuning_key
    instr = pc.get("valu_instr_per_launch") if same else None
    ach = instr / seg_avg_s / 1e9 if instr else None
    traffic = int(pc["hbm_bytes_per_launch"]) if same and "hbm_bytes_per_launch" in pc else None
    lds_counted = pc.get("lds_bytes_per_launch") if same else None
    w_hbm, hbm_parts = hbm_model_bytes(terms, primes, 32 * shape[0] * shape[1])
    out["roofline"] = {
        "bound": "valu", "achieved": round(ach, 1) if ach else None, "peak": round(VALU_PEAK_G, 1),
        "unit": "G wave64-instr/s", "frac": round(ach / VALU_PEAK_G, 4) if ach else None,
        "traffic": traffic,
        "kernel": "k_segments (persistent: hard-leaf + P2 segment tiles, easy-leaf and P2-lookup tiles)",
        "kernel_ms": round(seg_avg_s * 1000, 3),
        "valu_instr_per_launch": instr,
        "peak_source": "2 wave64 32-bit VALU instructions per CU-clock (MI355X FP32 vector peak 157.3 TFLOP/s "
                       "= 256 CUs x 2.4 GHz x 128 FMA lanes)",
        "probe_peak": round(valu_probe_g, 1),
        "frac_vs_probe": round(ach / valu_probe_g, 4) if ach else None,
        "probe_source": "cpx_probe_peaks: 8 independent add/shift/xor chains per lane, 24 VALU per iteration "
                        "(SQ_INSTS_VALU matches; profiles/peaks.md), measured live",
        "counters_source": "profiles/k_segments_batch_pmc.md at tuning %s%s"
                           % (pc.get("tuning"), "" if same else " (differs from this run: not used)"),
        "components": {
            "lds": {"counted_bytes": lds_counted,
                    "counted_GBs": round(lds_counted / seg_avg_s / 1e9, 1) if lds_counted else None,
                    "model_bytes": w_lds,
                    "model": "W_LDS = B_h*ceil(z/2)/8 + 8*sum ceil((z/2)/p_b) per query group (one "
                             "popcount pass per stage + one RMW per cross-off); the kernel does less "
                             "(no per-stage pass), so W_LDS over time is not a bandwidth",
                    "peak_GBs": round(LDS_PEAK_GBS, 1), "probe_GBs": round(lds_probe / 1e9, 1),
                    "frac_counted": round(lds_counted / seg_avg_s / 1e9 / LDS_PEAK_GBS, 4) if lds_counted else None},
            "hbm": {"counted_bytes": traffic,
                    "counted_GBs": round(traffic / seg_avg_s / 1e9, 1) if traffic else None,
                    "model_bytes": w_hbm, "model_parts": hbm_parts,
                    "peak_GBs": HBM_PEAK_GBS,
                    "frac_counted": round(traffic / seg_avg_s / 1e9 / HBM_PEAK_GBS, 4) if traffic else None}}}
    if "single_1e16" in out:
        s16 = out["single_1e16"]
        s16["lds_model_GBs"] = round(s16["lds_algorithmic_GB"] / (s16["segments_ms"] / 1000.0), 1)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(xs, full=args.cpu_full)
        log("cpu baseline: %s" % out["cpu_baseline"])

    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
