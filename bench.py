"""Benchmark: error-corrected SHGEMM at m = n = k = 16384 on one MI355X.

One step = one mtk::shgemm call (op_n/op_n, fp16 compute type, alpha=1,
beta=0) on A FP32 U[-1,1] (16384 x 16384) and B FP16 U[-1,1], operands
resident in HBM.  Reported beside it, on the same inputs and GPU:
  * the library's plain FP32 SGEMM (the comparator of the metric), its
    throughput and the ratio;
  * ||C - C64||_F / ||C64||_F of both against an FP64 product over a sample
    of rows (numpy float64 on the host; FP64 is used only for this check);
  * roofline of the SHGEMM kernel from HIP events on its own stream;
  * the host FP32 GEMM (CPU baseline): the whole product when it fits the
    time budget, else a bounded block of rows; and the 1024^3 and sketch
    shapes timed in full;
  * a short sketch-serving run (p50 / p99 latency at 2000 req/s offered).

Prints one JSON line.  Multi-GPU runs are not part of this build.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP16_DENSE = 2516.6   # TFLOP/s, MI355X FP16 dense MFMA
PEAK_FP32_MATRIX = 157.3   # TFLOP/s, MI355X FP32 matrix
PEAK_HBM = 8.0             # TB/s
NOMINAL_CLOCK_MHZ = 2400.0  # GFX clock of the published MFMA peaks (256 CUs x 4 SIMDs x 1024 FP16 FLOP/clock)
P99_BOUND_MS = 10.0        # stated sketch-server latency bound (p99, ms) up to the reported rate


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["square", "grouped", "serve", "configs", "bands", "energy"], default="square",
                    help="square: m=n=k=--size (the headline line); grouped: 512 uneven problems in one group; "
                         "serve: sketch server under Poisson arrivals; configs: the 1024^3, sketch and "
                         "tall-skinny configurations, one JSON line each; bands: error by exponent band of A; "
                         "energy: W and J per GEMM from the card's energy counter")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=16384)
    ap.add_argument("--sample-rows", type=int, default=256, help="rows of C checked against FP64")
    ap.add_argument("--cpu-seconds", type=float, default=25.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="square: skip the short serving run")
    ap.add_argument("--no-power", action="store_true",
                    help="square, configs, grouped: skip the 2 s clock and power window after the timed calls")
    ap.add_argument("--sketch-data", action="store_true",
                    help="energy: only the sketch shape, on A = U[-1,1) and on A = 0 (the data dependence of the clock)")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="square: offered rate of the serving run")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="square: arrival window of the serving run")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 1 GiB device copies (reference for FETCH_SIZE/WRITE_SIZE scaling)")
    ap.add_argument("--problems", type=int, default=512, help="grouped workload: number of problems")
    ap.add_argument("--group-only", action="store_true", help="grouped: skip the per-problem comparison launches")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--rates", default="200,500,1000,2000", help="serve: offered request rates (req/s)")
    ap.add_argument("--seconds", type=float, default=3.0, help="serve: arrival window per rate; energy: measuring window per GEMM")
    ap.add_argument("--lanes", type=int, default=4, help="serve: concurrent lanes (handle + stream each)")
    ap.add_argument("--max-batch", type=int, default=1,
                    help="serve: requests admitted together to one lane run as one grouped launch (up to this many)")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serve: HBM admission budget")
    return ap.parse_args()


def _frac_at_clock(flops, byts, ms, power):
    clk = power.get("median_gfxclk_mhz") if power else None
    if not isinstance(clk, (int, float)) or clk <= 0:
        return None
    t_mfma = flops / (PEAK_FP16_DENSE / 2.0 * 1e12) * NOMINAL_CLOCK_MHZ / clk
    return round(max(t_mfma, byts / (PEAK_HBM * 1e12)) / (ms * 1e-3), 4)


def power_window(fn, seconds=2.0):
    """GFX clock, socket power and energy per call while `fn` repeats for
    about `seconds` (shgemm_amd.energy: the card's energy accumulator around
    the window, the metrics table sampled every 50 ms); None when ROCm SMI is
    not importable.  The metrics table refreshes too slowly for a window of a
    few calls, so this is a window of its own after the timed steps."""
    try:
        import torch

        from shgemm_amd.energy import EnergyMeter, measure

        meter = EnergyMeter()
    except Exception:
        return None
    try:
        r = measure(meter, fn, seconds, torch.cuda.synchronize)
    finally:
        meter.close()
    return {"median_gfxclk_mhz": r["median_gfxclk_mhz"], "median_socket_power_w": r["median_socket_power_w"],
            "avg_power_w": round(r["avg_power_w"], 1), "energy_per_call_j": round(r["energy_per_call_j"], 4),
            "calls": r["count"], "window_s": round(r["window_s"], 2),
            "source": "ROCm SMI, a window of the same call repeated after the timed steps: the card's energy "
                      "accumulator around it, GFX clock and socket power medians of the metrics table"}


def timed_launches(fn, steps, warmup, stream, after_warmup=None):
    """Run warmup untimed launches, then `steps` launches each bracketed by HIP
    events on `stream` (the stream the kernels run on).  Returns (event time
    of the whole timed region, its wall ms, mean per-launch ms from the
    events).  after_warmup() runs between the two phases."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    if after_warmup is not None:
        after_warmup()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for s, e in evs:
        s.record(stream)
        fn()
        e.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3
    per = [s.elapsed_time(e) for s, e in evs]
    region = evs[0][0].elapsed_time(evs[-1][1])
    return region, wall, sum(per) / len(per)


def fp64_sample_error(A_rows, B, C_rows):
    import numpy as np

    ref = A_rows.astype(np.float64) @ B.astype(np.float64)
    d = np.linalg.norm(C_rows.astype(np.float64) - ref)
    return float(d / np.linalg.norm(ref))


def _median_time(fn, reps):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2]


def cpu_baseline(A_np, B32_np, m, n, k, budget_s):
    """Host FP32 GEMM (oracle.sgemm_nn_packed: packed panels, 16 x 6 AVX2 FMA
    register tiles, every host thread) on the headline problem: the whole
    product when a first block of rows predicts it fits the budget, else row
    blocks until the budget is used (rate = 2 * rows * n * k / t)."""
    import numpy as np

    import oracle

    threads = oracle.max_threads()
    out = np.zeros(m * n, dtype=np.float32)
    run = lambda r0, r1: oracle.sgemm_nn_packed(m, n, k, 1.0, A_np, m, B32_np, k, 0.0, None, m, r0, r1, out=out,
                                                threads=threads)
    rows0 = min(m, 192 * threads)
    t0 = time.perf_counter()
    run(0, rows0)
    t = time.perf_counter() - t0
    done = rows0
    if t * m / rows0 <= budget_s:
        t0 = time.perf_counter()
        run(0, m)
        t, done = time.perf_counter() - t0, m
    else:
        while t < budget_s and done < m:
            r1 = min(m, done + rows0)
            t0 = time.perf_counter()
            run(done, r1)
            t += time.perf_counter() - t0
            done = r1
    return 2.0 * done * n * k / t / 1e12, threads, done, t


def cpu_configs(A_np, B32_np, N):
    """The same host GEMM timed in full on the 1024^3 configuration and on the
    sketch shape 16384 x 256 x 16384 (median of 3), on blocks of the headline
    operands (the time does not depend on the values)."""
    import numpy as np

    import oracle

    threads = oracle.max_threads()
    res = {}
    for name, m, n, k in (("cfg1_1024x1024x1024", 1024, 1024, 1024), ("sketch_16384x256x16384", N, 256, N)):
        A = A_np[:m * k]
        B = B32_np[:k * n]
        out = np.zeros(m * n, dtype=np.float32)
        fn = lambda: oracle.sgemm_nn_packed(m, n, k, 1.0, A, m, B, k, 0.0, None, m, out=out, threads=threads)
        fn()
        t = _median_time(fn, 3)
        res[name] = {"ms": round(t * 1e3, 3), "tflops": round(2.0 * m * n * k / t / 1e12, 4)}
    return res


def grouped_problems(count, seed):
    """The batched workload: m log-uniform in [256, 8192], n in {64, 128, 256,
    512}, k log-uniform in [256, 16384] (seeded)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    m = np.exp(rng.uniform(np.log(256), np.log(8192), count)).astype(int)
    k = np.exp(rng.uniform(np.log(256), np.log(16384), count)).astype(int)
    n = rng.choice([64, 128, 256, 512], count)
    return [(int(a), int(b), int(c)) for a, b, c in zip(m, n, k)]


def bench_grouped(args):
    """One step = the whole batch.  Timed three ways on the same operands:
    the persistent grouped launch (one graph replay), one shgemm launch per
    problem, and those per-problem launches captured in one graph."""
    import numpy as np
    import torch

    import shgemm_amd as sg

    dev = torch.device("cuda:0")
    shapes = grouped_problems(args.problems, args.seed)
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed)
    As, Bs, Cs = [], [], []
    for m, n, k in shapes:
        As.append(torch.rand(m * k, device=dev, generator=g) * 2 - 1)
        Bs.append((torch.randn(k * n, device=dev, generator=g) / float(np.sqrt(n))).half())
        Cs.append(torch.empty(m * n, device=dev))
    probs = [(m, n, k, A, m, B, k, C, m, 1.0, 0.0) for (m, n, k), A, B, C in zip(shapes, As, Bs, Cs)]
    flops = sum(2.0 * m * n * k for m, n, k in shapes)
    byts = sum(4.0 * m * k + 2.0 * k * n + 4.0 * m * n for m, n, k in shapes)

    stream = torch.cuda.current_stream()
    h = sg.create()
    sg.set_cuda_stream(h, stream)
    grp = sg.Group(h, sg.op_n, sg.op_n, probs, sg.fp16)
    cls, slices, units = grp.plan()
    run_group = lambda: grp.run(stream)
    g_region, g_wall, g_ms = timed_launches(run_group, args.steps, args.warmup, stream)
    g_power = None if args.no_power else power_window(run_group)

    def per_problem():
        for m, n, k, A, lda, B, ldb, C, ldc, al, be in probs:
            sg.shgemm(h, sg.op_n, sg.op_n, m, n, k, al, A, lda, B, ldb, be, C, ldc, sg.fp16)

    p_region, pg_ms = None, None
    if not args.group_only:
        p_region, _, p_ms = timed_launches(per_problem, max(2, args.steps // 2), 1, stream)
    # the same per-problem launches replayed from one captured graph
    try:
        if args.group_only:
            raise RuntimeError("skipped (--group-only)")
        cs = torch.cuda.Stream()
        sg.set_cuda_stream(h, cs)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=cs):
            per_problem()
        sg.set_cuda_stream(h, stream)
        _, _, pg_ms = timed_launches(lambda: graph.replay(), max(2, args.steps // 2), 1, stream)
    except Exception as e:  # capture unsupported: report as not measured
        print(f"bench.py: per-problem graph capture failed: {e}", file=sys.stderr)
        sg.set_cuda_stream(h, stream)

    # accuracy: FP64 error of sampled rows of a few problems, beside SGEMM's
    run_group()
    torch.cuda.synchronize()
    rng = np.random.default_rng(1)
    errs = []
    for i in rng.choice(len(shapes), size=min(8, len(shapes)), replace=False):
        m, n, k = shapes[i]
        rows = np.sort(rng.choice(m, size=min(64, m), replace=False))
        ridx = torch.from_numpy(rows).to(dev)
        B32 = Bs[i].float()
        Cs32 = torch.empty(m * n, device=dev)
        sg.sgemm_fp32(h, sg.op_n, sg.op_n, m, n, k, 1.0, As[i], m, B32, k, 0.0, Cs32, m)
        torch.cuda.synchronize()
        A_rows = As[i].view(k, m).t()[ridx].cpu().numpy()
        Bh = B32.view(n, k).t().cpu().numpy()
        e = fp64_sample_error(A_rows, Bh, Cs[i].view(n, m).t()[ridx].cpu().numpy())
        es = fp64_sample_error(A_rows, Bh, Cs32.view(n, m).t()[ridx].cpu().numpy())
        errs.append(e / es)
    grp.destroy()
    sg.destroy(h)

    ms_per_step = g_region / args.steps
    value = flops / (ms_per_step * 1e-3) / 1e12
    t_roof = max(flops / (PEAK_FP16_DENSE / 2.0 * 1e12), byts / (PEAK_HBM * 1e12))
    hist = {c: int(sum(1 for x in cls if x == c)) for c in (0, 1, 2, 3)}
    per_class = {}
    for c in (0, 1, 2, 3):
        sel = [i for i, x in enumerate(cls) if x == c]
        per_class[c] = {"problems": len(sel),
                        "gflop": round(sum(2.0 * shapes[i][0] * shapes[i][1] * shapes[i][2] for i in sel) / 1e9, 1),
                        "gbytes": round(sum(4.0 * shapes[i][0] * shapes[i][2] + 2.0 * shapes[i][2] * shapes[i][1]
                                            + 4.0 * shapes[i][0] * shapes[i][1] for i in sel) / 1e9, 3)}
    line = {
        "metric": "eff. TFLOP/s, batched uneven SHGEMM (one grouped launch)",
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16 MFMA (f32 split hi+lo) / f32 accumulate",
        "data": "synthetic (seeded torch RNG on device)",
        "config": {"workload": f"grouped shgemm, {len(shapes)} problems, m logU[256,8192], n in {{64..512}}, "
                               f"k logU[256,16384], A U[-1,1], B N(0,1/n), op_n/op_n fp16",
                   "problems": len(shapes), "total_flops": flops, "total_bytes": byts,
                   "work_units": units, "tile_classes": hist, "per_class": per_class,
                   "classes_side_by_side": os.environ.get("MTK_SHGEMM_GROUP_CONCURRENT", "0") == "1",
                   "split_problems": int(sum(1 for s in slices if s > 1)), "parallelism": "single GPU"},
        "comparator": None if p_region is None else {
            "per_problem_launch_ms": round(p_region / max(2, args.steps // 2), 4),
            "per_problem_launch_tflops": round(flops / (p_region / max(2, args.steps // 2) * 1e-3) / 1e12, 3),
            "per_problem_graph_ms": round(pg_ms, 4) if pg_ms else None,
            "grouped_speedup_vs_per_problem": round(p_region / max(2, args.steps // 2) / ms_per_step, 3)},
        "error": {"max_ratio_vs_sgemm": max(errs), "sampled_problems": len(errs), "rows_per_problem": 64},
        "roofline": {"bound": "mfma" if flops / (PEAK_FP16_DENSE / 2.0 * 1e12) > byts / (PEAK_HBM * 1e12) else "hbm",
                     "achieved": round(flops / (g_ms * 1e-3) / 1e12, 3), "peak": round(PEAK_FP16_DENSE / 2.0, 1),
                     "unit": "TFLOP/s", "frac": round(t_roof * 1e3 / g_ms, 4),
                     "frac_at_measured_clock": _frac_at_clock(flops, byts, g_ms, g_power),
                     "traffic": None,
                     "note": "frac = batch roofline time max(flops/1258 TF, bytes/8 TB/s) / measured graph time; "
                             "frac_at_measured_clock scales the MFMA term to the GFX clock of the batch's power window"},
        "clock_and_power": g_power,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


class _ServeSetup:
    """Device-resident request pools (four column-major 32768 x 16384 FP32
    matrices, U[-1,1]) and the lanes, with a request maker: m logU[1024,32768],
    k logU[1024,16384], l in {64..1024}, A a block of one pool."""

    ROWS, COLS, NPOOL = 32768, 16384, 4

    def __init__(self, seed, lanes_n, max_batch=1):
        import numpy as np
        import torch

        from shgemm_amd.serve import GpuLanes

        dev = torch.device("cuda:0")
        g = torch.Generator(device=dev)
        g.manual_seed(seed)
        self.pool = [torch.rand(self.ROWS * self.COLS, device=dev, generator=g) * 2 - 1 for _ in range(self.NPOOL)]
        self.rng = np.random.default_rng(seed)
        self.max_batch = max_batch
        self.lanes = GpuLanes(lanes_n, max_batch=max_batch if max_batch > 1 else 0)
        self.lanes.reserve(self.ROWS, self.COLS, 1024)

    def make(self, n, rid0):
        from shgemm_amd.serve import SketchRequest, sketch_shapes

        out = []
        for i, (m, k, l) in enumerate(sketch_shapes(n, self.rng)):
            p = (rid0 + i) % self.NPOOL
            r0 = int(self.rng.integers(0, self.ROWS - m + 1)) // 4 * 4
            c0 = int(self.rng.integers(0, self.COLS - k + 1))
            addr = self.pool[p].data_ptr() + 4 * (c0 * self.ROWS + r0)
            out.append(SketchRequest(rid=rid0 + i, m=m, k=k, l=l, seed=rid0 + i, a=addr, lda=self.ROWS))
        return out

    def warm(self, slots, budget):
        import torch

        from shgemm_amd.serve import SketchServer, run_arrivals

        srv = SketchServer(hbm_budget=budget, slots=slots, lanes=self.lanes, max_batch=self.max_batch)
        run_arrivals(srv, self.make(4 * slots, 10 ** 6), [0.0] * (4 * slots))
        torch.cuda.synchronize()

    def close(self):
        self.lanes.close()
        self.pool = []


def serve_summary(seed=4, rate=2000.0, seconds=2.0, lanes_n=4, budget_gb=200.0, max_batch=1):
    """A short serving run for the headline line: Poisson arrivals at `rate`
    for `seconds`; p50 / p99 latency (arrival to observed completion) and the
    sustained rate."""
    from shgemm_amd.serve import SketchServer, poisson_offsets, run_arrivals

    su = _ServeSetup(seed, lanes_n, max_batch)
    su.warm(lanes_n, budget_gb * 1e9)
    n = max(20, int(rate * seconds))
    srv = SketchServer(hbm_budget=budget_gb * 1e9, slots=lanes_n, lanes=su.lanes, max_batch=max_batch)
    st = run_arrivals(srv, su.make(n, 0), poisson_offsets(n, rate, su.rng))
    su.close()
    return {"workload": "sketch server, Poisson arrivals, m logU[1024,32768], k logU[1024,16384], "
                        "l in {64..1024}, fp16, Omega from the request seed on device",
            "offered_req_s": rate, "requests": st["requests"], "sustained_req_s": round(st["sustained_req_s"], 1),
            "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
            "p99_bound_ms": P99_BOUND_MS, "p99_within_bound": bool(st["p99_ms"] <= P99_BOUND_MS),
            "lanes": lanes_n, "max_batch": max_batch, "mean_batch": round(st["mean_batch"], 2),
            "hbm_budget_gb": budget_gb}


def bench_serve(args):
    """Sketch serving: Poisson arrivals at each offered rate; per request an
    FP32 A (m logU[1024,32768] x k logU[1024,16384]) resident in HBM (a block
    of one of four 32768 x 16384 matrices), l in {64..1024}, Omega generated on
    the device from the request seed; p50 / p99 latency (arrival to observed
    completion) and the sustained rate."""
    import torch

    from shgemm_amd.serve import SketchServer, poisson_offsets, run_arrivals

    su = _ServeSetup(args.seed, args.lanes, args.max_batch)
    lanes, make, rng = su.lanes, su.make, su.rng
    # warm-up: every lane, plan lookups and the allocator
    su.warm(args.lanes, args.budget_gb * 1e9)
    results = []
    for rate in [float(x) for x in args.rates.split(",") if x]:
        n = max(20, int(rate * args.seconds))
        reqs = make(n, len(results) * 10 ** 5)
        srv = SketchServer(hbm_budget=args.budget_gb * 1e9, slots=args.lanes, lanes=lanes, max_batch=args.max_batch)
        st = run_arrivals(srv, reqs, poisson_offsets(n, rate, rng))
        st["offered_rate"] = rate
        st["mean_request_gflop"] = sum(r.flops for r in reqs) / n / 1e9
        results.append(st)
        print(f"bench.py serve: rate {rate:.0f}/s -> p50 {st['p50_ms']:.2f} ms p99 {st['p99_ms']:.2f} ms "
              f"sustained {st['sustained_req_s']:.0f}/s peak lanes {st['peak_inflight']}", file=sys.stderr, flush=True)
        del reqs
    # saturation: every request present at time zero
    n = max(40, int(max(r["offered_rate"] for r in results) * args.seconds / 2))
    reqs = make(n, 9 * 10 ** 5)
    srv = SketchServer(hbm_budget=args.budget_gb * 1e9, slots=args.lanes, lanes=lanes, max_batch=args.max_batch)
    sat = run_arrivals(srv, reqs, [0.0] * n)
    su.close()
    ok = [r for r in results if r["sustained_req_s"] >= 0.95 * r["offered_rate"]]
    head = max(ok, key=lambda r: r["offered_rate"]) if ok else results[0]
    line = {
        "metric": "sketch p99 latency at sustained req/s (Y = A*Omega, continuous batching)",
        "value": round(head["p99_ms"], 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": int(sum(r["requests"] for r in results)),
        "warmup": 4 * args.lanes,
        "ms_per_step": round(head["mean_ms"], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16 MFMA (f32 split hi+lo) / f32 accumulate",
        "data": "synthetic: A resident (seeded torch RNG), Omega from request seed on device",
        "config": {"workload": "sketch server, Poisson arrivals, m logU[1024,32768], k logU[1024,16384], "
                               "l in {64,128,256,512,1024}, fp16 compute type",
                   "lanes": args.lanes, "max_batch": args.max_batch, "hbm_budget_gb": args.budget_gb,
                   "seconds_per_rate": args.seconds,
                   "parallelism": "single GPU"},
        "sustained_req_s": round(head["sustained_req_s"], 1),
        "offered_req_s": head["offered_rate"],
        "p50_ms": round(head["p50_ms"], 3),
        # the stated bound: p99 within P99_BOUND_MS at every offered rate the
        # server sustains (>= 95% of the offered rate)
        "p99_bound_ms": P99_BOUND_MS,
        "p99_within_bound": all(r["p99_ms"] <= P99_BOUND_MS for r in ok) and bool(ok),
        "per_rate": [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in r.items()} for r in results],
        "saturation": {"requests": sat["requests"], "req_s": round(sat["sustained_req_s"], 1),
                       "mean_batch": round(sat["mean_batch"], 2),
                       "tflops": round(sat["flops_per_s"] / 1e12, 2)},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def _time_ms(fn, iters, stream, warm_ms=100.0):
    """Median per-call milliseconds over 3 batches of `iters` calls (HIP events
    on `stream`), after untimed batches of `iters` calls until at least
    `warm_ms` of GPU time has passed: the GPU idles while the host computes
    the previous line's errors, and its clock takes tens of milliseconds of
    work to come back (a 0.34 ms call measured 30% slower after 10 calls)."""
    import statistics

    import torch

    warm = 0.0
    while True:
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        for _ in range(iters):
            fn()
        e.record(stream)
        torch.cuda.synchronize()
        warm += s.elapsed_time(e)
        if warm >= warm_ms:
            break
    out = []
    for _ in range(3):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        for _ in range(iters):
            fn()
        e.record(stream)
        torch.cuda.synchronize()
        out.append(s.elapsed_time(e) / iters)
    return statistics.median(out)


def bench_configs(args):
    """The remaining configurations, each timed on its own inputs with the
    plan the library picks, its FP64 error on sampled rows beside the
    library SGEMM's, and its roofline fraction:
      cfg1    1024^3 op_n/op_n, A U[-1,1], B FP16 U[-1,1]          (MFMA-bound, launch-latency dominated)
      sketch  A 16384^2 = sign*U[1,2)*2^e, e in [-15,15]; Omega 16384x256 FP16 N(0,1)/16
              shgemm nn, shgemm tn (A stored transposed), hsgemm (Y^T = Omega^T A^T); fp16 and tf32
      cfg3    m = n = 128, k = 2^20, A U[-1,1], B FP16 N(0,1)      (HBM-bound: 4mk + 2kn + 4mn bytes)"""
    import numpy as np
    import torch

    import shgemm_amd as sg

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    h = sg.create()
    sg.set_cuda_stream(h, stream)
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed)
    rng = np.random.default_rng(args.seed)

    def emit(name, m, n, k, ms, err, err_s, extra, power=None):
        flops = 2.0 * m * n * k
        byts = 4.0 * m * k + 2.0 * k * n + 4.0 * m * n
        t_mfma = flops / (PEAK_FP16_DENSE / 2.0 * 1e12) * (2.5 if extra.get("compute_type") == "tf32" else 1.0)
        t_hbm = byts / (PEAK_HBM * 1e12)
        line = {"config": name, "m": m, "n": n, "k": k, "ms": round(ms, 5),
                "tflops_eff": round(flops / (ms * 1e-3) / 1e12, 2),
                "gbps_alg": round(byts / (ms * 1e-3) / 1e9, 1),
                "bound": "mfma" if t_mfma > t_hbm else "hbm",
                "roofline_frac": round(max(t_mfma, t_hbm) / (ms * 1e-3), 4),
                "err": err, "sgemm_err": err_s, "err_ratio": err / err_s}
        clk = power.get("median_gfxclk_mhz") if power else None
        if isinstance(clk, (int, float)) and clk > 0:
            # the MFMA term at the clock the card ran this call at (HBM's does not scale)
            line["roofline_frac_at_measured_clock"] = round(
                max(t_mfma * NOMINAL_CLOCK_MHZ / clk, t_hbm) / (ms * 1e-3), 4)
            line["clock_and_power"] = power
        line.update(extra)
        print(json.dumps(line), flush=True)

    def rows_err(A_rows, B64, C_rows):
        ref = A_rows.astype(np.float64) @ B64
        return float(np.linalg.norm(C_rows.astype(np.float64) - ref) / np.linalg.norm(ref))

    # ---- cfg1 ---------------------------------------------------------------
    m = n = k = 1024
    A = torch.rand(m * k, device=dev, generator=g) * 2 - 1
    B = (torch.rand(k * n, device=dev, generator=g) * 2 - 1).half()
    C = torch.empty(m * n, device=dev)
    Cs = torch.empty(m * n, device=dev)
    ms = _time_ms(lambda: sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, C, m, sg.fp16), 50, stream)
    pw = None if args.no_power else power_window(lambda: sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, C, m, sg.fp16))
    sg.sgemm_fp32(h, 0, 0, m, n, k, 1.0, A, m, B.float(), k, 0.0, Cs, m)
    torch.cuda.synchronize()
    An = A.view(k, m).t().cpu().numpy()
    Bn = B.float().view(n, k).t().cpu().numpy().astype(np.float64)
    plan = sg.query_plan(h, 0, 0, m, n, k, sg.fp16)
    emit("cfg1 1024^3 nn fp16", m, n, k, ms, rows_err(An, Bn, C.view(n, m).t().cpu().numpy()),
         rows_err(An, Bn, Cs.view(n, m).t().cpu().numpy()),
         {"compute_type": "fp16", "plan": [sg.PLAN_VARIANTS[plan[0]], plan[1]]}, pw)
    del A, B, C, Cs

    # ---- sketch (wide exponent) -----------------------------------------------
    m, k, l = 16384, 16384, 256
    e = torch.randint(-15, 16, (m * k,), device=dev, generator=g).float()
    sgn = torch.where(torch.rand(m * k, device=dev, generator=g) < 0.5, -1.0, 1.0)
    A = sgn * (1.0 + torch.rand(m * k, device=dev, generator=g)) * torch.exp2(e)
    del e, sgn
    At = A.view(k, m).t().contiguous().view(-1)   # A stored transposed (k x m col-major = m x k row-major)
    Om = (torch.randn(k * l, device=dev, generator=g) / 16.0).half()
    Om32 = Om.float()
    rows = np.sort(rng.choice(m, size=256, replace=False))
    ridx = torch.from_numpy(rows).to(dev)
    A_rows = A.view(k, m).t()[ridx].cpu().numpy()
    Om64 = Om32.view(l, k).t().cpu().numpy().astype(np.float64)
    Ys = torch.empty(m * l, device=dev)
    sg.sgemm_fp32(h, 0, 0, m, l, k, 1.0, A, m, Om32, k, 0.0, Ys, m)
    torch.cuda.synchronize()
    err_s = rows_err(A_rows, Om64, Ys.view(l, m).t()[ridx].cpu().numpy())
    Y = torch.empty(m * l, device=dev)
    for ct, ctn in ((sg.fp16, "fp16"), (sg.tf32, "tf32")):
        for variant in ("shgemm nn", "shgemm tn", "hsgemm"):
            if variant == "shgemm nn":
                call = lambda: sg.shgemm(h, 0, 0, m, l, k, 1.0, A, m, Om, k, 0.0, Y, m, ct)
                q = sg.query_plan(h, 0, 0, m, l, k, ct)
            elif variant == "shgemm tn":
                call = lambda: sg.shgemm(h, 1, 0, m, l, k, 1.0, At, k, Om, k, 0.0, Y, m, ct)
                q = sg.query_plan(h, 1, 0, m, l, k, ct)
            else:  # Y^T (l x m) = Omega^T A^T: hsgemm(op_t, op_t) with A16 = Omega, B32 = A
                call = lambda: sg.hsgemm(h, 1, 1, l, m, k, 1.0, Om, k, A, m, 0.0, Y, l, ct)
                q = sg.query_plan(h, 0, 0, m, l, k, ct, 1)
            ms = _time_ms(call, 10, stream)
            pw = None if args.no_power or variant != "shgemm nn" else power_window(call)
            torch.cuda.synchronize()
            Yh = Y.view(m, l)[ridx] if variant == "hsgemm" else Y.view(l, m).t()[ridx]
            emit(f"sketch {variant} {ctn}", m, l, k, ms, rows_err(A_rows, Om64, Yh.cpu().numpy()), err_s,
                 {"compute_type": ctn, "plan": [sg.PLAN_VARIANTS[q[0]], q[1]], "prescaled": ctn == "fp16"}, pw)
    del A, At, Om, Om32, Y, Ys
    torch.cuda.empty_cache()

    # ---- cfg3 tall-skinny ------------------------------------------------------
    m = n = 128
    k = 1 << 20
    A = torch.rand(m * k, device=dev, generator=g) * 2 - 1
    B = torch.randn(k * n, device=dev, generator=g).half()
    C = torch.empty(m * n, device=dev)
    Cs = torch.empty(m * n, device=dev)
    for scaling in (True, False):
        sg.set_exponent_scaling(h, scaling)
        ms = _time_ms(lambda: sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, C, m, sg.fp16), 20, stream)
        pw = None if args.no_power or not scaling else power_window(
            lambda: sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, C, m, sg.fp16))
        runs = []
        for _ in range(10):
            sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, C, m, sg.fp16)
            runs.append(C.clone())
        torch.cuda.synchronize()
        identical = all(torch.equal(runs[0].view(torch.int32), r.view(torch.int32)) for r in runs[1:])
        if scaling:
            sg.sgemm_fp32(h, 0, 0, m, n, k, 1.0, A, m, B.float(), k, 0.0, Cs, m)
            torch.cuda.synchronize()
            An = A.view(k, m).t().cpu().numpy()
            Bn = B.float().view(n, k).t().cpu().numpy().astype(np.float64)
            err_s = rows_err(An, Bn, Cs.view(n, m).t().cpu().numpy())
        q = sg.query_plan(h, 0, 0, m, n, k, sg.fp16)
        emit(f"cfg3 128x128x2^20 fp16 {'prescaled' if scaling else 'classic split'}", m, n, k, ms,
             rows_err(An, Bn, C.view(n, m).t().cpu().numpy()), err_s,
             {"compute_type": "fp16", "plan": [sg.PLAN_VARIANTS[q[0]], q[1]], "bit_identical_10_runs": identical}, pw)
    sg.set_exponent_scaling(h, True)
    sg.destroy(h)


def bench_bands(args):
    """Accuracy and speed by exponent band of A (the reference's exponent
    study): A = 2^(u*(hi-lo)+lo), u ~ U[0,1), bands [-15,14], [-36,-15],
    [-100,-36]; B = FP16(U[0,1)); m = n = k = 4096 op_n/op_n.  Per band and
    compute type: relative Frobenius and max relative error against FP64 on
    256 sampled rows, beside the library SGEMM's, and TFLOP/s."""
    import numpy as np
    import torch

    import shgemm_amd as sg

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    h = sg.create()
    sg.set_cuda_stream(h, stream)
    m = n = k = 4096
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed)
    B = torch.rand(k * n, device=dev, generator=g).half()
    B32 = B.float()
    Bh = B32.view(n, k).t().cpu().numpy().astype(np.float64)
    C = torch.empty(m * n, device=dev)
    rows = np.sort(np.random.default_rng(3).choice(m, 256, replace=False))
    ridx = torch.from_numpy(rows).to(dev)
    for lo, hi in [(-15, 14), (-36, -15), (-100, -36)]:
        A = torch.exp2(torch.rand(m * k, device=dev, generator=g) * (hi - lo) + lo)
        Ar = A.view(k, m).t()[ridx].cpu().numpy().astype(np.float64)
        ref = Ar @ Bh

        def errs():
            torch.cuda.synchronize()
            got = C.view(n, m).t()[ridx].cpu().numpy().astype(np.float64)
            rel = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
            mx = float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)))
            return rel, mx

        sg.sgemm_fp32(h, 0, 0, m, n, k, 1.0, A, m, B32, k, 0.0, C, m)
        s_rel, s_max = errs()
        for name, ct, scaling in (("fp16", sg.fp16, True), ("fp16 classic", sg.fp16, False), ("tf32", sg.tf32, True)):
            sg.set_exponent_scaling(h, scaling)
            call = lambda: sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, C, m, ct)
            ms = _time_ms(call, 5, stream)
            call()
            rel, mx = errs()
            print(json.dumps({"band": f"{lo}:{hi}", "compute_type": name, "m": m, "n": n, "k": k,
                              "rel_fro": rel, "max_rel": mx, "sgemm_rel_fro": s_rel, "sgemm_max_rel": s_max,
                              "ratio": rel / s_rel, "tflops_eff": round(2.0 * m * n * k / (ms * 1e-3) / 1e12, 2)}),
                  flush=True)
        sg.set_exponent_scaling(h, True)
        del A
    sg.destroy(h)


def bench_energy(args):
    """Power and energy per GEMM (the reference's power study,
    test/power_consumption_test.cu:55-126, 230-301): m = n = k = 2^10..2^14,
    op_n/op_n, A U[0,1), B FP16(U[0,1)); for tf32, fp16 and the library SGEMM,
    time 16 calls, then repeat the call for about --seconds and read the
    card's energy counter around that window.  Prints the reference's columns
    (throughput, average W, J per GEMM, call count) plus GFLOP/J and the
    median GFX clock and socket power sampled during the window; the FP16
    GEMM of the platform library (torch.mm, hipBLASLt) is timed beside them as
    a reference for what the card sustains under its power cap."""
    import torch

    import shgemm_amd as sg

    from shgemm_amd.energy import EnergyMeter, measure

    try:
        meter = EnergyMeter()
    except RuntimeError as e:  # the energy workload needs the SMI library
        raise SystemExit(f"energy workload: {e}")
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    h = sg.create()
    sg.set_cuda_stream(h, stream)
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed)
    if args.sketch_data:
        # The sketch shape's fp16 call on A = U[-1,1) and on A = 0 (same plan,
        # same instructions, operands that toggle nothing): under the power cap
        # the clock, and with it the call's time, follows the data.
        m, n, k = 16384, 256, 16384
        Bs = (torch.randn(k * n, device=dev, generator=g) / 16.0).half()
        Cs = torch.empty(m * n, device=dev)
        for name in ("uniform[-1,1)", "zeros", "uniform[-1,1)"):
            As = (torch.rand(m * k, device=dev, generator=g) * 2 - 1) if name != "zeros" else torch.zeros(m * k, device=dev)
            fn = lambda: sg.shgemm(h, 0, 0, m, n, k, 1.0, As, m, Bs, k, 0.0, Cs, m, sg.fp16)
            r = measure(meter, fn, args.seconds, torch.cuda.synchronize)
            for _ in range(20):
                fn()
            sg.enable_kernel_timing(h, 50)
            _, _, call_ms = timed_launches(fn, 50, 0, stream)
            nl, kt = sg.kernel_timing(h)
            sg.enable_kernel_timing(h, 0)
            print(json.dumps({"config": "sketch shgemm nn fp16", "m": m, "n": n, "k": k, "A": name,
                              "call_us": round(call_ms * 1e3, 1), "gemm_kernel_us": round(kt / max(nl, 1) * 1e3, 1),
                              "avg_power_w": round(r["avg_power_w"], 1), "energy_per_call_j": round(r["energy_per_call_j"], 4),
                              "median_gfxclk_mhz": r["median_gfxclk_mhz"],
                              "median_socket_power_w": r["median_socket_power_w"]}), flush=True)
            del As
        meter.close()
        sg.destroy(h)
        return
    top = 1 << 14
    A = torch.rand(top * top, device=dev, generator=g)
    B = torch.rand(top * top, device=dev, generator=g).half()
    B32 = B.float()
    A16 = A.half()
    C = torch.empty(top * top, device=dev)
    for log_n in range(10, 15):
        m = n = k = 1 << log_n
        calls = [("tf32", lambda: sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, C, m, sg.tf32)),
                 ("fp16", lambda: sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, C, m, sg.fp16)),
                 ("sgemm_fp32", lambda: sg.sgemm_fp32(h, 0, 0, m, n, k, 1.0, A, m, B32, k, 0.0, C, m)),
                 ("hipblaslt_hgemm_fp16", lambda: torch.mm(A16[:m * k].view(m, k), B[:k * n].view(k, n)))]
        for name, fn in calls:
            r = measure(meter, fn, args.seconds, torch.cuda.synchronize)
            flop = 2.0 * m * n * k
            print(json.dumps({"compute_type": name, "m": m, "n": n, "k": k, "op_a": "N", "op_b": "N",
                              "throughput_tflops": round(flop / r["per_call_s"] / 1e12, 2),
                              "avg_power_w": round(r["avg_power_w"], 1),
                              "energy_per_gemm_j": r["energy_per_call_j"], "test_count": r["count"],
                              "gflop_per_j": round(flop * r["count"] / r["joules"] / 1e9, 1),
                              "median_gfxclk_mhz": r["median_gfxclk_mhz"],
                              "median_socket_power_w": r["median_socket_power_w"]}), flush=True)
    meter.close()
    sg.destroy(h)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (got --gpus {args.gpus}); the library runs every "
              "problem on one GPU and does not shard across GPUs.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "grouped":
        bench_grouped(args)
        return
    if args.workload == "serve":
        bench_serve(args)
        return
    if args.workload == "configs":
        bench_configs(args)
        return
    if args.workload == "bands":
        bench_bands(args)
        return
    if args.workload == "energy":
        bench_energy(args)
        return

    import numpy as np
    import torch

    if args.copy_calibration:
        src = torch.ones(1 << 28, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 1 << 30, "copies": 3}))
        return

    import shgemm_amd as sg

    N = args.size
    m = n = k = N
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(20241203)
    A = torch.rand(m * k, device=dev, generator=g) * 2 - 1                  # FP32, column-major m x k
    B16 = (torch.rand(k * n, device=dev, generator=g) * 2 - 1).half()      # FP16, column-major k x n
    B32 = B16.float()
    C = torch.empty(m * n, device=dev)
    Cs = torch.empty(m * n, device=dev)

    stream = torch.cuda.current_stream()
    h = sg.create()
    sg.set_cuda_stream(h, stream)

    # The library brackets its main GEMM kernel with HIP events on `stream`
    # (enable_kernel_timing), so kern_ms is that kernel alone: the
    # row-exponent pass of each call is in ms_per_step but not in kern_ms.
    def kernel_ms():
        launches, total = sg.kernel_timing(h)
        sg.enable_kernel_timing(h, 0)
        return total / max(1, launches)

    step = lambda: sg.shgemm(h, sg.op_n, sg.op_n, m, n, k, 1.0, A, m, B16, k, 0.0, C, m, sg.fp16)
    region_ms, wall_ms, _ = timed_launches(step, args.steps, args.warmup, stream,
                                           after_warmup=lambda: sg.enable_kernel_timing(h, args.steps))
    kern_ms = kernel_ms()
    # The card's GFX clock and power on this call (ROCm SMI, when importable):
    # under the card's power cap the clock, not the nominal 2.4 GHz behind the
    # MFMA peak, bounds the kernel.
    clock = None if args.no_power else power_window(step)
    s_steps = max(2, args.steps // 2)
    sstep = lambda: sg.sgemm_fp32(h, sg.op_n, sg.op_n, m, n, k, 1.0, A, m, B32, k, 0.0, Cs, m)
    s_region_ms, _, _ = timed_launches(sstep, s_steps, 1, stream, after_warmup=lambda: sg.enable_kernel_timing(h, s_steps))
    s_kern_ms = kernel_ms()
    torch.cuda.synchronize()
    sg.destroy(h)

    flops = 2.0 * m * n * k
    ms_per_step = region_ms / args.steps
    value = flops / (ms_per_step * 1e-3) / 1e12
    s_value = flops / (s_region_ms / max(2, args.steps // 2) * 1e-3) / 1e12

    # Error vs FP64 on a sample of rows (every C row depends on its own A row
    # only), beside the library SGEMM's and a round-to-nearest host FP32 GEMM's
    # on the same rows.
    import oracle

    rows = np.sort(np.random.default_rng(7).choice(m, size=min(args.sample_rows, m), replace=False))
    ridx = torch.from_numpy(rows).to(dev)
    A_rows = A.view(k, m).t()[ridx].cpu().numpy()            # (rows, k)
    B_host = B32.view(n, k).t().cpu().numpy()                 # (k, n)
    C_rows = C.view(n, m).t()[ridx].cpu().numpy()
    Cs_rows = Cs.view(n, m).t()[ridx].cpu().numpy()
    err = fp64_sample_error(A_rows, B_host, C_rows)
    err_s = fp64_sample_error(A_rows, B_host, Cs_rows)
    nr = len(rows)
    At = np.ascontiguousarray(A_rows.T).ravel()               # column-major nr x k
    Bt = np.ascontiguousarray(B_host.T).ravel()               # column-major k x n
    rn = oracle.sgemm_nn_packed(nr, n, k, 1.0, At, nr, Bt, k, 0.0, None, nr).reshape(n, nr).T
    err_rn = fp64_sample_error(A_rows, B_host, rn)
    del At, Bt, rn
    # the platform's FP32 GEMM (torch.mm: hipBLASLt, no TF32) on the same rows
    torch.backends.cuda.matmul.allow_tf32 = False
    err_plat = fp64_sample_error(A_rows, B_host, (A.view(k, m).t()[ridx] @ B32.view(n, k).t()).cpu().numpy())

    serving = None
    cpu = None
    cpu_cfg = None
    if not args.no_cpu:
        A_np = A.cpu().numpy()
        B_np = B32.cpu().numpy()
    del A, B16, B32, C, Cs
    torch.cuda.empty_cache()
    if not args.no_serve:
        serving = serve_summary(seed=args.seed, rate=args.serve_rate, seconds=args.serve_seconds,
                                max_batch=args.max_batch)
    if not args.no_cpu:
        c_tf, threads, crows, ct = cpu_baseline(A_np, B_np, m, n, k, args.cpu_seconds)
        what = "the whole" if crows == m else f"rows 0..{crows} of the"
        cpu = {"value": round(c_tf, 6), "unit": "TFLOP/s", "cores": threads, "kind": "port",
               "sample": f"host FP32 GEMM (oracle/oracle.c sgemm_nn_packed: packed panels, 16x6 AVX2 FMA tiles, "
                         f"OpenMP, {threads} threads) on {what} {m}x{n}x{k} problem, {ct:.1f} s; "
                         f"rate = 2*rows*n*k/t"}
        cpu_cfg = cpu_configs(A_np, B_np, N)
        del A_np, B_np

    achieved = flops / (kern_ms * 1e-3) / 1e12
    peak_eff = PEAK_FP16_DENSE / 2.0
    clk = clock.get("median_gfxclk_mhz") if clock else None
    traffic, traffic_src = None, None
    tf = ROOT / "profiles" / "traffic_shgemm_16384.json"
    if tf.exists():
        try:
            tj = json.loads(tf.read_text())
            traffic = tj.get("hbm_bytes_per_launch")
            traffic_src = (f"stored, not measured in this run: {tf.relative_to(ROOT)} "
                           f"(rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of {tj.get('kernel', 'the GEMM kernel')}, "
                           f"build {tj.get('build', 'unrecorded')})")
        except Exception:
            traffic = None
    line = {
        "metric": "eff. TFLOP/s vs FP32 SGEMM at equal FP64-oracle error; sketch p99 latency",
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16 MFMA (f32 split hi+lo) / f32 accumulate",
        "data": "synthetic (seeded torch RNG on device)",
        "config": {"workload": f"shgemm op_n/op_n fp16 m=n=k={N}, A f32 U[-1,1], B f16 U[-1,1], alpha=1 beta=0",
                   "m": m, "n": n, "k": k, "parallelism": "single GPU"},
        "comparator": {"sgemm_fp32_tflops": round(s_value, 3), "speedup_vs_sgemm": round(value / s_value, 3),
                       "sgemm_kernel_ms": round(s_kern_ms, 4),
                       "sgemm_roofline_frac": round(s_value / PEAK_FP32_MATRIX, 4)},
        "error": {"shgemm_rel_fro": err, "sgemm_rel_fro": err_s, "ratio": err / err_s,
                  "platform_fp32_gemm_rel_fro": err_plat,
                  "host_rn_fp32_rel_fro": err_rn, "ratio_vs_host_rn_fp32": err / err_rn,
                  "sample_rows": int(nr),
                  "oracle": "numpy float64 product of the sampled rows; platform_fp32_gemm = torch.mm in FP32 "
                            "(hipBLASLt) on those rows; host_rn_fp32 = the packed host FP32 GEMM (round-to-nearest "
                            "FMA chain over k) on the same rows"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": round(peak_eff, 1), "unit": "TFLOP/s",
                     "frac": round(achieved / peak_eff, 4),
                     "frac_of_fp16_dense_peak": round(achieved / PEAK_FP16_DENSE, 4),
                     "frac_at_measured_clock": (round(achieved / (peak_eff * clk / NOMINAL_CLOCK_MHZ), 4)
                                                if isinstance(clk, (int, float)) and clk > 0 else None),
                     "traffic": traffic, "traffic_source": traffic_src,
                     "note": "achieved = 2mnk / mean duration of the main GEMM kernel alone (HIP events the library "
                             "records around it on the launch stream); left out of that duration and counted in "
                             "ms_per_step: the row-exponent pass over A and, on the pre-split plan, the pack "
                             "pass of A (split into FP16 hi/lo) and, for op_t B only, of B; peak = FP16 dense 2516.6 / 2, since "
                             "every algorithmic FMA costs two FP16 MFMA products (hi and lo); "
                             "frac_of_fp16_dense_peak = achieved / 2516.6; frac_at_measured_clock = achieved / "
                             "(peak x median GFX clock of the call's power window / 2400 MHz, the clock of the "
                             "nominal peak)",
                     "kernel_ms": round(kern_ms, 4)},
        "cpu_baseline": cpu,
        "cpu_configs": cpu_cfg,
        "sketch_serving": serving,
        "clock_and_power": clock,
        "wall_ms_timed_region": round(wall_ms, 3),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
