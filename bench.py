"""Benchmark: error-corrected SHGEMM at m = n = k = 16384 on one MI355X.

One step = one mtk::shgemm call (op_n/op_n, fp16 compute type, alpha=1,
beta=0) on A FP32 U[-1,1] (16384 x 16384) and B FP16 U[-1,1], operands
resident in HBM.  Reported beside it, on the same inputs and GPU:
  * the library's plain FP32 SGEMM (the comparator of the metric), its
    throughput and the ratio;
  * ||C - C64||_F / ||C64||_F of both against an FP64 product over a sample
    of rows (numpy float64 on the host; FP64 is used only for this check);
  * roofline of the SHGEMM kernel from HIP events on its own stream;
  * the host FP32 GEMM (CPU baseline) on a bounded sample of rows.

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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=16384)
    ap.add_argument("--sample-rows", type=int, default=256, help="rows of C checked against FP64")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three 1 GiB device copies (reference for FETCH_SIZE/WRITE_SIZE scaling)")
    return ap.parse_args()


def timed_launches(fn, steps, warmup, stream):
    """Run warmup untimed launches, then `steps` launches each bracketed by HIP
    events on `stream` (the stream the kernels run on).  Returns (wall ms of the
    whole timed region, mean per-launch ms from the events)."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
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


def cpu_baseline(A_np, B32_np, m, n, k, budget_s):
    """Host FP32 GEMM (oracle/liboracle.so) on row blocks of the same problem
    until the time budget is used; returns (TFLOP/s, threads, rows done)."""
    import oracle

    threads = oracle.max_threads()
    rows_per_block = 32
    done, t = 0, 0.0
    while t < budget_s and done < m:
        r1 = min(m, done + rows_per_block)
        t0 = time.perf_counter()
        oracle.gemm_f32_rows(0, 0, 0, m, n, k, 1.0, A_np, m, B32_np, k, 0.0, None, m, done, r1,
                             out=_cpu_out(m, n), threads=threads)
        t += time.perf_counter() - t0
        done = r1
    return 2.0 * done * n * k / t / 1e12, threads, done, t


_OUT = {}


def _cpu_out(m, n):
    import numpy as np

    if "c" not in _OUT:
        _OUT["c"] = np.zeros(m * n, dtype=np.float32)
    return _OUT["c"]


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {args.gpus}); "
              "multi-GPU sharding is described in DESIGN.md and not run.", file=sys.stderr)
        sys.exit(2)

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

    step = lambda: sg.shgemm(h, sg.op_n, sg.op_n, m, n, k, 1.0, A, m, B16, k, 0.0, C, m, sg.fp16)
    region_ms, wall_ms, kern_ms = timed_launches(step, args.steps, args.warmup, stream)
    sstep = lambda: sg.sgemm_fp32(h, sg.op_n, sg.op_n, m, n, k, 1.0, A, m, B32, k, 0.0, Cs, m)
    s_region_ms, _, s_kern_ms = timed_launches(sstep, max(2, args.steps // 2), 1, stream)
    torch.cuda.synchronize()
    sg.destroy(h)

    flops = 2.0 * m * n * k
    ms_per_step = region_ms / args.steps
    value = flops / (ms_per_step * 1e-3) / 1e12
    s_value = flops / (s_region_ms / max(2, args.steps // 2) * 1e-3) / 1e12

    # Error vs FP64 on a sample of rows (every C row depends on its own A row only).
    rows = np.sort(np.random.default_rng(7).choice(m, size=min(args.sample_rows, m), replace=False))
    ridx = torch.from_numpy(rows).to(dev)
    A_rows = A.view(k, m).t()[ridx].cpu().numpy()            # (rows, k)
    B_host = B32.view(n, k).t().cpu().numpy()                 # (k, n)
    C_rows = C.view(n, m).t()[ridx].cpu().numpy()
    Cs_rows = Cs.view(n, m).t()[ridx].cpu().numpy()
    err = fp64_sample_error(A_rows, B_host, C_rows)
    err_s = fp64_sample_error(A_rows, B_host, Cs_rows)

    cpu = None
    if not args.no_cpu:
        A_np = A.cpu().numpy()
        B_np = B32.cpu().numpy()
        del A, B16, B32, C, Cs
        torch.cuda.empty_cache()
        c_tf, threads, crows, ct = cpu_baseline(A_np, B_np, m, n, k, args.cpu_seconds)
        cpu = {"value": round(c_tf, 6), "unit": "TFLOP/s", "cores": threads, "kind": "port",
               "sample": f"host FP32 GEMM (oracle/oracle.c, OpenMP, {threads} threads) on rows 0..{crows} of the "
                         f"same {m}x{n}x{k} problem, {ct:.1f} s of CPU work; rate = 2*rows*n*k/t"}

    achieved = flops / (kern_ms * 1e-3) / 1e12
    peak_eff = PEAK_FP16_DENSE / 2.0
    traffic = None
    tf = ROOT / "profiles" / "traffic_shgemm_16384.json"
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get("hbm_bytes_per_launch")
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
                  "sample_rows": int(len(rows)), "oracle": "numpy float64 product of the sampled rows"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": round(peak_eff, 1), "unit": "TFLOP/s",
                     "frac": round(achieved / peak_eff, 4), "traffic": traffic,
                     "note": "achieved = 2mnk / mean kernel time (HIP events); peak = FP16 dense 2516.6 / 2, "
                             "since every algorithmic FMA costs two FP16 MFMA products (hi and lo)",
                     "kernel_ms": round(kern_ms, 4)},
        "cpu_baseline": cpu,
        "wall_ms_timed_region": round(wall_ms, 3),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
