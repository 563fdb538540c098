"""Autotuner: measures candidate kernel plans on the GPU and writes the plan
table the runtime loads (plans_gfx950.tsv beside libshgemm_amd.so).

Replaces the reference's static per-architecture tile tables and its
level / split-K heuristic (src/shgemm.cu:427-489, 532, 643-656) with
measurements on the device at hand.

Two kinds of rows (format in csrc/plans.hpp):
  single   one shgemm call at shape point (2^lm, 2^ln, 2^lk): every plan
           variant (P0, P1, P2, P2 ring, P1 three-stage, P2 on the pre-split
           A, P0 on whole-row exponents, and the split-K variants, among them
           the deep 128x128x64 tile and the large tile of 64x128 wavefronts,
           at 2..256 slices) is timed alone; the fastest is kept.
  grouped  one problem inside a grouped launch: tile class (0 small, 1 128x128,
           2 large, 3 128x512 in the prescaled fp16 mode) x k-slices {1, 2, 4}
           are timed on a group of identical copies large enough to fill the
           GPU several times over, so the choice is the best throughput inside
           a busy group.

Usage (on a GPU):
  python -m shgemm_amd.tune --out shgemm_amd/plans_gfx950.tsv
  python -m shgemm_amd.tune --quick --out /tmp/plans.tsv     # reduced grid
  python -m shgemm_amd.tune --single '' --keep shgemm_amd/plans_gfx950.tsv \
      --out shgemm_amd/plans_gfx950.tsv                       # retune the grouped rows only
"""
from __future__ import annotations

import argparse
import datetime
import math
import statistics
import sys
import time
from pathlib import Path

import shgemm_amd as sg

MODES = {  # plan-table mode -> (compute type, exponent scaling)
    "fp16s": (sg.fp16, True),
    "fp16": (sg.fp16, False),
    "bf16": (sg.tf32, True),
}
OPS = {"N": sg.op_n, "T": sg.op_t}

SINGLE_M = [128, 512, 1024, 2048, 4096, 8192, 16384]
SINGLE_N = [64, 128, 256, 1024, 2048, 4096, 8192, 16384]
SINGLE_K = [256, 1024, 2048, 4096, 16384, 65536, 262144, 1048576]
QUICK_M = [128, 2048, 16384]
QUICK_N = [64, 1024, 16384]
QUICK_K = [1024, 16384, 1048576]
MAX_BYTES = 3.0e9       # operands + C of one tuning shape
MAX_FLOPS = 2.0 ** 43   # 2mnk of one tuning shape (m = n = k = 16384 included)

GROUP_M = [256, 1024, 4096, 8192]
GROUP_N = [64, 128, 256, 512]
GROUP_K = [256, 1024, 4096, 16384]
QUICK_GROUP_M = [256, 4096]
QUICK_GROUP_N = [64, 512]
QUICK_GROUP_K = [1024, 16384]
GROUP_SLICES = [1, 2, 4]
GROUP_UNITS = 8192      # work units per timed group (about 16 per workgroup slot)
GROUP_MAX_COPIES = 64


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise SystemExit("tune.py needs a GPU")
    return torch


def time_fn(fn, min_ms: float = 2.0, reps: int = 3) -> float:
    """Median per-call milliseconds of fn over `reps` timed batches, each batch
    long enough (>= min_ms) to hide launch jitter."""
    torch = _torch()
    fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    fn()
    e.record()
    torch.cuda.synchronize()
    one = max(s.elapsed_time(e), 1e-3)
    iters = max(1, min(200, int(math.ceil(min_ms / one))))
    out = []
    for _ in range(reps):
        s.record()
        for _ in range(iters):
            fn()
        e.record()
        torch.cuda.synchronize()
        out.append(s.elapsed_time(e) / iters)
    return statistics.median(out)


def single_shapes(quick: bool):
    ms, ns, ks = (QUICK_M, QUICK_N, QUICK_K) if quick else (SINGLE_M, SINGLE_N, SINGLE_K)
    for m in ms:
        for n in ns:
            for k in ks:
                if 4.0 * m * k + 2.0 * k * n + 4.0 * m * n > MAX_BYTES or 2.0 * m * n * k > MAX_FLOPS:
                    continue
                yield m, n, k


SPLIT_SLICES = (2, 4, 8, 16, 32, 64, 128, 256)


def _kept_single(path) -> dict:
    """(mode, ops, lm, ln, lk) -> (choice, slices, row text) of the single rows of a table."""
    out = {}
    for line in Path(path).read_text().splitlines():
        tok = line.split()
        if len(tok) >= 10 and tok[0] == "single":
            out[(tok[1], "".join(tok[2:5]), int(tok[5]), int(tok[6]), int(tok[7]))] = (int(tok[8]), int(tok[9]), line)
    return out


def tune_single(h, mode: str, ops: str, quick: bool, log, only=None, kept=None) -> list[str]:
    """Rows for one mode and layout.  With `only` (plan ids), each shape point
    times just those plans (the k-sliced tiles 10 and 11 only where m*n <=
    2^24, as they lose wherever the output tiles alone fill the GPU) against the kept
    table's choice for the point, and keeps the faster."""
    torch = _torch()
    ct, scaling = MODES[mode]
    oa, ob, oc = (OPS[c] for c in ops)
    sg.set_exponent_scaling(h, scaling)
    rows = []
    for m, n, k in single_shapes(quick):
        lm, ln, lk = (int(round(math.log2(x))) for x in (m, n, k))
        if only is None:
            candidates = [(v, 1) for v in (0, 1, 5, 6, 7, 8, 9)] + [(v, s) for v in (2, 3, 4, 10, 11) for s in SPLIT_SLICES]
        else:
            new = []
            for v in only:
                if v in (10, 11) and m * n > (1 << 24):
                    continue
                new += [(v, s) for s in SPLIT_SLICES] if v in (2, 3, 4, 10, 11) else [(v, 1)]
            old = kept.get((mode, ops, lm, ln, lk))
            if not new:
                if old is not None:
                    rows.append(old[2])  # nothing new to time at this point
                continue
            candidates = ([old[:2]] if old is not None else []) + new
        a_rows, b_rows, c_rows = (m if oa == sg.op_n else k), (k if ob == sg.op_n else n), (m if oc == sg.op_n else n)
        A = torch.rand(m * k, device="cuda") * 2 - 1
        B = torch.randn(k * n, device="cuda").half()
        C = torch.empty(m * n, device="cuda")
        seen, best = set(), None
        for v, s in candidates:
            sg.force_plan(h, v, s)
            eff = sg.query_plan(h, oa, ob, m, n, k, ct, oc)[:2]
            if eff in seen:
                continue
            seen.add(eff)
            call = lambda: sg.shgemm(h, oa, ob, m, n, k, 1.0, A, a_rows, B, b_rows, 0.0, C, c_rows, ct, oc)
            t = time_fn(call)
            if best is None or t < best[0]:
                best = (t, eff)
        sg.force_plan(h, -1, 1)
        t, (v, s) = best
        rows.append(f"single\t{mode}\t{ops[0]}\t{ops[1]}\t{ops[2]}\t{lm}\t{ln}\t{lk}\t{v}\t{s}\t{t * 1e3:.1f}")
        log(f"single {mode} {ops} m={m} n={n} k={k}: {sg.PLAN_VARIANTS[v]} x{s} {t * 1e3:.1f} us "
            f"({2.0 * m * n * k / t / 1e9:.1f} TFLOP/s, {len(seen)} plans)")
        del A, B, C
        torch.cuda.empty_cache()
    return rows


def tune_grouped(h, mode: str, quick: bool, log) -> list[str]:
    torch = _torch()
    ct, scaling = MODES[mode]
    sg.set_exponent_scaling(h, scaling)
    gm, gn, gk = (QUICK_GROUP_M, QUICK_GROUP_N, QUICK_GROUP_K) if quick else (GROUP_M, GROUP_N, GROUP_K)
    rows = []
    for m in gm:
        for n in gn:
            for k in gk:
                # distinct operands per copy (up to ~4 GB of A), so copies running
                # side by side do not share cache lines the real batch would not
                n_a = max(1, min(GROUP_MAX_COPIES, int(4e9 // (4 * m * k))))
                As = [torch.rand(m * k, device="cuda") * 2 - 1 for _ in range(n_a)]
                Bs = [torch.randn(k * n, device="cuda").half() for _ in range(n_a)]
                best = None
                for cls in GROUP_CLASSES[mode]:
                    for s in GROUP_SLICES:
                        # copies so the group has about GROUP_UNITS work units
                        tiles = _tiles(cls, mode, m, n)
                        copies = max(1, min(GROUP_MAX_COPIES, GROUP_UNITS // max(1, tiles * s)))
                        C = torch.empty(copies * m * n, device="cuda")
                        probs = [(m, n, k, As[i % n_a], m, Bs[i % n_a], k, C[i * m * n:(i + 1) * m * n], m, 1.0, 0.0)
                                 for i in range(copies)]
                        with sg.Group(h, sg.op_n, sg.op_n, probs, ct, tile_class=[cls] * copies,
                                      slices=[s] * copies) as grp:
                            _, used, _ = grp.plan()
                            if used[0] != s:
                                continue  # k too short for this many slices
                            t = time_fn(lambda: grp.run()) / copies
                        del C
                        if best is None or t < best[0]:
                            best = (t, cls, s)
                t, cls, s = best
                lm, ln, lk = (int(round(math.log2(x))) for x in (m, n, k))
                rows.append(f"grouped\t{mode}\tN\tN\tN\t{lm}\t{ln}\t{lk}\t{cls}\t{s}\t{t * 1e3:.2f}")
                log(f"grouped {mode} m={m} n={n} k={k}: class {cls} x{s} {t * 1e3:.2f} us/problem "
                    f"({2.0 * m * n * k / t / 1e9:.1f} TFLOP/s)")
                del As, Bs
                torch.cuda.empty_cache()
    return rows


# Tile shapes of the grouped classes (csrc/grouped_kernel.hpp GroupCfg; the
# prescaled mode's large tile is 128 x 256, the others' 128 x 128; class 3,
# 128 x 512, exists in the prescaled mode only).
GROUP_CLASSES = {"fp16s": (0, 1, 2, 3), "fp16": (0, 1, 2), "bf16": (0, 1, 2)}


def _tiles(cls: int, mode: str, m: int, n: int) -> int:
    bm, bn = {0: (64, 64), 1: (128, 128), 2: (128, 256 if mode == "fp16s" else 128), 3: (128, 512)}[cls]
    return -(-m // bm) * -(-n // bn)


def sweep_tiles(candidates, sizes, log) -> None:
    """Times the compiled tile candidates of the op_n/op_n fp16 kernel
    (csrc/inst_tune.hip; MTK_SHGEMM_TUNE_P1=<i> is read when a handle is
    created and replaces the unsplit large-tile plan, which is forced here) at
    m = n = k = each size; -1 is the library's own P1 tile."""
    import os

    torch = _torch()
    # bring the GPU to its sustained clock before the first timed candidate
    x = torch.randn(4096, 4096, device="cuda")
    t_end = time.time() + 1.0
    while time.time() < t_end:
        x = x @ x
        x /= x.norm()
    torch.cuda.synchronize()
    for N in sizes:
        g = torch.Generator(device="cuda")
        g.manual_seed(1)
        A = torch.rand(N * N, device="cuda", generator=g) * 2 - 1
        B = (torch.rand(N * N, device="cuda", generator=g) * 2 - 1).half()
        C = torch.empty(N * N, device="cuda")
        ref = None
        for v in candidates:
            saved = {k: os.environ.get(k) for k in ("MTK_SHGEMM_TUNE_P1", "MTK_SHGEMM_EXP_SCALING")}
            if v >= 0:
                os.environ["MTK_SHGEMM_TUNE_P1"] = str(v)
            else:
                os.environ.pop("MTK_SHGEMM_TUNE_P1", None)
            os.environ["MTK_SHGEMM_EXP_SCALING"] = "0" if v == 0 else "1"  # candidate 0 is the two-accumulator mode
            h = sg.create()
            sg.force_plan(h, 1, 1)  # the candidate replaces the unsplit large tile
            for k_, val in saved.items():
                if val is None:
                    os.environ.pop(k_, None)
                else:
                    os.environ[k_] = val
            t = time_fn(lambda: sg.shgemm(h, sg.op_n, sg.op_n, N, N, N, 1.0, A, N, B, N, 0.0, C, N, sg.fp16))
            torch.cuda.synchronize()
            if ref is None:
                ref = C.clone()
                d = 0.0
            else:
                d = ((C - ref).norm() / ref.norm()).item()
            log(f"N={N} candidate={v:3d} {t:.3f} ms {2.0 * N ** 3 / t / 1e9:7.1f} TFLOP/s eff  reldiff_vs_first={d:.2e}")
            sg.destroy(h)
        del A, B, C


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", help="plan table to write")
    ap.add_argument("--sweep-tiles", default="", help="comma list of tile candidates to time instead of tuning")
    ap.add_argument("--sizes", default="4096,8192,16384", help="sizes for --sweep-tiles")
    ap.add_argument("--quick", action="store_true", help="reduced shape grid")
    ap.add_argument("--single", default="fp16s:NNN,fp16s:TNN,fp16s:TTT,fp16:NNN,bf16:NNN",
                    help="mode:ops list for single-call rows ('' for none)")
    ap.add_argument("--grouped", default="fp16s,bf16", help="modes for grouped rows ('' for none)")
    ap.add_argument("--keep", default="", help="existing table whose rows of the kinds and modes not tuned "
                                               "now are kept")
    ap.add_argument("--only", default="", help="comma list of plan ids: time only these against each point's "
                                               "choice in the --keep table (single rows)")
    args = ap.parse_args(argv)
    torch = _torch()
    t0 = time.time()
    log = lambda msg: print(f"[tune {time.time() - t0:6.1f}s] {msg}", flush=True)
    if args.sweep_tiles:
        sweep_tiles([int(x) for x in args.sweep_tiles.split(",")], [int(x) for x in args.sizes.split(",")], log)
        return 0
    if not args.out:
        ap.error("--out is required unless --sweep-tiles is given")
    only = [int(x) for x in args.only.split(",")] if args.only else None
    if only is not None and not args.keep:
        ap.error("--only needs --keep (the table whose choices the plans are timed against)")
    kept_single = _kept_single(args.keep) if only is not None else None
    h = sg.create()
    sg.load_plans(h, None)  # measure from the cost model's state, not an older table
    rows = []
    try:
        for item in filter(None, args.single.split(",")):
            mode, ops = item.split(":")
            rows += tune_single(h, mode, ops, args.quick, log, only, kept_single)
        for mode in filter(None, args.grouped.split(",")):
            rows += tune_grouped(h, mode, args.quick, log)
    finally:
        sg.destroy(h)
    tuned = {("single", item.split(":")[0], item.split(":")[1]) for item in filter(None, args.single.split(","))}
    tuned |= {("grouped", mode, None) for mode in filter(None, args.grouped.split(","))}
    kept = []
    if args.keep:
        for line in Path(args.keep).read_text().splitlines():
            tok = line.split()
            if not tok or tok[0].startswith("#"):
                continue
            if (tok[0], tok[1], "".join(tok[2:5])) in tuned or (tok[0], tok[1], None) in tuned:
                continue
            kept.append(line)
        log(f"kept {len(kept)} rows of {args.keep}")
    rows = kept + rows
    props = torch.cuda.get_device_properties(0)
    head = [
        "# shgemm_amd plan table (kind mode op_a op_b op_c log2m log2n log2k choice slices time_us)",
        f"# device: {props.name} ({getattr(props, 'gcnArchName', '?')}, {props.multi_processor_count} CUs)",
        f"# made by: python -m shgemm_amd.tune {' '.join(argv if argv is not None else sys.argv[1:])}",
        f"# date: {datetime.datetime.now(datetime.timezone.utc).strftime('%Y-%m-%d %H:%M UTC')}",
        "# single: choice = plan variant (0 P0, 1 P1, 2 SK128, 3 P1-SK, 4 P0-SK, 5 P2 256x256, 6 P2 ring, 7 P1 three-stage,"
        " 8 P2 on pre-split A, 9 P0 on whole-row exponents, 10 deep 128x128x64 split-K, 11 large tile of 64x128 wavefronts split-K);"
        " time_us per call",
        "# grouped: choice = tile class (0 64x64, 1 128x128, 2 large, 3 128x512); time_us per problem in a full group",
    ]
    if args.keep:
        prev = [l for l in Path(args.keep).read_text().splitlines() if l.startswith(("# made by:", "# date:", "# kept"))]
        head += ["# kept rows (not re-measured, or re-measured only against --only) come from: " + l[2:]
                 if not l.startswith("# kept") else l for l in prev]
    out = Path(args.out)
    out.write_text("\n".join(head + rows) + "\n")
    n = sg.check_plan_file(out)
    log(f"wrote {out} ({n} rows)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
