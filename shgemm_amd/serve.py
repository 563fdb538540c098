"""Sketch serving: random-projection sketches Y = A * Omega with continuous
batching on one GPU.

A request names a device-resident FP32 matrix A (m x k, column-major), a
sketch width l and a seed.  The server generates Omega (k x l, FP16
Gaussian / sqrt(l)) on the device from the seed (Philox4x32-10 + Box-Muller,
shgemm_amd.gaussian_f16), keeps it in HBM for the life of the request, and
computes Y = A * Omega with shgemm (fp16 compute type, error-corrected).

Continuous batching: the server owns `slots` independent lanes, each a
library handle bound to its own HIP stream (default priority).  At every
scheduling step the waiting requests are admitted first come, first served
while a lane is free and the in-flight footprint plus theirs (4mk + 2kl + 4ml
bytes each) fits the HBM budget: with max_batch > 1 up to max_batch of them
go to one free lane together and run as ONE grouped launch of problems of
different sizes (shgemm_amd.Batch: each request one problem, planned by its
own shape), otherwise each takes a lane and runs as one shgemm call.  Work
admitted while other lanes run overlaps with theirs on the device; a lane and
its requests' budget are freed when its launch completes.  A request larger
than the whole budget is rejected when it is submitted.

Results do not depend on the mix: Omega depends only on (seed, index), and
every kernel of the product is deterministic for a given shape (fixed plans,
fixed-order split-K reduction; in a batch each problem's plan depends on its
own shape only), so a request returns the same bits alone and among others.

There is no serving component in the reference; the sketch use case is the
one its README names (README.md:64-72) and its tests seed inputs with a
counter-based Philox generator (test/shgemm_test.cu:247-252).
"""
from __future__ import annotations

import collections
import dataclasses
import math
import time
from typing import Callable

import shgemm_amd as sg


@dataclasses.dataclass
class SketchRequest:
    rid: int
    m: int
    k: int
    l: int
    seed: int
    a: object            # device tensor or address of A (FP32, column-major m x k)
    lda: int
    arrival: float | None = None   # server clock, seconds
    admitted: float | None = None
    done: float | None = None
    slot: int = -1
    y: object = None       # result: FP32 m x l, column-major (ld = m)
    omega: object = None   # FP16 k x l, column-major (ld = k)

    @property
    def footprint(self) -> int:
        """HBM bytes the request works on: A, Omega and Y."""
        return 4 * self.m * self.k + 2 * self.k * self.l + 4 * self.m * self.l

    @property
    def flops(self) -> float:
        return 2.0 * self.m * self.k * self.l

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class GpuLanes:
    """The device side of the server: per lane a handle and a HIP stream."""

    def __init__(self, slots: int, compute_type=sg.fp16, device: str = "cuda:0", max_batch: int = 0,
                 batch_bytes: int = 1 << 30):
        """max_batch > 0: every lane also owns a shgemm_amd.Batch for up to
        max_batch requests with batch_bytes of device scratch per slot."""
        import torch

        self.torch = torch
        self.device = torch.device(device)
        self.compute_type = compute_type
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(slots)]
        self.handles = []
        self.batches = []
        for s in self.streams:
            h = sg.create()
            sg.set_cuda_stream(h, s)
            # lanes run side by side: keep CUs shareable (measured: p99 8.8
            # vs 10.8 ms at 2000 req/s with the whole-CU 256x256 plans)
            sg.set_shared_device(h, slots > 1)
            self.handles.append(h)
            if max_batch > 0:
                self.batches.append(sg.Batch(h, sg.op_n, sg.op_n, compute_type, sg.op_n, max_problems=max_batch,
                                             bytes=batch_bytes))

    def reserve(self, max_m: int, max_k: int = 0, max_l: int = 0) -> None:
        """Prepare every lane for requests up to (max_m, max_k, max_l) before
        serving, so no device allocation lands inside a request's latency:
        grow each handle's scratch (shgemm_amd.reserve: row exponents for
        max_m rows), and, when max_k and max_l are given, leave room for two
        requests' Omega and Y in the lane stream's pool of the caching
        allocator (pools are per stream)."""
        torch = self.torch
        for h in self.handles:
            sg.reserve(h, max_m, max(1, max_l), max(1, max_k))
        if max_k > 0 and max_l > 0:
            for s in self.streams:
                with torch.cuda.stream(s):
                    blocks = [torch.empty(max_k * max_l, dtype=torch.float16, device=self.device) for _ in range(2)]
                    blocks += [torch.empty(max_m * max_l, dtype=torch.float32, device=self.device) for _ in range(2)]
                del blocks
        torch.cuda.synchronize(self.device)

    def launch(self, req: SketchRequest, slot: int):
        torch = self.torch
        s = self.streams[slot]
        with torch.cuda.stream(s):
            req.omega = torch.empty(req.k * req.l, dtype=torch.float16, device=self.device)
            req.y = torch.empty(req.m * req.l, dtype=torch.float32, device=self.device)
        sg.gaussian_f16(req.omega, req.k * req.l, req.seed, 1.0 / math.sqrt(req.l), stream=s)
        sg.shgemm(self.handles[slot], sg.op_n, sg.op_n, req.m, req.l, req.k, 1.0, req.a, req.lda, req.omega, req.k,
                  0.0, req.y, req.m, self.compute_type)
        ev = torch.cuda.Event()
        ev.record(s)
        return ev

    def launch_batch(self, reqs: list, slot: int):
        """Omega of every request, then one grouped launch of all of them on
        the lane's stream (split in halves if the batch's scratch would not
        fit the lane's batch slots; each problem's plan depends on its shape
        only, so the bits are the same)."""
        torch = self.torch
        s = self.streams[slot]
        for req in reqs:
            with torch.cuda.stream(s):
                req.omega = torch.empty(req.k * req.l, dtype=torch.float16, device=self.device)
                req.y = torch.empty(req.m * req.l, dtype=torch.float32, device=self.device)
            sg.gaussian_f16(req.omega, req.k * req.l, req.seed, 1.0 / math.sqrt(req.l), stream=s)

        def run(part):
            probs = [(r.m, r.l, r.k, r.a, r.lda, r.omega, r.k, r.y, r.m, 1.0, 0.0) for r in part]
            try:
                self.batches[slot].run(probs, stream=s)
            except sg.ShgemmError as e:
                if e.status != sg.STATUS_EINVAL or len(part) == 1:
                    raise
                run(part[:len(part) // 2])
                run(part[len(part) // 2:])

        run(reqs)
        ev = torch.cuda.Event()
        ev.record(s)
        return ev

    @staticmethod
    def finished(token) -> bool:
        return token.query()

    def close(self) -> None:
        self.torch.cuda.synchronize(self.device)
        for b in self.batches:
            b.destroy()
        self.batches = []
        for h in self.handles:
            sg.destroy(h)
        self.handles = []


class SketchServer:
    """FIFO admission under an HBM budget onto `slots` concurrent lanes; with
    max_batch > 1 the requests admitted to one lane in a step run as one
    grouped launch."""

    def __init__(self, hbm_budget: float = 200e9, slots: int = 4, lanes=None,
                 clock: Callable[[], float] = time.perf_counter, keep_omega: bool = False, max_batch: int = 1):
        if slots < 1:
            raise ValueError("slots must be >= 1")
        if max_batch < 1:
            raise ValueError("max_batch must be >= 1")
        self.budget = int(hbm_budget)
        self.max_batch = max_batch
        self.lanes = lanes if lanes is not None else GpuLanes(slots, max_batch=max_batch if max_batch > 1 else 0)
        self.clock = clock
        self.keep_omega = keep_omega
        self.free = collections.deque(range(slots))
        self.queue: collections.deque[SketchRequest] = collections.deque()
        self.inflight: dict[int, tuple[list, object]] = {}  # lane -> (requests, completion token)
        self.batch_sizes: list[int] = []
        self.in_use = 0
        self.peak_in_use = 0
        self.peak_inflight = 0
        self.admit_order: list[int] = []

    def submit(self, req: SketchRequest) -> None:
        if req.footprint > self.budget:
            raise ValueError(f"request {req.rid} needs {req.footprint} bytes, above the {self.budget}-byte budget")
        if req.arrival is None:
            req.arrival = self.clock()
        self.queue.append(req)

    def poll(self) -> list[SketchRequest]:
        """Completes every request whose work has finished; frees its lane and budget."""
        done = []
        for slot, (reqs, token) in list(self.inflight.items()):
            if self.lanes.finished(token):
                t = self.clock()
                del self.inflight[slot]
                self.free.append(slot)
                for req in reqs:
                    req.done = t
                    self.in_use -= req.footprint
                    if not self.keep_omega:
                        req.omega = None
                    done.append(req)
        return done

    def admit(self) -> int:
        """Admits waiting requests in arrival order while a lane and budget
        allow: up to max_batch of them per free lane, as one launch."""
        n = 0
        while self.queue and self.free and self.in_use + self.queue[0].footprint <= self.budget:
            slot = self.free.popleft()
            batch = []
            while (self.queue and len(batch) < self.max_batch
                   and self.in_use + self.queue[0].footprint <= self.budget):
                req = self.queue.popleft()
                req.slot = slot
                req.admitted = self.clock()
                self.in_use += req.footprint
                self.admit_order.append(req.rid)
                batch.append(req)
            self.peak_in_use = max(self.peak_in_use, self.in_use)
            token = (self.lanes.launch_batch(batch, slot) if self.max_batch > 1
                     else self.lanes.launch(batch[0], slot))
            self.inflight[slot] = (batch, token)
            self.batch_sizes.append(len(batch))
            self.peak_inflight = max(self.peak_inflight, sum(len(b) for b, _ in self.inflight.values()))
            n += len(batch)
        return n

    def step(self) -> list[SketchRequest]:
        done = self.poll()
        self.admit()
        return done

    def idle(self) -> bool:
        return not self.queue and not self.inflight

    def drain(self) -> list[SketchRequest]:
        done = []
        while not self.idle():
            done += self.step()
        return done

    def close(self) -> None:
        self.lanes.close()


def percentile(xs, q: float) -> float:
    """Nearest-rank percentile (q in [0, 100])."""
    s = sorted(xs)
    if not s:
        return float("nan")
    i = max(0, min(len(s) - 1, math.ceil(q / 100.0 * len(s)) - 1))
    return s[i]


def run_arrivals(server: SketchServer, requests: list[SketchRequest], offsets: list[float]) -> dict:
    """Submits request i at start + offsets[i] (its arrival time is the planned
    one, so host lateness counts as latency), serving continuously until all
    complete.  Returns latency percentiles and the sustained rate."""
    order = sorted(range(len(requests)), key=lambda i: offsets[i])
    t0 = server.clock()
    nxt = 0
    done: list[SketchRequest] = []
    while len(done) < len(requests):
        now = server.clock()
        while nxt < len(order) and t0 + offsets[order[nxt]] <= now:
            r = requests[order[nxt]]
            r.arrival = t0 + offsets[order[nxt]]
            server.submit(r)
            nxt += 1
        done += server.step()
    lat = [r.latency for r in done]
    first = min(r.arrival for r in done)
    last = max(r.done for r in done)
    return {
        "requests": len(done),
        "p50_ms": percentile(lat, 50) * 1e3,
        "p99_ms": percentile(lat, 99) * 1e3,
        "mean_ms": sum(lat) / len(lat) * 1e3,
        "max_ms": max(lat) * 1e3,
        "sustained_req_s": len(done) / (last - first),
        "offered_req_s": ((len(offsets) - 1) / (max(offsets) - min(offsets))
                          if len(offsets) > 1 and max(offsets) > min(offsets) else float("inf")),
        "flops_per_s": sum(r.flops for r in done) / (last - first),
        "peak_inflight": server.peak_inflight,
        "mean_batch": sum(server.batch_sizes) / max(1, len(server.batch_sizes)),
        "max_batch_seen": max(server.batch_sizes, default=0),
        "peak_budget_bytes": server.peak_in_use,
    }


def poisson_offsets(n: int, rate: float, rng) -> list[float]:
    """Arrival offsets (seconds) of a Poisson process of the given rate."""
    gaps = rng.exponential(1.0 / rate, n)
    gaps[0] = 0.0
    t, out = 0.0, []
    for g in gaps:
        t += g
        out.append(t)
    return out


def sketch_shapes(n: int, rng, m_range=(1024, 32768), k_range=(1024, 16384), widths=(64, 128, 256, 512, 1024)):
    """Request shapes: m and k log-uniform in their ranges, l uniform over widths."""
    m = [int(math.exp(x)) for x in rng.uniform(math.log(m_range[0]), math.log(m_range[1]), n)]
    k = [int(math.exp(x)) for x in rng.uniform(math.log(k_range[0]), math.log(k_range[1]), n)]
    l = [int(x) for x in rng.choice(widths, n)]
    return list(zip(m, k, l))
