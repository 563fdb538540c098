"""GPU: the sketch generator and the sketch server.

Generator bars: Philox4x32-10 words bit-identical to the oracle's (whose
Philox is pinned by the published known-answer vectors in test_oracle.py);
FP16 Gaussians within one FP16 ulp of the oracle's FP64 Box-Muller values
(and equal for all but a few per million: the device evaluates log/sin/cos in
FP32); element i depends only on (seed, i).
Server bars: each request's Y is bit-identical alone and in a concurrent mix;
Y's error against an FP64 product of A and the generated Omega is within 2x
of the library SGEMM's; the HBM budget is never exceeded.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    require_gpu()
    import torch

    return torch


@pytest.mark.parametrize("seed", [0, 1, (1 << 40) + 7, (1 << 64) - 1])
def test_philox_words_bit_exact(torch_mod, seed):
    import shgemm_amd as sg

    torch = torch_mod
    blocks = 1000
    out = torch.empty(4 * blocks, dtype=torch.int32, device="cuda")
    sg.philox_words(out, blocks, seed)
    got = out.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, oracle.philox_words(blocks, seed))


def _f16_ulps(a16: np.ndarray, b16: np.ndarray) -> np.ndarray:
    """Distance in FP16 representable steps (sign-magnitude order)."""
    def key(x):
        u = x.view(np.uint16).astype(np.int32)
        return np.where(u & 0x8000, -(u & 0x7FFF), u)
    return np.abs(key(a16) - key(b16))


@pytest.mark.parametrize("seed,scale", [(3, 1.0), (12345, 1.0 / 16.0), ((1 << 63) + 5, 0.125)])
def test_gaussian_matches_oracle(torch_mod, seed, scale):
    import shgemm_amd as sg

    torch = torch_mod
    n = 1 << 18
    out = torch.empty(n + 3, dtype=torch.float16, device="cuda")
    sg.gaussian_f16(out, n + 3, seed, scale)
    got = out.cpu().numpy()
    ref = oracle.gaussian(n + 3, seed, scale).astype(np.float16)
    d = _f16_ulps(got, ref)
    assert d.max() <= 1
    assert (d == 0).mean() > 0.999
    # element i depends only on (seed, i): a shorter or offset launch agrees
    short = torch.empty(1001, dtype=torch.float16, device="cuda")
    sg.gaussian_f16(short, 1001, seed, scale)
    assert np.array_equal(short.cpu().numpy().view(np.uint16), got[:1001].view(np.uint16))
    z = got.astype(np.float64) / scale
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01


def test_gaussian_seeds_differ(torch_mod):
    import shgemm_amd as sg

    torch = torch_mod
    a = torch.empty(4096, dtype=torch.float16, device="cuda")
    b = torch.empty(4096, dtype=torch.float16, device="cuda")
    sg.gaussian_f16(a, 4096, 1, 1.0)
    sg.gaussian_f16(b, 4096, 2, 1.0)
    assert (a != b).float().mean().item() > 0.99


def _pool(torch, rows=4096, cols=4096, seed=9):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.rand(rows * cols, device="cuda", generator=g) * 2 - 1, rows


def _requests(torch, pool, ld, n, rng):
    from shgemm_amd.serve import SketchRequest

    reqs = []
    for i in range(n):
        m = int(rng.integers(100, 1500))
        k = int(rng.integers(64, 2500))
        l = int(rng.choice([64, 128, 256]))
        r0 = int(rng.integers(0, ld - m)) // 4 * 4
        c0 = int(rng.integers(0, ld - k))
        addr = pool.data_ptr() + 4 * (c0 * ld + r0)
        reqs.append(SketchRequest(rid=i, m=m, k=k, l=l, seed=1000 + i, a=addr, lda=ld))
    return reqs


def test_server_results_same_alone_and_in_mix(torch_mod):
    from shgemm_amd.serve import GpuLanes, SketchServer, poisson_offsets, run_arrivals

    torch = torch_mod
    pool, ld = _pool(torch)
    rng = np.random.default_rng(21)
    alone_reqs = _requests(torch, pool, ld, 24, rng)
    rng = np.random.default_rng(21)
    mix_reqs = _requests(torch, pool, ld, 24, rng)

    # one request at a time on the same four-lane configuration (the lanes'
    # handles choose kernels for a shared device, as in the mix below)
    srv = SketchServer(hbm_budget=8e9, slots=4, lanes=GpuLanes(4))
    alone = []
    for r in alone_reqs:
        srv.submit(r)
        srv.drain()
        torch.cuda.synchronize()
        alone.append(r.y.cpu().numpy().copy())
    srv.close()

    lanes = GpuLanes(4)
    lanes.reserve(2048)
    srv = SketchServer(hbm_budget=8e9, slots=4, lanes=lanes)
    offs = poisson_offsets(len(mix_reqs), 20000.0, np.random.default_rng(2))
    st = run_arrivals(srv, mix_reqs, offs)
    torch.cuda.synchronize()
    assert st["requests"] == 24 and st["peak_inflight"] >= 2   # requests did overlap
    for r, y in zip(mix_reqs, alone):
        assert np.array_equal(r.y.cpu().numpy().view(np.uint32), y.view(np.uint32)), r.rid
    srv.close()


def test_server_error_vs_fp64_and_budget(torch_mod):
    import shgemm_amd as sg
    from shgemm_amd.serve import GpuLanes, SketchServer

    torch = torch_mod
    pool, ld = _pool(torch, seed=4)
    rng = np.random.default_rng(5)
    reqs = _requests(torch, pool, ld, 10, rng)
    budget = 2 * max(r.footprint for r in reqs)
    srv = SketchServer(hbm_budget=budget, slots=4, lanes=GpuLanes(4), keep_omega=True)
    for r in reqs:
        srv.submit(r)
    srv.drain()
    torch.cuda.synchronize()
    assert srv.peak_in_use <= budget
    h = sg.create()
    P = pool.view(ld, ld)  # P[c, r] = A_pool[r, c] (column-major pool)
    for r in reqs[:5]:
        r0 = (r.a - pool.data_ptr()) // 4 % ld
        c0 = (r.a - pool.data_ptr()) // 4 // ld
        A = P[c0:c0 + r.k, r0:r0 + r.m].t().cpu().numpy().astype(np.float64)        # m x k
        Om = r.omega.view(r.l, r.k).t().cpu().numpy().astype(np.float64)           # k x l
        ref = A @ Om
        Y = r.y.view(r.l, r.m).t().cpu().numpy().astype(np.float64)
        Om32 = r.omega.float()
        Ys = torch.empty(r.m * r.l, device="cuda")
        sg.sgemm_fp32(h, sg.op_n, sg.op_n, r.m, r.l, r.k, 1.0, r.a, r.lda, Om32, r.k, 0.0, Ys, r.m)
        torch.cuda.synchronize()
        Ys_h = Ys.view(r.l, r.m).t().cpu().numpy().astype(np.float64)
        e = np.linalg.norm(Y - ref) / np.linalg.norm(ref)
        es = np.linalg.norm(Ys_h - ref) / np.linalg.norm(ref)
        assert e <= 2.0 * es + 2.0 ** -22, (r.rid, e, es)
        # Omega is the generator's: N(0,1)/sqrt(l) for this seed
        want = oracle.gaussian(r.k * r.l, r.seed, 1.0 / math.sqrt(r.l)).astype(np.float16)
        assert _f16_ulps(r.omega.cpu().numpy(), want).max() <= 1
    sg.destroy(h)
    srv.close()


def _stated_requests(torch, pools, rows, cols, n, rng, rid0=0):
    """Requests from the serving workload's distribution: m log-uniform in
    [1024, 32768], k log-uniform in [1024, 16384], l in {64, ..., 1024}; A is
    a block of one of the column-major rows x cols pools (lda = rows)."""
    from shgemm_amd.serve import SketchRequest, sketch_shapes

    reqs = []
    for i, (m, k, l) in enumerate(sketch_shapes(n, rng)):
        p = (rid0 + i) % len(pools)
        r0 = int(rng.integers(0, rows - m + 1)) // 4 * 4
        c0 = int(rng.integers(0, cols - k + 1))
        addr = pools[p].data_ptr() + 4 * (c0 * rows + r0)
        reqs.append(SketchRequest(rid=rid0 + i, m=m, k=k, l=l, seed=5000 + rid0 + i, a=addr, lda=rows))
    return reqs


def test_server_stated_distribution_alone_mix_and_budget(torch_mod):
    """Full-size requests (A blocks of four 32768 x 16384 pools): Y is
    bit-identical alone, in a Poisson mix at 2000 req/s under the 200 GB
    budget, and in a mix whose budget admits only about two requests at a
    time (the budget is never exceeded); sampled rows of Y are within 2x of
    SGEMM's FP64 error."""
    import shgemm_amd as sg
    from shgemm_amd.serve import GpuLanes, SketchServer, poisson_offsets, run_arrivals

    torch = torch_mod
    ROWS, COLS, NPOOL, NREQ = 32768, 16384, 4, 24
    g = torch.Generator(device="cuda")
    g.manual_seed(17)
    pools = [torch.rand(ROWS * COLS, device="cuda", generator=g) * 2 - 1 for _ in range(NPOOL)]

    def draw():
        return _stated_requests(torch, pools, ROWS, COLS, NREQ, np.random.default_rng(33))

    lanes = GpuLanes(4)
    lanes.reserve(ROWS, COLS, 1024)
    alone_reqs = draw()
    srv = SketchServer(hbm_budget=200e9, slots=4, lanes=lanes, keep_omega=True)
    alone = []
    for r in alone_reqs:
        srv.submit(r)
        srv.drain()
        torch.cuda.synchronize()
        alone.append(r.y.cpu().numpy().copy())
    assert max(r.m for r in alone_reqs) > 8192 and max(r.k for r in alone_reqs) > 4096
    # accuracy of sampled rows against FP64 (A rows from the pool, Omega as generated)
    h = sg.create()
    rng = np.random.default_rng(8)
    for r in alone_reqs[:8]:
        pool = next(p for p in pools if 0 <= (r.a - p.data_ptr()) // 4 < ROWS * COLS)
        off = (r.a - pool.data_ptr()) // 4
        r0, c0 = off % ROWS, off // ROWS
        rows = np.sort(rng.choice(r.m, min(r.m, 16), replace=False))
        idx = torch.from_numpy(rows + r0).cuda()
        A_rows = pool.view(COLS, ROWS)[c0:c0 + r.k][:, idx].t().double().cpu().numpy()
        Om = r.omega.view(r.l, r.k).t().double().cpu().numpy()
        ref = A_rows @ Om
        Y = alone[alone_reqs.index(r)].reshape(r.l, r.m).T[rows].astype(np.float64)
        Ys = torch.empty(r.m * r.l, device="cuda")
        sg.sgemm_fp32(h, 0, 0, r.m, r.l, r.k, 1.0, r.a, r.lda, r.omega.float(), r.k, 0.0, Ys, r.m)
        Ys_rows = Ys.view(r.l, r.m).t()[torch.from_numpy(rows).cuda()].double().cpu().numpy()
        e = np.linalg.norm(Y - ref) / np.linalg.norm(ref)
        es = np.linalg.norm(Ys_rows - ref) / np.linalg.norm(ref)
        assert e <= 2.0 * es, (r.rid, r.m, r.k, r.l, e, es)
    sg.destroy(h)
    srv.lanes = None  # the lanes are reused below

    # the same requests in a Poisson mix at 2000 req/s, 200 GB budget
    mix = draw()
    srv = SketchServer(hbm_budget=200e9, slots=4, lanes=lanes)
    st = run_arrivals(srv, mix, poisson_offsets(NREQ, 2000.0, np.random.default_rng(3)))
    torch.cuda.synchronize()
    assert st["requests"] == NREQ and st["peak_inflight"] >= 2
    for r, y in zip(mix, alone):
        assert np.array_equal(r.y.cpu().numpy().view(np.uint32), y.view(np.uint32)), r.rid
    # a budget that admits about two of the largest requests at a time
    tight = draw()
    budget = int(2.2 * max(r.footprint for r in tight))
    srv = SketchServer(hbm_budget=budget, slots=4, lanes=lanes)
    st = run_arrivals(srv, tight, [0.0] * NREQ)
    torch.cuda.synchronize()
    assert st["requests"] == NREQ and srv.peak_in_use <= budget
    for r, y in zip(tight, alone):
        assert np.array_equal(r.y.cpu().numpy().view(np.uint32), y.view(np.uint32)), r.rid
    lanes.close()
    del pools
    torch.cuda.empty_cache()


def test_batched_server_alone_and_in_mix(torch_mod):
    """max_batch = 4: the requests admitted together run as one grouped
    launch.  Requests from the stated distribution (A blocks of four
    32768 x 16384 pools) give the same bits alone (batches of one) and in a
    burst where batches of several form; sampled rows stay within 2x of
    SGEMM's FP64 error."""
    import shgemm_amd as sg
    from shgemm_amd.serve import GpuLanes, SketchServer, poisson_offsets, run_arrivals

    torch = torch_mod
    ROWS, COLS, NPOOL, NREQ = 32768, 16384, 4, 20
    g = torch.Generator(device="cuda")
    g.manual_seed(23)
    pools = [torch.rand(ROWS * COLS, device="cuda", generator=g) * 2 - 1 for _ in range(NPOOL)]

    def draw():
        return _stated_requests(torch, pools, ROWS, COLS, NREQ, np.random.default_rng(44))

    lanes = GpuLanes(2, max_batch=4)
    alone_reqs = draw()
    srv = SketchServer(hbm_budget=200e9, slots=2, lanes=lanes, keep_omega=True, max_batch=4)
    alone = []
    for r in alone_reqs:
        srv.submit(r)
        srv.drain()
        torch.cuda.synchronize()
        alone.append(r.y.cpu().numpy().copy())
    assert srv.batch_sizes == [1] * NREQ
    h = sg.create()
    rng = np.random.default_rng(9)
    for r, y in list(zip(alone_reqs, alone))[:6]:
        pool = next(p for p in pools if 0 <= (r.a - p.data_ptr()) // 4 < ROWS * COLS)
        off = (r.a - pool.data_ptr()) // 4
        r0, c0 = off % ROWS, off // ROWS
        rows = np.sort(rng.choice(r.m, min(r.m, 16), replace=False))
        A_rows = pool.view(COLS, ROWS)[c0:c0 + r.k][:, torch.from_numpy(rows + r0).cuda()].t().double().cpu().numpy()
        ref = A_rows @ r.omega.view(r.l, r.k).t().double().cpu().numpy()
        Ys = torch.empty(r.m * r.l, device="cuda")
        sg.sgemm_fp32(h, 0, 0, r.m, r.l, r.k, 1.0, r.a, r.lda, r.omega.float(), r.k, 0.0, Ys, r.m)
        Ys_rows = Ys.view(r.l, r.m).t()[torch.from_numpy(rows).cuda()].double().cpu().numpy()
        Y = y.reshape(r.l, r.m).T[rows].astype(np.float64)
        e = np.linalg.norm(Y - ref) / np.linalg.norm(ref)
        es = np.linalg.norm(Ys_rows - ref) / np.linalg.norm(ref)
        assert e <= 2.0 * es, (r.rid, e, es)
    sg.destroy(h)

    mix = draw()
    srv = SketchServer(hbm_budget=200e9, slots=2, lanes=lanes, max_batch=4)
    st = run_arrivals(srv, mix, poisson_offsets(NREQ, 50000.0, np.random.default_rng(5)))
    torch.cuda.synchronize()
    assert st["requests"] == NREQ and st["max_batch_seen"] >= 2   # batches did form
    for r, y in zip(mix, alone):
        assert np.array_equal(r.y.cpu().numpy().view(np.uint32), y.view(np.uint32)), r.rid
    lanes.close()
    del pools
    torch.cuda.empty_cache()
