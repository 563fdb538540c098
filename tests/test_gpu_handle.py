"""GPU: the handle's scratch under several streams and under graph capture,
the deep-prefetch plan's 32-bit offset guard, and debug mode.

Bars:
  * calls through one handle that alternate between two streams, each using
    the handle's split-K workspace, row-exponent scratch or pre-split
    buffer, give the same bits as the same calls run one at a time;
  * graph capture: after reserve(), captured calls (pre-split, 256x256 on
    whole-row exponents, split-K; fp16 and tf32) replay to the eager results
    on every replay, with eager work between replays; without reserve a pre-split
    call runs as the 256x256 tile (bit-identical) and a call whose row
    scratch would have to grow raises before anything is enqueued;
  * plan 7 (deep prefetch) resolves to plan 1 where its per-lane byte
    offsets would not fit 32 bits (op_b = N with ldb = k >= 2^23, op_t A
    with lda = k >= 2^23), and stays plan 7 below that (query only: nothing
    of that size is launched);
  * plan 8 (pre-split) reads op_n B where it lies while the tile's column
    byte offsets fit 32 bits (BN * ldb * 2 < 2^32; this shape runs the
    128 x 128 form) and packs it otherwise; on both sides of that limit it
    gives plan 5's bits (B of three columns at ldb = 2^24 - 8 and 2^24: a
    64 MiB buffer);
  * set_debug_mode prints the reference's debug line for each call.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu

FP16 = 0
PLAN_P1, PLAN_P2, PLAN_P1SK, PLAN_BD, PLAN_PS = 1, 5, 3, 7, 8


@pytest.fixture(scope="module")
def env():
    require_gpu()
    import torch

    import shgemm_amd as sg

    return torch, sg


def _operands(torch, m, n, k, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    A = torch.rand(m * k, device="cuda", generator=g) * 2 - 1
    B = (torch.rand(k * n, device="cuda", generator=g) * 2 - 1).half()
    return A, B


@pytest.mark.parametrize("k_log,op_a,want", [(22, 0, PLAN_BD), (23, 0, PLAN_P1), (24, 0, PLAN_P1),
                                             (22, 1, PLAN_BD), (23, 1, PLAN_P1)])
def test_deep_prefetch_offset_guard(env, k_log, op_a, want):
    torch, sg = env
    h = sg.create()
    try:
        sg.force_plan(h, PLAN_BD, 1)
        v, s, src = sg.query_plan(h, op_a, 0, 256, 256, 1 << k_log, FP16)
        assert (v, s, src) == (want, 1, 2), (k_log, op_a, v, s, src)
    finally:
        sg.destroy(h)


# (plan, slices, m, n, k): the three kinds of handle scratch
SCRATCH_CASES = [(PLAN_P1SK, 8, 512, 384, 1 << 15), (PLAN_P2, 1, 1024, 768, 2048), (PLAN_PS, 1, 768, 512, 4096)]


def test_two_streams_share_scratch_in_order(env):
    torch, sg = env
    h = sg.create()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    cases = []
    for i, (plan, sl, m, n, k) in enumerate(SCRATCH_CASES * 2):
        A, B = _operands(torch, m, n, k, 100 + i)
        cases.append((plan, sl, m, n, k, A, B, torch.empty(m * n, device="cuda")))
    try:
        # one at a time
        want = []
        for plan, sl, m, n, k, A, B, C in cases:
            sg.force_plan(h, plan, sl)
            sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, C, m, FP16)
            torch.cuda.synchronize()
            want.append(C.clone())
        # alternating streams, no host synchronisation in between
        for C in (c[7] for c in cases):
            C.fill_(float("nan"))
        torch.cuda.synchronize()
        for rep in range(3):
            for i, (plan, sl, m, n, k, A, B, C) in enumerate(cases):
                sg.set_cuda_stream(h, s1 if (i + rep) % 2 == 0 else s2)
                sg.force_plan(h, plan, sl)
                sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, C, m, FP16)
            torch.cuda.synchronize()
            for (plan, sl, m, n, k, A, B, C), w in zip(cases, want):
                assert torch.equal(C.view(torch.int32), w.view(torch.int32)), (rep, plan, sl, m, n, k)
    finally:
        torch.cuda.synchronize()
        sg.destroy(h)


def _capture(torch, sg, h, fn):
    """Captures fn() on a side stream (the handle's stream during capture)."""
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        sg.set_cuda_stream(h, side)
        with torch.cuda.graph(graph, stream=side):
            fn()
    torch.cuda.current_stream().wait_stream(side)
    sg.set_cuda_stream(h, torch.cuda.current_stream())
    return graph


@pytest.mark.parametrize("ct", [0, 1])  # fp16, tf32 (its pre-split plan packs three BF16 parts of A, two of B)
def test_graph_capture_after_reserve(env, ct):
    torch, sg, = env
    h = sg.create()
    sg.set_cuda_stream(h, torch.cuda.current_stream())
    try:
        items = []
        for i, (plan, sl, m, n, k) in enumerate(SCRATCH_CASES):
            A, B = _operands(torch, m, n, k, 200 + i)
            items.append((plan, sl, m, n, k, A, B, torch.empty(m * n, device="cuda")))
        sg.reserve(h, max(c[2] for c in SCRATCH_CASES), max(c[3] for c in SCRATCH_CASES),
                   max(c[4] for c in SCRATCH_CASES))

        def calls():
            for plan, sl, m, n, k, A, B, C in items:
                sg.force_plan(h, plan, sl)
                sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, C, m, ct)

        calls()
        torch.cuda.synchronize()
        eager = [it[7].clone() for it in items]
        for it in items:
            it[7].fill_(float("nan"))
        graph = _capture(torch, sg, h, calls)
        for rep in range(4):
            for it in items:
                it[7].fill_(float("nan"))
            graph.replay()
            torch.cuda.synchronize()
            for it, want in zip(items, eager):
                assert torch.equal(it[7].view(torch.int32), want.view(torch.int32)), (rep, it[:5])
            # eager work through the same handle between replays
            calls()
            torch.cuda.synchronize()
        del graph
    finally:
        torch.cuda.synchronize()
        sg.destroy(h)


def test_graph_capture_without_reserve(env):
    torch, sg = env
    h = sg.create()
    sg.set_cuda_stream(h, torch.cuda.current_stream())
    try:
        m, n, k = 768, 512, 4096
        A, B = _operands(torch, m, n, k, 300)
        C = torch.empty(m * n, device="cuda")
        # the row scratch exists for m rows (an eager 256x256 call), the
        # pre-split buffer does not: the captured pre-split call runs the
        # 256x256 tile, whose bits are the pre-split plan's
        sg.force_plan(h, PLAN_P2, 1)
        sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, C, m, FP16)
        torch.cuda.synchronize()
        want = C.clone()
        C.fill_(float("nan"))

        def ps_call():
            sg.force_plan(h, PLAN_PS, 1)
            sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, C, m, FP16)

        graph = _capture(torch, sg, h, ps_call)
        for _ in range(2):
            C.fill_(float("nan"))
            graph.replay()
            torch.cuda.synchronize()
            assert torch.equal(C.view(torch.int32), want.view(torch.int32))
        del graph
        # a call needing more row scratch than the handle holds raises
        m2 = 4 * m
        A2, B2 = _operands(torch, m2, n, k, 301)
        C2 = torch.empty(m2 * n, device="cuda")
        raised = []

        def grow_call():
            sg.force_plan(h, PLAN_P2, 1)
            try:
                sg.shgemm(h, 0, 0, m2, n, k, 1.0, A2, m2, B2, k, 0.0, C2, m2, FP16)
            except sg.ShgemmError as e:
                raised.append(str(e))
            # keep the capture non-empty
            C2.zero_()

        g2 = _capture(torch, sg, h, grow_call)
        del g2
        assert raised and "reserve" in raised[0], raised
    finally:
        torch.cuda.synchronize()
        sg.force_plan(h, -1, 1)
        sg.destroy(h)


def test_debug_mode_prints_call_line(env, capfd):
    torch, sg = env
    h = sg.create()
    try:
        m, n, k = 300, 200, 100
        A, B = _operands(torch, m, n, k, 5)
        C = torch.empty(m * n, device="cuda")
        sg.set_debug_mode(h, 1)
        sg.shgemm(h, 1, 0, m, n, k, 1.0, A, k, B, k, 0.0, C, m, FP16)
        torch.cuda.synchronize()
        sg.set_debug_mode(h, 0)
        sg.shgemm(h, 1, 0, m, n, k, 1.0, A, k, B, k, 0.0, C, m, FP16)
        torch.cuda.synchronize()
        out = capfd.readouterr().out
        lines = [l for l in out.splitlines() if l.startswith("[shape=")]
        assert len(lines) == 1, out
        assert lines[0].startswith("[shape=(300,200,100),op_a=T,op_b=N,compute_type=FP16]"), lines[0]
    finally:
        sg.destroy(h)


@pytest.mark.parametrize("ldb", [(1 << 24) - 8, 1 << 24])
def test_presplit_reads_b_in_place_below_the_offset_limit(env, capfd, ldb):
    torch, sg = env
    h = sg.create()
    try:
        m, n, k = 300, 3, 128
        rng = np.random.default_rng(ldb % 97)
        A = torch.from_numpy(rng.uniform(-1, 1, m * k).astype(np.float32)).cuda()
        cols = torch.from_numpy(rng.standard_normal((n, k)).astype(np.float16)).cuda()
        B = torch.zeros((n - 1) * ldb + k, dtype=torch.float16, device="cuda")
        for j in range(n):
            B[j * ldb:j * ldb + k] = cols[j]
        out = {}
        for plan in (PLAN_PS, PLAN_P2):
            C = torch.full((m * n,), float("nan"), device="cuda")
            sg.force_plan(h, plan, 1)
            sg.set_debug_mode(h, 1)
            sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B, ldb, 0.0, C, m, FP16)
            torch.cuda.synchronize()
            sg.set_debug_mode(h, 0)
            out[plan] = C.cpu().numpy()
        lines = [l for l in capfd.readouterr().out.splitlines() if l.startswith("[shape=")]
        assert len(lines) == 2, lines
        assert "tile=128x128x32" in lines[0], lines[0]
        in_place = 128 * ldb * 2 < (1 << 32)
        assert ("pre-split A, B in place" in lines[0]) == in_place, lines[0]
        assert "pre-split" in lines[0], lines[0]
        assert np.array_equal(out[PLAN_PS].view(np.uint32), out[PLAN_P2].view(np.uint32))
        ref = A.cpu().numpy().astype(np.float64).reshape(k, m).T @ cols.cpu().numpy().astype(np.float64).T
        got = out[PLAN_PS].reshape(n, m).T
        assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()
    finally:
        sg.force_plan(h, -1, 1)
        sg.destroy(h)
