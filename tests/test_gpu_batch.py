"""GPU: batches (shgemm_amd.Batch / shgemm_amd_batch_*) and the 128x512
grouped tile class.

Bars:
  * a batch's problems are bit-identical to the same problems in a group
    (both take tile class and slices from the plan table's grouped rows) and
    to single shgemm calls with the matching plan, over several runs that
    reuse the batch's slots with different problems;
  * problems in tile class 3 (128 x 512, prescaled fp16) give the bits of
    the single-call large tile (unsplit: plan 1; k-sliced: plan 3) and stay
    within 2x of the library SGEMM's FP64 error; class 3 is rejected for the
    compute types that do not compile it;
  * a batch needing more scratch than its slots hold, or more problems than
    it was created for, is rejected before anything is enqueued.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu

FP16, TF32 = 0, 1
CLASS_PLAN = {0: (0, 4), 1: (1, 2), 2: (1, 3), 3: (1, 3)}  # tile class -> (unsplit, split) single-call variant


@pytest.fixture(scope="module")
def env():
    require_gpu()
    import torch

    import shgemm_amd as sg

    h = sg.create()
    sg.set_cuda_stream(h, torch.cuda.current_stream())
    yield torch, sg, h
    torch.cuda.synchronize()
    sg.destroy(h)


def _problems(torch, shapes, seed, op_a=0, op_b=0, op_c=0):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    out = []
    for m, n, k in shapes:
        A = torch.rand(m * k, device="cuda", generator=g) * 2 - 1
        B = (torch.randn(k * n, device="cuda", generator=g) / np.sqrt(n)).half()
        C = torch.full((m * n,), float("nan"), device="cuda")
        lda, ldb, ldc = (m if op_a == 0 else k), (k if op_b == 0 else n), (m if op_c == 0 else n)
        out.append([m, n, k, A, lda, B, ldb, C, ldc, 1.0, 0.0])
    return out


def _single(torch, sg, h, q, cls, s, ct, ops):
    m, n, k, A, lda, B, ldb, C, ldc, al, be = q
    out = torch.full((m * n,), float("nan"), device="cuda")
    sg.force_plan(h, CLASS_PLAN[cls][1 if s > 1 else 0], s)
    try:
        sg.shgemm(h, ops[0], ops[1], m, n, k, al, A, lda, B, ldb, be, out, ldc, ct, ops[2])
    finally:
        sg.force_plan(h, -1, 1)
    return out


RUNS = [
    [(700, 512, 3000), (256, 64, 256), (2048, 256, 9000), (300, 128, 40000), (1500, 512, 700)],
    [(4000, 512, 16000), (128, 1000, 2000), (64, 64, 64)],
    [(8192, 256, 2500), (1024, 512, 1024), (333, 77, 5000), (2000, 128, 300)],
]


@pytest.mark.parametrize("ops", [(0, 0, 0), (1, 0, 1)])
def test_batch_runs_match_group_and_single_calls(env, ops):
    torch, sg, h = env
    with sg.Batch(h, ops[0], ops[1], FP16, ops[2], max_problems=8, bytes=512 << 20, slots=2) as bt:
        for r, shapes in enumerate(RUNS):
            probs = _problems(torch, shapes, 10 + r, *ops)
            bt.run(probs, torch.cuda.current_stream())
            torch.cuda.synchronize()
            cls, slices, units = bt.plan()
            assert units > 0 and all(c >= 0 for c in cls)
            got = [q[7].clone() for q in probs]
            for q in probs:
                q[7].fill_(float("nan"))
            with sg.Group(h, ops[0], ops[1], probs, FP16, ops[2]) as grp:
                gc, gs, _ = grp.plan()
                assert (gc, gs) == (cls, slices)
                grp.run(torch.cuda.current_stream())
                torch.cuda.synchronize()
            for i, q in enumerate(probs):
                assert torch.equal(q[7].view(torch.int32), got[i].view(torch.int32)), (r, i, shapes[i])
                ref = _single(torch, sg, h, q, cls[i], slices[i], FP16, ops)
                assert torch.equal(ref.view(torch.int32), got[i].view(torch.int32)), (r, i, cls[i], slices[i])


def test_class3_matches_large_tile_and_fp64(env):
    torch, sg, h = env
    shapes = [(1000, 512, 4000), (4096, 500, 2048), (130, 300, 700), (2000, 512, 30000)]
    slices = [1, 1, 1, 4]
    probs = _problems(torch, shapes, 3)
    with sg.Group(h, 0, 0, probs, FP16, 0, tile_class=[3] * len(shapes), slices=slices) as grp:
        cls, used, _ = grp.plan()
        assert cls == [3] * len(shapes) and used == slices
        grp.run(torch.cuda.current_stream())
        torch.cuda.synchronize()
    rng = np.random.default_rng(4)
    Cs = torch.empty(max(m * n for m, n, _ in shapes), device="cuda")
    for i, (q, s) in enumerate(zip(probs, slices)):
        m, n, k, A, lda, B, ldb, C, ldc = q[:9]
        ref = _single(torch, sg, h, q, 3, s, FP16, (0, 0, 0))
        assert torch.equal(ref.view(torch.int32), C.view(torch.int32)), (i, shapes[i])
        rows = np.sort(rng.choice(m, min(m, 64), replace=False))
        idx = torch.from_numpy(rows).cuda()
        B32 = B.float()
        want = A.view(k, m).t()[idx].double().cpu().numpy() @ B32.view(n, k).t().double().cpu().numpy()
        sg.sgemm_fp32(h, 0, 0, m, n, k, 1.0, A, m, B32, k, 0.0, Cs, m)
        err = oracle.rel_fro_error(C.view(n, m).t()[idx].cpu().numpy(), want)
        err_s = oracle.rel_fro_error(Cs[:m * n].view(n, m).t()[idx].cpu().numpy(), want)
        assert err <= 2.0 * err_s, (i, err, err_s)


def test_class3_rejected_without_it(env):
    torch, sg, h = env
    probs = _problems(torch, [(256, 512, 256)], 5)
    with pytest.raises(sg.ShgemmError) as ei:
        sg.Group(h, 0, 0, probs, TF32, 0, tile_class=[3])
    assert ei.value.status == sg.STATUS_EINVAL


def test_batch_limits(env):
    torch, sg, h = env
    small = _problems(torch, [(64, 64, 64)] * 3, 7)
    with sg.Batch(h, 0, 0, FP16, 0, max_problems=2, bytes=1 << 20, slots=1) as bt:
        with pytest.raises(sg.ShgemmError) as ei:
            bt.run(small)
        assert ei.value.status == sg.STATUS_EINVAL
        bt.run(small[:1])
        torch.cuda.synchronize()
        assert not torch.isnan(small[0][7]).any().item()
    # a problem the plan table splits along k needs its partials (slices x m x
    # n floats) in the slot: 4096 bytes cannot hold them
    for shape in ((256, 64, 16384), (128, 128, 65536), (256, 256, 262144)):
        probs = _problems(torch, [shape], 8)
        with sg.Batch(h, 0, 0, FP16, 0, max_problems=1, bytes=64 << 20, slots=1) as bt:
            bt.run(probs)
            torch.cuda.synchronize()
            _, slices, _ = bt.plan()
        if slices[0] > 1:
            break
    assert slices[0] > 1, "no split problem found"
    with sg.Batch(h, 0, 0, FP16, 0, max_problems=1, bytes=4096, slots=1) as bt:
        with pytest.raises(sg.ShgemmError) as ei:
            bt.run(probs)
        assert ei.value.status == sg.STATUS_EINVAL and "bytes" in str(ei.value)
