"""GPU: the allocation-free entry point (shgemm_amd_shgemm_no_alloc) against
shgemm_amd_shgemm, both through the C ABI.

Bar: bit-identical results (integer, byte-exact comparison of the FP32
outputs) for two calls in a row on the static scratch, the second with a
smaller input than the first, for plans on every kind of scratch: k-sliced
(split-K partials), the 256x256 tile on whole-row exponents (row pass), the
pre-split plan (which runs as the 256x256 tile there) and the table's own
choice; a plan that needs row exponents for more rows than the scratch holds
is rejected before anything is enqueued.
"""
from __future__ import annotations

import pytest

from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu

FP16, TF32 = 0, 1


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


def _case(torch, m, n, k, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    A = torch.rand(m * k, device="cuda", generator=g) * 2 - 1
    B = (torch.rand(k * n, device="cuda", generator=g) * 2 - 1).half()
    return A, B


# (plan or -1 for the table's choice, slices, compute type, first shape, second (smaller) shape)
CASES = [
    (3, 16, FP16, (640, 384, 1 << 16), (256, 128, 1 << 14)),     # large-tile split-K
    (2, 32, TF32, (128, 128, 1 << 18), (96, 64, 50000)),         # 128x128 split-K, BF16 split
    (10, 64, FP16, (128, 128, 1 << 18), (96, 64, 50000)),        # deep 128x128x64 split-K
    (11, 4, FP16, (4096, 256, 8192), (1000, 200, 3000)),         # large tile of 64x128 wavefronts, split-K
    (5, 1, FP16, (2048, 1536, 4096), (1000, 700, 1500)),         # 256x256 tile, row pass
    (8, 1, FP16, (2048, 2048, 4096), (768, 512, 2048)),          # pre-split plan -> 256x256 tile
    (-1, 1, FP16, (4096, 4096, 4096), (1024, 1024, 1024)),       # the table's plans
    (-1, 1, FP16, (128, 128, 1 << 20), (128, 128, 1 << 16)),     # tall-skinny: table's split-K
]


@pytest.mark.parametrize("plan,slices,ct,first,second", CASES)
def test_no_alloc_matches_shgemm(env, plan, slices, ct, first, second):
    torch, sg, h = env
    if plan >= 0:
        sg.force_plan(h, plan, slices)
    try:
        for i, (m, n, k) in enumerate((first, second)):
            A, B = _case(torch, m, n, k, 7 * i + plan + 100)
            C = torch.empty(m * n, device="cuda")
            C_na = torch.empty(m * n, device="cuda")
            C.fill_(float("nan"))
            C_na.fill_(float("nan"))
            lvl = sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, C, m, ct)
            lvl_na = sg.shgemm_no_alloc(h, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, C_na, m, ct)
            torch.cuda.synchronize()
            assert lvl == lvl_na
            assert torch.equal(C_na.view(torch.int32), C.view(torch.int32)), (plan, slices, ct, m, n, k)
    finally:
        sg.force_plan(h, -1, 1)


def test_no_alloc_beta_and_transposes(env):
    """op_t A, op_t B, C stored transposed, beta != 0, on the k-sliced plan."""
    torch, sg, h = env
    m, n, k = 300, 200, 40000
    A, B = _case(torch, m, n, k, 5)
    C0 = torch.rand(m * n, device="cuda")
    C, C_na = C0.clone(), C0.clone()
    sg.force_plan(h, 3, 8)
    try:
        sg.shgemm(h, 1, 1, m, n, k, -0.5, A, k, B, n, 0.25, C, n, FP16, 1)
        sg.shgemm_no_alloc(h, 1, 1, m, n, k, -0.5, A, k, B, n, 0.25, C_na, n, FP16, 1)
        torch.cuda.synchronize()
    finally:
        sg.force_plan(h, -1, 1)
    assert torch.equal(C_na.view(torch.int32), C.view(torch.int32))


def test_no_alloc_rejects_rows_beyond_scratch(env):
    torch, sg, h = env
    m, n, k = (1 << 21) + 64, 8, 64
    A = torch.zeros(m * k, device="cuda")
    B = torch.zeros(k * n, device="cuda", dtype=torch.float16)
    C = torch.full((m * n,), 3.0, device="cuda")
    sg.force_plan(h, 5, 1)   # whole-row exponents
    try:
        with pytest.raises(sg.ShgemmError) as ei:
            sg.shgemm_no_alloc(h, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, C, m, FP16)
        assert ei.value.status == sg.STATUS_EINVAL and "rows" in str(ei.value)
        torch.cuda.synchronize()
        assert bool((C == 3.0).all())   # nothing was enqueued
    finally:
        sg.force_plan(h, -1, 1)
