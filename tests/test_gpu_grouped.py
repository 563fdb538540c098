"""GPU: grouped SHGEMM (persistent grouped launch captured in one HIP graph).

Bars:
  * every problem of a group is bit-identical to the same problem run alone
    through shgemm with the matching plan (same k-slices and tile depth);
    unsplit problems match any unsplit tile, since the k accumulation order of
    a C element does not depend on the tile shape;
  * graph replays (with eager work between them) and the eager launch
    sequence give identical results;
  * with the plans the library picks itself, each problem's error against an
    FP64 product is within 2x of the library's FP32 SGEMM on the same inputs;
  * on rows that grow along k (running exponents rising at later k-tiles)
    the grouped result still equals the single call bit for bit;
  * invalid problems are rejected before anything is enqueued.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests.gpu_util import FP16, FP16_CLASSIC, TF32, dev_buffer, read_buffer, require_gpu

pytestmark = pytest.mark.gpu

# tile class -> (unsplit plan variant, split plan variant) of the single call
CLASS_PLAN = {0: (0, 4), 1: (1, 2), 2: (1, 3)}


@pytest.fixture(scope="module")
def h():
    require_gpu()
    import shgemm_amd as sg

    handle = sg.create()
    yield handle
    sg.destroy(handle)


class Prob:
    """Host and device operands of one problem (column-major, padded ld)."""

    def __init__(self, rng, op_a, op_b, op_c, m, n, k, alpha=1.0, beta=0.0, pad=3):
        self.m, self.n, self.k, self.alpha, self.beta = m, n, k, alpha, beta
        self.lda = (m if op_a == 0 else k) + pad
        self.ldb = (k if op_b == 0 else n) + pad
        self.ldc = (m if op_c == 0 else n) + pad
        a_cols = k if op_a == 0 else m
        b_cols = n if op_b == 0 else k
        c_cols = n if op_c == 0 else m
        self.A = oracle.gen_uniform(rng, max(1, self.lda * a_cols))
        self.B16 = oracle.gen_f16_normal(rng, max(1, self.ldb * b_cols))
        self.C0 = oracle.gen_uniform(rng, max(1, self.ldc * c_cols))
        self.a_t, self.a_p = dev_buffer(self.A)
        self.b_t, self.b_p = dev_buffer(self.B16)
        self.c_t, self.c_p = dev_buffer(self.C0)

    def reset_c(self):
        import torch

        self.c_t[:self.C0.size].copy_(torch.from_numpy(self.C0).to(self.c_t.device))

    def c(self):
        return read_buffer(self.c_t, 0, self.C0.size).copy()

    def tup(self):
        return (self.m, self.n, self.k, self.a_p, self.lda, self.b_p, self.ldb, self.c_p, self.ldc, self.alpha,
                self.beta)


def _ct(variant):
    return 0 if variant == FP16_CLASSIC else variant


def _single(h, p, op_a, op_b, op_c, variant, plan):
    import shgemm_amd as sg

    p.reset_c()
    sg.force_plan(h, *plan)
    try:
        sg.shgemm(h, op_a, op_b, p.m, p.n, p.k, p.alpha, p.a_p, p.lda, p.b_p, p.ldb, p.beta, p.c_p, p.ldc,
                  _ct(variant), op_c)
    finally:
        sg.force_plan(h, -1, 1)
    return p.c()


SPECS = [  # m, n, k, alpha, beta, tile class, slices
    (200, 96, 300, 1.0, 0.0, 1, 1),
    (64, 64, 64, 1.0, 0.0, 0, 1),
    (1, 1, 1, 1.0, 0.0, 0, 1),
    (129, 257, 1000, -0.5, 0.0, 2, 1),
    (300, 70, 33, 1.0, 0.25, 1, 1),
    (0, 5, 5, 1.0, 0.0, 1, 1),
    (17, 0, 3, 1.0, 0.0, 1, 1),
    (40, 50, 0, 1.0, 0.5, 1, 1),
    (33, 44, 55, 0.0, 2.0, 1, 1),
    (256, 128, 4096, 1.0, 0.0, 1, 4),
    (130, 260, 2048, 2.0, -1.0, 2, 2),
    (100, 90, 8192, 1.0, 0.0, 0, 8),
    (520, 64, 700, 1.0, 0.0, 0, 1),
]


@pytest.mark.parametrize("variant", [FP16, TF32, FP16_CLASSIC])
@pytest.mark.parametrize("ops", [(0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 0)])
def test_group_matches_single_calls_bitwise(h, variant, ops):
    import shgemm_amd as sg

    op_a, op_b, op_c = ops
    rng = np.random.default_rng(1000 + 10 * variant + 4 * op_a + 2 * op_b + op_c)
    probs = [Prob(rng, op_a, op_b, op_c, m, n, k, al, be) for m, n, k, al, be, _, _ in SPECS]
    classes = [s[5] for s in SPECS]
    slices = [s[6] for s in SPECS]
    if variant == FP16_CLASSIC:
        sg.set_exponent_scaling(h, False)
    try:
        for p in probs:
            p.reset_c()
        with sg.Group(h, op_a, op_b, [p.tup() for p in probs], _ct(variant), op_c, classes, slices) as g:
            got_cls, got_sl, units = g.plan()
            g.run()
            grouped = [p.c() for p in probs]
        for i, (p, s) in enumerate(zip(probs, SPECS)):
            if p.m == 0 or p.n == 0 or p.k == 0 or p.alpha == 0.0:
                assert got_cls[i] == -1 and got_sl[i] == 0
            else:
                assert got_cls[i] == s[5] and got_sl[i] == s[6], (i, got_cls[i], got_sl[i])
            unsplit, split = CLASS_PLAN[s[5]]
            plan = (split, s[6]) if s[6] > 1 else (unsplit, 1)
            alone = _single(h, p, op_a, op_b, op_c, variant, plan)
            assert np.array_equal(grouped[i].view(np.uint32), alone.view(np.uint32)), f"problem {i} {s}"
        assert units > 0
    finally:
        sg.set_exponent_scaling(h, True)


@pytest.mark.parametrize("concurrent", ["0", "1"])
def test_graph_replays_match_eager(h, monkeypatch, concurrent):
    """Three graph replays with eager torch work between them, and the eager
    launch sequence, all give the same bits; the graph chains the tile
    classes (the default) or runs them as parallel branches
    (MTK_SHGEMM_GROUP_CONCURRENT=1, read when the group is created)."""
    import torch

    import shgemm_amd as sg

    monkeypatch.setenv("MTK_SHGEMM_GROUP_CONCURRENT", concurrent)
    rng = np.random.default_rng(77)
    probs = [Prob(rng, 0, 0, 0, m, n, k, al, be) for m, n, k, al, be, _, _ in SPECS]
    with sg.Group(h, 0, 0, [p.tup() for p in probs], 0, 0, [s[5] for s in SPECS], [s[6] for s in SPECS]) as g:
        for p in probs:
            p.reset_c()
        g.run(use_graph=False)
        eager = [p.c() for p in probs]
        for rep in range(3):
            for p in probs:
                p.reset_c()
            junk = torch.randn(1 << 20, device="cuda")
            (junk * 2).sum().item()  # unrelated eager work between replays
            g.run(use_graph=True)
            torch.cuda.synchronize()
            for i, p in enumerate(probs):
                assert np.array_equal(p.c().view(np.uint32), eager[i].view(np.uint32)), f"replay {rep} problem {i}"


def _fp64_err(p, C, op_a=0, op_b=0, op_c=0):
    m, n, k = p.m, p.n, p.k
    A = p.A[:p.lda * k].reshape(k, p.lda)[:, :m].T.astype(np.float64)
    B = p.B16.astype(np.float64)[:p.ldb * n].reshape(n, p.ldb)[:, :k].T
    ref = p.alpha * (A @ B) + p.beta * p.C0[:p.ldc * n].reshape(n, p.ldc)[:, :m].T.astype(np.float64)
    got = C[:p.ldc * n].reshape(n, p.ldc)[:, :m].T.astype(np.float64)
    return np.linalg.norm(got - ref) / np.linalg.norm(ref)


@pytest.mark.parametrize("variant", [FP16, TF32])
def test_group_auto_plan_error_vs_fp64(h, variant):
    """Problems drawn like the batched benchmark (log-uniform m and k, n in
    {64..512}), smaller; library-chosen plans; error within 2x of SGEMM."""
    import torch

    import shgemm_amd as sg

    rng = np.random.default_rng(5 + variant)
    specs = []
    for _ in range(24):
        m = int(np.exp(rng.uniform(np.log(64), np.log(1024))))
        k = int(np.exp(rng.uniform(np.log(64), np.log(4096))))
        n = int(rng.choice([64, 128, 256, 512]))
        specs.append((m, n, k))
    probs = [Prob(rng, 0, 0, 0, m, n, k, 1.0, 0.0, pad=0) for m, n, k in specs]
    with sg.Group(h, 0, 0, [p.tup() for p in probs], variant) as g:
        g.run()
        outs = [p.c() for p in probs]
    for p, C in zip(probs, outs):
        B32 = p.B16.astype(np.float32)
        b_t, b_p = dev_buffer(B32)
        cs_t, cs_p = dev_buffer(np.zeros_like(p.C0))
        sg.sgemm_fp32(h, 0, 0, p.m, p.n, p.k, 1.0, p.a_p, p.lda, b_p, p.ldb, 0.0, cs_p, p.ldc)
        Cs = read_buffer(cs_t, 0, p.C0.size)
        e, es = _fp64_err(p, C), _fp64_err(p, Cs)
        assert e <= 2.0 * es + 2.0 ** -22, (p.m, p.n, p.k, e, es)
        del b_t, cs_t
    torch.cuda.synchronize()


def test_group_rejects_invalid(h):
    import shgemm_amd as sg

    rng = np.random.default_rng(3)
    good = Prob(rng, 0, 0, 0, 64, 64, 64)
    bad = list(good.tup())
    bad[4] = 10  # lda < m
    with pytest.raises(sg.ShgemmError) as ei:
        sg.Group(h, 0, 0, [good.tup(), tuple(bad)], 0)
    assert ei.value.status == sg.STATUS_EINVAL
    with pytest.raises(sg.ShgemmError) as ei:
        sg.Group(h, 0, 0, [good.tup()], 0, 0, [4], [1])
    assert ei.value.status == sg.STATUS_EINVAL
    with pytest.raises(sg.ShgemmError) as ei:
        sg.Group(h, 0, 0, [], 0)
    assert ei.value.status == sg.STATUS_EINVAL


class GrowingProb(Prob):
    """A problem on oracle.gen_growing_rows operands (op_n / op_n / op_n):
    rows of A that grow along k, so running exponents rise at later k-tiles."""

    def __init__(self, rng, m, n, k, step, width):
        self.m, self.n, self.k, self.alpha, self.beta = m, n, k, 1.0, 0.0
        self.lda, self.ldb, self.ldc = m, k, m
        self.A, self.B16 = oracle.gen_growing_rows(rng, m, n, k, step=step, width=width)
        self.C0 = np.zeros(m * n, np.float32)
        self.a_t, self.a_p = dev_buffer(self.A)
        self.b_t, self.b_p = dev_buffer(self.B16)
        self.c_t, self.c_p = dev_buffer(self.C0)


def test_group_with_growing_rows_matches_single_calls(h):
    """Running exponents inside the persistent kernel (re-split and rescale
    when a row's maximum grows, state re-initialised per work unit): every
    problem equals the same plan run alone, bit for bit, and stays within 2x
    of the library SGEMM's FP64 error."""
    import shgemm_amd as sg
    from tests.gpu_util import run_sgemm

    rng = np.random.default_rng(4242)
    specs = [  # m, n, k, step, width, tile class, slices
        (300, 200, 320, 3, 32, 0, 1), (301, 150, 320, 3, 32, 1, 1), (256, 256, 320, 3, 32, 2, 1),
        (190, 64, 1024, 4, 128, 0, 2), (200, 128, 1024, 4, 128, 1, 2), (128, 256, 1024, 4, 128, 2, 2),
    ]
    probs = [GrowingProb(rng, m, n, k, st, w) for m, n, k, st, w, _, _ in specs]
    classes = [s[5] for s in specs]
    slices = [s[6] for s in specs]
    with sg.Group(h, 0, 0, [p.tup() for p in probs], sg.fp16, 0, classes, slices) as g:
        got_cls, got_sl, _ = g.plan()
        g.run()
        grouped = [p.c() for p in probs]
    for i, (p, s) in enumerate(zip(probs, specs)):
        assert (got_cls[i], got_sl[i]) == (s[5], s[6])
        unsplit, split = CLASS_PLAN[s[5]]
        alone = _single(h, p, 0, 0, 0, FP16, (split, s[6]) if s[6] > 1 else (unsplit, 1))
        assert np.array_equal(grouped[i].view(np.uint32), alone.view(np.uint32)), f"problem {i}"
        B32 = p.B16.astype(np.float32)
        ref = p.A.reshape(p.k, p.m).T.astype(np.float64) @ B32.reshape(p.n, p.k).T.astype(np.float64)
        Cs = run_sgemm(h, 0, 0, 0, p.m, p.n, p.k, 1.0, p.A, p.m, B32, p.k, 0.0, np.zeros(p.m * p.n, np.float32), p.m)
        err = oracle.rel_fro_error(grouped[i].reshape(p.n, p.m).T, ref)
        err_s = oracle.rel_fro_error(Cs.reshape(p.n, p.m).T, ref)
        assert err <= 2.0 * err_s, (i, err, err_s)
