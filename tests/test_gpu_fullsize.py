"""GPU: the reference's own sweeps at their sizes and the workload
configurations at their stated sizes, against an FP64 product of sampled rows.

Sweeps of the reference's tests:
  * test/shgemm_test.cu:12-15, 259-306: m = n = k = 2^10, 2^12, 2^14, every
    (op_a, op_b, op_c), compute types tf32 and fp16; A and B uniform on
    (0, 1) (the reference fills them with curand's uniform generator,
    :247-252), B rounded to FP16;
  * test/shgemm_k_slicing_test.cu:10-13, 153-170: m = n = 128 and
    k = 2^11 .. 2^22, op_n / op_n, the same input distribution.
Workloads at their stated sizes:
  * wide-exponent sketch: A 16384 x 16384 FP32 = sign * U[1,2) * 2^e with e
    uniform on [-15, 15], Omega 16384 x 256 FP16 N(0,1)/16; shgemm op_n and
    op_t and hsgemm with the roles transposed, fp16 and tf32;
  * 512 uneven problems in one grouped launch (m log-uniform [256, 8192],
    n in {64, 128, 256, 512}, k log-uniform [256, 16384]).

Bar: every row of C depends only on its own row of op(A), so the sampled
rows of C are an exact restriction of the product; their relative Frobenius
error against the numpy FP64 product of the same rows of op(A) and all of
op(B) must be within 2x of the library FP32 SGEMM's on the same rows.
Split-K results are also bit-identical from run to run.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu

FP16, TF32 = 0, 1
SAMPLE_ROWS = 256


@pytest.fixture(scope="module")
def env():
    require_gpu()
    import torch

    import shgemm_amd as sg

    h = sg.create()
    yield torch, sg, h
    torch.cuda.synchronize()
    sg.destroy(h)
    torch.cuda.empty_cache()


def _gen(torch, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return g


def _rel_err(C_rows: np.ndarray, ref: np.ndarray) -> float:
    return float(np.linalg.norm(C_rows.astype(np.float64) - ref) / np.linalg.norm(ref))


# Views of column-major storage as row-major torch tensors: a column-major
# r x c matrix with leading dimension r is the row-major c x r tensor, so
# op(A)[i, :] is a column of that view for op_n and a row for op_t.
def _op_rows(X, op, rows_op, cols_op, idx):
    """Rows `idx` of op(X), where op(X) is rows_op x cols_op and X is stored
    column-major with the smallest leading dimension."""
    if op == 0:
        return X.view(cols_op, rows_op)[:, idx].t()
    return X.view(rows_op, cols_op)[idx]


def _op_full(X, op, rows_op, cols_op):
    return X.view(cols_op, rows_op).t() if op == 0 else X.view(rows_op, cols_op)


def _c_rows(C, op_c, m, n, idx):
    return C.view(n, m)[:, idx].t() if op_c == 0 else C.view(m, n)[idx]


# ---- reference sweep: m = n = k = 2^10, 2^12, 2^14 ---------------------------
_SWEEP = {}


def _sweep_inputs(torch, N):
    if N not in _SWEEP:
        _SWEEP.clear()
        torch.cuda.empty_cache()
        g = _gen(torch, 10 + N)
        A = torch.rand(N * N, device="cuda", generator=g)
        B16 = torch.rand(N * N, device="cuda", generator=g).half()
        _SWEEP[N] = (A, B16, B16.float())
    return _SWEEP[N]


@pytest.mark.parametrize("op_b", [0, 1])
@pytest.mark.parametrize("op_a", [0, 1])
@pytest.mark.parametrize("log_n", [10, 12, 14])
def test_reference_sweep(env, log_n, op_a, op_b):
    """Every op_c and both compute types for this (size, op_a, op_b)."""
    torch, sg, h = env
    N = 1 << log_n
    m = n = k = N
    A, B16, B32 = _sweep_inputs(torch, N)
    rows = np.sort(np.random.default_rng(log_n * 4 + op_a * 2 + op_b).choice(m, SAMPLE_ROWS, replace=False))
    idx = torch.from_numpy(rows).cuda()
    A_rows = _op_rows(A, op_a, m, k, idx).double().cpu().numpy()
    opB = _op_full(B32, op_b, k, n).double().cpu().numpy()
    ref = A_rows @ opB
    lda, ldb = (m if op_a == 0 else k), (k if op_b == 0 else n)
    C = torch.empty(m * n, device="cuda")
    sg.sgemm_fp32(h, op_a, op_b, m, n, k, 1.0, A, lda, B32, ldb, 0.0, C, m)
    err_s = _rel_err(_c_rows(C, 0, m, n, idx).cpu().numpy(), ref)
    for ct in (TF32, FP16):
        for op_c in (0, 1):
            ldc = m if op_c == 0 else n
            C.fill_(float("nan"))
            sg.shgemm(h, op_a, op_b, m, n, k, 1.0, A, lda, B16, ldb, 0.0, C, ldc, ct, op_c)
            err = _rel_err(_c_rows(C, op_c, m, n, idx).cpu().numpy(), ref)
            print(f"sweep N=2^{log_n} op_a={op_a} op_b={op_b} op_c={op_c} ct={ct}: err={err:.3e} "
                  f"sgemm={err_s:.3e}", flush=True)
            assert err <= 2.0 * err_s, (N, op_a, op_b, op_c, ct, err, err_s)
    del C


# ---- reference k-slicing sweep: m = n = 128, k = 2^11 .. 2^22 -----------------
KMAX = 1 << 22


@pytest.fixture(scope="module")
def kslice_inputs(env):
    torch, _, _ = env
    _SWEEP.clear()
    torch.cuda.empty_cache()
    g = _gen(torch, 11)
    A = torch.rand(128 * KMAX, device="cuda", generator=g)          # 128 x KMAX, lda = 128
    B16 = torch.rand(128 * KMAX, device="cuda", generator=g).half()  # KMAX x 128, ldb = KMAX
    yield A, B16
    del A, B16
    torch.cuda.empty_cache()


@pytest.mark.parametrize("log_k", list(range(11, 23)))
def test_reference_k_slicing_sweep(env, kslice_inputs, log_k):
    """The first k columns of A (lda = 128) and the first k rows of B
    (ldb = 2^22); at k = 2^22 also bit-identical over 10 runs."""
    torch, sg, h = env
    A, B16 = kslice_inputs
    m = n = 128
    k = 1 << log_k
    Bk = B16.view(128, KMAX)[:, :k]                     # n x k view: op(B)^T
    ref = A[:m * k].view(k, m).t().double().cpu().numpy() @ Bk.t().double().cpu().numpy()
    B32 = Bk.float().contiguous()                      # k x n with ldb = k for the SGEMM
    Cs = torch.empty(m * n, device="cuda")
    sg.sgemm_fp32(h, 0, 0, m, n, k, 1.0, A, m, B32, k, 0.0, Cs, m)
    err_s = _rel_err(Cs.view(n, m).t().cpu().numpy(), ref)
    runs = 10 if log_k == 22 else 2
    C = torch.empty(m * n, device="cuda")
    outs = []
    for _ in range(runs):
        C.fill_(float("nan"))
        sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B16, KMAX, 0.0, C, m, FP16)
        outs.append(C.cpu().numpy().copy())
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint32), outs[0].view(np.uint32))
    err = _rel_err(outs[0].reshape(n, m).T, ref)
    print(f"k-slicing k=2^{log_k}: plan={sg.query_plan(h, 0, 0, m, n, k, FP16)} err={err:.3e} sgemm={err_s:.3e}",
          flush=True)
    assert err <= 2.0 * err_s, (k, err, err_s)


# ---- wide-exponent sketch at 16384 x 16384 x 256 ------------------------------
SK_M = SK_K = 16384
SK_L = 256


@pytest.fixture(scope="module")
def sketch_inputs(env):
    torch, _, _ = env
    _SWEEP.clear()
    torch.cuda.empty_cache()
    g = _gen(torch, 2)
    u = torch.rand(SK_M * SK_K, device="cuda", generator=g) + 1.0
    e = torch.randint(-15, 16, (SK_M * SK_K,), device="cuda", generator=g, dtype=torch.int32)
    s = torch.randint(0, 2, (SK_M * SK_K,), device="cuda", generator=g, dtype=torch.int32) * 2 - 1
    A = torch.ldexp(u, e) * s
    del u, e, s
    Om16 = (torch.randn(SK_K * SK_L, device="cuda", generator=g) / 16.0).half()
    yield A, Om16
    del A, Om16
    torch.cuda.empty_cache()


@pytest.mark.parametrize("ct", [FP16, TF32])
def test_wide_exponent_sketch_full_size(env, sketch_inputs, ct):
    """Y = A * Omega as shgemm op_n, shgemm op_t (A stored k x m) and
    hsgemm (Y^T = Omega^T A^T); each within 2x of SGEMM's FP64 error."""
    torch, sg, h = env
    A, Om16 = sketch_inputs
    m, k, l = SK_M, SK_K, SK_L
    assert (A.abs() >= 65520).any().item()      # beyond FP16's range without prescaling
    Om32 = Om16.float()
    rows = np.sort(np.random.default_rng(3).choice(m, SAMPLE_ROWS, replace=False))
    idx = torch.from_numpy(rows).cuda()
    Om64 = Om32.view(l, k).t().double().cpu().numpy()
    Y = torch.empty(m * l, device="cuda")
    for op_a in (0, 1):
        A_rows = _op_rows(A, op_a, m, k, idx).double().cpu().numpy()
        ref = A_rows @ Om64
        lda = m if op_a == 0 else k
        sg.sgemm_fp32(h, op_a, 0, m, l, k, 1.0, A, lda, Om32, k, 0.0, Y, m)
        err_s = _rel_err(_c_rows(Y, 0, m, l, idx).cpu().numpy(), ref)
        Y.fill_(float("nan"))
        sg.shgemm(h, op_a, 0, m, l, k, 1.0, A, lda, Om16, k, 0.0, Y, m, ct)
        err = _rel_err(_c_rows(Y, 0, m, l, idx).cpu().numpy(), ref)
        print(f"sketch shgemm op_a={op_a} ct={ct}: err={err:.3e} sgemm={err_s:.3e}", flush=True)
        assert err <= 2.0 * err_s, (op_a, ct, err, err_s)
        if op_a == 0:
            # hsgemm: C (l x m, ldc = l) = op_t(Omega as k x l) * op_t(A as m x k) = Y^T
            Y.fill_(float("nan"))
            sg.hsgemm(h, 1, 1, l, m, k, 1.0, Om16, k, A, m, 0.0, Y, l, ct)
            err_h = _rel_err(Y.view(m, l)[idx].cpu().numpy(), ref)
            print(f"sketch hsgemm ct={ct}: err={err_h:.3e} sgemm={err_s:.3e}", flush=True)
            assert err_h <= 2.0 * err_s, (ct, err_h, err_s)


# ---- 512 uneven problems in one grouped launch --------------------------------
def _grouped_shapes(count, seed):
    """m log-uniform in [256, 8192], n in {64, 128, 256, 512}, k log-uniform
    in [256, 16384] (the same draw as the grouped benchmark)."""
    rng = np.random.default_rng(seed)
    m = np.exp(rng.uniform(np.log(256), np.log(8192), count)).astype(int)
    k = np.exp(rng.uniform(np.log(256), np.log(16384), count)).astype(int)
    n = rng.choice([64, 128, 256, 512], count)
    return [(int(a), int(b), int(c)) for a, b, c in zip(m, n, k)]


CLASS_PLAN = {0: (0, 4), 1: (1, 2), 2: (1, 3), 3: (1, 3)}  # tile class -> (unsplit, split) single-call variant


def test_grouped_512_full_size(env):
    torch, sg, h = env
    _SWEEP.clear()
    torch.cuda.empty_cache()
    shapes = _grouped_shapes(512, 4)
    g = _gen(torch, 4)
    probs = []
    for m, n, k in shapes:
        A = torch.rand(m * k, device="cuda", generator=g) * 2 - 1
        B16 = (torch.randn(k * n, device="cuda", generator=g) / np.sqrt(n)).half()
        C = torch.full((m * n,), float("nan"), device="cuda")
        probs.append((m, n, k, A, B16, C))
    tup = [(m, n, k, A, m, B16, k, C, m, 1.0, 0.0) for m, n, k, A, B16, C in probs]
    with sg.Group(h, 0, 0, tup, FP16, 0) as grp:
        classes, slices, units = grp.plan()
        grp.run(torch.cuda.current_stream())
        torch.cuda.synchronize()
    assert units > 0
    rng = np.random.default_rng(5)
    worst = 0.0
    Cs = torch.empty(8192 * 512, device="cuda")
    for i, (m, n, k, A, B16, C) in enumerate(probs):
        rows = np.sort(rng.choice(m, min(m, 32), replace=False))
        idx = torch.from_numpy(rows).cuda()
        B32 = B16.float()
        ref = _op_rows(A, 0, m, k, idx).double().cpu().numpy() @ B32.view(n, k).t().double().cpu().numpy()
        sg.sgemm_fp32(h, 0, 0, m, n, k, 1.0, A, m, B32, k, 0.0, Cs, m)
        err_s = _rel_err(_c_rows(Cs[:m * n], 0, m, n, idx).cpu().numpy(), ref)
        err = _rel_err(_c_rows(C, 0, m, n, idx).cpu().numpy(), ref)
        worst = max(worst, err / err_s)
        assert err <= 2.0 * err_s, (i, m, n, k, classes[i], slices[i], err, err_s)
    print(f"grouped 512: units={units} worst error ratio to SGEMM {worst:.3f}", flush=True)
    # a sample of problems is bit-identical to the single call with the matching plan
    single = torch.empty(8192 * 512, device="cuda")
    for i in rng.choice(len(probs), 24, replace=False):
        m, n, k, A, B16, C = probs[i]
        c, s = classes[i], slices[i]
        sg.force_plan(h, CLASS_PLAN[c][1 if s > 1 else 0], s)
        try:
            sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B16, k, 0.0, single, m, FP16)
        finally:
            sg.force_plan(h, -1, 1)
        assert torch.equal(single[:m * n].view(torch.int32), C.view(torch.int32)), (i, m, n, k, c, s)
    del probs, Cs, single
    torch.cuda.empty_cache()
