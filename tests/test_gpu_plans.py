"""GPU: kernel-plan selection (forced plan, fixed level, measured plan table,
cost model) and that a plan from the table runs the kernel it names."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests.gpu_util import dev_buffer, read_buffer, require_gpu, run_sgemm

pytestmark = pytest.mark.gpu
COST, TABLE, FORCED, FIXED = 0, 1, 2, 3


@pytest.fixture()
def h():
    require_gpu()
    import shgemm_amd as sg

    handle = sg.create()
    yield handle
    sg.destroy(handle)


def test_plan_sources_and_precedence(h, tmp_path):
    import shgemm_amd as sg

    table = tmp_path / "plans.tsv"
    table.write_text("# test table\n"
                     "single fp16s N N N 10 10 10 0 1\n"
                     "single fp16s N N N 7 7 16 2 4 12.5\n"
                     "single bf16 N N N 10 10 10 1 1\n")
    sg.load_plans(h, table)
    assert sg.query_plan(h, 0, 0, 1024, 1024, 1024, sg.fp16) == (0, 1, TABLE)
    assert sg.query_plan(h, 0, 0, 1100, 900, 1000, sg.fp16) == (0, 1, TABLE)   # nearest point
    assert sg.query_plan(h, 0, 0, 128, 128, 1 << 16, sg.fp16) == (2, 4, TABLE)
    assert sg.query_plan(h, 0, 0, 1024, 1024, 1024, sg.tf32) == (1, 1, TABLE)
    # layout not in the table: falls back to the N/N/N rows of the mode
    assert sg.query_plan(h, 1, 1, 1024, 1024, 1024, sg.fp16, 1) == (0, 1, TABLE)
    # far from every point: cost model
    assert sg.query_plan(h, 0, 0, 16384, 16384, 16384, sg.fp16)[2] == COST
    # forced beats the table, fixed level beats the table, forced beats fixed
    sg.enable_kernel_level_fixing(h, 1)
    assert sg.query_plan(h, 0, 0, 1024, 1024, 1024, sg.fp16) == (1, 1, FIXED)
    sg.force_plan(h, 4, 8)
    v, s, src = sg.query_plan(h, 0, 0, 1024, 1024, 1 << 16, sg.fp16)
    assert (v, s, src) == (4, 8, FORCED)
    sg.force_plan(h, -1, 1)
    sg.disable_kernel_level_fixing(h)
    sg.load_plans(h, None)
    assert sg.query_plan(h, 0, 0, 1024, 1024, 1024, sg.fp16)[2] == COST


def test_forced_split_is_clamped(h):
    import shgemm_amd as sg

    sg.force_plan(h, 2, 256)
    v, s, src = sg.query_plan(h, 0, 0, 128, 128, 4096, sg.fp16)
    assert src == FORCED and v == 2 and 2 <= s <= 32  # every slice keeps >= 4 k-tiles of 32
    # a split variant with k too short for two slices runs unsplit
    v, s, _ = sg.query_plan(h, 0, 0, 128, 128, 200, sg.fp16)
    assert s == 1 and v == 1
    sg.force_plan(h, -1, 1)


def test_prescaled_split_plans_resolve(h):
    """Plan 10 (the deep 128x128x64 k-sliced tile) runs in the prescaled fp16
    mode, as plan 2 in the others, and unsplit (plan 1) where k is too short;
    its slices keep at least 4 k-tiles of 64.  Plan 11 (the large tile of
    64x128 wavefronts, k-sliced) likewise, as plan 3 in the other modes."""
    import shgemm_amd as sg

    sg.force_plan(h, 10, 256)
    try:
        assert sg.query_plan(h, 0, 0, 128, 128, 1 << 20, sg.fp16) == (10, 256, FORCED)
        assert sg.query_plan(h, 0, 0, 128, 128, 1 << 14, sg.fp16)[:2] == (10, 64)
        assert sg.query_plan(h, 0, 0, 128, 128, 1 << 20, sg.tf32)[0] == 2
        sg.set_exponent_scaling(h, False)
        assert sg.query_plan(h, 0, 0, 128, 128, 1 << 20, sg.fp16)[0] == 2
        sg.set_exponent_scaling(h, True)
        assert sg.query_plan(h, 0, 0, 128, 128, 300, sg.fp16)[:2] == (1, 1)
        sg.force_plan(h, 11, 4)
        assert sg.query_plan(h, 0, 0, 16384, 256, 16384, sg.fp16) == (11, 4, FORCED)
        assert sg.query_plan(h, 0, 0, 16384, 256, 16384, sg.tf32)[0] == 3
        sg.set_exponent_scaling(h, False)
        assert sg.query_plan(h, 0, 0, 16384, 256, 16384, sg.fp16)[0] == 3
        sg.set_exponent_scaling(h, True)
        assert sg.query_plan(h, 0, 0, 128, 128, 200, sg.fp16)[:2] == (1, 1)
    finally:
        sg.set_exponent_scaling(h, True)
        sg.force_plan(h, -1, 1)


def test_bad_plan_table_is_rejected(h, tmp_path):
    import shgemm_amd as sg

    bad = tmp_path / "bad.tsv"
    bad.write_text("single fp16s N N N 10 10 10 12 1\n")  # variant 12 does not exist
    with pytest.raises(sg.ShgemmError) as ei:
        sg.load_plans(h, bad)
    assert ei.value.status == sg.STATUS_EINVAL and "bad.tsv:1" in str(ei.value)
    with pytest.raises(sg.ShgemmError):
        sg.force_plan(h, 12, 1)


def test_table_plan_runs_named_kernel(h, tmp_path):
    """A table row selecting split-K gives the same bits as forcing that plan,
    and a different row (unsplit) stays within the error bar."""
    import shgemm_amd as sg

    m, n, k = 128, 128, 1 << 14
    rng = np.random.default_rng(11)
    A = oracle.gen_uniform(rng, m * k)
    B16 = oracle.gen_f16_normal(rng, k * n)
    a_t, a_p = dev_buffer(A)
    b_t, b_p = dev_buffer(B16)
    c_t, c_p = dev_buffer(np.zeros(m * n, np.float32))
    table = tmp_path / "plans.tsv"
    table.write_text("single fp16s N N N 7 7 14 2 8\n")
    sg.load_plans(h, table)
    assert sg.query_plan(h, 0, 0, m, n, k, sg.fp16) == (2, 8, TABLE)
    sg.shgemm(h, 0, 0, m, n, k, 1.0, a_p, m, b_p, k, 0.0, c_p, m, sg.fp16)
    from_table = read_buffer(c_t, 0, m * n).copy()
    sg.load_plans(h, None)
    sg.force_plan(h, 2, 8)
    sg.shgemm(h, 0, 0, m, n, k, 1.0, a_p, m, b_p, k, 0.0, c_p, m, sg.fp16)
    forced = read_buffer(c_t, 0, m * n).copy()
    sg.force_plan(h, -1, 1)
    assert np.array_equal(from_table.view(np.uint32), forced.view(np.uint32))
    ref = A.reshape(k, m).T.astype(np.float64) @ B16.astype(np.float64).reshape(n, k).T
    err = np.linalg.norm(from_table.reshape(n, m).T - ref) / np.linalg.norm(ref)
    assert err < 1e-6


def test_shipped_table_loads_at_create(h):
    import shgemm_amd as sg

    shipped = sg.LIB_PATH.parent / "plans_gfx950.tsv"
    if not shipped.exists():
        pytest.skip("no shipped plan table")
    assert sg.check_plan_file(shipped) > 0
    assert sg.query_plan(h, 0, 0, 1024, 1024, 1024, sg.fp16)[2] == TABLE


@pytest.mark.parametrize("ct,scaling", [(0, True), (1, True), (0, False)])
@pytest.mark.parametrize("ops", [(0, 0, 0), (1, 0, 1), (0, 1, 0), (1, 1, 1)])
def test_large_tiles_agree_bitwise(h, ct, scaling, ops):
    """The 256x256 four-wavefront tile (plan 5), its direct-to-LDS ring form
    (plan 6, on whole tiles: the 512x768x256 case), the large tile (plan 1)
    and the three-stage pipeline (plan 7, where it applies) accumulate every C
    element over k in the same order, so unsplit results are bit-identical;
    odd sizes exercise every edge path.  With per-row prescaling the large
    tile keeps running row exponents (its k-tiles are scaled by the maximum
    so far, not by the row's final maximum): plans 5-7 still agree bitwise
    and plan 1 is held to the FP64 bar (2x the library SGEMM's error).  The
    small tile on whole-row exponents (plan 9) joins plans 5-7 there."""
    import shgemm_amd as sg

    op_a, op_b, op_c = ops
    running = ct == 0 and scaling
    rng = np.random.default_rng(100 + 8 * ct + 4 * op_a + 2 * op_b + op_c)
    for m, n, k in [(300, 517, 333), (1000, 600, 704), (257, 255, 96), (512, 768, 256)]:
        lda, ldb, ldc = (m if op_a == 0 else k) + 1, (k if op_b == 0 else n) + 2, (m if op_c == 0 else n) + 3
        A = oracle.gen_uniform(rng, lda * (k if op_a == 0 else m))
        B16 = oracle.gen_f16_normal(rng, ldb * (n if op_b == 0 else k))
        a_t, a_p = dev_buffer(A)
        b_t, b_p = dev_buffer(B16)
        c_t, c_p = dev_buffer(np.zeros(ldc * (n if op_c == 0 else m), np.float32))
        sg.set_exponent_scaling(h, scaling)
        outs = []
        for plan in (1, 5, 6, 7) + ((9,) if running else ()):
            sg.force_plan(h, plan, 1)
            sg.shgemm(h, op_a, op_b, m, n, k, 1.0, a_p, lda, b_p, ldb, 0.0, c_p, ldc, ct, op_c)
            outs.append(read_buffer(c_t, 0, c_t.numel() - 8).copy())
        sg.force_plan(h, -1, 1)
        sg.set_exponent_scaling(h, True)
        base = outs[1] if running else outs[0]
        for o in outs[1:]:
            assert np.array_equal(o.view(np.uint32), base.view(np.uint32)), (m, n, k)
        if running:
            B32 = B16.astype(np.float32)
            ref = oracle.op_view(A.astype(np.float64), op_a, m, k, lda) @ oracle.op_view(
                B32.astype(np.float64), op_b, k, n, ldb)
            Cs = run_sgemm(h, op_a, op_b, op_c, m, n, k, 1.0, A, lda, B32, ldb, 0.0,
                           np.zeros(c_t.numel(), np.float32), ldc)
            err = oracle.rel_fro_error(oracle.c_matrix(outs[0], op_c, m, n, ldc), ref)
            err_s = oracle.rel_fro_error(oracle.c_matrix(Cs, op_c, m, n, ldc), ref)
            assert err <= 2.0 * err_s, (m, n, k, err, err_s)


def _growing_rows(rng, m, k):
    """A (m x k, column-major) whose rows change scale along k: row r has
    magnitude 2^(s_r * t) in k-tile t (32 columns), s_r in {-3, 0, 3, 7};
    every fifth row starts with two all-zero k-tiles and one row is zero."""
    t = np.arange(k) // 32
    slope = np.array([-3, 0, 3, 7])[np.arange(m) % 4]
    expo = np.clip(slope[:, None] * t[None, :] - 20, -60, 60)
    A = rng.uniform(-1.0, 1.0, (m, k)) * np.exp2(expo)
    A[::5, :64] = 0.0
    A[m // 2, :] = 0.0
    return np.asfortranarray(A.astype(np.float32)).ravel(order="F")


@pytest.mark.parametrize("op_a,op_c", [(0, 0), (1, 1), (0, 1)])
@pytest.mark.parametrize("plan,slices", [(0, 1), (1, 1), (2, 3), (3, 2), (4, 4), (10, 2), (11, 3)])
def test_running_exponents_follow_growing_rows(h, op_a, op_c, plan, slices):
    """The kernels with running row exponents rescale their accumulators when
    a row's maximum grows along k: rows growing, shrinking or starting with
    zero k-tiles stay within 2x of the library SGEMM's FP64 error, unsplit
    and split over k."""
    import shgemm_amd as sg

    m, n, k = 301, 200, 544
    rng = np.random.default_rng(40 + 10 * plan + 2 * op_a + op_c)
    A0 = _growing_rows(rng, m, k)                      # column-major m x k
    A = A0 if op_a == 0 else np.ascontiguousarray(A0.reshape(k, m).T).ravel()  # op_t: stored k x m
    lda = m if op_a == 0 else k
    ldc = m if op_c == 0 else n
    B16 = oracle.gen_f16_normal(rng, k * n)
    B32 = B16.astype(np.float32)
    a_t, a_p = dev_buffer(A)
    b_t, b_p = dev_buffer(B16)
    c_t, c_p = dev_buffer(np.zeros(m * n, np.float32))
    sg.force_plan(h, plan, slices)
    try:
        sg.shgemm(h, op_a, 0, m, n, k, 1.0, a_p, lda, b_p, k, 0.0, c_p, ldc, sg.fp16, op_c)
    finally:
        sg.force_plan(h, -1, 1)
    got = oracle.c_matrix(read_buffer(c_t, 0, m * n), op_c, m, n, ldc)
    ref = oracle.op_view(A.astype(np.float64), op_a, m, k, lda) @ B32.astype(np.float64).reshape(n, k).T
    Cs = run_sgemm(h, op_a, 0, op_c, m, n, k, 1.0, A, lda, B32, k, 0.0, np.zeros(m * n, np.float32), ldc)
    # row-relative errors: the rows differ by up to 2^400 in scale
    diff = np.abs(got - ref).max(axis=1)
    diff_s = np.abs(oracle.c_matrix(Cs, op_c, m, n, ldc) - ref).max(axis=1)
    scale = np.abs(ref).max(axis=1)
    live = scale > 0
    assert np.all(got[~live] == 0.0)
    r, r_s = diff[live] / scale[live], diff_s[live] / scale[live]
    assert np.all(r <= 2.0 * r_s + 2.0 ** -22), (plan, slices, float(r.max()), float(r_s.max()))


KTILE = {0: 64, 1: 32, 2: 32, 3: 32, 4: 64, 10: 64, 11: 32}  # k-tile depth of each running-exponent plan's kernel


@pytest.mark.parametrize("op_a,op_c", [(0, 0), (1, 1)])
@pytest.mark.parametrize("plan,slices,k,step,width", [(0, 1, 320, 3, 32), (1, 1, 320, 3, 32), (2, 2, 1024, 4, 128),
                                                       (3, 2, 1024, 4, 128), (4, 2, 1024, 4, 128),
                                                       (10, 2, 1024, 4, 128), (11, 2, 1024, 4, 128)])
def test_running_exponents_match_the_oracle(h, op_a, op_c, plan, slices, k, step, width):
    """Balanced growing rows (oracle.gen_growing_rows): the running-exponent kernels
    agree with the oracle's running-exponent emulation (k-tiles of their
    depth, restarting at each k-slice) to FP32 accumulation accuracy, and not
    with the whole-row emulation, which misses the FP64 product by far more
    than 2x the library SGEMM's error; the kernel stays within that bar."""
    import shgemm_amd as sg

    m, n = 301, 200
    rng = np.random.default_rng(900 + 10 * plan + 2 * op_a + op_c)
    A0, B16 = oracle.gen_growing_rows(rng, m, n, k, step=step, width=width)   # column-major m x k, k x n
    A = A0 if op_a == 0 else np.ascontiguousarray(A0.reshape(k, m).T).ravel()  # op_t: stored k x m
    lda = m if op_a == 0 else k
    ldc = m if op_c == 0 else n
    B32 = B16.astype(np.float32)
    a_t, a_p = dev_buffer(A)
    b_t, b_p = dev_buffer(B16)
    c_t, c_p = dev_buffer(np.zeros(m * n, np.float32))
    sg.force_plan(h, plan, slices)
    try:
        v, s, _ = sg.query_plan(h, op_a, 0, m, n, k, sg.fp16, op_c)
        assert (v, s) == (plan, slices)
        sg.shgemm(h, op_a, 0, m, n, k, 1.0, a_p, lda, b_p, k, 0.0, c_p, ldc, sg.fp16, op_c)
    finally:
        sg.force_plan(h, -1, 1)
    got = oracle.c_matrix(read_buffer(c_t, 0, m * n), op_c, m, n, ldc)
    kt = KTILE[plan]
    kslice = ((k + slices - 1) // slices + kt - 1) // kt * kt if slices > 1 else None
    emu = lambda ct: oracle.c_matrix(oracle.emulate_shgemm(op_a, 0, op_c, m, n, k, 1.0, A, lda, B32, k, 0.0, None,
                                                           ldc, ct, ktile=kt, kslice=kslice), op_c, m, n, ldc)
    run, whole = emu(3), emu(2)
    ref = oracle.op_view(A.astype(np.float64), op_a, m, k, lda) @ B32.astype(np.float64).reshape(n, k).T
    Cs = run_sgemm(h, op_a, 0, op_c, m, n, k, 1.0, A, lda, B32, k, 0.0, np.zeros(m * n, np.float32), ldc)
    err_s = oracle.rel_fro_error(oracle.c_matrix(Cs, op_c, m, n, ldc), ref)
    d_run, d_whole = oracle.rel_fro_error(got, run), oracle.rel_fro_error(got, whole)
    print(f"plan {plan}x{slices}: vs running {d_run:.2e}, vs whole-row {d_whole:.2e}, sgemm {err_s:.2e}")
    assert d_run <= 2.0 * err_s
    assert oracle.rel_fro_error(whole, ref) > 8.0 * err_s
    assert d_whole > 8.0 * d_run
    assert oracle.rel_fro_error(got, ref) <= 2.0 * err_s


@pytest.mark.parametrize("plan,slices", [(0, 1), (1, 1), (3, 2), (5, 1), (8, 1), (9, 1), (10, 2), (11, 2)])
def test_non_finite_inputs_stay_in_their_rows_and_columns(h, plan, slices):
    """A NaN or an infinity in A makes its row of C non-finite and one in B
    its column (the FP16 residual of an infinity is NaN, as in the
    reference's split); every other element keeps the FP64 bar, in the
    running-exponent kernels (plans 0, 1, 3) and the whole-row ones (5, 8, 9)."""
    import shgemm_amd as sg

    m, n, k = 300, 260, 1024
    rng = np.random.default_rng(70 + plan)
    A = oracle.gen_uniform(rng, m * k).reshape(k, m)   # column-major m x k: A[j, i] is element (i, j)
    A[100, 7] = np.nan
    A[900, 20] = np.inf
    A[0, 33] = -np.inf
    A = A.ravel()
    B16 = oracle.gen_f16_normal(rng, k * n).reshape(n, k)
    B16[3, 5] = np.inf
    B16 = B16.ravel()
    a_t, a_p = dev_buffer(A)
    b_t, b_p = dev_buffer(B16)
    c_t, c_p = dev_buffer(np.zeros(m * n, np.float32))
    sg.force_plan(h, plan, slices)
    try:
        assert sg.query_plan(h, 0, 0, m, n, k, sg.fp16)[:2] == (plan, slices)
        sg.shgemm(h, 0, 0, m, n, k, 1.0, a_p, m, b_p, k, 0.0, c_p, m, sg.fp16)
    finally:
        sg.force_plan(h, -1, 1)
    got = read_buffer(c_t, 0, m * n).reshape(n, m).T
    bad_r, bad_c = [7, 20, 33], [3]
    assert not np.isfinite(got[bad_r, :]).any()
    assert not np.isfinite(got[:, bad_c]).any()
    ok_r = np.setdiff1d(np.arange(m), bad_r)
    ok_c = np.setdiff1d(np.arange(n), bad_c)
    B32 = B16.astype(np.float32)
    Am = A.reshape(k, m).T.astype(np.float64)
    Bm = B32.reshape(n, k).T.astype(np.float64)
    ref = Am[ok_r] @ Bm[:, ok_c]
    Cs = run_sgemm(h, 0, 0, 0, m, n, k, 1.0, A, m, B32, k, 0.0, np.zeros(m * n, np.float32), m).reshape(n, m).T
    err = oracle.rel_fro_error(got[np.ix_(ok_r, ok_c)], ref)
    err_s = oracle.rel_fro_error(Cs[np.ix_(ok_r, ok_c)], ref)
    assert np.isfinite(got[np.ix_(ok_r, ok_c)]).all()
    assert err <= 2.0 * err_s, (plan, err, err_s)


def test_kernel_timing_counts_main_kernel_launches(h):
    """enable_kernel_timing brackets the main GEMM kernel of the next N calls;
    each query returns the launches since the previous one."""
    import shgemm_amd as sg

    m = n = k = 512
    rng = np.random.default_rng(5)
    a_t, a_p = dev_buffer(oracle.gen_uniform(rng, m * k))
    b_t, b_p = dev_buffer(oracle.gen_f16_normal(rng, k * n))
    c_t, c_p = dev_buffer(np.zeros(m * n, np.float32))
    call = lambda: sg.shgemm(h, 0, 0, m, n, k, 1.0, a_p, m, b_p, k, 0.0, c_p, m, sg.fp16)
    sg.enable_kernel_timing(h, 3)
    for _ in range(5):
        call()
    launches, ms = sg.kernel_timing(h)
    assert launches == 3 and 0.0 < ms < 1e4
    call()  # the pairs are free again after a query
    assert sg.kernel_timing(h)[0] == 1
    sg.force_plan(h, 2, 4)  # split-K: the main kernel is timed, the reduction is not
    call()
    assert sg.kernel_timing(h)[0] == 1
    sg.force_plan(h, -1, 1)
    sg.enable_kernel_timing(h, 0)
    call()
    assert sg.kernel_timing(h) == (0, 0.0)
    with pytest.raises(sg.ShgemmError):
        sg.enable_kernel_timing(h, -1)


@pytest.mark.parametrize("ct", ["fp16", "tf32", "classic"])
@pytest.mark.parametrize("op_a,op_b,op_c", [(0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 1), (0, 1, 0)])
def test_presplit_plan_matches_p2_bitwise(h, op_a, op_b, op_c, ct):
    """Plan 8 (A split once per call and packed in tile order -- prescaled
    FP16 hi/lo, or tf32's three BF16 parts -- B read in place or packed, the
    256x256 tile then runs without a split) accumulates each C element in the
    256x256 tile's order: bit-identical to plan 5.  Cases cover both B
    layouts (aligned and unaligned ldb), partial tiles in m and n, the
    shortest k (two k-tiles) and a padded lda / ldc; the smaller shapes run
    the 128x128 form (fewer 256x256 tiles than CUs), 4096 x 4096 the
    256x256 one.  With prescaling off ("classic": FP16 hi and 2^11-scaled lo
    in two accumulators) plan 8 runs the 128x128 form at every size and
    matches plan 1, whose 64x64 wavefronts take the same products in the same
    order."""
    import shgemm_amd as sg

    ctype = sg.fp16 if ct == "classic" else getattr(sg, ct)
    ref_plan = 1 if ct == "classic" else 5
    rng = np.random.default_rng(40 + 4 * op_a + 2 * op_b + op_c + {"fp16": 0, "tf32": 100, "classic": 200}[ct])
    sg.set_exponent_scaling(h, ct != "classic")
    try:
        for m, n, k in [(512, 768, 256), (1000, 600, 704), (257, 255, 128), (300, 517, 64), (64, 1100, 2048),
                        (4096, 4096, 128)]:
            lda, ldc = (m if op_a == 0 else k) + 3, (m if op_c == 0 else n) + 1
            ldb = (k if op_b == 0 else n) + (0 if op_a == 0 else 5)  # aligned and unaligned B
            A = oracle.gen_uniform(rng, lda * (k if op_a == 0 else m))
            B16 = oracle.gen_f16_normal(rng, ldb * (n if op_b == 0 else k))
            a_t, a_p = dev_buffer(A)
            b_t, b_p = dev_buffer(B16)
            c_t, c_p = dev_buffer(np.zeros(ldc * (n if op_c == 0 else m), np.float32))
            outs = []
            for plan in (ref_plan, 8):
                sg.force_plan(h, plan, 1)
                assert sg.query_plan(h, op_a, op_b, m, n, k, ctype, op_c)[0] == plan
                sg.shgemm(h, op_a, op_b, m, n, k, 1.0, a_p, lda, b_p, ldb, 0.0, c_p, ldc, ctype, op_c)
                outs.append(read_buffer(c_t, 0, c_t.numel() - 8).copy())
            sg.force_plan(h, -1, 1)
            assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32)), (m, n, k)
    finally:
        sg.force_plan(h, -1, 1)
        sg.set_exponent_scaling(h, True)


@pytest.mark.parametrize("ct", ["fp16", "tf32", "classic"])
def test_presplit_ring_repeats_p2_bits_on_random_shapes(h, ct):
    """The pre-split kernels move every operand byte through an LDS ring whose
    barriers wait only for the older k-tile: a load that landed late would
    show as different bits on some call.  24 random shapes (k a multiple of
    64, edges in m and n, both op_a), each run three times with plan 8 and
    once with plan 5 (with prescaling off, "classic", plan 5 runs as plan 1):
    every run bit-identical to plan 5."""
    import shgemm_amd as sg

    ctype = sg.fp16 if ct == "classic" else getattr(sg, ct)
    rng = np.random.default_rng(1234 + {"fp16": 0, "tf32": 7, "classic": 11}[ct])
    sg.set_exponent_scaling(h, ct != "classic")
    try:
        _random_shape_runs(sg, h, ctype, rng)
    finally:
        sg.force_plan(h, -1, 1)
        sg.set_exponent_scaling(h, True)


def _random_shape_runs(sg, h, ctype, rng):
    for _ in range(24):
        m, n = (int(x) for x in rng.integers(200, 1400, 2))
        k = 64 * int(rng.integers(2, 40))
        op_a = int(rng.integers(0, 2))
        lda = m if op_a == 0 else k
        A = oracle.gen_uniform(rng, lda * (k if op_a == 0 else m))
        B16 = oracle.gen_f16_normal(rng, k * n)
        a_t, a_p = dev_buffer(A)
        b_t, b_p = dev_buffer(B16)
        c_t, c_p = dev_buffer(np.zeros(m * n, np.float32))
        outs = []
        for plan in (5, 8, 8, 8):
            sg.force_plan(h, plan, 1)
            sg.shgemm(h, op_a, 0, m, n, k, 1.0, a_p, lda, b_p, k, 0.0, c_p, m, ctype)
            outs.append(read_buffer(c_t, 0, m * n).copy())
        sg.force_plan(h, -1, 1)
        for o in outs[1:]:
            assert np.array_equal(outs[0].view(np.uint32), o.view(np.uint32)), (m, n, k, op_a)


def test_presplit_plan_falls_back_where_it_does_not_apply(h):
    import shgemm_amd as sg

    sg.force_plan(h, 8, 1)
    assert sg.query_plan(h, 0, 0, 1024, 1024, 1024, sg.fp16)[0] == 8
    assert sg.query_plan(h, 0, 1, 1024, 1024, 1024, sg.fp16)[0] == 8   # op_b = T: B is packed too
    assert sg.query_plan(h, 0, 0, 1024, 1024, 1000, sg.fp16)[0] == 5   # k not a multiple of 32
    assert sg.query_plan(h, 0, 0, 1024, 1024, 32, sg.fp16)[0] == 5     # one k-tile
    assert sg.query_plan(h, 0, 0, 1024, 1024, 96, sg.fp16)[0] == 5     # odd number of k-tiles
    assert sg.query_plan(h, 0, 0, 1024, 1024, 1024, sg.tf32)[0] == 8   # BF16 parts packed
    assert sg.query_plan(h, 0, 0, 1024, 1024, 96, sg.tf32)[0] == 5     # odd number of k-tiles
    sg.set_exponent_scaling(h, False)
    assert sg.query_plan(h, 0, 0, 1024, 1024, 1024, sg.fp16)[0] == 8   # two-accumulator mode: 128x128 form
    assert sg.query_plan(h, 0, 0, 1024, 1024, 96, sg.fp16)[0] == 1     # ... falls back to its large tile
    sg.set_exponent_scaling(h, True)
    sg.force_plan(h, -1, 1)


def test_whole_row_small_tile_falls_back_outside_prescaled_fp16(h):
    import shgemm_amd as sg

    sg.force_plan(h, 9, 1)
    assert sg.query_plan(h, 1, 1, 300, 200, 1000, sg.fp16, 1)[0] == 9
    assert sg.query_plan(h, 0, 0, 1024, 1024, 1024, sg.tf32)[0] == 0   # BF16 split
    sg.set_exponent_scaling(h, False)
    assert sg.query_plan(h, 0, 0, 1024, 1024, 1024, sg.fp16)[0] == 0   # two-accumulator mode
    sg.set_exponent_scaling(h, True)
    sg.force_plan(h, -1, 1)


def test_shared_device_avoids_whole_cu_plans(h):
    import shgemm_amd as sg

    sg.force_plan(h, 8, 1)
    assert sg.query_plan(h, 0, 0, 16384, 16384, 16384, sg.fp16)[0] == 8
    sg.set_shared_device(h, True)
    for v in (5, 6, 8):
        sg.force_plan(h, v, 1)
        assert sg.query_plan(h, 0, 0, 16384, 16384, 16384, sg.fp16)[0] == 1
    sg.force_plan(h, 9, 1)  # the small tile holds no whole CU
    assert sg.query_plan(h, 0, 0, 16384, 16384, 16384, sg.fp16)[0] == 9
    sg.force_plan(h, -1, 1)
    assert sg.query_plan(h, 0, 0, 16384, 16384, 16384, sg.fp16)[0] not in (5, 6, 8)
    sg.load_plans(h, None)  # cost model
    assert sg.query_plan(h, 0, 0, 16384, 16384, 16384, sg.fp16)[0] != 5
    sg.set_shared_device(h, False)
