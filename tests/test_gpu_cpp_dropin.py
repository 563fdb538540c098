"""GPU: the reference's C++ usage example (README.md:15-56) built against the
drop-in header include/shgemm/shgemm.hpp and linked to libshgemm_amd.so
(tests/cpp/readme_usage, built by tests/build_native.py), run as its own
process.

Bars: shgemm and hsgemm results within 2x of the library FP32 SGEMM's error
against the oracle's FP64 GEMM on the same inputs (2^-22 slack at this size),
for fp16 and tf32; debug mode (set_debug_mode, and MTK_SHGEMM_DEBUG=1 at
create) prints the reference's call line (src/shgemm.cu:583-592) for both
calls; MTK_SHGEMM_KERNEL_LEVEL fixes the returned kernel level
(src/shgemm.cu:407-414, 655).
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from tests.gpu_util import require_gpu, run_sgemm

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent / "cpp" / "readme_usage"
M, N, K = 300, 260, 500


def _run(tmp_path, ct, extra=(), env_add=None):
    if not EXE.exists():
        pytest.fail(f"{EXE} is missing: run __graft_entry__.build() first")
    env = {k: v for k, v in os.environ.items() if not k.startswith("MTK_SHGEMM_")}
    env.update(env_add or {})
    r = subprocess.run([str(EXE), str(M), str(N), str(K), ct, str(tmp_path), *extra], capture_output=True, text=True,
                       env=env, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    return r.stdout


def _load(tmp_path, name, dtype):
    return np.fromfile(tmp_path / name, dtype=dtype)


@pytest.fixture(scope="module")
def sgemm_handle():
    require_gpu()
    import shgemm_amd as sg

    h = sg.create()
    yield h
    sg.destroy(h)


@pytest.mark.parametrize("ct", ["fp16", "tf32"])
def test_readme_usage_results(tmp_path, sgemm_handle, ct):
    out = _run(tmp_path, ct)
    assert "levels" in out
    a32 = _load(tmp_path, "a_fp32.bin", np.float32)
    b16 = _load(tmp_path, "b_fp16.bin", np.float16)
    a16 = _load(tmp_path, "a_fp16.bin", np.float16)
    b32 = _load(tmp_path, "b_fp32.bin", np.float32)
    c_sh = _load(tmp_path, "c_shgemm.bin", np.float32)
    c_hs = _load(tmp_path, "c_hsgemm.bin", np.float32)
    zero = np.zeros(M * N, np.float32)
    for A, B, C in ((a32, b16.astype(np.float32), c_sh), (a16.astype(np.float32), b32, c_hs)):
        ref = oracle.gemm_f64(0, 0, 0, M, N, K, 1.0, A, M, B, K, 0.0, None, M)
        Cs = run_sgemm(sgemm_handle, 0, 0, 0, M, N, K, 1.0, A, M, B, K, 0.0, zero, M)
        err, err_s = oracle.rel_fro_error(C, ref), oracle.rel_fro_error(Cs, ref)
        print(f"readme usage {ct}: err={err:.3e} sgemm={err_s:.3e}")
        assert err <= 2.0 * err_s + 2.0 ** -22, (ct, err, err_s)


def test_readme_usage_debug_line_and_env_flags(tmp_path):
    require_gpu()
    want_sh = f"[shape=({M},{N},{K}),op_a=N,op_b=N,compute_type=FP16]"
    want_hs = f"[shape=({N},{M},{K}),op_a=T,op_b=T,compute_type=FP16]"   # hsgemm runs the swapped product
    quiet = _run(tmp_path, "fp16")
    assert "[shape=" not in quiet
    for out in (_run(tmp_path, "fp16", ["debug"]), _run(tmp_path, "fp16", env_add={"MTK_SHGEMM_DEBUG": "1"})):
        lines = [l for l in out.splitlines() if l.startswith("[shape=")]
        assert len(lines) == 2 and lines[0].startswith(want_sh) and lines[1].startswith(want_hs), out
    for lvl in ("0", "1"):
        out = _run(tmp_path, "fp16", env_add={"MTK_SHGEMM_KERNEL_LEVEL": lvl})
        assert f"levels {lvl} {lvl}" in out, out
