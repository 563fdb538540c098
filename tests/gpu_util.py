"""Helpers for the GPU tests: device buffers from numpy, calls through the C ABI."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import shgemm_amd as sg

DEV = "cuda:0"


def dev_buffer(host: np.ndarray, offset: int = 0) -> tuple[torch.Tensor, int]:
    """Copy a flat numpy array to the GPU starting `offset` elements into a
    fresh allocation (offset > 0 gives a misaligned base pointer).  Returns the
    owning tensor and the device address of element 0."""
    t = torch.from_numpy(np.ascontiguousarray(host.ravel()))
    buf = torch.zeros(t.numel() + offset + 8, dtype=t.dtype, device=DEV)
    buf[offset:offset + t.numel()].copy_(t.to(DEV))
    return buf, buf.data_ptr() + offset * buf.element_size()


def read_buffer(buf: torch.Tensor, offset: int, size: int) -> np.ndarray:
    torch.cuda.synchronize()
    return buf[offset:offset + size].cpu().numpy()


# Test-side compute variants: 0 = fp16 (default: per-row prescaled, one
# accumulator), 1 = tf32 (BF16 split), 2 = fp16 with prescaling off (the
# reference's hi + 2^-11 lo with two accumulators).
FP16, TF32, FP16_CLASSIC = 0, 1, 2
VARIANTS = [FP16, TF32, FP16_CLASSIC]


def run_shgemm(h, op_a, op_b, op_c, m, n, k, alpha, A, lda, B16, ldb, beta, C0, ldc, compute_type, off=(0, 0, 0)):
    """A float32 flat, B16 float16 flat, C0 float32 flat (C's storage).  Returns (C, level)."""
    a_t, a_p = dev_buffer(A, off[0])
    b_t, b_p = dev_buffer(B16.view(np.float16) if B16.dtype == np.uint16 else B16, off[1])
    c_t, c_p = dev_buffer(C0, off[2])
    ct = 0 if compute_type == FP16_CLASSIC else compute_type
    if compute_type == FP16_CLASSIC:
        sg.set_exponent_scaling(h, False)
    try:
        lvl = sg.shgemm(h, op_a, op_b, m, n, k, alpha, a_p, lda, b_p, ldb, beta, c_p, ldc, ct, op_c)
    finally:
        if compute_type == FP16_CLASSIC:
            sg.set_exponent_scaling(h, True)
    return read_buffer(c_t, off[2], C0.size), lvl


def run_hsgemm(h, op_a, op_b, op_c, m, n, k, alpha, A16, lda, B, ldb, beta, C0, ldc, compute_type):
    a_t, a_p = dev_buffer(A16)
    b_t, b_p = dev_buffer(B)
    c_t, c_p = dev_buffer(C0)
    sg.hsgemm(h, op_a, op_b, m, n, k, alpha, a_p, lda, b_p, ldb, beta, c_p, ldc, compute_type, op_c)
    return read_buffer(c_t, 0, C0.size)


def run_sgemm(h, op_a, op_b, op_c, m, n, k, alpha, A, lda, B32, ldb, beta, C0, ldc):
    a_t, a_p = dev_buffer(A)
    b_t, b_p = dev_buffer(B32)
    c_t, c_p = dev_buffer(C0)
    sg.sgemm_fp32(h, op_a, op_b, m, n, k, alpha, a_p, lda, b_p, ldb, beta, c_p, ldc, op_c)
    return read_buffer(c_t, 0, C0.size)


def lds(op_a, op_b, op_c, m, n, k, pad=0):
    return ((m if op_a == 0 else k) + pad, (k if op_b == 0 else n) + pad, (m if op_c == 0 else n) + pad)


def sizes(op_a, op_b, op_c, m, n, k, lda, ldb, ldc):
    return (lda * (k if op_a == 0 else m), ldb * (n if op_b == 0 else k), ldc * (n if op_c == 0 else m))


def require_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU in this environment")
    # The native library must be present on a GPU box: no fallback.
    sg.load_library()
