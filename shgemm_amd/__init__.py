"""shgemm_amd — Python binding of the MI355X SHGEMM library (C ABI via ctypes).

Mirrors the reference's mtk::shgemm interface (include/shgemm/shgemm.hpp:94-155):
create / destroy / set_cuda_stream / enable_kernel_level_fixing /
disable_kernel_level_fixing / set_debug_mode / shgemm / hsgemm, with the same
argument order and meaning.  Device operands are HIP device pointers given as
ints or as torch tensors on a ROCm device (their data_ptr() is used); alpha and
beta are Python floats.  Failures raise ShgemmError.

The library is libshgemm_amd.so next to this file (built by
shgemm_amd/build.py).  There is no fallback: importing works without it, but
every call raises if the native library is missing.
"""
from __future__ import annotations

import ctypes
import enum
import re
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libshgemm_amd.so"
HEADER_PATH = _HERE.parent / "include" / "shgemm_amd.h"

_lib = None


class ShgemmError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"shgemm_amd error {status}: {message}")
        self.status = status


class operation_t(enum.IntEnum):  # reference include/shgemm/shgemm.hpp:36-39
    op_n = 0
    op_t = 1


class tc_t(enum.IntEnum):  # reference include/shgemm/shgemm.hpp:41-44
    fp16 = 0
    tf32 = 1


class kernel_level(enum.IntEnum):  # reference include/shgemm/shgemm.hpp:30-34
    P0 = 0
    P1 = 1


op_n, op_t = operation_t.op_n, operation_t.op_t
fp16, tf32 = tc_t.fp16, tc_t.tf32
STATUS_OK, STATUS_EINVAL, STATUS_EHIP, STATUS_EINTERNAL = 0, -1, -2, -3


def exported_symbols_in_header() -> list[str]:
    """Names of every function declared in include/shgemm_amd.h."""
    text = HEADER_PATH.read_text()
    return sorted(set(re.findall(r"\b(shgemm_amd_[a-z0-9_]+)\s*\(", text)))


def load_library(path: Path | str | None = None) -> ctypes.CDLL:
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise ShgemmError(STATUS_EINTERNAL, f"native library {p} is missing; run shgemm_amd/build.py")
    L = ctypes.CDLL(str(p))
    vp, sz, i32, u32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint
    fp = ctypes.POINTER(ctypes.c_float)
    ip = ctypes.POINTER(ctypes.c_int)
    L.shgemm_amd_create.argtypes = [ctypes.POINTER(vp)]
    L.shgemm_amd_destroy.argtypes = [vp]
    L.shgemm_amd_set_stream.argtypes = [vp, vp]
    L.shgemm_amd_enable_kernel_level_fixing.argtypes = [vp, i32]
    L.shgemm_amd_disable_kernel_level_fixing.argtypes = [vp]
    L.shgemm_amd_set_debug_mode.argtypes = [vp, u32]
    L.shgemm_amd_shgemm.argtypes = [vp, i32, i32, sz, sz, sz, fp, vp, sz, vp, sz, fp, vp, sz, i32, i32, ip]
    L.shgemm_amd_hsgemm.argtypes = [vp, i32, i32, sz, sz, sz, fp, vp, sz, vp, sz, fp, vp, sz, i32, i32, ip]
    L.shgemm_amd_shgemm_no_alloc.argtypes = [vp, i32, i32, sz, sz, sz, fp, vp, sz, vp, sz, fp, vp, sz, i32, i32, ip]
    L.shgemm_amd_sgemm_fp32.argtypes = [vp, i32, i32, sz, sz, sz, fp, vp, sz, vp, sz, fp, vp, sz, i32]
    L.shgemm_amd_set_exponent_scaling.argtypes = [vp, i32]
    L.shgemm_amd_set_shared_device.argtypes = [vp, i32]
    L.shgemm_amd_enable_kernel_timing.argtypes = [vp, i32]
    L.shgemm_amd_kernel_timing.argtypes = [vp, ip, ctypes.POINTER(ctypes.c_float)]
    L.shgemm_amd_force_plan.argtypes = [vp, i32, i32]
    L.shgemm_amd_reserve.argtypes = [vp, sz, sz, sz]
    L.shgemm_amd_load_plans.argtypes = [vp, ctypes.c_char_p]
    L.shgemm_amd_query_plan.argtypes = [vp, i32, i32, sz, sz, sz, i32, i32, ip, ip, ip]
    L.shgemm_amd_check_plan_file.argtypes = [ctypes.c_char_p, ip]
    L.shgemm_amd_group_create.argtypes = [vp, i32, i32, i32, i32, ctypes.POINTER(Problem), sz, ip, ip,
                                          ctypes.POINTER(vp)]
    L.shgemm_amd_group_run.argtypes = [vp, vp, i32]
    L.shgemm_amd_group_plan.argtypes = [vp, ip, ip, ctypes.POINTER(sz)]
    L.shgemm_amd_group_destroy.argtypes = [vp]
    L.shgemm_amd_batch_create.argtypes = [vp, i32, i32, i32, i32, sz, sz, i32, ctypes.POINTER(vp)]
    L.shgemm_amd_batch_run.argtypes = [vp, ctypes.POINTER(Problem), sz, vp]
    L.shgemm_amd_batch_plan.argtypes = [vp, ip, ip, ctypes.POINTER(sz)]
    L.shgemm_amd_batch_destroy.argtypes = [vp]
    L.shgemm_amd_gaussian_f16.argtypes = [vp, sz, ctypes.c_uint64, ctypes.c_float, vp]
    L.shgemm_amd_philox_words.argtypes = [vp, sz, ctypes.c_uint64, vp]
    L.shgemm_amd_last_error.restype = ctypes.c_char_p
    L.shgemm_amd_version.restype = ctypes.c_char_p
    for name in exported_symbols_in_header():
        if name not in ("shgemm_amd_last_error", "shgemm_amd_version"):
            getattr(L, name).restype = i32
    if path is None:
        _lib = L
    return L


def _check(status: int) -> None:
    if status != STATUS_OK:
        msg = load_library().shgemm_amd_last_error().decode(errors="replace")
        raise ShgemmError(status, msg)


def _dptr(x) -> int | None:
    """Device address of a torch tensor or an int."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        if hasattr(x, "is_cuda") and not x.is_cuda:
            raise ShgemmError(STATUS_EINVAL, "operand tensor is not on the GPU")
        return int(x.data_ptr())
    raise TypeError(f"expected a device tensor or an int address, got {type(x)}")


def _fptr(v: float):
    return ctypes.pointer(ctypes.c_float(float(v)))


class Problem(ctypes.Structure):
    """One problem of a group (shgemm_amd_problem in include/shgemm_amd.h)."""

    _fields_ = [("m", ctypes.c_size_t), ("n", ctypes.c_size_t), ("k", ctypes.c_size_t),
                ("a", ctypes.c_void_p), ("lda", ctypes.c_size_t),
                ("b", ctypes.c_void_p), ("ldb", ctypes.c_size_t),
                ("c", ctypes.c_void_p), ("ldc", ctypes.c_size_t),
                ("alpha", ctypes.c_float), ("beta", ctypes.c_float)]


class shgemmHandle_t:
    """Owns one native handle (reference shgemmHandle_t, shgemm.hpp:46-92)."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        self.cuda_stream = None

    @property
    def ptr(self):
        if not self._h:
            raise ShgemmError(STATUS_EINVAL, "handle not created")
        return self._h

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        destroy(self)


def create(handle: shgemmHandle_t | None = None) -> shgemmHandle_t:
    h = handle or shgemmHandle_t()
    _check(load_library().shgemm_amd_create(ctypes.byref(h._h)))
    return h


def destroy(handle: shgemmHandle_t) -> None:
    if handle._h:
        _check(load_library().shgemm_amd_destroy(handle._h))
        handle._h = ctypes.c_void_p()


def set_cuda_stream(handle: shgemmHandle_t, stream) -> None:
    """stream: a hipStream_t as int, a torch.cuda.Stream, or None (default)."""
    raw = stream
    if stream is not None and hasattr(stream, "cuda_stream"):
        raw = stream.cuda_stream
    _check(load_library().shgemm_amd_set_stream(handle.ptr, raw))
    handle.cuda_stream = stream


def enable_kernel_level_fixing(handle: shgemmHandle_t, level: int) -> None:
    _check(load_library().shgemm_amd_enable_kernel_level_fixing(handle.ptr, int(level)))


def disable_kernel_level_fixing(handle: shgemmHandle_t) -> None:
    _check(load_library().shgemm_amd_disable_kernel_level_fixing(handle.ptr))


def set_debug_mode(handle: shgemmHandle_t, on: int) -> None:
    _check(load_library().shgemm_amd_set_debug_mode(handle.ptr, int(on)))


def set_exponent_scaling(handle: shgemmHandle_t, on: bool) -> None:
    _check(load_library().shgemm_amd_set_exponent_scaling(handle.ptr, 1 if on else 0))


def set_shared_device(handle: shgemmHandle_t, on: bool) -> None:
    """Calls on `handle` run beside concurrent kernels: plans with whole-CU
    256x256 workgroups run as the 128x256 tile instead."""
    _check(load_library().shgemm_amd_set_shared_device(handle.ptr, 1 if on else 0))


def shgemm(handle, op_a, op_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, compute_type, op_c=op_n) -> kernel_level:
    """C = alpha op(A32) op(B16) + beta C (reference shgemm.hpp:105-119)."""
    lvl = ctypes.c_int(-1)
    _check(load_library().shgemm_amd_shgemm(handle.ptr, int(op_a), int(op_b), m, n, k, _fptr(alpha), _dptr(a), lda,
                                            _dptr(b), ldb, _fptr(beta), _dptr(c), ldc, int(compute_type), int(op_c),
                                            ctypes.byref(lvl)))
    return kernel_level(lvl.value)


def shgemm_no_alloc(handle, op_a, op_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, compute_type,
                    op_c=op_n) -> kernel_level:
    """shgemm on the library's static device scratch: allocates nothing
    (for graph capture and callers that cannot allocate); same results."""
    lvl = ctypes.c_int(-1)
    _check(load_library().shgemm_amd_shgemm_no_alloc(handle.ptr, int(op_a), int(op_b), m, n, k, _fptr(alpha),
                                                     _dptr(a), lda, _dptr(b), ldb, _fptr(beta), _dptr(c), ldc,
                                                     int(compute_type), int(op_c), ctypes.byref(lvl)))
    return kernel_level(lvl.value)


def hsgemm(handle, op_a, op_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, compute_type, op_c=op_n) -> kernel_level:
    """C = alpha op(A16) op(B32) + beta C (reference shgemm.hpp:121-155)."""
    lvl = ctypes.c_int(-1)
    _check(load_library().shgemm_amd_hsgemm(handle.ptr, int(op_a), int(op_b), m, n, k, _fptr(alpha), _dptr(a), lda,
                                            _dptr(b), ldb, _fptr(beta), _dptr(c), ldc, int(compute_type), int(op_c),
                                            ctypes.byref(lvl)))
    return kernel_level(lvl.value)


def sgemm_fp32(handle, op_a, op_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, op_c=op_n) -> None:
    """Plain FP32 SGEMM (the comparator), same conventions."""
    _check(load_library().shgemm_amd_sgemm_fp32(handle.ptr, int(op_a), int(op_b), m, n, k, _fptr(alpha), _dptr(a),
                                                lda, _dptr(b), ldb, _fptr(beta), _dptr(c), ldc, int(op_c)))


# ---- kernel plans -------------------------------------------------------------
PLAN_VARIANTS = {0: "P0", 1: "P1", 2: "SK128", 3: "P1-SK", 4: "P0-SK", 5: "P2", 6: "P2-ring", 7: "P1-3stage",
                 8: "P2-presplit", 9: "P0-rowexp", 10: "SK-deep", 11: "P1W-SK"}
PLAN_SOURCES = {0: "cost model", 1: "table", 2: "forced", 3: "fixed level"}


def force_plan(handle, variant: int, slices: int = 1) -> None:
    """Run every later call on `handle` with this plan (variant < 0 clears)."""
    _check(load_library().shgemm_amd_force_plan(handle.ptr, int(variant), int(slices)))


def enable_kernel_timing(handle, max_launches: int) -> None:
    """Bracket the main GEMM kernel of the next `max_launches` calls on
    `handle` with HIP events on its stream (0 stops)."""
    _check(load_library().shgemm_amd_enable_kernel_timing(handle.ptr, int(max_launches)))


def kernel_timing(handle) -> tuple[int, float]:
    """(launches, total kernel ms) of the bracketed launches since the last query."""
    n = ctypes.c_int(0)
    ms = ctypes.c_float(0.0)
    _check(load_library().shgemm_amd_kernel_timing(handle.ptr, ctypes.byref(n), ctypes.byref(ms)))
    return n.value, float(ms.value)


def reserve(handle, m: int, n: int, k: int) -> None:
    """Grow the handle's scratch for calls up to m x n x k (so such calls
    allocate nothing, e.g. under graph capture)."""
    _check(load_library().shgemm_amd_reserve(handle.ptr, m, n, k))


def load_plans(handle, path) -> None:
    """Replace the handle's plan table (None or "": no table, cost model only)."""
    raw = None if path is None else str(path).encode()
    _check(load_library().shgemm_amd_load_plans(handle.ptr, raw))


def query_plan(handle, op_a, op_b, m, n, k, compute_type, op_c=op_n) -> tuple[int, int, int]:
    """(variant, slices, source) that shgemm would use for these arguments."""
    v, s, src = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(load_library().shgemm_amd_query_plan(handle.ptr, int(op_a), int(op_b), m, n, k, int(compute_type),
                                                int(op_c), ctypes.byref(v), ctypes.byref(s), ctypes.byref(src)))
    return v.value, s.value, src.value


def check_plan_file(path) -> int:
    """Parse a plan table on the host; returns its row count or raises."""
    rows = ctypes.c_int()
    _check(load_library().shgemm_amd_check_plan_file(str(path).encode(), ctypes.byref(rows)))
    return rows.value


# ---- grouped SHGEMM -------------------------------------------------------------
class Group:
    """A captured grouped SHGEMM (shgemm_amd_group_*).  `problems` is a list of
    dicts or tuples (m, n, k, A, lda, B, ldb, C, ldc, alpha, beta) with device
    tensors or int addresses; operands must stay alive while the group is used."""

    def __init__(self, handle, op_a, op_b, problems, compute_type=fp16, op_c=op_n, tile_class=None, slices=None):
        n = len(problems)
        arr = (Problem * max(n, 1))()
        keys = ("m", "n", "k", "a", "lda", "b", "ldb", "c", "ldc", "alpha", "beta")
        for i, q in enumerate(problems):
            d = dict(zip(keys, q)) if not isinstance(q, dict) else q
            arr[i] = Problem(d["m"], d["n"], d["k"], _dptr(d["a"]), d["lda"], _dptr(d["b"]), d["ldb"],
                             _dptr(d["c"]), d["ldc"], float(d.get("alpha", 1.0)), float(d.get("beta", 0.0)))
        tc = (ctypes.c_int * n)(*tile_class) if tile_class is not None else None
        sl = (ctypes.c_int * n)(*slices) if slices is not None else None
        self._g = ctypes.c_void_p()
        self.count = n
        _check(load_library().shgemm_amd_group_create(handle.ptr, int(op_a), int(op_b), int(op_c), int(compute_type),
                                                      arr, n, tc, sl, ctypes.byref(self._g)))

    def run(self, stream=None, use_graph: bool = True) -> None:
        raw = stream.cuda_stream if stream is not None and hasattr(stream, "cuda_stream") else stream
        _check(load_library().shgemm_amd_group_run(self._g, raw, 1 if use_graph else 0))

    def plan(self) -> tuple[list[int], list[int], int]:
        """Per-problem tile class and slice count, and the total work units."""
        tc = (ctypes.c_int * self.count)()
        sl = (ctypes.c_int * self.count)()
        units = ctypes.c_size_t()
        _check(load_library().shgemm_amd_group_plan(self._g, tc, sl, ctypes.byref(units)))
        return list(tc), list(sl), units.value

    def destroy(self) -> None:
        if self._g:
            _check(load_library().shgemm_amd_group_destroy(self._g))
            self._g = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()


class Batch:
    """Grouped launches of problems that change from call to call, with no
    allocation per call (shgemm_amd_batch_*): `slots` device regions of
    `bytes` each hold one batch in flight; run() plans and enqueues one batch
    of problems (same tuples as Group) on `stream`."""

    def __init__(self, handle, op_a, op_b, compute_type=fp16, op_c=op_n, max_problems=64, bytes=256 << 20,
                 slots=2):
        self._b = ctypes.c_void_p()
        self.max_problems = int(max_problems)
        self._arr = (Problem * self.max_problems)()
        self._last = 0
        _check(load_library().shgemm_amd_batch_create(handle.ptr, int(op_a), int(op_b), int(op_c), int(compute_type),
                                                      self.max_problems, int(bytes), int(slots),
                                                      ctypes.byref(self._b)))

    def run(self, problems, stream=None) -> None:
        n = len(problems)
        if n > self.max_problems:
            raise ShgemmError(STATUS_EINVAL, f"{n} problems exceed max_problems {self.max_problems}")
        keys = ("m", "n", "k", "a", "lda", "b", "ldb", "c", "ldc", "alpha", "beta")
        for i, q in enumerate(problems):
            d = dict(zip(keys, q)) if not isinstance(q, dict) else q
            self._arr[i] = Problem(d["m"], d["n"], d["k"], _dptr(d["a"]), d["lda"], _dptr(d["b"]), d["ldb"],
                                   _dptr(d["c"]), d["ldc"], float(d.get("alpha", 1.0)), float(d.get("beta", 0.0)))
        self._last = n
        _check(load_library().shgemm_amd_batch_run(self._b, self._arr, n, _raw_stream(stream)))

    def plan(self) -> tuple[list[int], list[int], int]:
        """Per-problem tile class and slice count of the last run, and its work units."""
        tc = (ctypes.c_int * max(1, self._last))()
        sl = (ctypes.c_int * max(1, self._last))()
        units = ctypes.c_size_t()
        _check(load_library().shgemm_amd_batch_plan(self._b, tc, sl, ctypes.byref(units)))
        return list(tc)[:self._last], list(sl)[:self._last], units.value

    def destroy(self) -> None:
        if self._b:
            _check(load_library().shgemm_amd_batch_destroy(self._b))
            self._b = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()


# ---- sketch matrices -----------------------------------------------------------
def _raw_stream(stream):
    return stream.cuda_stream if stream is not None and hasattr(stream, "cuda_stream") else stream


def gaussian_f16(out, count: int, seed: int, scale: float = 1.0, stream=None) -> None:
    """out[i] = FP16(scale * N(0,1)) from Philox4x32-10(seed) + Box-Muller;
    element i depends only on (seed, i)."""
    _check(load_library().shgemm_amd_gaussian_f16(_dptr(out), count, seed & (2**64 - 1), float(scale),
                                                  _raw_stream(stream)))


def philox_words(out, blocks: int, seed: int, stream=None) -> None:
    """out[4b + j] = word j of Philox4x32-10 block b (key = seed)."""
    _check(load_library().shgemm_amd_philox_words(_dptr(out), blocks, seed & (2**64 - 1), _raw_stream(stream)))


def version() -> str:
    return load_library().shgemm_amd_version().decode()
