"""The C-ABI library loads, exports every declared entry point, and its host
logic fails cleanly without a GPU (no compute calls here).  CPU only."""
from __future__ import annotations

import ctypes
import subprocess

import pytest

import shgemm_amd as sg
from tests.conftest import gpu_available


def test_library_exports_every_declared_symbol():
    lib = sg.load_library()
    names = sg.exported_symbols_in_header()
    assert "shgemm_amd_shgemm" in names and "shgemm_amd_hsgemm" in names and len(names) >= 12
    for name in names:
        assert hasattr(lib, name), name
    # the dynamic symbol table agrees (no C++ mangling on the ABI)
    out = subprocess.run(["nm", "-D", "--defined-only", str(sg.LIB_PATH)], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert set(names) <= exported


def test_version_and_last_error():
    assert "gfx950" in sg.version()
    assert isinstance(sg.load_library().shgemm_amd_last_error(), bytes)


def test_null_handle_is_einval():
    lib = sg.load_library()
    one = ctypes.pointer(ctypes.c_float(1.0))
    st = lib.shgemm_amd_shgemm(None, 0, 0, 4, 4, 4, one, None, 4, None, 4, one, None, 4, 0, 0, None)
    assert st == sg.STATUS_EINVAL
    assert b"handle" in lib.shgemm_amd_last_error()
    assert lib.shgemm_amd_set_stream(None, None) == sg.STATUS_EINVAL
    assert lib.shgemm_amd_destroy(None) == sg.STATUS_OK
    assert lib.shgemm_amd_create(None) == sg.STATUS_EINVAL


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_create_without_gpu_reports_hip_error():
    with pytest.raises(sg.ShgemmError) as ei:
        sg.create()
    assert ei.value.status == sg.STATUS_EHIP


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(sg.ShgemmError):
        sg.load_library(tmp_path / "absent.so")
