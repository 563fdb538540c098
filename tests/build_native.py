"""Builds the tests' native helpers (run by __graft_entry__.build() after the
library): tests/native/libshgemm_test_kernels.so (single MFMA instructions and
the library's split on the device) and tests/cpp/readme_usage (the reference
README's usage example against include/shgemm/shgemm.hpp, linked to
shgemm_amd/libshgemm_amd.so)."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = ["--offload-arch=gfx950"]
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
KERNELS = ROOT / "tests" / "native" / "libshgemm_test_kernels.so"
USAGE = ROOT / "tests" / "cpp" / "readme_usage"


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        log = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"build failed: {' '.join(map(str, cmd))}\n{log}")


def build() -> None:
    _run([HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", *ARCH, *DEVICE_FLAGS,
          str(ROOT / "tests" / "native" / "test_kernels.hip"), "-o", str(KERNELS)])
    _run([HIPCC, "-O2", "-std=c++17", *ARCH, *DEVICE_FLAGS, f"-I{ROOT / 'include'}",
          str(ROOT / "tests" / "cpp" / "readme_usage.cpp"), f"-L{ROOT / 'shgemm_amd'}", "-lshgemm_amd",
          "-Wl,-rpath,$ORIGIN/../../shgemm_amd", "-o", str(USAGE)])


if __name__ == "__main__":
    build()
