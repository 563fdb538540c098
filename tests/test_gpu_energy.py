"""GPU: energy per GEMM from the card's counters (the reference's power
study, test/power_consumption_test.cu:55-126, sampling at :94-112, with ROCm
SMI in place of NVML).

Bars: the energy accumulator is readable and advances over a window of
repeated calls; J per GEMM is finite and positive, and the average power
lies between 20 W and 3 kW (the card's range); the metrics table gives
clock and power samples; results computed while the sampler runs are
bit-identical to those computed without it.
"""
from __future__ import annotations

import math

import pytest

from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu


def test_energy_per_gemm_counters(capsys):
    require_gpu()
    import torch

    import shgemm_amd as sg
    from shgemm_amd.energy import EnergyMeter, measure

    try:
        meter = EnergyMeter()
    except RuntimeError as e:
        pytest.skip(f"ROCm SMI unavailable: {e}")
    h = sg.create()
    try:
        m = n = k = 4096
        g = torch.Generator(device="cuda")
        g.manual_seed(30)
        A = torch.rand(m * k, device="cuda", generator=g)
        B = torch.rand(k * n, device="cuda", generator=g).half()
        C = torch.empty(m * n, device="cuda")
        sg.set_cuda_stream(h, torch.cuda.current_stream())
        fn = lambda: sg.shgemm(h, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, C, m, sg.fp16)
        fn()
        torch.cuda.synchronize()
        quiet = C.clone()
        r = measure(meter, fn, 1.5, torch.cuda.synchronize)
        assert torch.equal(C.view(torch.int32), quiet.view(torch.int32))
        with capsys.disabled():
            print(f"\nenergy 4096^3 fp16: {r['energy_per_call_j']:.4f} J/GEMM, {r['avg_power_w']:.0f} W, "
                  f"{r['count']} calls, clock {r['median_gfxclk_mhz']} MHz, socket {r['median_socket_power_w']} W")
        assert r["joules"] > 0 and math.isfinite(r["energy_per_call_j"]) and r["energy_per_call_j"] > 0
        assert 20.0 < r["avg_power_w"] < 3000.0, r
        assert r["samples"] >= 3
        clk, pw = r["median_gfxclk_mhz"], r["median_socket_power_w"]
        assert clk is not None and clk > 0 and pw is not None and pw > 0, r
    finally:
        sg.destroy(h)
        meter.close()
