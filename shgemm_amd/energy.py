"""Energy and power of GEMM calls from the card's own counters (ROCm SMI).

The reference measures W and J per GEMM with NVML around a timed loop
(test/power_consumption_test.cu:55-126, sampling at :94-112).  Here the
card's energy accumulator (amdsmi_get_energy_count: accumulator x resolution
in microjoules) is read before and after a window of repeated calls, and the
GFX clock and socket power are sampled from the GPU metrics table every
50 ms on a thread during the window.  The accumulator covers the whole card.
"""
from __future__ import annotations

import statistics
import threading
import time


class EnergyMeter:
    """Readers of one card's energy accumulator and metrics table."""

    def __init__(self, index: int = 0):
        try:
            import amdsmi
        except ImportError as e:
            raise RuntimeError(f"amdsmi is not importable ({e})") from e
        self._smi = amdsmi
        amdsmi.amdsmi_init()
        self.dev = amdsmi.amdsmi_get_processor_handles()[index]

    def joules(self) -> float:
        r = self._smi.amdsmi_get_energy_count(self.dev)
        return r["energy_accumulator"] * r["counter_resolution"] * 1e-6

    def sample(self) -> tuple:
        """(GFX clock MHz, socket power W) from the GPU metrics table."""
        m = self._smi.amdsmi_get_gpu_metrics_info(self.dev)
        return m.get("current_gfxclk"), m.get("current_socket_power")

    def close(self) -> None:
        try:
            self._smi.amdsmi_shut_down()
        except Exception:
            pass


class Sampler:
    """Samples (clock, power) every `period` seconds on a thread while active."""

    def __init__(self, sample, period: float = 0.05):
        self.sample, self.period, self.out = sample, period, []

    def __enter__(self):
        self.stop = threading.Event()

        def run():
            while not self.stop.is_set():
                self.out.append(self.sample())
                time.sleep(self.period)

        self.th = threading.Thread(target=run, daemon=True)
        self.th.start()
        return self

    def __exit__(self, *exc):
        self.stop.set()
        self.th.join()

    def medians(self):
        tail = self.out[len(self.out) // 3:]  # past the ramp
        med = lambda i: statistics.median(v[i] for v in tail if isinstance(v[i], (int, float))) if tail else None
        return med(0), med(1)


def measure(meter: EnergyMeter, fn, seconds: float, sync) -> dict:
    """Times 16 calls of fn, then repeats it for about `seconds` with the
    energy counter read around the window; `sync` waits for the device."""
    fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(16):
        fn()
    sync()
    per_call = (time.perf_counter() - t0) / 16
    count = max(1, int(seconds / per_call))
    with Sampler(meter.sample) as smp:
        e0, t0 = meter.joules(), time.perf_counter()
        for _ in range(count):
            fn()
        sync()
        e1, t1 = meter.joules(), time.perf_counter()
    clk, pw = smp.medians()
    return {"per_call_s": per_call, "count": count, "window_s": t1 - t0, "joules": e1 - e0,
            "energy_per_call_j": (e1 - e0) / count, "avg_power_w": (e1 - e0) / (t1 - t0),
            "median_gfxclk_mhz": clk, "median_socket_power_w": pw, "samples": len(smp.out)}
