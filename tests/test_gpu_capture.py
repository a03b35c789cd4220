"""The one-shot FIR forms that derive a table from the taps (fft_ola's tap spectrum, the MFMA
variants' split taps) as the very first calls of a process, made inside a stream capture: the
constant tables are compiled into the library and the capture stream's workspace block is
allocated outside the capture, so the capture succeeds, and every replay of the graph gives the
eager bits. It runs in a child process, so that nothing earlier in the test session has used
these variants or this stream; the stream is created with the HIP runtime directly (a stream
made by rf_stream_create already has its block).
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

CHILD = r"""
import ctypes, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from _common import load_rfhip
rf = load_rfhip()
hip = ctypes.CDLL("libamdhip64.so")
raw = ctypes.c_void_p()
assert hip.hipStreamCreateWithFlags(ctypes.byref(raw), 1) == 0  # hipStreamNonBlocking
stream = raw.value

def kaiser(num_taps):
    n = np.arange(num_taps) - (num_taps - 1) / 2
    h = 0.2 * np.sinc(0.2 * n) * np.kaiser(num_taps, 8.0)
    return (h / h.sum()).astype(np.float32)

rng = np.random.default_rng(3)
x = (rng.standard_normal(30011) + 1j * rng.standard_normal(30011)).astype(np.complex64)
cases = []
for impl, num_taps in [("fft_ola", 1025), ("mfma_f16x3", 777), ("mfma_bf16x3", 2049),
                       ("fft_ola", 226), ("fft_ola", 63)]:
    h = kaiser(num_taps)
    cases.append((impl, h, rf.DeviceBuffer.from_array(h), rf.DeviceBuffer(x.size)))
d_x = rf.DeviceBuffer.from_array(x)
print("capturing the first fft_ola / mfma calls of the process", flush=True)
rf.check(rf._lib.rf_graph_begin_capture(stream))
status = rf.RF_OK
try:
    for impl, h, d_h, d_y in cases:
        status = rf._lib.rf_fir(rf.fir_impl_index(impl), d_x.ptr, d_h.ptr, d_y.ptr, x.size,
                                h.size, 0, stream)
        if status != rf.RF_OK:
            break
finally:
    graph = ctypes.c_void_p()
    end = rf._lib.rf_graph_end_capture(stream, ctypes.byref(graph))
rf.check(status)
rf.check(end)
want = [rf.fir(x, h, impl) for impl, h, _, _ in cases]
for rep in range(3):
    for _, _, _, d_y in cases:
        d_y.zero()
    rf.check(rf._lib.rf_stream_synchronize(None))
    rf.check(rf._lib.rf_graph_launch(graph.value, stream))
    rf.check(rf._lib.rf_stream_synchronize(stream))
    for (impl, h, _, d_y), w in zip(cases, want):
        np.testing.assert_array_equal(d_y.download(), w, err_msg=f"{impl} replay {rep}")
    rf.fir(x[:5000], cases[rep % len(cases)][1], cases[rep % len(cases)][0])  # eager work
rf.check(rf._lib.rf_graph_destroy(graph.value))
rf.check(rf._lib.rf_stream_destroy(stream))
print("capture ok", flush=True)
"""


def test_first_transform_fir_calls_inside_a_capture() -> None:
    tests_dir = Path(__file__).resolve().parent
    proc = subprocess.run([sys.executable, "-c", CHILD, str(tests_dir)], capture_output=True,
                          text=True, timeout=300)
    assert proc.returncode == 0 and "capture ok" in proc.stdout, proc.stdout + proc.stderr


SPECTRUM_CHILD = r"""
import ctypes, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from _common import load_rfhip
rf = load_rfhip()
hip = ctypes.CDLL("libamdhip64.so")
raw = ctypes.c_void_p()
assert hip.hipStreamCreateWithFlags(ctypes.byref(raw), 1) == 0  # hipStreamNonBlocking
stream = raw.value
rng = np.random.default_rng(5)
cases = []
for nfft, frames in [(4, 1000), (100, 37), (1024, 9), (16384, 3), (1 << 18, 1)]:
    x = (rng.standard_normal(nfft * frames)
         + 1j * rng.standard_normal(nfft * frames)).astype(np.complex64)
    plan = rf.FftPlan(nfft, frames)
    d_mag = rf.DeviceBuffer(x.size, np.float32)
    cases.append((nfft, x, plan, rf.DeviceBuffer.from_array(x), d_mag))
print("capturing rf_spectrum_magnitude at 4, 100, 1024, 16384 and 2^18 points", flush=True)
rf.check(rf._lib.rf_graph_begin_capture(stream))
status = rf.RF_OK
try:
    for _, _, plan, d_x, d_mag in cases:
        status = rf._lib.rf_spectrum_magnitude(plan.handle, d_x.ptr, d_mag.ptr, stream)
        if status != rf.RF_OK:
            break
finally:
    graph = ctypes.c_void_p()
    end = rf._lib.rf_graph_end_capture(stream, ctypes.byref(graph))
rf.check(status)
rf.check(end)
want = [rf.spectrum_magnitude(x, nfft) for nfft, x, _, _, _ in cases]
for rep in range(3):
    for *_, d_mag in cases:
        d_mag.zero()
    rf.check(rf._lib.rf_stream_synchronize(None))
    rf.check(rf._lib.rf_graph_launch(graph.value, stream))
    rf.check(rf._lib.rf_stream_synchronize(stream))
    for (nfft, *_, d_mag), w in zip(cases, want):
        assert d_mag.download().tobytes() == w.tobytes(), f"nfft {nfft} replay {rep}"
    rf.spectrum_magnitude(cases[rep][1], cases[rep][0])  # eager work between replays
rf.check(rf._lib.rf_graph_destroy(graph.value))
for _, _, plan, _, _ in cases:
    plan.destroy()
rf.check(rf._lib.rf_stream_destroy(stream))
print("capture ok", flush=True)
"""


def test_spectrum_magnitude_inside_a_capture() -> None:
    """rf_spectrum_magnitude allocates nothing for any plan kind (the last pass stores |X|), so
    it records into a graph at every length, and each replay gives the eager bits."""
    tests_dir = Path(__file__).resolve().parent
    proc = subprocess.run([sys.executable, "-c", SPECTRUM_CHILD, str(tests_dir)],
                          capture_output=True, text=True, timeout=300)
    assert proc.returncode == 0 and "capture ok" in proc.stdout, proc.stdout + proc.stderr
