"""rf_fir_noalloc, the FIR entry point that allocates nothing (its tap spectrum / split taps go
to a fixed scratch array compiled into the library): called through the C ABI, two calls in a
row per variant, the second with a smaller input and other taps, each bit-equal to rf_fir with
the same arguments (the same kernels run; only where the derived taps are kept differs)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from _common import load_rfhip

pytestmark = pytest.mark.gpu
rf = load_rfhip()


def kaiser(num_taps: int, cutoff: float) -> np.ndarray:
    n = np.arange(num_taps) - (num_taps - 1) / 2
    h = 2 * cutoff * np.sinc(2 * cutoff * n) * np.kaiser(num_taps, 8.0)
    return (h / h.sum()).astype(np.float32)


def noise(seed: int, n: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)


def call(fn, impl: str, x: np.ndarray, h: np.ndarray, stream) -> np.ndarray:
    d_x, d_h = rf.DeviceBuffer.from_array(x), rf.DeviceBuffer.from_array(h)
    d_y = rf.DeviceBuffer(x.size)
    status = fn(rf.fir_impl_index(impl), ctypes.c_void_p(d_x.ptr), ctypes.c_void_p(d_h.ptr),
                ctypes.c_void_p(d_y.ptr), x.size, h.size, 0, ctypes.c_void_p(stream))
    rf.check(status)
    y = d_y.download(stream=stream)
    for b in (d_x, d_h, d_y):
        b.free()
    return y


@pytest.mark.parametrize("impl", ["fft_ola", "mfma_f16x3", "mfma_bf16x3", "register", "naive"])
def test_noalloc_matches_rf_fir_on_two_calls(impl: str) -> None:
    lib = rf.lib()
    stream = rf.Stream()
    # first call: long capture and the longest transform filter; second: shorter capture, other
    # taps (the scratch keeps the first call's contents and must be overwritten, not reused);
    # a third, shorter still, with a filter of the 512-point overlap-save range
    for seed, n, num_taps, cutoff in [(1, 300007, 2049, 0.1), (2, 40001, 777, 0.2),
                                      (3, 20011, 63, 0.2)]:
        x, h = noise(seed, n), kaiser(num_taps, cutoff)
        print(f"rf_fir_noalloc {impl} n={n} taps={num_taps}", flush=True)
        got = call(lib.rf_fir_noalloc, impl, x, h, stream.handle)
        want = call(lib.rf_fir, impl, x, h, stream.handle)
        np.testing.assert_array_equal(got, want)
    stream.destroy()


def test_noalloc_argument_errors() -> None:
    lib = rf.lib()
    d = rf.DeviceBuffer(16)
    assert lib.rf_fir_noalloc(99, d.ptr, d.ptr, d.ptr, 1, 1, 0, None) == rf.RF_INVALID_ARG
    assert lib.rf_fir_noalloc(0, d.ptr, d.ptr, d.ptr, 1, 1, 48, None) == rf.RF_INVALID_ARG
    assert lib.rf_fir_noalloc(0, None, None, None, 0, 1, 0, None) == rf.RF_OK
    d.free()
