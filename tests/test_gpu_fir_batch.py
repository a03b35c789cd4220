"""Persistent batched FIR (rf_fir_batch_*): hundreds of captures of uneven length, one launch.

Every capture's output equals rf_fir with the same impl bit for bit (the same segment / block
arithmetic), and a sample of captures equals the float64 oracle at the reference's FIR
tolerance (rtol = atol = 1e-5, tests/test_pipeline.py:113-117 of the reference). Replays of a
captured graph of the run give the eager bits.
"""
from __future__ import annotations

import numpy as np
import pytest

from _common import load_rfhip, oracle, taps

pytestmark = pytest.mark.gpu
rf = load_rfhip()


def captures(seed: int, count: int, longest: int) -> list[np.ndarray]:
    rng = np.random.default_rng(seed)
    lengths = np.exp(rng.uniform(0.0, np.log(longest), count)).astype(int)
    # edge lengths: empty, one sample, shorter than the halo, one 4096-point segment of a
    # 1025-tap filter (3072 outputs) and around it, one 512-point segment of a 63-tap filter (450
    # outputs) and around it
    lengths[:9] = [0, 1, 62, 1024, 3072, 3073, 449, 450, 451]
    return [(rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
            for n in lengths]


@pytest.mark.parametrize("impl,taps_name", [("fft_ola", "lowpass_long"), ("fft_ola", "lowpass"),
                                            ("register", "lowpass"),
                                            ("register", "lowpass_long")])
def test_batch_equals_rf_fir_per_capture(impl: str, taps_name: str) -> None:
    h = taps(taps_name)
    xs = captures(7, 300, 1 << 17)
    d_h = rf.DeviceBuffer.from_array(h)
    d_x = [rf.DeviceBuffer.from_array(x) for x in xs]
    d_y = [rf.DeviceBuffer(x.size) for x in xs]
    batch = rf.FirBatch(impl, d_h.ptr, h.size, [x.size for x in xs])
    batch.bind([b.ptr for b in d_x], [b.ptr for b in d_y])
    print(f"fir batch {impl} {taps_name}: {len(xs)} captures", flush=True)
    batch.run()
    rf.synchronize()
    for k, (x, y) in enumerate(zip(xs, d_y)):
        got = y.download()
        np.testing.assert_array_equal(got, rf.fir(x, h, impl), err_msg=f"capture {k}")
        if k < 12:
            np.testing.assert_allclose(got, oracle.fir(x, h), rtol=1e-5, atol=1e-5)
    batch.destroy()


@pytest.mark.parametrize("taps_name", ["lowpass_long", "lowpass"])
def test_batch_graph_replays(taps_name: str) -> None:
    h = taps(taps_name)
    xs = captures(8, 64, 1 << 16)
    stream = rf.Stream()
    d_h = rf.DeviceBuffer.from_array(h)
    d_x = [rf.DeviceBuffer.from_array(x) for x in xs]
    d_y = [rf.DeviceBuffer(x.size) for x in xs]
    batch = rf.FirBatch("fft_ola", d_h.ptr, h.size, [x.size for x in xs], stream)
    batch.bind([b.ptr for b in d_x], [b.ptr for b in d_y])
    batch.run(stream)
    stream.synchronize()
    want = [y.download() for y in d_y]
    graph = rf.Graph(stream, lambda: batch.run(stream))
    for rep in range(3):
        for y in d_y:
            y.zero(stream)
        graph.launch(stream)
        stream.synchronize()
        for y, w in zip(d_y, want):
            np.testing.assert_array_equal(y.download(stream=stream), w, err_msg=f"replay {rep}")
        rf.fir(xs[1], h, "fft_ola")  # eager work between replays
    graph.destroy()
    batch.destroy()
    stream.destroy()


def test_batch_argument_errors() -> None:
    d = rf.DeviceBuffer(16)
    with pytest.raises(rf.RfError) as e:
        rf.FirBatch("naive", d.ptr, 4, [10])
    assert e.value.status == rf.RF_INVALID_ARG
    b = rf.FirBatch("register", d.ptr, 4, [10, 0])
    with pytest.raises(rf.RfError):
        b.run()  # before bind
    b.destroy()
