"""The multithreaded FP32 CPU baseline (oracle/cpu_fast.cpp) against the float64 oracle: the
numbers bench.py reports beside the GPU's must come from a correct computation. FP32 tolerances
(the reference's rtol = atol = 1e-5 scaled by the output's size), exact peak lags on the
fixtures, and results that do not depend on the thread count."""
from __future__ import annotations

import json
import sys

import numpy as np
import pytest

from _common import FIXTURES, ROOT, cf32, oracle, taps

sys.path.insert(0, str(ROOT / "oracle"))
import cpu_fast  # noqa: E402


def noise(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)


@pytest.mark.parametrize("n,t", [(1, 1), (15, 63), (1000, 1), (4097, 63), (20000, 1025),
                                 (3000, 2049)])
def test_fir_matches_oracle(n: int, t: int) -> None:
    x = noise(n, n + t)
    h = (np.random.default_rng(t).standard_normal(t) / np.sqrt(t)).astype(np.float32)
    want = oracle.fir(x, h)
    got = cpu_fast.fir(x, h)
    scale = np.abs(oracle.fir(np.abs(x), np.abs(h)))
    assert np.all(np.abs(got - want) <= 1e-5 + 1e-5 * scale)
    assert np.array_equal(cpu_fast.fir(x, h, 1), cpu_fast.fir(x, h, 5))


def test_fir_fixture_identity_and_zeros() -> None:
    x = cf32("tone_odd")
    assert np.array_equal(cpu_fast.fir(x, taps("identity")), x)
    assert not np.any(cpu_fast.fir(cf32("zeros"), taps("lowpass")))


@pytest.mark.parametrize("channels", [1, 12, 16, 64])
def test_channelize_matches_oracle(channels: int) -> None:
    x = noise(5000, channels)
    h = taps("channelizer_proto")
    want = oracle.channelize(x, h, channels)
    got = cpu_fast.channelize(x, h, channels)
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-4)
    assert np.array_equal(cpu_fast.channelize(x, h, channels, 1),
                          cpu_fast.channelize(x, h, channels, 3))


def test_gcc_phat_peaks_on_fixtures() -> None:
    a, b = cf32("ch0"), cf32("ch1")
    true = json.loads((FIXTURES / "ch1.json").read_text())["true_delay_samples"]
    pairs = [(a, b)] + [(cf32("array0"), cf32(f"array{k}")) for k in (1, 2, 3)]
    peaks = cpu_fast.gcc_phat_peaks(pairs)
    lags = [int(p) - (x.size - 1) for p, (x, _) in zip(peaks, pairs)]
    assert lags[0] == true
    assert lags[1:] == [json.loads((FIXTURES / f"array{k}.json").read_text())["delay_samples"]
                        for k in (1, 2, 3)]
    for p, (x, y) in zip(peaks, pairs):
        assert int(p) == oracle.peak(oracle.gcc_phat(x, y))[0]


def test_gcc_phat_peaks_uneven_pairs() -> None:
    rng = np.random.default_rng(3)
    pairs = []
    for k in range(12):
        na, nb = int(rng.integers(1, 3000)), int(rng.integers(1, 3000))
        src = noise(max(na, nb) + 300, k)
        shift = int(rng.integers(0, 200))
        pairs.append((src[300:300 + na], src[300 - shift:300 - shift + nb]))
    got = cpu_fast.gcc_phat_peaks(pairs, 4)
    want = [oracle.peak(oracle.gcc_phat(a, b))[0] for a, b in pairs]
    assert list(got) == want
