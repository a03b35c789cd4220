"""Pins the float64 oracle (oracle/oracle.c) before anything is compared with it.

Each definition is checked against something independent of the oracle's own code: numpy's
library routines (np.convolve, np.correlate, np.fft), brute-force loops written from the
reference's definitions (tests/test_reference.py:38-59,198-216 of the reference), and the
ground truth stored in the reference's fixture sidecars (known delays, array angle, tone bins).
No GPU is needed.
"""
from __future__ import annotations

import hashlib
import re

import numpy as np
import pytest

from _common import FIXTURES, cf32, meta, oracle, taps

RTOL = 1e-9
ATOL = 1e-9


def random_complex(seed: int, n: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def fir_brute(x, h):
    y = np.zeros(len(x), dtype=np.complex128)
    for n in range(len(x)):
        for k in range(len(h)):
            if n - k >= 0:
                y[n] += h[k] * x[n - k]
    return y


def xcorr_brute(a, b):
    r = np.zeros(len(a) + len(b) - 1, dtype=np.complex128)
    for i, k in enumerate(range(-(len(a) - 1), len(b))):
        for n in range(len(a)):
            if 0 <= n + k < len(b):
                r[i] += b[n + k] * np.conj(a[n])
    return r


def channelize_brute(x, h, m):
    frames = -(-len(x) // m)
    z = np.zeros((frames, m), dtype=np.complex128)
    for t in range(frames):
        for k in range(m):
            z[t, k] = sum(h[l] * x[t * m - l] * np.exp(2j * np.pi * k * l / m)
                          for l in range(len(h)) if t * m - l >= 0)
    return z


# --- fixtures ------------------------------------------------------------------------------

def test_fixture_checksums_match_provenance() -> None:
    text = (FIXTURES.parent / "PROVENANCE.md").read_text()
    listed = dict((name, digest) for digest, name in re.findall(r"([0-9a-f]{64})  (\S+)", text))
    assert len(listed) == 30
    for name, digest in listed.items():
        assert hashlib.sha256((FIXTURES / name).read_bytes()).hexdigest() == digest, name


# --- FIR -----------------------------------------------------------------------------------

@pytest.mark.parametrize("n,t", [(50, 7), (50, 1), (5, 12), (1, 1), (300, 63)])
def test_fir_matches_brute_force_and_numpy(n: int, t: int) -> None:
    x = random_complex(3, n)
    h = np.random.default_rng(4).standard_normal(t)
    y = oracle.fir(x, h)
    np.testing.assert_allclose(y, fir_brute(x, h), rtol=RTOL, atol=ATOL)
    np.testing.assert_allclose(y, np.convolve(x, h)[:n], rtol=RTOL, atol=ATOL)


@pytest.mark.parametrize("name", ["tone", "tone_odd", "short"])
def test_fir_identity_returns_the_input(name: str) -> None:
    x = cf32(name).astype(np.complex128)
    np.testing.assert_array_equal(oracle.fir(x, taps("identity")), x)


def test_fir_of_zeros_is_zero() -> None:
    assert not np.any(oracle.fir(cf32("zeros"), taps("lowpass_long")))


def test_fir_impulse_gives_the_taps() -> None:
    h = taps("lowpass").astype(np.float64)
    x = np.zeros(100, dtype=np.complex128)
    x[0] = 1.0
    np.testing.assert_array_equal(oracle.fir(x, h)[: h.size].real, h)


@pytest.mark.parametrize("taps_name", ["lowpass", "lowpass_long"])
def test_fir_lowpass_keeps_passband_and_removes_stopband(taps_name: str) -> None:
    # Reference tests/test_reference.py:131-141: bin 40 keeps its level, bin -300 loses > 40 dB.
    y = oracle.fir(cf32("tone"), taps(taps_name))
    mag = np.abs(oracle.spectrum(y, 1024))[2]
    np.testing.assert_allclose(mag[40], 1024.0, rtol=0.02)
    assert mag[-300 % 1024] < 512.0 / 100.0


# --- FFT / spectrum ------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 3, 4, 8, 12, 16, 100, 1024, 4096, 8192, 1 << 15])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_matches_numpy(n: int, sign: int) -> None:
    x = random_complex(n, n)
    want = np.fft.fft(x) if sign == -1 else np.fft.ifft(x) * n
    np.testing.assert_allclose(oracle.fft(x, sign), want, rtol=1e-9, atol=1e-9 * np.sqrt(n))


def test_spectrum_matches_brute_force_dft() -> None:
    x = random_complex(2, 40)
    frames = oracle.spectrum(x, 16)
    assert frames.shape == (3, 16)
    padded = np.concatenate([x, np.zeros(8)])
    k = np.arange(16)
    for f in range(3):
        frame = padded[16 * f: 16 * (f + 1)]
        dft = np.array([np.sum(frame * np.exp(-2j * np.pi * m * k / 16)) for m in range(16)])
        np.testing.assert_allclose(frames[f], dft, rtol=RTOL, atol=ATOL)


def test_spectrum_peaks_on_the_fixture_tone_bins() -> None:
    m = meta("tone")
    mag = np.abs(oracle.spectrum(cf32("tone"), 1024))
    expected = sorted(t["bin"] % 1024 for t in m["tones"])
    for frame in mag:
        assert sorted(np.argsort(frame)[-len(expected):]) == expected
    for tone in m["tones"]:
        np.testing.assert_allclose(mag[:, tone["bin"] % 1024], tone["amplitude"] * 1024, rtol=0.02)


@pytest.mark.parametrize("name,frames", [("tone", 64), ("tone_odd", 10), ("short", 1)])
def test_spectrum_frame_count(name: str, frames: int) -> None:
    assert oracle.spectrum(cf32(name), 1024).shape == (frames, 1024)


def test_magnitude_is_hypot() -> None:
    x = random_complex(9, 33)
    np.testing.assert_allclose(oracle.magnitude(x), np.abs(x), rtol=1e-15)


# --- cross-correlation ---------------------------------------------------------------------

@pytest.mark.parametrize("na,nb", [(20, 20), (9, 31), (31, 9), (1, 5), (1, 1)])
def test_xcorr_matches_brute_force_and_numpy(na: int, nb: int) -> None:
    a = random_complex(5, na)
    b = random_complex(6, nb)
    r = oracle.xcorr(a, b)
    np.testing.assert_allclose(r, xcorr_brute(a, b), rtol=RTOL, atol=ATOL)
    np.testing.assert_allclose(r, np.correlate(b, a, "full"), rtol=RTOL, atol=ATOL)


def test_xcorr_recovers_the_fixture_delay() -> None:
    a, b = cf32("ch0"), cf32("ch1")
    r = oracle.xcorr(a, b)
    lags = oracle.xcorr_lags(a.size, b.size)
    assert lags[np.argmax(np.abs(r))] == meta("ch1")["true_delay_samples"] == 37


def test_xcorr_of_zeros_is_zero() -> None:
    z = cf32("zeros")
    assert not np.any(oracle.xcorr(z, z))


# --- channelizer ---------------------------------------------------------------------------

@pytest.mark.parametrize("n,t,m", [(100, 37, 4), (64, 16, 16), (7, 63, 4), (33, 1, 1),
                                   (50, 24, 8), (130, 40, 6)])
def test_channelize_matches_brute_force(n: int, t: int, m: int) -> None:
    rng = np.random.default_rng(n + t + m)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    h = rng.standard_normal(t)
    np.testing.assert_allclose(oracle.channelize(x, h, m), channelize_brute(x, h, m), atol=1e-10)


def test_channelize_one_channel_is_the_fir() -> None:
    x, h = cf32("tone_odd"), taps("lowpass")
    np.testing.assert_allclose(oracle.channelize(x, h, 1).ravel(), oracle.fir(x, h), atol=1e-12)


def test_channelize_puts_a_tone_in_its_channel() -> None:
    # Reference tests/test_pipeline.py:386-398 on the oracle itself.
    n = np.arange(16 * 512)
    x = np.exp(2j * np.pi * 3 * n / 16)
    z = oracle.channelize(x, taps("channelizer_proto"), 16)[16:]
    power = np.mean(np.abs(z) ** 2, axis=0)
    assert np.argmax(power) == 3
    assert power[3] > 1000 * np.max(np.delete(power, 3))
    np.testing.assert_allclose(np.abs(z[:, 3]), 1.0, atol=0.02)


# --- GCC-PHAT, peak, DOA -------------------------------------------------------------------

def test_gcc_phat_is_xcorr_of_whitened_spectra() -> None:
    a = random_complex(31, 50)
    b = random_complex(32, 70)
    length = 128
    cross = np.fft.fft(b, length) * np.conj(np.fft.fft(a, length))
    circular = np.fft.ifft(cross / np.abs(cross))
    lags = oracle.xcorr_lags(50, 70)
    np.testing.assert_allclose(oracle.gcc_phat(a, b), circular[lags % length], atol=1e-12)


def test_gcc_phat_of_a_pure_delay_is_one_at_that_lag() -> None:
    a = random_complex(33, 256)
    b = np.roll(a, 9)
    r = oracle.gcc_phat(a, b)
    idx, mag = oracle.peak(r)
    assert oracle.xcorr_lags(256, 256)[idx] == 9
    assert np.max(np.abs(r)) <= 1.0 + 1e-12


def test_gcc_phat_peaks_at_the_fixture_delay() -> None:
    r = oracle.gcc_phat(cf32("ch0"), cf32("ch1"))
    mags = np.abs(r)
    idx, _ = oracle.peak(r)
    assert oracle.xcorr_lags(16384, 16384)[idx] == 37
    assert mags[idx] > 20 * np.sort(mags)[-2]


def test_gcc_phat_of_zeros_is_exactly_zero() -> None:
    z = cf32("zeros")
    r = oracle.gcc_phat(z, z)
    assert r.size == 2 * 4096 - 1 and not np.any(r)


def test_peak_takes_the_first_maximum() -> None:
    r = np.array([1, 3, 2, 3, -3], dtype=np.complex128)
    assert oracle.peak(r) == (1, 3.0)


def test_doa_recovers_the_array_fixture_angle() -> None:
    m = meta("array0")
    chans = [cf32(f"array{i}") for i in range(m["channels"])]
    assert [oracle.tdoa(chans[0], c) for c in chans[1:]] == [2, 4, 6]
    angle = oracle.doa(chans, m["spacing_m"], m["speed_m_s"], m["sample_rate"])
    assert angle == pytest.approx(m["angle_deg"], abs=1e-9)
