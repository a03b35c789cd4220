"""Pins the float64 oracle of the learned delay search (oracle/delay.py) by what fixes it
independently: published known-answer vectors, closed forms, an identity against the C
oracle's GCC-PHAT, and finite differences. CPU only."""
from __future__ import annotations

import numpy as np
import pytest
from scipy.special import erf

from _common import oracle  # noqa: F401  (puts oracle/ on sys.path)
import delay as D


# --- Philox4x32-10: Random123's published known-answer vectors ----------------------------

@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox_known_answers(ctr, key, expect) -> None:
    out = D.philox4x32_10(np.array([ctr], dtype=np.uint32), *key)[0]
    assert tuple(int(v) for v in out) == expect


def test_box_muller_closed_form_and_moments() -> None:
    # u1 = ((x0 >> 8) + 1) 2^-24, u2 = (x1 >> 8) 2^-24
    x0 = np.uint32(((1 << 23) - 1) << 8)   # u1 = 1/2
    x1 = np.uint32((1 << 22) << 8)         # u2 = 1/4
    z = D.complex_normal(x0, x1)
    assert z == pytest.approx(np.sqrt(np.log(2.0)) * 1j, abs=1e-15)
    zs = D._stream(7, 3, 1, 1 << 16)
    assert abs(zs.mean()) < 0.02
    assert np.mean(np.abs(zs) ** 2) == pytest.approx(1.0, abs=0.02)
    assert abs(np.mean(zs * zs)) < 0.02    # circular: E[z^2] = 0


# --- generator --------------------------------------------------------------------------

def test_pair_parameters_stay_in_their_ranges() -> None:
    for p in range(200):
        q = D.pair_params(0xC5, p)
        assert -512.0 <= q["delay"] < 512.0
        assert -10.0 <= q["snr_db"] <= 20.0
        assert 0.0 <= q["rho"][0] < 0.9 and 0.0 <= q["rho"][1] < 0.9
        assert q["delay"] == float(np.float32(q["delay"]))


def test_fractional_delay_filter() -> None:
    h0 = D.fractional_delay_taps(0.0)
    expect = np.zeros(65)
    expect[32] = 1.0
    np.testing.assert_allclose(h0, expect, atol=1e-15)
    # delays a slow tone by f: sum_j h[j] x[n - j] = x(n - f)
    f = 0.3
    h = D.fractional_delay_taps(f)
    omega = 0.05 * 2 * np.pi
    j = np.arange(-32, 33)
    response = np.sum(h * np.exp(-1j * omega * j))
    assert response == pytest.approx(np.exp(-1j * omega * f), abs=2e-3)


def test_ar1_lag_one_correlation() -> None:
    e = D._stream(11, 0, 2, 1 << 16)
    for rho in (0.0, 0.5, 0.85):
        w = D.ar1(e, rho)
        c1 = np.vdot(w[:-1], w[1:]).real / np.vdot(w, w).real
        assert c1 == pytest.approx(rho, abs=0.02)
        assert np.mean(np.abs(w) ** 2) == pytest.approx(1.0, abs=0.05)
    w = D.ar1(e[:5], 0.5)
    assert w[0] == e[0]
    assert w[1] == pytest.approx(0.5 * e[0] + np.sqrt(0.75) * e[1])


def test_gcc_phat_of_synthetic_pairs_peaks_at_the_rounded_delay() -> None:
    n = 1 << 13
    for p in range(6):
        a, b, q = D.synth_pair(0xC5, p, n)
        r = oracle.gcc_phat(a, b)
        idx, _ = oracle.peak(r)
        assert abs((idx - (n - 1)) - q["delay"]) <= 1.0, (p, q)


# --- features ---------------------------------------------------------------------------

def test_coarse_correlation_is_the_windowed_aliased_gcc_phat() -> None:
    """c[m] = (1/B) sum_{l = m mod 1024} r(l) D(l), D(l) = sum_{u<B} exp(-2 pi i u l / L),
    with r from the C oracle (its own FFT)."""
    n = 1 << 14
    a, b, _ = D.synth_pair(0xC5, 2, n)
    L = D.fft_length(n, n)
    B = L // D.BANDS
    c = D.coarse_correlation(D.phat_spectrum(a, b), n)
    r = oracle.gcc_phat(a, b)                   # index i = lag + n - 1
    lags = np.arange(-(n - 1), n)
    u = np.arange(B)
    dk = np.exp(-2j * np.pi * np.outer(lags, u) / L).sum(axis=1)
    expect = np.zeros(D.BANDS, dtype=np.complex128)
    np.add.at(expect, lags % D.BANDS, r * dk / B)
    np.testing.assert_allclose(c, expect, atol=1e-9)


def test_phat_spectrum_inverts_to_the_gcc_phat_output() -> None:
    a, b, _ = D.synth_pair(0xC5, 4, 3000)
    w = D.phat_spectrum(a, b)
    r = np.fft.ifft(w) * w.size
    np.testing.assert_allclose(r[:5999], oracle.gcc_phat(a, b), atol=1e-12)


def test_features_are_normalised_pooled_magnitudes() -> None:
    n = 1 << 14
    a, b, q = D.synth_pair(0xC5, 5, n)
    f, coarse = D.features(a, b)
    assert f.shape == (256,) and f.max() == 1.0 and f.min() >= 0.0
    assert coarse == D.label_of(q["delay"])


# --- network ----------------------------------------------------------------------------

def test_bf16_rounding() -> None:
    assert D.bf16(1.0 + 2.0 ** -8) == 1.0                       # tie, even mantissa
    assert D.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6       # tie, rounds up to even
    assert D.bf16(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    assert D.bf16(-3.0) == -3.0
    assert D.bf16(np.float32(1e-3)) == pytest.approx(1e-3, rel=2 ** -8)


def _permutation_params(scale: float = 1.0) -> np.ndarray:
    w1 = np.zeros((256, 256))
    w1[np.arange(256), (np.arange(256) * 3 + 1) % 256] = scale  # h1[o] = gelu(x[(3o+1)%256])
    w2 = np.eye(256)
    w3 = np.zeros((64, 256))
    w3[np.arange(64), np.arange(64) * 4] = 1.0                   # z3[o] = h2[4o]
    b3 = np.zeros(64)
    b3[5] = np.log(2.0)
    return np.concatenate([w1.ravel(), np.zeros(256), w2.ravel(), np.zeros(256), w3.ravel(), b3])


def test_forward_closed_form() -> None:
    x = np.zeros((2, 256))
    x[0, 1] = 1.0        # reaches h1[0] -> h2[0] -> z3[0]
    x[1, 13] = -1.0      # (3*4+1) = 13: h1[4] -> z3[1]
    out = D.mlp_forward(_permutation_params(), x, round_bf16=False)
    g1 = 0.5 * (1 + erf(1 / np.sqrt(2)))         # gelu(1) = Phi(1)
    gm1 = -0.5 * (1 + erf(-1 / np.sqrt(2)))      # gelu(-1) = -Phi(-1)
    assert out["h1"][0, 0] == pytest.approx(g1)
    assert out["z3"][0, 0] == pytest.approx(0.5 * g1 * (1 + erf(g1 / np.sqrt(2))))
    assert out["z3"][1, 1] == pytest.approx(0.5 * gm1 * (1 + erf(gm1 / np.sqrt(2))))
    # all other logits of row 1 are 0 except z3[5] = ln 2
    z = np.zeros(64)
    z[1] = out["z3"][1, 1]
    z[5] = np.log(2.0)
    np.testing.assert_allclose(out["p"][1], np.exp(z) / np.exp(z).sum(), rtol=1e-12)
    np.testing.assert_allclose(out["p"].sum(axis=1), 1.0, rtol=1e-12)


def test_backward_matches_finite_differences() -> None:
    rng = np.random.default_rng(3)
    params = rng.normal(0, 0.08, D.PARAM_COUNT)
    x = rng.uniform(0, 1, (3, 256))
    labels = np.array([4, 17, 63])
    g = D.mlp_backward(params, x, labels, round_bf16=False)
    idx = np.concatenate([rng.integers(0, D.PARAM_COUNT, 40),
                          [65536, 65536 + 255, 131328, 147968, 148031]])  # biases too
    eps = 1e-6
    for i in idx:
        p1 = params.copy()
        p2 = params.copy()
        p1[i] += eps
        p2[i] -= eps
        num = (D.mlp_loss(p1, x, labels) - D.mlp_loss(p2, x, labels)) / (2 * eps)
        assert g[i] == pytest.approx(num, rel=1e-5, abs=1e-9), i


def test_adam_first_step_is_a_signed_learning_rate() -> None:
    g = np.array([0.3, -2.0, 1e-3])
    p, m, v = D.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01)
    np.testing.assert_allclose(p, -0.01 * np.sign(g), rtol=1e-4)


def test_refine_finds_the_window_peak() -> None:
    n = 1 << 14
    a, b, q = D.synth_pair(0xC5, 7, n)
    w = D.phat_spectrum(a, b)
    r = np.fft.ifft(w) * w.size
    exact = int(np.argmax(np.abs(r[:2 * n - 1])))
    out = D.refine(w, n, D.label_of(q["delay"]))
    assert out is not None and out[0] == exact
    assert out[1] == pytest.approx(abs(r[exact]), rel=1e-9)


def test_refine_rejects_a_window_of_pure_noise() -> None:
    # Independent captures: the PHAT output holds no signal. Its lags have E|r|^2 = 1/L (unit
    # weights of 1/L each), no lag reaches the noise floor sqrt((ln L + 5) / L), and the refine
    # rejects every window it is offered.
    n = 1 << 12
    rng = np.random.default_rng(17)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    w = D.phat_spectrum(a, b)
    L = w.size
    r = np.fft.ifft(w) * L
    assert np.mean(np.abs(r) ** 2) == pytest.approx(1.0 / L, rel=1e-9)
    assert np.max(np.abs(r)) < D.noise_floor(L)
    assert all(D.refine(w, n, b) is None for b in (0, 17, 31, 63))
