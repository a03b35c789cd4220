"""float64 restatement of the learned delay search (librfhip's delay_synth.hip, delay_search.hip
and delay_mlp.hip), which has no counterpart in the reference.

TEST INFRASTRUCTURE ONLY: the oracle the GPU path is compared with. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.

What pins each part (tests/test_oracle_delay.py):
  philox4x32_10      the published known-answer vectors of Philox4x32-10 (Random123);
  complex_normal     the Box-Muller closed form at chosen words, moments of 2^16 samples;
  synth_pairs        AR(1) lag-1 correlation, the windowed-sinc filter's impulse and
                     delayed-tone responses, GCC-PHAT peak at the rounded true delay;
  features           the identity c[m] = (1/B) sum_{l = m mod 1024} r(l) D(l) against the C
                     oracle's GCC-PHAT output (an independent FFT);
  mlp_forward        closed-form outputs for permutation weights (GELU(1) = Phi(1), ...);
  mlp_backward       central finite differences of mlp_loss in float64;
  bf16               round-to-nearest-even ties and known values.
"""
from __future__ import annotations

import numpy as np

SOURCE_OFFSET = 576
SOURCE_EXTRA = 1152
SPAN = 512
BANDS = 1024
FEATURES = 256
BINS = 64
WINDOW = 64
WINDOW_MARGIN = 24
N_IN, N_HIDDEN, N_OUT = 256, 256, 64
PARAM_COUNT = N_HIDDEN * N_IN + N_HIDDEN + N_HIDDEN * N_HIDDEN + N_HIDDEN + N_OUT * N_HIDDEN + N_OUT

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr: np.ndarray, k0: int, k1: int) -> np.ndarray:
    """Philox4x32-10 of counters ctr[..., 4] (uint32) under key (k0, k1)."""
    c = [np.asarray(ctr[..., i], dtype=np.uint64) for i in range(4)]
    key0, key1 = k0 & 0xFFFFFFFF, k1 & 0xFFFFFFFF
    for _ in range(10):
        p0 = _M0 * c[0]
        p1 = _M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c = [hi1 ^ c[1] ^ np.uint64(key0), lo1, hi0 ^ c[3] ^ np.uint64(key1), lo0]
        key0 = (key0 + _W0) & 0xFFFFFFFF
        key1 = (key1 + _W1) & 0xFFFFFFFF
    return np.stack([x.astype(np.uint32) for x in c], axis=-1)


def uniform24(x) -> np.ndarray:
    return (np.asarray(x, dtype=np.uint32) >> np.uint32(8)).astype(np.float64) * 2.0 ** -24


def complex_normal(x0, x1) -> np.ndarray:
    """CN(0, 1) from two words: u1 in (0, 1], u2 in [0, 1), sqrt(-ln u1) exp(2 pi i u2)."""
    u1 = ((np.asarray(x0, dtype=np.uint32) >> np.uint32(8)).astype(np.float64) + 1.0) * 2.0 ** -24
    u2 = uniform24(x1)
    return np.sqrt(-np.log(u1)) * np.exp(2j * np.pi * u2)


def _stream(seed: int, pair: int, stream: int, count: int) -> np.ndarray:
    ctr = np.zeros((count, 4), dtype=np.uint32)
    ctr[:, 0] = np.arange(count, dtype=np.uint32)
    ctr[:, 1] = pair
    ctr[:, 2] = stream
    x = philox4x32_10(ctr, seed & 0xFFFFFFFF, seed >> 32)
    return complex_normal(x[:, 0], x[:, 1])


def pair_params(seed: int, pair: int) -> dict:
    """Delay, SNR and AR(1) coefficients of one pair, with the float32 roundings of the GPU."""
    x = philox4x32_10(np.array([[0, pair, 0, 0]], dtype=np.uint32), seed & 0xFFFFFFFF, seed >> 32)[0]
    f32 = np.float32
    u = [f32(uniform24(v)) for v in x]
    delay = float(f32(-512.0) + f32(1024.0) * u[0])
    snr_db = float(f32(-10.0) + f32(30.0) * u[1])
    return {"delay": delay, "snr_db": snr_db, "sigma": 10.0 ** (-snr_db / 20.0),
            "rho": (float(f32(0.9) * u[2]), float(f32(0.9) * u[3]))}


def fractional_delay_taps(frac: float) -> np.ndarray:
    """h[j + 32], j = -32..32: sinc(j - f) Kaiser(beta 8), normalised to unit sum."""
    j = np.arange(-32, 33, dtype=np.float64)
    h = np.sinc(j - frac) * np.i0(8.0 * np.sqrt(1.0 - (j / 32.0) ** 2)) / np.i0(8.0)
    return h / h.sum()


def ar1(e: np.ndarray, rho: float) -> np.ndarray:
    """w[0] = e[0], w[i] = rho w[i-1] + sqrt(1 - rho^2) e[i]."""
    from scipy.signal import lfilter

    c = np.sqrt(1.0 - rho * rho)
    w = np.empty_like(e)
    w[0] = e[0]
    if e.size > 1:
        w[1:] = lfilter([c], [1.0, -rho], e[1:], zi=[rho * e[0]])[0]
    return w


def synth_pair(seed: int, pair: int, n: int):
    """(a, b, params) of pair `pair`: a = s + sigma w_a, b = s delayed by d + sigma w_b."""
    q = pair_params(seed, pair)
    s = _stream(seed, pair, 1, n + SOURCE_EXTRA)
    d = q["delay"]
    di = int(np.floor(d))
    h = fractional_delay_taps(d - di)
    idx = np.arange(n) + SOURCE_OFFSET
    a_sig = s[idx]
    # b_sig[i] = sum_j h[j] s[i + 576 - di - j], j = -32..32
    b_sig = np.zeros(n, dtype=np.complex128)
    for jj in range(65):
        b_sig += h[jj] * s[idx - di - (jj - 32)]
    wa = ar1(_stream(seed, pair, 2, n), q["rho"][0])
    wb = ar1(_stream(seed, pair, 3, n), q["rho"][1])
    return a_sig + q["sigma"] * wa, b_sig + q["sigma"] * wb, q


def synth_pairs(seed: int, first: int, count: int, n: int):
    a = np.empty((count, n), dtype=np.complex128)
    b = np.empty((count, n), dtype=np.complex128)
    d = np.empty(count)
    for k in range(count):
        a[k], b[k], q = synth_pair(seed, first + k, n)
        d[k] = q["delay"]
    return a, b, d


def fft_length(na: int, nb: int) -> int:
    lags = na + nb - 1
    return 1 << (lags - 1).bit_length()


def phat_spectrum(a, b) -> np.ndarray:
    """W[k] = X / |X| / L (0 where X = 0), X = FFT(b padded) conj(FFT(a rotated))
    (the spectrum whose inverse DFT is the GCC-PHAT output in xcorr's index order)."""
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    na, nb = a.size, b.size
    L = fft_length(na, nb)
    ar = np.zeros(L, dtype=np.complex128)
    m = (np.arange(L) + na - 1) % L
    ok = m < na
    ar[ok] = a[m[ok]]
    bp = np.zeros(L, dtype=np.complex128)
    bp[:nb] = b
    x = np.fft.fft(bp) * np.conj(np.fft.fft(ar))
    mag = np.abs(x)
    w = np.zeros(L, dtype=np.complex128)
    nz = mag > 0
    w[nz] = x[nz] / mag[nz] / L
    return w


def coarse_correlation(w: np.ndarray, na: int) -> np.ndarray:
    """c[m], m < 1024: inverse DFT of the band sums of W[k] exp(+2 pi i k (na-1) / L)."""
    L = w.size
    k = np.arange(L)
    wp = w * np.exp(2j * np.pi * ((k * (na - 1)) % L) / L)
    y = wp.reshape(BANDS, L // BANDS).sum(axis=1)
    return np.fft.ifft(y) * BANDS


def features(a, b):
    """(256 features, coarse bin) of one pair."""
    na = np.asarray(a).size
    c = np.abs(coarse_correlation(phat_spectrum(a, b), na))
    lags = np.arange(-SPAN, SPAN)
    mag = c[lags % BANDS]
    f = mag.reshape(FEATURES, 4).max(axis=1)
    top = f.max()
    feats = f / top if top > 0 else np.zeros(FEATURES)
    coarse_bin = int(np.argmax(mag)) >> 4
    return feats, coarse_bin


def label_of(delay: float) -> int:
    lag = int(np.rint(np.float32(delay)))
    return int(min(max((lag + SPAN) >> 4, 0), BINS - 1))


# --- network ------------------------------------------------------------------------------

def bf16(x) -> np.ndarray:
    """Round to bfloat16 (nearest, ties to even) through float32; returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    bits = f.view(np.uint32).astype(np.uint64)
    bias = np.uint64(0x7FFF) + ((bits >> np.uint64(16)) & np.uint64(1))
    out = ((bits + bias) & np.uint64(0xFFFF0000)).astype(np.uint32)
    return out.view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    o = 0

    def take(*shape):
        nonlocal o
        size = int(np.prod(shape))
        v = p[o:o + size].reshape(shape)
        o += size
        return v

    w1, b1 = take(N_HIDDEN, N_IN), take(N_HIDDEN)
    w2, b2 = take(N_HIDDEN, N_HIDDEN), take(N_HIDDEN)
    w3, b3 = take(N_OUT, N_HIDDEN), take(N_OUT)
    return w1, b1, w2, b2, w3, b3


def gelu(z):
    from scipy.special import erf

    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    from scipy.special import erf

    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def mlp_forward(params, x, round_bf16: bool = True) -> dict:
    """Activations of the 256-256-256-64 network on rows x; GEMM operands rounded to BF16
    (as the matrix cores see them) when round_bf16."""
    r = bf16 if round_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    w1, b1, w2, b2, w3, b3 = unpack(params)
    x = np.asarray(x, dtype=np.float64)
    z1 = r(x) @ r(w1).T + b1
    h1 = gelu(z1)
    z2 = r(h1) @ r(w2).T + b2
    h2 = gelu(z2)
    z3 = r(h2) @ r(w3).T + b3
    z = z3 - z3.max(axis=1, keepdims=True)
    p = np.exp(z)
    p /= p.sum(axis=1, keepdims=True)
    return {"x": x, "z1": z1, "h1": h1, "z2": z2, "h2": h2, "z3": z3, "p": p}


def mlp_loss(params, x, labels, round_bf16: bool = False) -> float:
    p = mlp_forward(params, x, round_bf16)["p"]
    return float(-np.mean(np.log(p[np.arange(len(labels)), labels])))


def mlp_backward(params, x, labels, round_bf16: bool = True) -> np.ndarray:
    """Gradient of the mean cross-entropy, in the parameter layout."""
    r = bf16 if round_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    w1, b1, w2, b2, w3, b3 = unpack(params)
    a = mlp_forward(params, x, round_bf16)
    batch = a["x"].shape[0]
    dz3 = a["p"].copy()
    dz3[np.arange(batch), labels] -= 1.0
    dz3 /= batch
    gw3 = r(dz3).T @ r(a["h2"])
    gb3 = dz3.sum(axis=0)
    dh2 = r(dz3) @ r(w3)
    dz2 = dh2 * gelu_grad(a["z2"])
    gw2 = r(dz2).T @ r(a["h1"])
    gb2 = dz2.sum(axis=0)
    dh1 = r(dz2) @ r(w2)
    dz1 = dh1 * gelu_grad(a["z1"])
    gw1 = r(dz1).T @ r(a["x"])
    gb1 = dz1.sum(axis=0)
    return np.concatenate([gw1.ravel(), gb1, gw2.ravel(), gb2, gw3.ravel(), gb3])


def adam_step(params, grads, m, v, step: int, lr: float):
    b1, b2, eps = 0.9, 0.999, 1e-8
    m = b1 * m + (1 - b1) * grads
    v = b2 * v + (1 - b2) * grads * grads
    p = params - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v


# --- window refine --------------------------------------------------------------------------

def window_lags(bin_index: int) -> np.ndarray:
    lo = -SPAN + 16 * bin_index - WINDOW_MARGIN
    return np.arange(lo, lo + WINDOW)


def noise_floor(L: int) -> float:
    """|r| level that the L lags of PHAT noise (E|r|^2 = 1/L, P(|r|^2 > t/L) = e^-t) all stay
    below with probability 1 - e^-5: sqrt((ln L + 5) / L)."""
    return float(np.sqrt((np.log(L) + 5.0) / L))


def refine(w: np.ndarray, na: int, bin_index: int):
    """(output index, |r|, frac) of the first largest |r| over the proposal's window, or None
    when it sits at either edge or below noise_floor(L); r(l) = sum_k W[k]
    exp(2 pi i k (l + na - 1) / L)."""
    L = w.size
    lags = window_lags(bin_index)
    n = (lags + na - 1) % L
    k = np.arange(L)
    r = np.array([np.sum(w * np.exp(2j * np.pi * ((k * int(nn)) % L) / L)) for nn in n])
    mag = np.abs(r)
    best = int(np.argmax(mag))
    if best in (0, WINDOW - 1) or mag[best] < noise_floor(L):
        return None
    m0, ml, mh = mag[best], mag[best - 1], mag[best + 1]
    frac = 0.0
    if m0 > 0:
        frac = mh / (m0 + mh) if mh >= ml else -ml / (m0 + ml)
        frac = min(max(frac, -0.5), 0.5)
    return int(lags[best] + na - 1), float(m0), float(frac)
