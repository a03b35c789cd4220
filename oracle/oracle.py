"""numpy front end of the float64 C oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: the oracle the GPU path is compared with. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.

Every function takes and returns numpy arrays (complex128 / float64) and restates one
definition of the reference's tools/reference.py; the C file cites the lines.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

_c_double_p = ctypes.POINTER(ctypes.c_double)
_lib = None


def build() -> None:
    """Compile liboracle.so with the Makefile next to this file."""
    import subprocess

    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        L = _lib
        L.orc_fir.argtypes = [_c_double_p, ctypes.c_long, _c_double_p, ctypes.c_long, _c_double_p]
        L.orc_fir.restype = None
        L.orc_xcorr.argtypes = [_c_double_p, ctypes.c_long, _c_double_p, ctypes.c_long,
                                _c_double_p]
        L.orc_xcorr.restype = None
        L.orc_fft.argtypes = [_c_double_p, ctypes.c_long, ctypes.c_int]
        L.orc_fft.restype = ctypes.c_int
        L.orc_spectrum.argtypes = [_c_double_p, ctypes.c_long, ctypes.c_long, _c_double_p]
        L.orc_spectrum.restype = ctypes.c_int
        L.orc_magnitude.argtypes = [_c_double_p, ctypes.c_long, _c_double_p]
        L.orc_magnitude.restype = None
        L.orc_channelize.argtypes = [_c_double_p, ctypes.c_long, _c_double_p, ctypes.c_long,
                                     ctypes.c_long, _c_double_p]
        L.orc_channelize.restype = ctypes.c_int
        L.orc_gcc_phat.argtypes = [_c_double_p, ctypes.c_long, _c_double_p, ctypes.c_long,
                                   _c_double_p]
        L.orc_gcc_phat.restype = ctypes.c_int
        L.orc_peak.argtypes = [_c_double_p, ctypes.c_long, _c_double_p]
        L.orc_peak.restype = ctypes.c_long
        L.orc_doa.argtypes = [ctypes.POINTER(ctypes.c_long), ctypes.c_long, ctypes.c_double,
                              ctypes.c_double, ctypes.c_double]
        L.orc_doa.restype = ctypes.c_double
    return _lib


def _cplx(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.complex128))


def _real(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _p(a: np.ndarray):
    return a.ctypes.data_as(_c_double_p)


def fir(x, taps) -> np.ndarray:
    x = _cplx(x)
    taps = _real(taps)
    if taps.size < 1:
        raise ValueError("need at least one tap")
    y = np.zeros(x.size, dtype=np.complex128)
    if x.size:
        lib().orc_fir(_p(x), x.size, _p(taps), taps.size, _p(y))
    return y


def xcorr(a, b) -> np.ndarray:
    a = _cplx(a)
    b = _cplx(b)
    if a.size == 0 or b.size == 0:
        raise ValueError("both inputs need at least one sample")
    r = np.zeros(a.size + b.size - 1, dtype=np.complex128)
    lib().orc_xcorr(_p(a), a.size, _p(b), b.size, _p(r))
    return r


def xcorr_lags(num_a: int, num_b: int) -> np.ndarray:
    return np.arange(-(num_a - 1), num_b)


def fft(x, sign: int = -1) -> np.ndarray:
    """Unnormalised DFT; sign -1 forward, +1 inverse."""
    y = _cplx(x).copy()
    if y.size:
        rc = lib().orc_fft(_p(y), y.size, sign)
        if rc != 0:
            raise ValueError(f"orc_fft failed ({rc})")
    return y


def spectrum(x, nfft: int) -> np.ndarray:
    x = _cplx(x)
    if nfft < 1:
        raise ValueError("nfft must be at least 1")
    frames = -(-x.size // nfft)
    out = np.zeros(frames * nfft, dtype=np.complex128)
    if frames:
        lib().orc_spectrum(_p(x), x.size, nfft, _p(out))
    return out.reshape(frames, nfft)


def magnitude(x) -> np.ndarray:
    x = _cplx(x)
    out = np.zeros(x.size, dtype=np.float64)
    lib().orc_magnitude(_p(x), x.size, _p(out))
    return out.reshape(np.shape(x))


def channelize(x, taps, channels: int) -> np.ndarray:
    x = _cplx(x)
    taps = _real(taps)
    if channels < 1 or taps.size < 1:
        raise ValueError("need at least one channel and one tap")
    frames = -(-x.size // channels)
    out = np.zeros(frames * channels, dtype=np.complex128)
    if frames:
        rc = lib().orc_channelize(_p(x), x.size, _p(taps), taps.size, channels, _p(out))
        if rc != 0:
            raise ValueError(f"orc_channelize failed ({rc})")
    return out.reshape(frames, channels)


def gcc_phat(a, b) -> np.ndarray:
    a = _cplx(a)
    b = _cplx(b)
    if a.size == 0 or b.size == 0:
        raise ValueError("both inputs need at least one sample")
    r = np.zeros(a.size + b.size - 1, dtype=np.complex128)
    rc = lib().orc_gcc_phat(_p(a), a.size, _p(b), b.size, _p(r))
    if rc != 0:
        raise ValueError(f"orc_gcc_phat failed ({rc})")
    return r


def peak(r) -> tuple[int, float]:
    r = _cplx(r)
    mag = ctypes.c_double(0.0)
    idx = lib().orc_peak(_p(r), r.size, ctypes.byref(mag))
    return int(idx), float(mag.value)


def tdoa(a, b) -> int:
    idx, _ = peak(gcc_phat(a, b))
    return int(idx - (len(a) - 1))


def doa(channels, spacing: float, speed: float, rate: float) -> float:
    if len(channels) < 2:
        raise ValueError("need at least two channels")
    lags = (ctypes.c_long * (len(channels) - 1))(*[tdoa(channels[0], ch) for ch in channels[1:]])
    return float(lib().orc_doa(lags, len(channels) - 1, spacing, speed, rate))
