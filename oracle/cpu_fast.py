"""ctypes binding of the multithreaded FP32 CPU baseline (cpu_fast.cpp, built into
_build/libcpufast.so by `make -C oracle`). Measurement infrastructure: bench.py's cpu_baseline
leg and the tests call it; the library never does."""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB = None


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = HERE / "_build" / "libcpufast.so"
        if not path.exists():
            subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        cdll = ctypes.CDLL(str(path))
        f32p = ctypes.POINTER(ctypes.c_float)
        cdll.cpu_threads.restype = ctypes.c_int
        cdll.cpu_fir_f32.argtypes = [f32p, ctypes.c_long, f32p, ctypes.c_long, f32p, ctypes.c_int]
        cdll.cpu_channelize_f32.argtypes = [f32p, ctypes.c_long, f32p, ctypes.c_long,
                                            ctypes.c_long, f32p, ctypes.c_int]
        cdll.cpu_gcc_phat_peaks.argtypes = [ctypes.c_int, ctypes.POINTER(f32p),
                                            ctypes.POINTER(ctypes.c_long), ctypes.POINTER(f32p),
                                            ctypes.POINTER(ctypes.c_long),
                                            ctypes.POINTER(ctypes.c_long), ctypes.c_int]
        _LIB = cdll
    return _LIB


def threads() -> int:
    return int(lib().cpu_threads())


def _f32p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _cf(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.complex64)


def fir(x, taps, nthreads: int | None = None) -> np.ndarray:
    x = _cf(x)
    h = np.ascontiguousarray(taps, dtype=np.float32)
    y = np.empty_like(x)
    rc = lib().cpu_fir_f32(_f32p(x.view(np.float32)), x.size, _f32p(h), h.size,
                           _f32p(y.view(np.float32)), nthreads or threads())
    if rc != 0:
        raise ValueError(f"cpu_fir_f32 failed ({rc})")
    return y


def channelize(x, taps, channels: int, nthreads: int | None = None) -> np.ndarray:
    x = _cf(x)
    h = np.ascontiguousarray(taps, dtype=np.float32)
    frames = -(-x.size // channels)
    z = np.empty(frames * channels, dtype=np.complex64)
    rc = lib().cpu_channelize_f32(_f32p(x.view(np.float32)), x.size, _f32p(h), h.size, channels,
                                  _f32p(z.view(np.float32)), nthreads or threads())
    if rc != 0:
        raise ValueError(f"cpu_channelize_f32 failed ({rc})")
    return z.reshape(frames, channels)


def gcc_phat_peaks(pairs, nthreads: int | None = None) -> np.ndarray:
    """pairs: list of (a, b) complex arrays; returns the peak index of |GCC-PHAT| per pair
    (lag = index - (len(a) - 1))."""
    arrays = [(_cf(a), _cf(b)) for a, b in pairs]
    count = len(arrays)
    f32p = ctypes.POINTER(ctypes.c_float)
    pa = (f32p * count)(*[_f32p(a.view(np.float32)) for a, _ in arrays])
    pb = (f32p * count)(*[_f32p(b.view(np.float32)) for _, b in arrays])
    na = (ctypes.c_long * count)(*[a.size for a, _ in arrays])
    nb = (ctypes.c_long * count)(*[b.size for _, b in arrays])
    peak = (ctypes.c_long * count)()
    rc = lib().cpu_gcc_phat_peaks(count, pa, na, pb, nb, peak, nthreads or threads())
    if rc != 0:
        raise ValueError(f"cpu_gcc_phat_peaks failed ({rc})")
    return np.array(peak[:], dtype=np.int64)
