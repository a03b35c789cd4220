"""Python binding of librfhip.so, the MI355X (gfx950) hot path of gpu-rf-pipeline.

The package directory holds the HIP sources (csrc/), the C++ host mirror of the reference's
operator surface (host/) and this ctypes binding, which mirrors the same surface for Python
callers and tests:

  * host-vector operations with the reference's names and semantics (src/ops.hpp:16-52):
    fir_impls(), xcorr_impls() (first entry = default), fir(), xcorr(), gcc_phat(),
    spectrum(), magnitude(), channelize(), doa(); each takes a Bench that times the
    upload / compute / download phases with device events (src/bench.hpp:1-46);
  * device-pointer operations (src/device_ops.hpp:17-106): fir_device(), fir_batched_device(),
    magnitude_device(), xcorr_direct_device(), XcorrFft, FftPlan, channelize_device(),
    spectrum_device(), peak_device();
  * DeviceBuffer / Stream / Event over the library's own runtime utilities.

There is no CPU fallback: if librfhip.so is missing or cannot be loaded, importing raises.
Build it with `make -C gpu-rf-pipeline_amd` (or __graft_entry__.build()).
"""
from __future__ import annotations

import atexit
import ctypes
import math
import os
from dataclasses import dataclass, field
from pathlib import Path
from typing import Callable, Sequence

import numpy as np

PACKAGE_DIR = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("RFHIP_LIB", PACKAGE_DIR / "librfhip.so"))
HEADER_PATH = PACKAGE_DIR.parent / "include" / "rfhip.h"

RF_OK, RF_INVALID_ARG, RF_LIMIT, RF_HIP_ERROR, RF_ALLOC = range(5)
RF_FFT_FORWARD, RF_FFT_INVERSE = -1, 1
DEFAULT_FIR_BLOCK = 256
MAX_FIR_BATCH = 8
MAX_CONST_TAPS = 8192


class RfError(RuntimeError):
    """A non-OK rf_status, with the library's message (the reference throws
    std::runtime_error with the same kind of text, src/check.hpp:42-66)."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


if not LIB_PATH.exists():
    raise ImportError(f"{LIB_PATH} is not built: run `make -C {PACKAGE_DIR}`")

_lib = ctypes.CDLL(str(LIB_PATH))
_vp = ctypes.c_void_p
_int = ctypes.c_int
_size = ctypes.c_size_t
_st = ctypes.c_int


class PeakResult(ctypes.Structure):
    _fields_ = [("index", ctypes.c_int32), ("magnitude", ctypes.c_float), ("frac", ctypes.c_double)]


# Set by an atexit hook. Once the interpreter is shutting down, finalizers leave device objects
# to the end of the process instead of calling into the HIP runtime, whose own teardown (and a
# profiler's, under rocprofv3) may already be under way; the driver reclaims everything when
# the process ends. Code that wants its objects released calls free()/destroy() itself.
_exiting = False


def _at_exit() -> None:
    global _exiting
    _exiting = True


atexit.register(_at_exit)


def _finalize(obj, attr: str, destroy) -> None:
    """Body of every __del__: release obj.<attr> with destroy(), unless the process is exiting
    or it was released already."""
    handle = getattr(obj, attr, None)
    if not handle or _exiting:
        return
    try:
        destroy(handle)
    except Exception:
        pass
    setattr(obj, attr, None)


def _proto(name: str, restype, *argtypes) -> None:
    fn = getattr(_lib, name)
    fn.restype = restype
    fn.argtypes = list(argtypes)


_proto("rf_status_string", ctypes.c_char_p, _st)
_proto("rf_last_error", ctypes.c_char_p)
_proto("rf_abi_version", _int)
_proto("rf_device_count", _st, ctypes.POINTER(_int))
_proto("rf_device_name", _st, ctypes.c_char_p, _int)
_proto("rf_device_memory", _st, ctypes.POINTER(_size))
_proto("rf_malloc", _st, ctypes.POINTER(_vp), _size)
_proto("rf_free", _st, _vp)
_proto("rf_host_alloc", _st, ctypes.POINTER(_vp), _size)
_proto("rf_host_free", _st, _vp)
_proto("rf_host_alloc_mapped", _st, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _size)
for _n in ("rf_memcpy_h2d", "rf_memcpy_d2h", "rf_memcpy_d2d"):
    _proto(_n, _st, _vp, _vp, _size, _vp)
_proto("rf_memset", _st, _vp, _int, _size, _vp)
_proto("rf_stream_create", _st, ctypes.POINTER(_vp))
_proto("rf_stream_destroy", _st, _vp)
_proto("rf_stream_synchronize", _st, _vp)
_proto("rf_device_synchronize", _st)
_proto("rf_event_create", _st, ctypes.POINTER(_vp))
_proto("rf_event_destroy", _st, _vp)
_proto("rf_event_record", _st, _vp, _vp)
_proto("rf_event_synchronize", _st, _vp)
_proto("rf_stream_wait_event", _st, _vp, _vp)
_proto("rf_event_elapsed_ms", _st, _vp, _vp, ctypes.POINTER(ctypes.c_float))
_proto("rf_fir_impl_count", _int)
_proto("rf_fir_impl_name", ctypes.c_char_p, _int)
_proto("rf_fir_impl_summary", ctypes.c_char_p, _int)
_proto("rf_fir_parse_impl", _st, ctypes.c_char_p, ctypes.POINTER(_int))
_proto("rf_fir_kernel_name", ctypes.c_char_p, _int, _int)
_proto("rf_fir", _st, _int, _vp, _vp, _vp, _int, _int, _int, _vp)
_proto("rf_fir_noalloc", _st, _int, _vp, _vp, _vp, _int, _int, _int, _vp)
_proto("rf_fir_batched", _st, _int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
       ctypes.POINTER(_int), _int, _vp, _int, _vp)
_proto("rf_magnitude", _st, _vp, _vp, _int, _vp)


class FirConfig(ctypes.Structure):
    """rf_fir_config: variant, lanes per workgroup, outputs per lane, measured ms, overlap-save
    transform length (fft_ola: 512 or 4096; 0 = the library's choice for the tap count)."""
    _fields_ = [("impl", ctypes.c_int32), ("block", ctypes.c_int32),
                ("outputs_per_lane", ctypes.c_int32), ("ms", ctypes.c_float),
                ("transform_points", ctypes.c_int32)]


_proto("rf_fir_tune", _st, _vp, _vp, _vp, _int, _int, ctypes.c_uint, _vp,
       ctypes.POINTER(FirConfig))
_proto("rf_fir_run_config", _st, ctypes.POINTER(FirConfig), _vp, _vp, _vp, _int, _int, _vp)
_proto("rf_graph_begin_capture", _st, _vp)
_proto("rf_graph_end_capture", _st, _vp, ctypes.POINTER(_vp))
_proto("rf_graph_launch", _st, _vp, _vp)
_proto("rf_graph_destroy", _st, _vp)
_proto("rf_fir_plan_create", _st, _int, _vp, _int, _int, _vp, ctypes.POINTER(_vp))
_proto("rf_fir_plan_run", _st, _vp, _vp, _vp, _int, _vp)
_proto("rf_fir_plan_destroy", _st, _vp)
_proto("rf_fir_batch_create", _st, _int, _vp, _int, _int, ctypes.POINTER(_int), _vp,
       ctypes.POINTER(_vp))
_proto("rf_fir_batch_bind", _st, _vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp))
_proto("rf_fir_batch_run", _st, _vp, _vp)
_proto("rf_fir_batch_destroy", _st, _vp)
_proto("rf_fft_plan_create", _st, _int, _int, ctypes.POINTER(_vp))
_proto("rf_fft_plan_destroy", _st, _vp)
_proto("rf_fft_plan_bytes", _st, _vp, ctypes.POINTER(_size))
_proto("rf_fft_exec", _st, _vp, _vp, _vp, _int, _vp)
_proto("rf_spectrum", _st, _vp, _vp, _vp, _vp)
_proto("rf_spectrum_magnitude", _st, _vp, _vp, _vp, _vp)
_proto("rf_channelize_frames", _int, _int, _int)
_proto("rf_channelize", _st, _vp, _vp, _vp, _vp, _int, _int, _int, _vp)
_proto("rf_channelize_mfma", _st, _vp, _vp, _vp, _int, _int, _int, _vp)
_proto("rf_xcorr_impl_count", _int)
_proto("rf_xcorr_impl_name", ctypes.c_char_p, _int)
_proto("rf_xcorr_impl_summary", ctypes.c_char_p, _int)
_proto("rf_xcorr_parse_impl", _st, ctypes.c_char_p, ctypes.POINTER(_int))
_proto("rf_xcorr_direct", _st, _int, _vp, _vp, _vp, _int, _int, _vp)
_proto("rf_xcorr_fft_create", _st, _int, _int, _int, ctypes.POINTER(_vp))
_proto("rf_xcorr_fft_destroy", _st, _vp)
_proto("rf_xcorr_fft_bytes", _st, _vp, ctypes.POINTER(_size))
_proto("rf_xcorr_fft_length", _st, _vp, ctypes.POINTER(_int))
_proto("rf_xcorr_fft_run", _st, _vp, _vp, _vp, _vp, _vp)
_proto("rf_peak", _st, _vp, _int, _vp, _vp)
_proto("rf_gcc_batch_create", _st, _int, ctypes.POINTER(_int), ctypes.POINTER(_int),
       ctypes.POINTER(_vp))
_proto("rf_gcc_batch_destroy", _st, _vp)
_proto("rf_gcc_batch_bytes", _st, _vp, ctypes.POINTER(_size))
_proto("rf_gcc_batch_bind", _st, _vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
       ctypes.POINTER(_vp))
_proto("rf_gcc_batch_run", _st, _vp, _vp, _vp)
_proto("rf_gcc_batch_launch_graph", _st, _vp, _vp, _vp)
_proto("rf_gcc_batch_tune", _st, _vp, _vp, _vp, ctypes.POINTER(_int))
_proto("rf_gcc_batch_set_group_bytes", _st, _vp, _size)
_proto("rf_gcc_batch_group_bytes", _st, _vp, ctypes.POINTER(_size), ctypes.POINTER(_int))
_ull = ctypes.c_ulonglong
_fp = ctypes.c_float
_proto("rf_synth_pairs", _st, _ull, _int, _int, _int, _vp, _vp, _vp, _vp)
_proto("rf_delay_mlp_create", _st, _int, _ull, ctypes.POINTER(_vp))
_proto("rf_delay_mlp_destroy", _st, _vp)
_proto("rf_delay_mlp_params", _st, _vp, ctypes.POINTER(_vp), ctypes.POINTER(_size))
_proto("rf_delay_mlp_forward", _st, _vp, _vp, _int, _vp, _vp)
_proto("rf_delay_mlp_gradients", _st, _vp, _vp, _vp, _int, _vp, _vp, _vp)
_proto("rf_delay_mlp_train_step", _st, _vp, _vp, _vp, _int, _fp, _vp, _vp)
_proto("rf_delay_mlp_train", _st, _vp, _ull, _int, _int, _int, _fp, _vp, _vp)
_proto("rf_delay_search_create", _st, _int, _int, ctypes.POINTER(_vp))
_proto("rf_delay_search_destroy", _st, _vp)
_proto("rf_delay_search_bytes", _st, _vp, ctypes.POINTER(_size))
_proto("rf_delay_search_features", _st, _vp, _vp, _vp, _vp, _vp, _vp)
_proto("rf_delay_search_run", _st, _vp, _vp, _vp, _vp, _fp, _vp, _vp)


def lib() -> ctypes.CDLL:
    """The loaded librfhip.so."""
    return _lib


def check(status: int) -> None:
    if status != RF_OK:
        name = _lib.rf_status_string(status).decode()
        raise RfError(status, f"{name}: {_lib.rf_last_error().decode()}")


def status_string(status: int) -> str:
    return _lib.rf_status_string(status).decode()


def last_error() -> str:
    return _lib.rf_last_error().decode()


# --- runtime ------------------------------------------------------------------------------

def device_count() -> int:
    n = _int(0)
    check(_lib.rf_device_count(ctypes.byref(n)))
    return n.value


def device_name() -> str:
    buf = ctypes.create_string_buffer(256)
    check(_lib.rf_device_name(buf, 256))
    return buf.value.decode()


def synchronize() -> None:
    check(_lib.rf_device_synchronize())


class DeviceBuffer:
    """Owning device allocation of `count` elements of `dtype` (src/device_buffer.hpp:11-76)."""

    def __init__(self, count: int, dtype=np.complex64):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        p = _vp()
        check(_lib.rf_malloc(ctypes.byref(p), self.nbytes))
        self._ptr = p.value or 0

    @classmethod
    def from_array(cls, a: np.ndarray, stream=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        buf = cls(a.size, a.dtype)
        buf.upload(a, stream)
        return buf

    @property
    def ptr(self) -> int:
        return self._ptr

    def offset(self, elements: int) -> int:
        return self._ptr + elements * self.dtype.itemsize

    def upload(self, a: np.ndarray, stream=None, offset: int = 0) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size + offset > self.count:
            raise ValueError("upload larger than the buffer")
        if a.size:
            check(_lib.rf_memcpy_h2d(self.offset(offset), a.ctypes.data, a.nbytes, _sp(stream)))
            check(_lib.rf_stream_synchronize(_sp(stream)))

    def download(self, count: int | None = None, stream=None) -> np.ndarray:
        count = self.count if count is None else count
        out = np.empty(count, dtype=self.dtype)
        if count:
            check(_lib.rf_memcpy_d2h(out.ctypes.data, self._ptr, out.nbytes, _sp(stream)))
            check(_lib.rf_stream_synchronize(_sp(stream)))
        return out

    def zero(self, stream=None) -> None:
        if self.nbytes:
            check(_lib.rf_memset(self._ptr, 0, self.nbytes, _sp(stream)))

    def free(self) -> None:
        if self._ptr:
            check(_lib.rf_free(self._ptr))
            self._ptr = 0

    def __del__(self):
        _finalize(self, "_ptr", _lib.rf_free)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()


class Stream:
    def __init__(self):
        p = _vp()
        check(_lib.rf_stream_create(ctypes.byref(p)))
        self.handle = p.value

    def synchronize(self) -> None:
        check(_lib.rf_stream_synchronize(self.handle))

    def destroy(self) -> None:
        if self.handle:
            check(_lib.rf_stream_destroy(self.handle))
            self.handle = None

    def __del__(self):
        _finalize(self, "handle", _lib.rf_stream_destroy)


def _sp(stream) -> int | None:
    if stream is None:
        return None
    return stream.handle if isinstance(stream, Stream) else stream


class Event:
    def __init__(self):
        p = _vp()
        check(_lib.rf_event_create(ctypes.byref(p)))
        self.handle = p.value

    def record(self, stream=None) -> None:
        check(_lib.rf_event_record(self.handle, _sp(stream)))

    def synchronize(self) -> None:
        check(_lib.rf_event_synchronize(self.handle))

    def ms_since(self, start: "Event") -> float:
        ms = ctypes.c_float(0.0)
        check(_lib.rf_event_elapsed_ms(start.handle, self.handle, ctypes.byref(ms)))
        return float(ms.value)

    def __del__(self):
        _finalize(self, "handle", _lib.rf_event_destroy)


# --- timing harness (src/bench.hpp:1-46, src/bench.cpp:65-93) ------------------------------

@dataclass
class PhaseTime:
    mean_ms: float = 0.0
    std_ms: float = 0.0


@dataclass
class BenchResult:
    h2d: PhaseTime = field(default_factory=PhaseTime)
    kernel: PhaseTime = field(default_factory=PhaseTime)
    d2h: PhaseTime = field(default_factory=PhaseTime)


def _summarize(samples: list[float]) -> PhaseTime:
    if not samples:
        return PhaseTime()
    mean = sum(samples) / len(samples)
    std = math.sqrt(sum((s - mean) ** 2 for s in samples) / (len(samples) - 1)) if len(samples) > 1 else 0.0
    return PhaseTime(mean, std)


class Bench:
    """`warmup` untimed passes, then `runs` passes with events around upload, compute and
    download on the default stream; mean and sample standard deviation per phase."""

    def __init__(self, warmup: int = 0, runs: int = 1):
        self.warmup = warmup
        self.runs = runs
        self.result = BenchResult()

    def run(self, upload: Callable[[], None], compute: Callable[[], None],
            download: Callable[[], None]) -> None:
        for _ in range(self.warmup):
            upload()
            compute()
            download()
        ev = [Event() for _ in range(4)]
        h2d, kern, d2h = [], [], []
        for _ in range(self.runs):
            ev[0].record()
            upload()
            ev[1].record()
            compute()
            ev[2].record()
            download()
            ev[3].record()
            ev[3].synchronize()
            h2d.append(ev[1].ms_since(ev[0]))
            kern.append(ev[2].ms_since(ev[1]))
            d2h.append(ev[3].ms_since(ev[2]))
        self.result = BenchResult(_summarize(h2d), _summarize(kern), _summarize(d2h))


# --- device-pointer operations (src/device_ops.hpp) -----------------------------------------

def fir_device(impl: int, x: int, taps: int, y: int, num_samples: int, num_taps: int,
               block: int = 0, stream=None) -> None:
    check(_lib.rf_fir(impl, x, taps, y, num_samples, num_taps, block, _sp(stream)))


def fir_batched_device(impl: int, xs: Sequence[int], ys: Sequence[int],
                       num_samples: Sequence[int], taps: int, num_taps: int,
                       stream=None) -> None:
    count = len(xs)
    xa = (_vp * max(count, 1))(*xs)
    ya = (_vp * max(count, 1))(*ys)
    na = (_int * max(count, 1))(*num_samples)
    check(_lib.rf_fir_batched(impl, xa, ya, na, count, taps, num_taps, _sp(stream)))


def magnitude_device(x: int, mag: int, num_samples: int, stream=None) -> None:
    check(_lib.rf_magnitude(x, mag, num_samples, _sp(stream)))


def xcorr_direct_device(impl: int, a: int, b: int, r: int, num_a: int, num_b: int,
                        stream=None) -> None:
    check(_lib.rf_xcorr_direct(impl, a, b, r, num_a, num_b, _sp(stream)))


def peak_device(r: int, n: int, result: int, stream=None) -> None:
    check(_lib.rf_peak(r, n, result, _sp(stream)))


def channelize_frames(num_samples: int, channels: int) -> int:
    return _lib.rf_channelize_frames(num_samples, channels)


class FftPlan:
    """`batch` transforms of `nfft` points back to back (src/fft_plan.hpp:11-37)."""

    def __init__(self, nfft: int, batch: int):
        p = _vp()
        check(_lib.rf_fft_plan_create(nfft, batch, ctypes.byref(p)))
        self.handle = p.value
        self.nfft = nfft
        self.batch = batch

    def bytes(self) -> int:
        n = _size(0)
        check(_lib.rf_fft_plan_bytes(self.handle, ctypes.byref(n)))
        return n.value

    def exec(self, src: int, dst: int, direction: int, stream=None) -> None:
        check(_lib.rf_fft_exec(self.handle, src, dst, direction, _sp(stream)))

    def destroy(self) -> None:
        if self.handle:
            check(_lib.rf_fft_plan_destroy(self.handle))
            self.handle = None

    def __del__(self):
        _finalize(self, "handle", _lib.rf_fft_plan_destroy)


def spectrum_device(plan: FftPlan, src: int, dst: int, stream=None) -> None:
    check(_lib.rf_spectrum(plan.handle, src, dst, _sp(stream)))


def channelize_device(plan: FftPlan | None, x: int, taps: int, out: int, num_samples: int,
                      channels: int, num_taps: int, stream=None) -> None:
    check(_lib.rf_channelize(plan.handle if plan else None, x, taps, out, num_samples, channels,
                             num_taps, _sp(stream)))


def channelize_mfma_device(x: int, taps: int, out: int, num_samples: int, channels: int,
                           num_taps: int, stream=None) -> None:
    """Channelizer with the DFT as a split-FP16 matrix-core GEMM (64 channels, <= 1024 taps)."""
    check(_lib.rf_channelize_mfma(x, taps, out, num_samples, channels, num_taps, _sp(stream)))


def channelize_mfma(x, taps, channels: int) -> np.ndarray:
    """Host-array form of rf_channelize_mfma; flat output, frame after frame."""
    x = _cf32(x)
    taps = _f32(taps)
    frames = channelize_frames(x.size, channels)
    if frames == 0:
        return np.zeros(0, dtype=np.complex64)
    d_x = DeviceBuffer.from_array(x)
    d_t = DeviceBuffer.from_array(taps)
    d_z = DeviceBuffer(frames * channels)
    channelize_mfma_device(d_x.ptr, d_t.ptr, d_z.ptr, x.size, channels, taps.size)
    return d_z.download()


class XcorrFft:
    """FFT cross-correlation for fixed lengths; `phat` gives GCC-PHAT
    (src/device_ops.hpp:70-88)."""

    def __init__(self, num_a: int, num_b: int, phat: bool = False):
        p = _vp()
        check(_lib.rf_xcorr_fft_create(num_a, num_b, 1 if phat else 0, ctypes.byref(p)))
        self.handle = p.value
        self.num_a = num_a
        self.num_b = num_b

    def run(self, a: int, b: int, r: int, stream=None) -> None:
        check(_lib.rf_xcorr_fft_run(self.handle, a, b, r, _sp(stream)))

    def bytes(self) -> int:
        n = _size(0)
        check(_lib.rf_xcorr_fft_bytes(self.handle, ctypes.byref(n)))
        return n.value

    def fft_length(self) -> int:
        n = _int(0)
        check(_lib.rf_xcorr_fft_length(self.handle, ctypes.byref(n)))
        return n.value

    def destroy(self) -> None:
        if self.handle:
            check(_lib.rf_xcorr_fft_destroy(self.handle))
            self.handle = None

    def __del__(self):
        _finalize(self, "handle", _lib.rf_xcorr_fft_destroy)


# --- host-vector operations (src/ops.hpp:16-52) --------------------------------------------

@dataclass(frozen=True)
class Impl:
    name: str
    summary: str
    fn: Callable


def _cf32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.complex64))


def _f32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32))


def _fir_with(impl: int):
    def run(x, taps, block: int = DEFAULT_FIR_BLOCK, bench: Bench | None = None) -> np.ndarray:
        x = _cf32(x)
        taps = _f32(taps)
        if taps.size == 0:
            raise RfError(RF_INVALID_ARG, "no taps")
        if x.size == 0:
            return np.zeros(0, dtype=np.complex64)
        bench = bench or Bench()
        d_x = DeviceBuffer(x.size)
        d_t = DeviceBuffer(taps.size, np.float32)
        d_y = DeviceBuffer(x.size)
        out: list[np.ndarray] = []
        bench.run(lambda: (d_x.upload(x), d_t.upload(taps)),
                  lambda: fir_device(impl, d_x.ptr, d_t.ptr, d_y.ptr, x.size, taps.size, block),
                  lambda: out.append(d_y.download()))
        return out[-1]
    return run


def fir_impls() -> list[Impl]:
    """The FIR variant table (src/fir.cu:725-734); the first entry is the default."""
    return [Impl(_lib.rf_fir_impl_name(i).decode(), _lib.rf_fir_impl_summary(i).decode(),
                 _fir_with(i)) for i in range(_lib.rf_fir_impl_count())]


def fir_impl_index(name: str) -> int:
    idx = _int(0)
    check(_lib.rf_fir_parse_impl(name.encode(), ctypes.byref(idx)))
    return idx.value


def fir_kernel(impl: str, num_taps: int) -> str:
    """Name of the kernel rf_fir runs for this variant and tap count (fft_ola: 512-point
    segments for 17..225 taps, 4096-point for 226..2049; the matrix-core variants: 66..2049;
    the register kernel outside those ranges)."""
    name = _lib.rf_fir_kernel_name(fir_impl_index(impl), num_taps)
    if name is None:
        raise RfError(RF_INVALID_ARG, f"no FIR kernel for {impl} with {num_taps} taps")
    return name.decode()


def fir(x, taps, impl: str = "naive", block: int = DEFAULT_FIR_BLOCK,
        bench: Bench | None = None) -> np.ndarray:
    return _fir_with(fir_impl_index(impl))(x, taps, block, bench)


def _xcorr_direct_with(impl: int):
    def run(a, b, bench: Bench | None = None) -> np.ndarray:
        a = _cf32(a)
        b = _cf32(b)
        if a.size == 0 or b.size == 0:
            raise RfError(RF_INVALID_ARG, "xcorr needs at least one sample in each input")
        bench = bench or Bench()
        d_a, d_b = DeviceBuffer(a.size), DeviceBuffer(b.size)
        d_r = DeviceBuffer(a.size + b.size - 1)
        out: list[np.ndarray] = []
        bench.run(lambda: (d_a.upload(a), d_b.upload(b)),
                  lambda: xcorr_direct_device(impl, d_a.ptr, d_b.ptr, d_r.ptr, a.size, b.size),
                  lambda: out.append(d_r.download()))
        return out[-1]
    return run


def _xcorr_fft_host(phat: bool):
    def run(a, b, bench: Bench | None = None) -> np.ndarray:
        a = _cf32(a)
        b = _cf32(b)
        if a.size == 0 or b.size == 0:
            raise RfError(RF_INVALID_ARG, "xcorr needs at least one sample in each input")
        bench = bench or Bench()
        d_a, d_b = DeviceBuffer(a.size), DeviceBuffer(b.size)
        d_r = DeviceBuffer(a.size + b.size - 1)
        h = XcorrFft(a.size, b.size, phat)
        out: list[np.ndarray] = []
        bench.run(lambda: (d_a.upload(a), d_b.upload(b)),
                  lambda: h.run(d_a.ptr, d_b.ptr, d_r.ptr),
                  lambda: out.append(d_r.download()))
        synchronize()
        h.destroy()
        return out[-1]
    return run


def xcorr_impls() -> list[Impl]:
    """The cross-correlation table (src/xcorr.cu:403-411); the first entry is the default."""
    fns = [_xcorr_direct_with(0), _xcorr_direct_with(1), _xcorr_fft_host(False)]
    return [Impl(_lib.rf_xcorr_impl_name(i).decode(), _lib.rf_xcorr_impl_summary(i).decode(),
                 fns[i]) for i in range(_lib.rf_xcorr_impl_count())]


def xcorr(a, b, impl: str = "naive", bench: Bench | None = None) -> np.ndarray:
    for entry in xcorr_impls():
        if entry.name == impl:
            return entry.fn(a, b, bench)
    raise RfError(RF_INVALID_ARG, f"unknown xcorr implementation '{impl}'")


def gcc_phat(a, b, bench: Bench | None = None) -> np.ndarray:
    """GCC-PHAT (src/xcorr.cu:379-399); same length and lag order as xcorr."""
    return _xcorr_fft_host(True)(a, b, bench)


def spectrum(x, nfft: int, bench: Bench | None = None) -> np.ndarray:
    """Framed forward FFT, last frame zero-padded (src/spectrum.cu:19-40). Flat output."""
    x = _cf32(x)
    if nfft < 1:
        raise RfError(RF_INVALID_ARG, "nfft must be at least 1")
    frames = -(-x.size // nfft)
    if frames == 0:
        return np.zeros(0, dtype=np.complex64)
    bench = bench or Bench()
    d_in = DeviceBuffer(frames * nfft)
    d_out = DeviceBuffer(frames * nfft)
    d_in.zero()
    plan = FftPlan(nfft, frames)
    out: list[np.ndarray] = []
    bench.run(lambda: d_in.upload(x),
              lambda: spectrum_device(plan, d_in.ptr, d_out.ptr),
              lambda: out.append(d_out.download()))
    synchronize()
    plan.destroy()
    return out[-1]


def spectrum_magnitude(x, nfft: int) -> np.ndarray:
    """|framed forward FFT| in one pass (rf_spectrum_magnitude). Flat float32 output."""
    x = _cf32(x)
    if nfft < 1:
        raise RfError(RF_INVALID_ARG, "nfft must be at least 1")
    frames = -(-x.size // nfft)
    if frames == 0:
        return np.zeros(0, dtype=np.float32)
    d_in = DeviceBuffer(frames * nfft)
    d_in.zero()
    d_in.upload(x)
    d_mag = DeviceBuffer(frames * nfft, np.float32)
    plan = FftPlan(nfft, frames)
    check(_lib.rf_spectrum_magnitude(plan.handle, d_in.ptr, d_mag.ptr, None))
    out = d_mag.download()
    plan.destroy()
    return out


def magnitude(x, bench: Bench | None = None) -> np.ndarray:
    x = _cf32(x)
    if x.size == 0:
        return np.zeros(0, dtype=np.float32)
    bench = bench or Bench()
    d_x = DeviceBuffer(x.size)
    d_m = DeviceBuffer(x.size, np.float32)
    out: list[np.ndarray] = []
    bench.run(lambda: d_x.upload(x),
              lambda: magnitude_device(d_x.ptr, d_m.ptr, x.size),
              lambda: out.append(d_m.download()))
    return out[-1]


def channelize(x, taps, channels: int, bench: Bench | None = None,
               use_plan: bool = False) -> np.ndarray:
    """Polyphase channelizer (src/channelize.cu:74-105). Flat output, frame after frame."""
    if channels < 1:
        raise RfError(RF_INVALID_ARG, "channelize: need at least one channel")
    x = _cf32(x)
    taps = _f32(taps)
    frames = channelize_frames(x.size, channels)
    if frames == 0:
        return np.zeros(0, dtype=np.complex64)
    bench = bench or Bench()
    d_x = DeviceBuffer(x.size)
    d_t = DeviceBuffer(taps.size, np.float32)
    d_z = DeviceBuffer(frames * channels)
    plan = FftPlan(channels, frames) if use_plan else None
    out: list[np.ndarray] = []
    bench.run(lambda: (d_x.upload(x), d_t.upload(taps)),
              lambda: channelize_device(plan, d_x.ptr, d_t.ptr, d_z.ptr, x.size, channels,
                                        taps.size),
              lambda: out.append(d_z.download()))
    if plan is not None:
        synchronize()
        plan.destroy()
    return out[-1]


class GccBatch:
    """Batched GCC-PHAT + peak search over uneven pairs (rf_gcc_batch_*)."""

    def __init__(self, num_a: Sequence[int], num_b: Sequence[int]):
        count = len(num_a)
        p = _vp()
        na = (_int * count)(*num_a)
        nb = (_int * count)(*num_b)
        check(_lib.rf_gcc_batch_create(count, na, nb, ctypes.byref(p)))
        self.handle = p.value
        self.count = count
        self.num_a = list(num_a)
        self.num_b = list(num_b)

    def bind(self, a: Sequence[int], b: Sequence[int], r: Sequence[int] | None = None) -> None:
        aa = (_vp * self.count)(*a)
        bb = (_vp * self.count)(*b)
        rr = (_vp * self.count)(*r) if r is not None else None
        check(_lib.rf_gcc_batch_bind(self.handle, aa, bb, rr))

    def run(self, results: int, stream=None, graph: bool = False) -> None:
        fn = _lib.rf_gcc_batch_launch_graph if graph else _lib.rf_gcc_batch_run
        check(fn(self.handle, results, _sp(stream)))

    def tune(self, results: int, stream=None) -> int:
        """Picks the persistent grid size (workgroups per CU) by timing; returns it."""
        best = _int(0)
        check(_lib.rf_gcc_batch_tune(self.handle, results, _sp(stream), ctypes.byref(best)))
        return best.value

    def set_group_bytes(self, nbytes: int) -> None:
        check(_lib.rf_gcc_batch_set_group_bytes(self.handle, nbytes))

    def group_bytes(self) -> tuple[int, int]:
        """(budget in bytes, number of groups) of the current schedule."""
        n = _size(0)
        g = _int(0)
        check(_lib.rf_gcc_batch_group_bytes(self.handle, ctypes.byref(n), ctypes.byref(g)))
        return n.value, g.value

    def bytes(self) -> int:
        n = _size(0)
        check(_lib.rf_gcc_batch_bytes(self.handle, ctypes.byref(n)))
        return n.value

    def destroy(self) -> None:
        if self.handle:
            check(_lib.rf_gcc_batch_destroy(self.handle))
            self.handle = None

    def __del__(self):
        _finalize(self, "handle", _lib.rf_gcc_batch_destroy)


def fir_tune(x: int, taps: int, y: int, n: int, num_taps: int, impls: Sequence[str],
             stream=None) -> FirConfig:
    """Autotunes the FIR variants named in `impls` on device buffers; returns the fastest."""
    mask = 0
    for name in impls:
        mask |= 1 << fir_impl_index(name)
    best = FirConfig()
    check(_lib.rf_fir_tune(x, taps, y, n, num_taps, mask, _sp(stream), ctypes.byref(best)))
    return best


def fir_run_config(config: FirConfig, x: int, taps: int, y: int, n: int, num_taps: int,
                   stream=None) -> None:
    check(_lib.rf_fir_run_config(ctypes.byref(config), x, taps, y, n, num_taps, _sp(stream)))


class Graph:
    """Work enqueued by `record()` on a created stream, captured once and replayed by launch()."""

    def __init__(self, stream: "Stream", record):
        check(_lib.rf_graph_begin_capture(stream.handle))
        try:
            record()
        finally:
            p = _vp()
            status = _lib.rf_graph_end_capture(stream.handle, ctypes.byref(p))
        check(status)
        self.handle = p.value

    def launch(self, stream=None) -> None:
        check(_lib.rf_graph_launch(self.handle, _sp(stream)))

    def destroy(self) -> None:
        if self.handle:
            check(_lib.rf_graph_destroy(self.handle))
            self.handle = None

    def __del__(self):
        _finalize(self, "handle", _lib.rf_graph_destroy)


class FirPlan:
    """A filter prepared once (device taps copied; fft_ola: the tap spectrum made) for many
    rf_fir_plan_run calls, each the same bits as rf_fir with that variant and block."""

    def __init__(self, impl: str | int, taps: int, num_taps: int, block: int = 0, stream=None):
        idx = fir_impl_index(impl) if isinstance(impl, str) else impl
        p = _vp()
        check(_lib.rf_fir_plan_create(idx, taps, num_taps, block, _sp(stream), ctypes.byref(p)))
        self.handle = p.value
        self.num_taps = num_taps

    def run(self, x: int, y: int, n: int, stream=None) -> None:
        check(_lib.rf_fir_plan_run(self.handle, x, y, n, _sp(stream)))

    def destroy(self) -> None:
        if self.handle:
            check(_lib.rf_fir_plan_destroy(self.handle))
            self.handle = None

    def __del__(self):
        _finalize(self, "handle", _lib.rf_fir_plan_destroy)


class FirBatch:
    """Persistent FIR over many uneven captures sharing one filter (rf_fir_batch_*): one launch
    per run, each capture's output the bits of rf_fir with the same impl."""

    def __init__(self, impl: str | int, taps: int, num_taps: int, lengths: Sequence[int],
                 stream=None):
        idx = fir_impl_index(impl) if isinstance(impl, str) else impl
        n = (_int * len(lengths))(*lengths)
        p = _vp()
        check(_lib.rf_fir_batch_create(idx, taps, num_taps, len(lengths), n, _sp(stream),
                                       ctypes.byref(p)))
        self.handle = p.value
        self.count = len(lengths)

    def bind(self, x: Sequence[int], y: Sequence[int]) -> None:
        check(_lib.rf_fir_batch_bind(self.handle, (_vp * self.count)(*x),
                                     (_vp * self.count)(*y)))

    def run(self, stream=None) -> None:
        check(_lib.rf_fir_batch_run(self.handle, _sp(stream)))

    def destroy(self) -> None:
        if self.handle:
            check(_lib.rf_fir_batch_destroy(self.handle))
            self.handle = None

    def __del__(self):
        _finalize(self, "handle", _lib.rf_fir_batch_destroy)


DELAY_FEATURES = 256
DELAY_BINS = 64
DELAY_SPAN = 512
DELAY_MLP_PARAMS = 148032


class DelayResult(ctypes.Structure):
    """rf_delay_result: peak (index, magnitude, frac) plus proposal, confidence and path."""
    _fields_ = [("index", ctypes.c_int32), ("magnitude", ctypes.c_float),
                ("frac", ctypes.c_double), ("proposal", ctypes.c_int32),
                ("confidence", ctypes.c_float), ("path", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


def synth_pairs(seed: int, first: int, count: int, n: int, a: int, b: int,
                delays: int | None = None, stream=None) -> None:
    """Synthetic capture pairs into device buffers a, b (count * n each) and delays."""
    check(_lib.rf_synth_pairs(seed, first, count, n, a, b, delays, _sp(stream)))


class DelayMlp:
    """Delay-proposal network (rf_delay_mlp_*): 256-256-256-64, GELU, BF16 GEMM operands."""

    def __init__(self, max_batch: int, seed: int = 0xC5A):
        p = _vp()
        check(_lib.rf_delay_mlp_create(max_batch, seed, ctypes.byref(p)))
        self.handle = p.value
        self.max_batch = max_batch

    def params_ptr(self) -> int:
        p = _vp()
        n = _size(0)
        check(_lib.rf_delay_mlp_params(self.handle, ctypes.byref(p), ctypes.byref(n)))
        return p.value

    def get_params(self) -> np.ndarray:
        out = np.empty(DELAY_MLP_PARAMS, dtype=np.float32)
        check(_lib.rf_memcpy_d2h(out.ctypes.data, self.params_ptr(), out.nbytes, None))
        check(_lib.rf_stream_synchronize(None))
        return out

    def set_params(self, params: np.ndarray) -> None:
        params = np.ascontiguousarray(params, dtype=np.float32)
        assert params.size == DELAY_MLP_PARAMS
        check(_lib.rf_memcpy_h2d(self.params_ptr(), params.ctypes.data, params.nbytes, None))
        check(_lib.rf_stream_synchronize(None))

    def forward(self, features: int, batch: int, probs: int, stream=None) -> None:
        check(_lib.rf_delay_mlp_forward(self.handle, features, batch, probs, _sp(stream)))

    def gradients(self, features: int, labels: int, batch: int, grads: int | None, loss: int,
                  stream=None) -> None:
        check(_lib.rf_delay_mlp_gradients(self.handle, features, labels, batch, grads, loss,
                                          _sp(stream)))

    def train_step(self, features: int, labels: int, batch: int, lr: float, loss: int,
                   stream=None) -> None:
        check(_lib.rf_delay_mlp_train_step(self.handle, features, labels, batch, lr, loss,
                                           _sp(stream)))

    def train(self, seed: int, pairs: int, batch: int, n: int, lr: float,
              stream=None) -> np.ndarray:
        """Trains on `pairs` synthesised pairs; returns the per-step mean losses."""
        steps = -(-pairs // batch)
        losses = np.zeros(steps, dtype=np.float32)
        check(_lib.rf_delay_mlp_train(self.handle, seed, pairs, batch, n, lr,
                                      losses.ctypes.data, _sp(stream)))
        return losses

    def destroy(self) -> None:
        if self.handle:
            check(_lib.rf_delay_mlp_destroy(self.handle))
            self.handle = None

    def __del__(self):
        _finalize(self, "handle", _lib.rf_delay_mlp_destroy)


class DelaySearch:
    """Learned delay search over `count` pairs of n samples (rf_delay_search_*)."""

    def __init__(self, count: int, n: int):
        p = _vp()
        check(_lib.rf_delay_search_create(count, n, ctypes.byref(p)))
        self.handle = p.value
        self.count = count
        self.n = n

    def features(self, a: int, b: int, features: int | None, coarse: int | None,
                 stream=None) -> None:
        check(_lib.rf_delay_search_features(self.handle, a, b, features, coarse, _sp(stream)))

    def run(self, a: int, b: int, mlp: "DelayMlp | None", threshold: float, results: int,
            stream=None) -> None:
        check(_lib.rf_delay_search_run(self.handle, a, b, mlp.handle if mlp else None,
                                       threshold, results, _sp(stream)))

    def bytes(self) -> int:
        n = _size(0)
        check(_lib.rf_delay_search_bytes(self.handle, ctypes.byref(n)))
        return n.value

    def destroy(self) -> None:
        if self.handle:
            check(_lib.rf_delay_search_destroy(self.handle))
            self.handle = None

    def __del__(self):
        _finalize(self, "handle", _lib.rf_delay_search_destroy)


def read_delay_results(d_results: "DeviceBuffer", count: int) -> list[DelayResult]:
    raw = d_results.download().tobytes()
    size = ctypes.sizeof(DelayResult)
    return [DelayResult.from_buffer_copy(raw[i * size:(i + 1) * size]) for i in range(count)]


def read_peaks(d_results: "DeviceBuffer", count: int) -> list[PeakResult]:
    raw = d_results.download().tobytes()
    size = ctypes.sizeof(PeakResult)
    return [PeakResult.from_buffer_copy(raw[i * size:(i + 1) * size]) for i in range(count)]


def peak(r_device: int, n: int) -> PeakResult:
    """Device peak search over n values; returns the host copy of the result."""
    d_res = DeviceBuffer(ctypes.sizeof(PeakResult), np.uint8)
    peak_device(r_device, n, d_res.ptr)
    raw = d_res.download()
    return PeakResult.from_buffer_copy(raw.tobytes())


def tdoa(a, b) -> int:
    """Delay of b behind a: the lag of the GCC-PHAT peak, found on the device."""
    a = _cf32(a)
    b = _cf32(b)
    d_a, d_b = DeviceBuffer.from_array(a), DeviceBuffer.from_array(b)
    d_r = DeviceBuffer(a.size + b.size - 1)
    h = XcorrFft(a.size, b.size, True)
    h.run(d_a.ptr, d_b.ptr, d_r.ptr)
    res = peak(d_r.ptr, a.size + b.size - 1)
    h.destroy()
    return int(res.index) - (a.size - 1)


def tdoa_batch(channels: Sequence[np.ndarray]) -> list[int]:
    """Lag of every channel 1.. behind channel 0 from one batched GCC-PHAT with the peak search
    fused (rf_gcc_batch); pairs too short for the batch (at most 2048 output lags) use tdoa."""
    ref = _cf32(channels[0])
    chans = [_cf32(c) for c in channels[1:]]
    lags = [0] * len(chans)
    batched = [i for i, c in enumerate(chans) if ref.size + c.size - 1 > 2048]
    for i, c in enumerate(chans):
        if i not in batched:
            lags[i] = tdoa(ref, c)
    if batched:
        d_ref = DeviceBuffer.from_array(ref)
        d_ch = [DeviceBuffer.from_array(chans[i]) for i in batched]
        h = GccBatch([ref.size] * len(batched), [chans[i].size for i in batched])
        h.bind([d_ref.ptr] * len(batched), [d.ptr for d in d_ch])
        res = DeviceBuffer(16 * len(batched), np.uint8)
        h.run(res.ptr)
        for i, p in zip(batched, read_peaks(res, len(batched))):
            lags[i] = int(p.index) - (ref.size - 1)
        h.destroy()
        for b in [d_ref, res, *d_ch]:
            b.free()
    return lags


def doa(channels: Sequence[np.ndarray], spacing: float, speed: float = 299792458.0,
        rate: float = 1.0) -> dict:
    """Direction of arrival for a uniform linear array (src/main.cpp:352-416): one batched
    GCC-PHAT of every channel against channel 0 with the peak searched on the GPU, then the
    least-squares angle on the host."""
    if len(channels) < 2:
        raise RfError(RF_INVALID_ARG, "doa needs at least two channels")
    lags = tdoa_batch(channels)
    sum_it = sum(i * lag / rate for i, lag in enumerate(lags, start=1))
    sum_ii = sum(float(i * i) for i in range(1, len(channels)))
    sine = max(-1.0, min(1.0, speed * (sum_it / sum_ii) / spacing))
    return {"tdoa_samples": lags, "angle_deg": math.degrees(math.asin(sine))}
