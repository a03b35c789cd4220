#!/usr/bin/env python3
"""Throughput of the hot path on one MI355X, as one JSON line.

Default workload (the metric's long-FIR configuration): 2^24 complex64 samples (complex
Gaussian noise of unit power plus 5 tones of amplitude sqrt(10) at random frequencies and
phases, seeded), filtered by a 1025-tap Kaiser (beta 8) low-pass with cutoff 0.1 fs and unit DC
gain. One step is one rf_fir call over the whole capture with inputs already in HBM.

  python bench.py [--steps K] [--warmup W] [--impl fft_ola] [--workload fir_long|fir_plain|
                   fir_short|channelize|gcc_batch|delay|copy|sigflow|stream] [--gpus 1]

fir_plain is 2^20 samples x 63-tap Hamming low-pass 0.2 fs (the CPU-runnable FIR), fir_short the
same filter over 2^24 samples. Input/output buffers rotate over enough sets (at least 320 MiB:
three at 2^24 samples, 20 at 2^20) that the 256 MB Infinity Cache cannot hold a step's input
from the previous visit. Timing: device events on the stream the kernels run on, bracketing all
K steps, after a device synchronisation. The roofline figure uses the algorithmic bytes of one
step (input + output + taps) over the time of the dominant kernel alone, measured the same way
over K launches of a prepared filter (rf_fir_plan_run, the same bits as rf_fir without the
per-call tap transform). The CPU baseline is the multithreaded FP32 CPU path
(oracle/cpu_fast.cpp) on the whole workload; the float64 oracle's own time is reported beside
it and its output is the checker.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector, published
METRIC = "Msamples/s + %HBM roofline (FIR, channelizer); GCC-PHAT pairs/s at fixed error"


def load_rfhip():
    pkg = ROOT / "gpu-rf-pipeline_amd"
    if not (pkg / "librfhip.so").exists():
        import subprocess
        subprocess.run(["make", "-s", "-C", str(pkg), "-j8", "librfhip.so"], check=True)
    spec = importlib.util.spec_from_file_location("rfhip", pkg / "__init__.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rfhip"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_oracle():
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle  # the float64 oracle: result checks and the CPU baseline
    return oracle


def load_cpu_fast():
    sys.path.insert(0, str(ROOT / "oracle"))
    import cpu_fast  # the multithreaded FP32 CPU path: the CPU baseline
    return cpu_fast


def cpu_host() -> str:
    """The host CPU's model name, for the baseline's record."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def kaiser_lowpass(num_taps: int, cutoff: float, beta: float) -> np.ndarray:
    n = np.arange(num_taps) - (num_taps - 1) / 2
    h = 2 * cutoff * np.sinc(2 * cutoff * n) * np.kaiser(num_taps, beta)
    return (h / h.sum()).astype(np.float32)


def hamming_lowpass(num_taps: int, cutoff: float) -> np.ndarray:
    n = np.arange(num_taps) - (num_taps - 1) / 2
    h = 2 * cutoff * np.sinc(2 * cutoff * n) * np.hamming(num_taps)
    return (h / h.sum()).astype(np.float32)


def fir_long_input(n: int = 1 << 24) -> tuple[np.ndarray, np.ndarray]:
    """Complex Gaussian noise (I, Q ~ N(0, 1/2)) plus 5 tones at 10 dB SNR, seed 0xC2."""
    rng = np.random.Generator(np.random.Philox(0xC2))
    x = (rng.standard_normal(n, dtype=np.float32)
         + 1j * rng.standard_normal(n, dtype=np.float32)) * np.float32(np.sqrt(0.5))
    t = np.arange(n, dtype=np.float64)
    freqs = rng.uniform(-0.5, 0.5, 5)
    phases = rng.uniform(0, 2 * np.pi, 5)
    tones = np.zeros(n, dtype=np.complex128)
    for f, p in zip(freqs, phases):
        tones += np.sqrt(10.0) * np.exp(1j * (2 * np.pi * f * t + p))
    return (x + tones).astype(np.complex64), kaiser_lowpass(1025, 0.1, 8.0)


def fir_plain_input(n: int = 1 << 20) -> tuple[np.ndarray, np.ndarray]:
    """I, Q ~ N(0, 1) plus unit tones at 0.05, 0.17, 0.31 fs; 63-tap Hamming sinc 0.2 fs."""
    rng = np.random.Generator(np.random.Philox(0xC1))
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    t = np.arange(n)
    for f in (0.05, 0.17, 0.31):
        x = x + np.exp(2j * np.pi * f * t)
    return x.astype(np.complex64), hamming_lowpass(63, 0.2)


def sampled_fir_error(x: np.ndarray, h: np.ndarray, y: np.ndarray, count: int = 2000):
    """Max abs / rel error of `count` outputs against float64 dot products."""
    rng = np.random.default_rng(11)
    idx = np.concatenate([np.arange(min(h.size + 8, x.size)), rng.integers(0, x.size, count)])
    k = np.arange(h.size)
    src = idx[:, None] - k[None, :]
    xv = np.where(src >= 0, x[np.clip(src, 0, None)].astype(np.complex128), 0)
    want = xv @ h.astype(np.float64)
    err = np.abs(y[idx].astype(np.complex128) - want)
    return float(err.max()), float((err / np.maximum(np.abs(want), 1e-30)).max()), \
        float((err - 1e-5 * np.abs(want)).max())


def release(rf, *objs) -> None:
    """Frees device buffers and destroys plans, graphs and streams explicitly, after the device
    has finished with them, so that nothing is left for finalizers at interpreter exit."""
    rf.synchronize()
    for o in objs:
        for item in (o if isinstance(o, (list, tuple)) else [o]):
            if item is None:
                continue
            for method in ("free", "destroy"):
                fn = getattr(item, method, None)
                if fn is not None:
                    fn()
                    break


def fir_errors(y: np.ndarray, want: np.ndarray) -> dict:
    """Error of every output against the float64 oracle output, with the reference's FIR
    tolerance (rtol = atol = 1e-5, tests/test_pipeline.py:113-117)."""
    err = np.abs(y.astype(np.complex128) - want)
    mag = np.abs(want)
    return {"max_abs_err": float(err.max()),
            "max_rel_err": float((err / np.maximum(mag, 1e-30)).max()),
            "within_rtol_atol_1e-5": bool(np.all(err <= 1e-5 + 1e-5 * mag)),
            "checked_outputs": f"all {y.size} outputs vs the float64 oracle"}


def time_steps(rf, step, steps: int, warmup: int, stream) -> float:
    for i in range(warmup):
        step(i)
    rf.synchronize()
    start, stop = rf.Event(), rf.Event()
    start.record(stream)
    for i in range(steps):
        step(warmup + i)
    stop.record(stream)
    stop.synchronize()
    rf.synchronize()
    return stop.ms_since(start)


def bench_fir(rf, args) -> dict:
    long_filter = args.workload == "fir_long"
    if long_filter:
        x, h = fir_long_input()
    else:
        x, h = fir_plain_input(1 << 24 if args.workload == "fir_short" else 1 << 20)
    n, t = x.size, h.size
    # Enough rotating input/output sets that a step's buffers were evicted from the 256 MB
    # Infinity Cache since their last visit: at least 320 MiB in all (3 sets at 2^24 samples,
    # 20 at 2^20).
    sets = max(3, -(-(320 << 20) // (16 * n)))
    stream = rf.Stream()
    d_h = rf.DeviceBuffer.from_array(h)
    d_x = [rf.DeviceBuffer.from_array(x) for _ in range(sets)]
    d_y = [rf.DeviceBuffer(n) for _ in range(sets)]
    impl = rf.fir_impl_index(args.impl)

    def step_for(impl_idx):
        def step(i):
            rf.fir_device(impl_idx, d_x[i % sets].ptr, d_h.ptr, d_y[i % sets].ptr, n, t, 0,
                          stream.handle)
        return step

    ms = time_steps(rf, step_for(impl), args.steps, args.warmup, stream.handle)
    ms_step = ms / args.steps
    y = d_y[(args.warmup + args.steps - 1) % sets].download()
    max_abs, max_rel, excess = sampled_fir_error(x, h, y)
    accuracy = {"max_abs_err": max_abs, "max_rel_err": max_rel,
                "within_rtol_atol_1e-5": excess <= 1e-5,
                "checked_outputs": "first taps+8 and 2000 random outputs vs float64"}
    bytes_step = 16 * n + 4 * t

    # The filtering kernel alone, timed the same way: a prepared filter (taps copied, their
    # spectrum made once) launches only that kernel per call; its bits equal rf_fir's.
    plan = rf.FirPlan(impl, d_h.ptr, t)

    def plan_step(i):
        plan.run(d_x[i % sets].ptr, d_y[i % sets].ptr, n, stream.handle)

    kernel_ms = time_steps(rf, plan_step, args.steps, args.warmup, stream.handle) / args.steps
    graph = None
    if n < 1 << 24:
        # A short capture is a few microseconds of GPU work per call, as long as a launch from
        # the host: the same K calls captured once into a HIP graph and replayed show the
        # GPU-side rate without per-call host gaps.
        g = rf.Graph(stream, lambda: [plan_step(i) for i in range(args.steps)])
        g.launch(stream)
        stream.synchronize()
        gms = time_steps(rf, lambda i: g.launch(stream), 1, 0, stream.handle) / args.steps
        g.destroy()
        graph = {"ms_per_step": round(gms, 5), "msamples_per_s": round(n / gms / 1e3, 1),
                 "hbm_frac": round(bytes_step / (gms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                 "calls_per_graph": args.steps}
    plan.destroy()
    # The timed one-shot calls themselves (tap transform and filtering kernel per call, as
    # `value`), K of them captured once into a HIP graph and replayed: the rate without the
    # host's launch gaps between the calls. Reported beside `value`, which stays the eager rate.
    g = rf.Graph(stream, lambda: [step_for(impl)(i) for i in range(args.steps)])
    g.launch(stream)
    stream.synchronize()
    gms = time_steps(rf, lambda i: g.launch(stream), 1, 0, stream.handle) / args.steps
    g.destroy()
    graph_one_shot = {"ms_per_step": round(gms, 5), "msamples_per_s": round(n / gms / 1e3, 1),
                      "calls_per_graph": args.steps}
    achieved = bytes_step / (kernel_ms * 1e-3) / 1e9

    # Arithmetic: the direct form costs 4 flop per tap and output (complex x real MAC); the
    # overlap-save form two 4096-point FFTs (5 P log2 P each) and 4096 complex products (6 flop)
    # per segment of 4097 - T outputs (512-point transforms and 513 - T outputs for 17..225 taps).
    direct_flop = 4 * n * t
    if args.impl == "fft_ola" and 225 < t <= 2049:
        segments = -(-n // (4097 - t))
        algo_flop = segments * (2 * 5 * 4096 * 12 + 6 * 4096)
    elif args.impl == "fft_ola" and 17 <= t <= 225:
        segments = -(-n // (513 - t))   # the 512-point form for short filters
        algo_flop = segments * (2 * 5 * 512 * 9 + 6 * 512)
    else:
        algo_flop = direct_flop
    compute = {"algorithm_flop_per_step": algo_flop, "direct_form_flop_per_step": direct_flop,
               "achieved_tflops": round(algo_flop / (kernel_ms * 1e-3) / 1e12, 2),
               "direct_form_equivalent_tflops": round(direct_flop / (kernel_ms * 1e-3) / 1e12, 1),
               "fp32_peak_tflops": FP32_PEAK_TFLOPS,
               "frac": round(algo_flop / (kernel_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4)}

    want = None
    cpu = None
    cpu_oracle = None
    if not args.no_cpu:
        oracle = load_oracle()
        t0 = time.perf_counter()
        want = oracle.fir(x, h)  # the whole capture: the checker, and a second CPU figure
        dt = time.perf_counter() - t0
        cpu_oracle = {"value": round(n / dt / 1e6, 4), "unit": "Msamples/s", "cores": 1,
                      "sample": f"float64 oracle (oracle/oracle.c orc_fir), one thread, all {n} "
                                f"samples x {t} taps, {dt:.2f} s"}
        accuracy = fir_errors(y, want)
        fast = load_cpu_fast()
        t0 = time.perf_counter()
        y_cpu = fast.fir(x, h)
        dt = time.perf_counter() - t0
        cpu = {"value": round(n / dt / 1e6, 4), "unit": "Msamples/s", "cores": fast.threads(),
               "kind": "port",
               "sample": f"FP32 direct FIR, {fast.threads()} threads (oracle/cpu_fast.cpp, AVX2), "
                         f"all {n} samples x {t} taps, {dt:.2f} s on {cpu_host()}",
               "max_abs_err_vs_oracle": float(np.abs(y_cpu.astype(np.complex128) - want).max())}

    variants = {}
    for name in [v.name for v in rf.fir_impls()]:
        if args.no_variants or name == args.impl or name == "naive" and long_filter:
            continue
        reps = max(2, args.steps // 5)
        vms = time_steps(rf, step_for(rf.fir_impl_index(name)), reps, 1, stream.handle) / reps
        vy = d_y[(1 + reps - 1) % sets].download()
        if want is not None:
            verr = fir_errors(vy, want)
            v_abs, v_rel = verr["max_abs_err"], verr["max_rel_err"]
        else:
            v_abs, v_rel, _ = sampled_fir_error(x, h, vy)
        variants[name] = {"kernel": rf.fir_kernel(name, t),
                          "ms_per_step": round(vms, 4), "msamples_per_s": round(n / vms / 1e3, 1),
                          "max_abs_err": v_abs, "max_rel_err": v_rel}
    tuned = None
    if not args.no_variants:
        # The library's autotuner over the variants that suit this filter (tile shapes of the
        # register kernel, fft_ola's transform length, the split-precision GEMMs), timed on
        # the same buffers: what it picks and its own timing of that configuration (one buffer
        # set, so partly resident in the Infinity Cache: faster than the rotating-set figures).
        names = ["register", "fft_ola", "mfma_f16x3", "mfma_bf16x3"]
        best = rf.fir_tune(d_x[0].ptr, d_h.ptr, d_y[0].ptr, n, t, names, stream.handle)
        tuned = {"impl": rf.fir_impls()[best.impl].name, "block": best.block,
                 "outputs_per_lane": best.outputs_per_lane,
                 "transform_points": best.transform_points, "ms": round(best.ms, 5),
                 "candidates": names,
                 "note": "the tuner's own timing on one buffer set (partly resident in the "
                         "256 MB Infinity Cache), not comparable with ms_per_step"}
    release(rf, d_h, d_x, d_y, stream)
    workload = (f"FIR {n} complex64 samples x {t}-tap "
                + ("Kaiser(beta=8) lowpass 0.1 fs" if long_filter else "Hamming lowpass 0.2 fs"))
    return {
        "metric": METRIC,
        "value": round(n / ms_step / 1e3, 1),
        "unit": "Msamples/s",
        "ms_per_step": round(ms_step, 5),
        "dtype": "f32",
        "config": {"workload": workload, "impl": args.impl, "kernel": rf.fir_kernel(args.impl, t),
                   "samples": n, "taps": t,
                   "buffer_sets": sets, "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(
                         f"fir_{args.impl}" if long_filter else f"{args.workload}_{args.impl}")},
        "roofline_kernel": {"kernel": rf.fir_kernel(args.impl, t),
                            "launches": "rf_fir_plan_run: the filtering kernel only, the kernel "
                                        "the timed rf_fir step runs after its tap transform",
                            "ms_per_launch": round(kernel_ms, 5),
                            "msamples_per_s": round(n / kernel_ms / 1e3, 1)},
        "step_roofline": {"kernels": "whole rf_fir call (tap transform + filtering kernel)",
                          "achieved": round(bytes_step / (ms_step * 1e-3) / 1e9, 1),
                          "frac": round(bytes_step / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "compute": compute,
        "graph_replay": graph,
        "graph_replay_one_shot": graph_one_shot,
        "accuracy": accuracy,
        "variants": variants,
        "tuned": tuned,
        "cpu_baseline": cpu,
        "cpu_oracle": cpu_oracle,
    }


def channelize_input(n: int = 1 << 24, seed: int = 0xC3):
    """40 QPSK carriers at distinct channel centres k fs / 64, powers U[-30, 0] dB, symbols
    (+-1 +-1j)/sqrt(2) held for 128 samples, plus complex AWGN at -40 dB (power 1e-4)."""
    rng = np.random.Generator(np.random.Philox(seed))
    centres = rng.choice(64, 40, replace=False)
    amps = 10.0 ** (rng.uniform(-30.0, 0.0, 40) / 20.0)
    hold = 128
    blocks = n // hold
    sym = ((rng.choice([-1.0, 1.0], (blocks, 40)) + 1j * rng.choice([-1.0, 1.0], (blocks, 40)))
           / np.sqrt(2.0))
    u = np.arange(hold)
    phasors = amps[:, None] * np.exp(2j * np.pi * np.outer(centres, u) / 64.0)  # 40 x 128
    x = (sym @ phasors).reshape(-1)                                               # n samples
    noise = (rng.standard_normal(n, dtype=np.float32)
             + 1j * rng.standard_normal(n, dtype=np.float32)) * np.float32(np.sqrt(0.5e-4))
    return (x + noise).astype(np.complex64)


def channelize_errors(got: np.ndarray, x: np.ndarray, h: np.ndarray, m: int,
                      frames: np.ndarray) -> dict:
    """Max abs error and max error / sum_l |h x| over the given frames, against float64."""
    xs = x.astype(np.complex128)
    l = np.arange(h.size)
    k = np.arange(m)
    dft = np.exp(2j * np.pi * np.outer(l, k) / m)
    worst_abs, worst_rel = 0.0, 0.0
    for f in frames:
        idx = f * m - l
        ok = idx >= 0
        terms = np.zeros(h.size, np.complex128)
        terms[ok] = h[ok] * xs[idx[ok]]
        want = terms @ dft
        err = np.max(np.abs(got[f] - want))
        worst_abs = max(worst_abs, float(err))
        worst_rel = max(worst_rel, float(err / np.sum(np.abs(terms))))
    return {"max_abs_err": worst_abs, "max_err_over_term_sum": worst_rel,
            "frames_checked": int(len(frames))}


def bench_channelize(rf, args) -> dict:
    n, m, t = 1 << 24, 64, 1024
    x = channelize_input(n)
    h = kaiser_lowpass(t, 1.0 / 128, 8.0)
    frames = rf.channelize_frames(n, m)
    sets = 3
    stream = rf.Stream()
    d_h = rf.DeviceBuffer.from_array(h)
    d_x = [rf.DeviceBuffer.from_array(x) for _ in range(sets)]
    d_z = [rf.DeviceBuffer(frames * m) for _ in range(sets)]

    def step(i):
        rf.channelize_device(None, d_x[i % sets].ptr, d_h.ptr, d_z[i % sets].ptr, n, m, t,
                             stream.handle)

    def step_mfma(i):
        rf.channelize_mfma_device(d_x[i % sets].ptr, d_h.ptr, d_z[i % sets].ptr, n, m, t,
                                  stream.handle)

    ms = time_steps(rf, step, args.steps, args.warmup, stream.handle)
    ms_step = ms / args.steps
    check_frames = np.sort(np.random.default_rng(5).choice(frames, 48, replace=False))
    z_fused = d_z[(args.warmup + args.steps - 1) % sets].download().reshape(-1, m)
    err_fused = channelize_errors(z_fused, x, h, m, check_frames)
    ms_mfma = time_steps(rf, step_mfma, args.steps, args.warmup, stream.handle) / args.steps
    z_mfma = d_z[(args.warmup + args.steps - 1) % sets].download().reshape(-1, m)
    err_mfma = channelize_errors(z_mfma, x, h, m, check_frames)
    release(rf, d_h, d_x, d_z, stream)
    bytes_step = 16 * n + 4 * t
    achieved = bytes_step / (ms_step * 1e-3) / 1e9
    cpu = None
    cpu_oracle = None
    if not args.no_cpu:
        oracle = load_oracle()
        sample = 1 << 22
        t0 = time.perf_counter()
        oracle.channelize(x[:sample], h, m)
        dt = time.perf_counter() - t0
        cpu_oracle = {"value": round(sample / dt / 1e6, 4), "unit": "Msamples/s", "cores": 1,
                      "sample": f"float64 oracle, one thread, first {sample} samples"}
        fast = load_cpu_fast()
        t0 = time.perf_counter()
        fast.channelize(x, h, m)
        dt = time.perf_counter() - t0
        cpu = {"value": round(n / dt / 1e6, 4), "unit": "Msamples/s", "cores": fast.threads(),
               "kind": "port",
               "sample": f"FP32 branch sums + radix-2 FFT, {fast.threads()} threads "
                         f"(oracle/cpu_fast.cpp), all {n} samples, {dt:.2f} s on {cpu_host()}"}
    return {
        "metric": METRIC, "value": round(n / ms_step / 1e3, 1), "unit": "Msamples/s",
        "ms_per_step": round(ms_step, 5), "dtype": "f32",
        "config": {"workload": f"polyphase channelizer {n} complex64 samples (40 QPSK carriers, "
                               f"-30..0 dB, AWGN -40 dB), M={m}, {t}-tap Kaiser prototype",
                   "samples": n, "channels": m, "taps": t, "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles("channelize_fused")},
        "accuracy": err_fused,
        "mfma_f16x3": {"ms_per_step": round(ms_mfma, 5),
                       "msamples_per_s": round(n / ms_mfma / 1e3, 1),
                       "hbm_frac": round(bytes_step / (ms_mfma * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       **err_mfma},
        "cpu_baseline": cpu,
        "cpu_oracle": cpu_oracle,
    }


def gcc_batch_input(count: int = 256, seed: int = 0xC4):
    """Capture pairs: N log-uniform in [2^12, 2^20]; a complex white Gaussian; b = a delayed by
    d ~ U[-200, 200] (integer part by indexing, fraction by a 65-tap Kaiser(beta=8)
    windowed-sinc fractional-delay filter) plus independent AWGN at SNR ~ U[0, 20] dB."""
    from scipy.signal import oaconvolve

    rng = np.random.Generator(np.random.Philox(seed))
    lengths = np.exp(rng.uniform(np.log(2.0 ** 12), np.log(2.0 ** 20), count)).astype(np.int64)
    delays = rng.uniform(-200.0, 200.0, count)
    snrs = rng.uniform(0.0, 20.0, count)
    taps_k = np.arange(65) - 32
    window = np.kaiser(65, 8.0)
    pairs = []
    for n, d, snr in zip(lengths, delays, snrs):
        n = int(n)
        di = int(np.floor(d))
        frac = d - di
        h = (np.sinc(taps_k - frac) * window).astype(np.float64)
        h /= h.sum()
        total = n + 400 + 64
        src = (rng.standard_normal(total, dtype=np.float32)
               + 1j * rng.standard_normal(total, dtype=np.float32)) * np.float32(np.sqrt(0.5))
        shifted = oaconvolve(src, h, mode="full")[32: 32 + total]   # src delayed by frac
        a = src[232: 232 + n]
        b = shifted[232 - di: 232 - di + n]                          # delayed by di + frac
        sigma = np.sqrt(0.5 * 10.0 ** (-snr / 10.0))
        a = a + sigma * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
        b = b + sigma * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
        pairs.append((a.astype(np.complex64), b.astype(np.complex64)))
    return pairs, delays


def bench_gcc_batch(rf, args) -> dict:
    pairs, delays = gcc_batch_input()
    count = len(pairs)
    stream = rf.Stream()
    d_a = [rf.DeviceBuffer.from_array(a) for a, _ in pairs]
    d_b = [rf.DeviceBuffer.from_array(b) for _, b in pairs]
    na = [int(a.size) for a, _ in pairs]
    nb = [int(b.size) for _, b in pairs]
    delays = [float(d) for d in delays]
    h = rf.GccBatch(na, nb)
    h.bind([x.ptr for x in d_a], [x.ptr for x in d_b])
    res = rf.DeviceBuffer(16 * count, np.uint8)
    # --no-variants: the library's default configuration only (no tuner, no group sweep, no
    # per-pair path), so that a profiler sees one configuration per launch
    tuned = None if args.no_variants else h.tune(res.ptr, stream.handle)
    group_bytes, groups = h.group_bytes()
    sweep = {}
    for budget in () if args.no_variants else (0, 48 << 20, 96 << 20, 160 << 20):
        h.set_group_bytes(budget)
        ms = time_steps(rf, lambda i: h.run(res.ptr, stream.handle), max(2, args.steps // 2), 1,
                        stream.handle) / max(2, args.steps // 2)
        sweep[f"{budget >> 20}MiB/{h.group_bytes()[1]}groups"] = round(ms, 4)
    h.set_group_bytes(group_bytes)
    ms_eager = time_steps(rf, lambda i: h.run(res.ptr, stream.handle), args.steps, args.warmup,
                          stream.handle) / args.steps
    ms_graph = time_steps(rf, lambda i: h.run(res.ptr, stream.handle, graph=True), args.steps,
                          args.warmup, stream.handle) / args.steps
    peaks = rf.read_peaks(res, count)

    # The plain FP32 path: one XcorrFft (GCC-PHAT) and one device peak search per pair.
    singles = [rf.XcorrFft(x, y, True) for x, y in zip(na, nb)]
    d_r = [rf.DeviceBuffer(x + y - 1) for x, y in zip(na, nb)]
    pres = rf.DeviceBuffer(16 * count, np.uint8)

    def plain(i):
        for k in range(count):
            singles[k].run(d_a[k].ptr, d_b[k].ptr, d_r[k].ptr, stream.handle)
            rf.peak_device(d_r[k].ptr, na[k] + nb[k] - 1, pres.offset(16 * k), stream.handle)

    if args.no_variants:
        ms_plain, plain_peaks = float("nan"), peaks
    else:
        ms_plain = time_steps(rf, plain, max(2, args.steps // 4), 1,
                              stream.handle) / max(2, args.steps // 4)
        plain_peaks = rf.read_peaks(pres, count)
    release(rf, h, singles, d_r, pres, res, d_a, d_b, stream)
    same = int(sum(p.index == q.index for p, q in zip(peaks, plain_peaks)))
    lag_err = [float(abs(p.index - (n - 1) + p.frac - d)) for p, n, d in zip(peaks, na, delays)]
    within = int(sum(e <= 0.1 for e in lag_err))
    checked = None
    if not args.no_cpu:
        oracle = load_oracle()
        order = [int(k) for k in np.argsort(na)[:12]]
        t0 = time.perf_counter()
        agree = 0
        for k in order:
            idx, _ = oracle.peak(oracle.gcc_phat(pairs[k][0], pairs[k][1]))
            agree += int(idx == peaks[k].index)
        dt = time.perf_counter() - t0
        cpu_oracle_pairs_per_s = len(order) / dt
        checked = {"pairs": len(order), "lag_equal_to_oracle": int(agree)}
        fast = load_cpu_fast()
        t0 = time.perf_counter()
        cpu_peaks = fast.gcc_phat_peaks(pairs)
        dt = time.perf_counter() - t0
        cpu_pairs_per_s = count / dt
        cpu_same = int(sum(int(c) == p.index for c, p in zip(cpu_peaks, peaks)))
        cpu_note = (f"FP32 GCC-PHAT + peak (radix-2 FFT), {fast.threads()} threads "
                    f"(oracle/cpu_fast.cpp), all {count} pairs, {dt:.2f} s on {cpu_host()}; "
                    f"same peak index as the GPU on {cpu_same}/{count} pairs")
    # Algorithmic bytes: both inputs read once (8 B per sample); the per-pair result is 16 B.
    bytes_step = sum(8 * (x + y) + 16 for x, y in zip(na, nb))
    # FFT work: three L-point transforms per pair (5 L log2 L each) plus the weighting.
    lengths = [1 << int(x + y - 2).bit_length() for x, y in zip(na, nb)]
    flop_step = sum(15 * n * (n.bit_length() - 1) + 12 * n for n in lengths)
    best = min(ms_eager, ms_graph)
    return {
        "metric": METRIC, "value": round(count / (best * 1e-3), 1), "unit": "pairs/s",
        "ms_per_step": round(best, 5), "dtype": "f32",
        "config": {"workload": f"batched GCC-PHAT, {count} pairs of uneven length "
                               f"(log-uniform 2^12..2^20 complex64), peak lag per pair",
                   "pairs": count, "samples": int(sum(na) + sum(nb)),
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": round(bytes_step / (best * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bytes_step / (best * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles("gcc_batch_c4")},
        "fft_valu": {"flop_per_step": int(flop_step),
                     "achieved_tflops": round(flop_step / (best * 1e-3) / 1e12, 2),
                     "peak_tflops": FP32_PEAK_TFLOPS,
                     "frac": round(flop_step / (best * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4)},
        "batched": {"eager_ms": round(ms_eager, 4), "graph_ms": round(ms_graph, 4),
                    "tuned_workgroups_per_cu": tuned,
                    "tuned_group_bytes": group_bytes, "groups": groups,
                    "group_budget_sweep_ms": sweep},
        "plain_per_pair": None if args.no_variants else {
            "ms": round(ms_plain, 4), "pairs_per_s": round(count / (ms_plain * 1e-3), 1),
            "speedup_of_batched": round(ms_plain / best, 2), "same_peak_index": same},
        "accuracy": {"pairs_within_0.1_sample": within, "max_abs_lag_error": round(max(lag_err), 4),
                     "oracle_check": checked},
        "cpu_baseline": None if args.no_cpu else {
            "value": round(cpu_pairs_per_s, 3), "unit": "pairs/s", "cores": fast.threads(),
            "kind": "port", "sample": cpu_note},
        "cpu_oracle": None if args.no_cpu else {
            "value": round(cpu_oracle_pairs_per_s, 3), "unit": "pairs/s", "cores": 1,
            "sample": "float64 oracle GCC-PHAT + peak, one thread, the 12 shortest pairs"},
    }


def bench_delay(rf, args) -> dict:
    """Learned delay search: 4096 pairs of 2^14 samples (delay U[-512, 512), SNR U[-10, 20] dB,
    AR(1) noise), network trained on 2^20 pairs synthesised on the GPU; compared with the exact
    batched GCC-PHAT on the same pairs and with the coarse-proposal ablation."""
    count, n = 4096, 1 << 14
    train_pairs, train_batch, lr = 1 << 20, 1024, 1e-3
    stream = rf.Stream()
    mlp = rf.DelayMlp(train_batch, seed=0xC5A)
    t0 = time.perf_counter()
    losses = mlp.train(0xC5A, train_pairs, train_batch, n, lr, stream)
    train_s = time.perf_counter() - t0
    a = rf.DeviceBuffer(count * n)
    b = rf.DeviceBuffer(count * n)
    d = rf.DeviceBuffer(count, np.float32)
    rf.synth_pairs(0xC5, 0, count, n, a.ptr, b.ptr, d.ptr, stream)
    rf.synchronize()
    delays = d.download()
    search = rf.DelaySearch(count, n)
    res = rf.DeviceBuffer(32 * count, np.uint8)

    def timed(mlp_or_none, threshold):
        ms = time_steps(rf, lambda i: search.run(a.ptr, b.ptr, mlp_or_none, threshold, res.ptr,
                                                 stream), args.steps, args.warmup,
                        stream.handle) / args.steps
        return ms, rf.read_delay_results(res, count)

    ms_mlp, r_mlp = timed(mlp, 0.9)
    ms_coarse, r_coarse = timed(None, 0.0)
    ms_full, r_full = timed(None, 2.0)
    batch = rf.GccBatch([n] * count, [n] * count)
    batch.bind([a.offset(k * n) for k in range(count)], [b.offset(k * n) for k in range(count)])
    pres = rf.DeviceBuffer(16 * count, np.uint8)
    batch.tune(pres.ptr, stream.handle)
    ms_exact = time_steps(rf, lambda i: batch.run(pres.ptr, stream.handle, graph=True),
                          args.steps, args.warmup, stream.handle) / args.steps
    exact = rf.read_peaks(pres, count)
    # The plain FP32 path the batch replaces: one XcorrFft (GCC-PHAT) and one device peak
    # search per pair, the pairs one after another on the stream.
    plain = rf.XcorrFft(n, n, True)
    d_r = rf.DeviceBuffer(2 * n - 1)
    pplain = rf.DeviceBuffer(16 * count, np.uint8)

    def plain_step(i):
        for k in range(count):
            plain.run(a.offset(k * n), b.offset(k * n), d_r.ptr, stream.handle)
            rf.peak_device(d_r.ptr, 2 * n - 1, pplain.offset(16 * k), stream.handle)

    ms_plain = time_steps(rf, plain_step, 2, 1, stream.handle) / 2
    plain_same = int(sum(p.index == e.index for p, e in zip(rf.read_peaks(pplain, count), exact)))
    release(rf, plain, d_r, pplain)

    labels = [int(min(max((int(np.rint(x)) + 512) >> 4, 0), 63)) for x in delays]
    proposals_right = sum(r.proposal == lab for r, lab in zip(r_mlp, labels))
    cpu = None
    cpu_oracle = None
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import delay as D
        oracle = load_oracle()
        ha = a.download(12 * n).reshape(12, n)
        hb = b.download(12 * n).reshape(12, n)
        t0 = time.perf_counter()
        for k in range(12):
            D.features(ha[k], hb[k])
            oracle.peak(oracle.gcc_phat(ha[k], hb[k]))
        dt = time.perf_counter() - t0
        cpu_oracle = {"value": round(12 / dt, 3), "unit": "pairs/s", "cores": 1,
                      "sample": "float64 oracle features + exact GCC-PHAT + peak, one thread, "
                                "12 pairs"}
        fast = load_cpu_fast()
        ha = a.download().reshape(count, n)
        hb = b.download().reshape(count, n)
        t0 = time.perf_counter()
        cpu_peaks = fast.gcc_phat_peaks(list(zip(ha, hb)))
        dt = time.perf_counter() - t0
        same = int(sum(int(c) == p.index for c, p in zip(cpu_peaks, exact)))
        cpu = {"value": round(count / dt, 1), "unit": "pairs/s", "cores": fast.threads(),
               "kind": "port",
               "sample": f"FP32 exact GCC-PHAT + peak (radix-2 FFT), {fast.threads()} threads "
                         f"(oracle/cpu_fast.cpp), all {count} pairs, {dt:.2f} s on {cpu_host()}; "
                         f"same peak index as the GPU's exact batch on {same}/{count} pairs"}
    release(rf, batch, search, pres, res, a, b, d)
    long_capture = delay_long_capture(rf, mlp, stream, args)
    release(rf, mlp, stream)

    def quality(results):
        agree = sum(r.index == e.index for r, e in zip(results, exact))
        fallback = sum(r.path == 1 for r in results)
        err = [abs(r.index - (n - 1) + r.frac - dd) for r, dd in zip(results, delays)]
        return {"lag_equal_to_exact": int(agree), "agreement": round(agree / count, 5),
                "fallback_rate": round(fallback / count, 5),
                "within_0.1_sample": int(sum(e <= 0.1 for e in err)),
                "max_abs_lag_error": round(float(max(err)), 4)}

    L = 1 << 15
    flop_features = 2 * 5 * L * 15 + 12 * L          # two forward FFTs + weighting
    flop_inverse = 5 * L * 15
    flop_mlp = 2 * (256 * 256 + 256 * 256 + 256 * 64)
    fallback = sum(r.path == 1 for r in r_mlp)
    flop_step = count * (flop_features + flop_mlp) + fallback * flop_inverse \
        + (count - fallback) * (L * 8 * (6 + 1) + L * 6)
    bytes_step = count * (16 * n + 32)
    best = ms_mlp
    return {
        "metric": METRIC, "value": round(count / (best * 1e-3), 1), "unit": "pairs/s",
        "ms_per_step": round(best, 5), "dtype": "f32 (network: bf16 operands, f32 accumulate)",
        "config": {"workload": f"learned delay search, {count} pairs of {n} complex64 samples "
                               "(delay U[-512,512), SNR U[-10,20] dB, AR(1) noise), "
                               "threshold 0.9",
                   "pairs": count, "samples_per_channel": n, "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": round(bytes_step / (best * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bytes_step / (best * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "fft_valu": {"flop_per_step": int(flop_step),
                     "achieved_tflops": round(flop_step / (best * 1e-3) / 1e12, 2),
                     "peak_tflops": FP32_PEAK_TFLOPS,
                     "frac": round(flop_step / (best * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4)},
        "training": {"pairs": train_pairs, "batch": train_batch, "steps": int(losses.size),
                     "learning_rate": lr, "seconds": round(train_s, 2),
                     "loss_first": round(float(losses[:8].mean()), 4),
                     "loss_last": round(float(losses[-8:].mean()), 4)},
        "mlp": {"ms": round(ms_mlp, 4), **quality(r_mlp),
                "proposal_equal_to_label": int(proposals_right)},
        "ablation_coarse_proposal": {"ms": round(ms_coarse, 4), **quality(r_coarse)},
        "full_inverse_path": {"ms": round(ms_full, 4), **quality(r_full)},
        "exact_gcc_batch": {"ms": round(ms_exact, 4),
                            "pairs_per_s": round(count / (ms_exact * 1e-3), 1),
                            "speedup_of_mlp_search": round(ms_exact / ms_mlp, 3)},
        "plain_per_pair": {"ms": round(ms_plain, 3),
                           "pairs_per_s": round(count / (ms_plain * 1e-3), 1),
                           "speedup_of_exact_batch": round(ms_plain / ms_exact, 2),
                           "speedup_of_mlp_search": round(ms_plain / ms_mlp, 2),
                           "same_peak_index_as_exact": plain_same},
        "long_capture": long_capture,
        "cpu_baseline": cpu,
        "cpu_oracle": cpu_oracle,
    }


def delay_long_capture(rf, mlp, stream, args, count: int = 1024, n: int = 1 << 18) -> dict:
    """The same search on longer captures (FFT length 2^19, where the exact path's inverse
    transform is a multi-pass four-step FFT): 1024 pairs of 2^18 samples of seed 0xC5, the
    network trained at 2^14 samples, against the exact batched GCC-PHAT."""
    a, b = rf.DeviceBuffer(count * n), rf.DeviceBuffer(count * n)
    rf.synth_pairs(0xC5, 0, count, n, a.ptr, b.ptr, None, stream)
    search = rf.DelaySearch(count, n)
    res = rf.DeviceBuffer(32 * count, np.uint8)
    exact = rf.GccBatch([n] * count, [n] * count)
    exact.bind([a.offset(k * n) for k in range(count)], [b.offset(k * n) for k in range(count)])
    pres = rf.DeviceBuffer(16 * count, np.uint8)
    reps = max(2, args.steps // 4)
    ms_exact = time_steps(rf, lambda i: exact.run(pres.ptr, stream.handle), reps, 1,
                          stream.handle) / reps
    want = [p.index for p in rf.read_peaks(pres, count)]
    out = {"pairs": count, "samples_per_channel": n, "exact_gcc_batch_ms": round(ms_exact, 4)}
    for name, net, threshold in [("mlp", mlp, 0.9), ("coarse", None, 0.0)]:
        ms = time_steps(rf, lambda i: search.run(a.ptr, b.ptr, net, threshold, res.ptr, stream),
                        reps, 1, stream.handle) / reps
        got = rf.read_delay_results(res, count)
        out[name] = {"ms": round(ms, 4), "speedup_over_exact": round(ms_exact / ms, 3),
                     "lag_equal_to_exact": int(sum(g.index == w for g, w in zip(got, want))),
                     "full_inverse_transforms": int(sum(g.path == 1 for g in got))}
    release(rf, exact, search, a, b, res, pres)
    return out


def bench_copy(rf, args) -> dict:
    """Achievable HBM bandwidth, for the roofline's context: device-to-device copies (read +
    write) of 1 GiB and of 256 MiB over rotating buffers, by the library's streaming copy
    kernel (four 16-byte loads in flight per lane) and by the HIP runtime's hipMemcpyAsync.
    value = the library kernel on 1 GiB."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_int, ctypes.c_void_p]
    stream = rf.Stream()
    results = {}
    for label, nbytes, sets in [("1GiB", 1 << 30, 1), ("256MiB_x4", 256 << 20, 4)]:
        src = [rf.DeviceBuffer(nbytes, np.uint8) for _ in range(sets)]
        dst = [rf.DeviceBuffer(nbytes, np.uint8) for _ in range(sets)]
        for b in src:
            b.zero()

        def kernel_copy(i):
            rf.check(rf.lib().rf_memcpy_d2d(dst[i % sets].ptr, src[i % sets].ptr, nbytes,
                                            stream.handle))

        def runtime_copy(i):
            err = hip.hipMemcpyAsync(dst[i % sets].ptr, src[i % sets].ptr, nbytes, 3,
                                     stream.handle)  # hipMemcpyDeviceToDevice
            if err != 0:
                raise RuntimeError(f"hipMemcpyAsync failed ({err})")

        for name, fn in [("copy_kernel", kernel_copy), ("hipMemcpyAsync", runtime_copy)]:
            ms = time_steps(rf, fn, args.steps, args.warmup, stream.handle) / args.steps
            results[f"{name}_{label}"] = {"ms": round(ms, 5),
                                          "GBps": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}
        release(rf, src, dst)
    release(rf, stream)
    main = results["copy_kernel_1GiB"]
    best = max(r["GBps"] for r in results.values())
    return {
        "metric": "copy GB/s (read + write)", "value": main["GBps"], "unit": "GB/s",
        "ms_per_step": main["ms"], "dtype": "u8",
        "config": {"workload": "device-to-device copy of 1 GiB", "bytes": 1 << 30,
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": main["GBps"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(main["GBps"] / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "copies": results,
        "best_GBps": best,
        "cpu_baseline": None,
    }


def bench_sigflow(rf, args) -> dict:
    """The SigFlow two_channel pipeline (two captures low-passed by one 63-tap filter, a
    1024-point magnitude spectrum of the first, and their cross-correlation) at 2^18 samples per
    channel, as sigflowc compiles it: the example program built for the register FIR and the FFT
    cross-correlation, with fusion (fir_batched + spectrum-magnitude) and without, run eagerly and
    as a replayed hipGraph (--no-variants: fused and eager only). value = samples of both
    channels per second of the faster fused kernel phase. The reference quotes 0.303 ms kernel phase and 1.82 ms end to end for its fastest
    variant of the same program and size (its docs/PROFILING_LOG.md:676-681)."""
    import subprocess
    import tempfile

    n = 1 << 18
    rng = np.random.Generator(np.random.Philox(0x5F))
    src = (rng.standard_normal(n + 64) + 1j * rng.standard_normal(n + 64)) * np.sqrt(0.5)
    noise = (rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))) * 0.3 * np.sqrt(0.5)
    ch0 = (src[64:] + noise[0]).astype(np.complex64)
    ch1 = (src[64 - 37:n + 64 - 37] + noise[1]).astype(np.complex64)  # ch0's source 37 later
    programs = ROOT / "gpu-rf-pipeline_amd" / "sigflow" / "_programs"
    taps = ROOT / "tests" / "golden" / "fixtures" / "lowpass.f32"
    runs = {}
    with tempfile.TemporaryDirectory() as tmp:
        work = Path(args.sigflow_dir) if args.sigflow_dir else Path(tmp)
        (work / "fixtures").mkdir(parents=True, exist_ok=True)
        ch0.tofile(work / "fixtures" / "ch0.cf32")
        ch1.tofile(work / "fixtures" / "ch1.cf32")
        (work / "fixtures" / "lowpass.f32").write_bytes(taps.read_bytes())
        configs = [("fused_eager", "register_fft", []),
                   ("fused_graph", "register_fft", ["--graph"]),
                   ("unfused_eager", "register_no_fuse", []),
                   ("unfused_graph", "register_no_fuse", ["--graph"])]
        for label, variant, extra in configs[:1] if args.no_variants else configs:
            exe = programs / f"two_channel.{variant}"
            if not exe.exists():
                raise RuntimeError(f"{exe} is not built (gpu-rf-pipeline_amd/sigflow/"
                                   "build_examples.py)")
            proc = subprocess.run([str(exe), "--bench", str(args.steps), "--warmup",
                                   str(args.warmup), *extra], cwd=work, capture_output=True,
                                  text=True, timeout=300, check=True)
            report = json.loads(proc.stdout.strip().splitlines()[-1])
            runs[label] = {k: report[k] for k in ("kernel_ms", "kernel_std_ms", "h2d_ms",
                                                  "d2h_ms", "calls", "calls_by_kind",
                                                  "allocated_bytes")}
            runs[label]["end_to_end_ms"] = round(report["h2d_ms"] + report["kernel_ms"]
                                                 + report["d2h_ms"], 4)
            if label == "fused_eager":
                spec = np.fromfile(work / "out" / "spectrum.f32", dtype=np.float32)
                corr = np.fromfile(work / "out" / "xcorr.cf32", dtype=np.complex64)
    oracle = load_oracle()
    h = np.fromfile(taps, dtype=np.float32)
    want = np.abs(np.fft.fft(oracle.fir(ch0, h).reshape(-1, 1024), axis=1)).ravel()
    err = np.abs(spec.astype(np.float64) - want)
    check = {"spectrum_max_abs_err": float(err.max()),
             "spectrum_within_rtol_1e-5_atol_1e-3": bool(np.all(err <= 1e-3 + 1e-5 * want)),
             "xcorr_peak_lag": int(np.argmax(np.abs(corr))) - (n - 1), "true_delay": 37}
    best = min((r for label, r in runs.items() if label.startswith("fused")),
               key=lambda r: r["kernel_ms"])
    return {
        "metric": "SigFlow two_channel pipeline: Msamples/s of kernel phase (both channels)",
        "value": round(2 * n / (best["kernel_ms"] * 1e-3) / 1e6, 1), "unit": "Msamples/s",
        "ms_per_step": best["kernel_ms"], "dtype": "f32",
        "config": {"workload": "two_channel.sf: 2 x 2^18 complex64 samples, 63-tap FIR each "
                               "(one batched call), 1024-point |FFT| of one, FFT xcorr of both",
                   "samples_per_channel": n, "parallelism": "single-gpu"},
        "runs": runs,
        "reference_published": {"kernel_ms": 0.303, "end_to_end_ms": 1.82,
                                "gpu": "RTX 5080 Laptop, --fir-impl=tiled_const --xcorr-impl=fft"},
        "check": check,
        "cpu_baseline": None,
    }


def bench_stream(rf, args) -> dict:
    """The streaming FIR (rfgpu stream: chunks uploaded, filtered and downloaded on three
    streams over a ring of page-locked slots) on 2^24 complex64 samples, PCIe included: chunk
    2^20, depth 4, each memory mode, 63 and 1025 taps, each with the register kernel and
    fft_ola. value = the 1025-tap fft_ola rate in copy mode, the mode with no help from the
    caller. The reference's best streaming times for 2^24 samples are 15.8 ms (63 taps) and
    21.6 ms (1025 taps) (its README.md:77-79, docs/PROFILING_LOG.md:464-471)."""
    import subprocess
    import tempfile

    n = 1 << 24
    x, h1025 = fir_long_input(n)
    h63 = hamming_lowpass(63, 0.2)
    rfgpu = ROOT / "gpu-rf-pipeline_amd" / "rfgpu"
    runs = {}
    with tempfile.TemporaryDirectory() as tmp:
        work = Path(tmp)
        x.tofile(work / "x.cf32")
        h63.tofile(work / "h63.f32")
        h1025.tofile(work / "h1025.f32")
        for taps, impl in [(63, "register"), (63, "fft_ola"), (1025, "register"),
                           (1025, "fft_ola")]:
            for memory in ["copy", "registered", "zero-copy"]:
                proc = subprocess.run(
                    [str(rfgpu), "stream", "--in", str(work / "x.cf32"), "--taps",
                     str(work / f"h{taps}.f32"), "--out", str(work / "y.cf32"), "--impl", impl,
                     "--memory", memory, "--chunk", str(1 << 20), "--depth", "4", "--rate", "1e6",
                     "--bench", str(max(args.steps, 10)), "--warmup", str(args.warmup)],
                    capture_output=True, text=True, timeout=300, check=True)
                report = json.loads(proc.stdout.strip().splitlines()[-1])
                runs[f"{taps}taps_{impl}_{memory}"] = {
                    "ms": report["total_ms"]["mean"], "std_ms": report["total_ms"]["std"],
                    "msamples_per_s": report["msps"]}
        y = np.fromfile(work / "y.cf32", dtype=np.complex64)  # the last run: 1025 taps fft_ola
    err = sampled_fir_error(x, h1025, y)
    main = runs["1025taps_fft_ola_copy"]
    return {
        "metric": "streaming FIR Msamples/s, host memory to host memory (PCIe included)",
        "value": main["msamples_per_s"], "unit": "Msamples/s", "ms_per_step": main["ms"],
        "dtype": "f32",
        "config": {"workload": "rfgpu stream: 2^24 complex64 samples, chunk 2^20, depth 4",
                   "samples": n, "parallelism": "single-gpu"},
        "runs": runs,
        "reference_published_ms": {"63taps": 15.8, "1025taps": 21.6},
        "accuracy_last_run": {"max_abs_err": err[0], "max_rel_err": err[1],
                              "checked_outputs": "sampled, against float64 dot products"},
        "cpu_baseline": None,
    }


def traffic_from_profiles(name: str):
    """HBM bytes per launch from a committed rocprofv3 counter summary, if one exists."""
    path = ROOT / "profiles" / f"{name}_traffic.json"
    if path.exists():
        try:
            return json.loads(path.read_text()).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            return None
    return None


def main() -> int:
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--impl", default="fft_ola")
    p.add_argument("--workload", default="fir_long",
                   choices=["fir_long", "fir_plain", "fir_short", "channelize", "gcc_batch",
                            "delay", "copy", "sigflow", "stream"])
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--sigflow-dir", default=None,
                   help="sigflow: write the 2^18-sample inputs and outputs here and keep them")
    p.add_argument("--no-variants", action="store_true",
                   help="FIR: time only --impl, not the other variants")
    args = p.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the hot path runs on one GPU; multi-GPU runs are not part of this "
              "library. Use --gpus 1.", file=sys.stderr)
        return 2
    rf = load_rfhip()
    if args.workload == "channelize":
        out = bench_channelize(rf, args)
    elif args.workload == "copy":
        out = bench_copy(rf, args)
    elif args.workload == "delay":
        out = bench_delay(rf, args)
    elif args.workload == "gcc_batch":
        out = bench_gcc_batch(rf, args)
    elif args.workload == "sigflow":
        out = bench_sigflow(rf, args)
    elif args.workload == "stream":
        out = bench_stream(rf, args)
    else:
        out = bench_fir(rf, args)
    line = {"metric": out.pop("metric"), "value": out.pop("value"), "unit": out.pop("unit"),
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": out.pop("ms_per_step"), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": out.pop("dtype"), "data": "synthetic",
            "config": out.pop("config")}
    line.update(out)
    line["gpu"] = rf.device_name()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
