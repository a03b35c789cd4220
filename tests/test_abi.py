"""The C ABI without a GPU: the library loads, exports every function include/rfhip.h
declares, keeps the reference's variant tables, and rejects bad arguments with the right
status before touching the device (src/main.cpp:541-548 maps these to exit codes)."""
from __future__ import annotations

import ctypes
import re
import subprocess

import pytest

from _common import LIB, ROOT, load_rfhip

rf = load_rfhip()


def declared_functions() -> list[str]:
    text = (ROOT / "include" / "rfhip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rf_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_function_is_exported() -> None:
    names = declared_functions()
    assert len(names) >= 40
    dynsym = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True,
                            text=True, check=True).stdout
    exported = set(re.findall(r"\s[TW]\s+(rf_\w+)$", dynsym, flags=re.M))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(rf.lib(), n)


def test_abi_version_and_status_strings() -> None:
    assert rf.lib().rf_abi_version() == 1
    assert [rf.status_string(s) for s in range(5)] == [
        "RF_OK", "RF_INVALID_ARG", "RF_LIMIT", "RF_HIP_ERROR", "RF_ALLOC"]


def test_fir_table_keeps_the_reference_order() -> None:
    # src/fir.cu:725-734, then this build's addition.
    assert [i.name for i in rf.fir_impls()] == ["naive", "tiled", "tiled_const", "register",
                                               "fft_ola", "mfma_f16x3", "mfma_bf16x3"]
    assert [i.name for i in rf.xcorr_impls()] == ["naive", "tiled", "fft"]
    assert rf.fir_impl_index("register") == 3


def test_unknown_impl_is_an_invalid_argument() -> None:
    with pytest.raises(rf.RfError) as e:
        rf.fir_impl_index("no_such_impl")
    assert e.value.status == rf.RF_INVALID_ARG and "no_such_impl" in str(e.value)
    idx = ctypes.c_int(0)
    assert rf.lib().rf_xcorr_parse_impl(b"fft", ctypes.byref(idx)) == 0 and idx.value == 2
    assert rf.lib().rf_xcorr_parse_impl(b"bogus", ctypes.byref(idx)) == rf.RF_INVALID_ARG


@pytest.mark.parametrize("block", [48, 2048, 16, -32])
def test_bad_block_size_is_rejected(block: int) -> None:
    st = rf.lib().rf_fir(0, 8, 8, 8, 100, 3, block, None)
    assert st == rf.RF_INVALID_ARG
    assert "--block" in rf.last_error()


def test_bad_fir_arguments_are_rejected() -> None:
    L = rf.lib()
    assert L.rf_fir(99, 8, 8, 8, 100, 3, 0, None) == rf.RF_INVALID_ARG
    assert L.rf_fir(0, 8, 8, 8, 100, 0, 0, None) == rf.RF_INVALID_ARG
    assert L.rf_fir(0, None, 8, 8, 100, 3, 0, None) == rf.RF_INVALID_ARG
    assert L.rf_fir(0, 8, 8, 8, -1, 3, 0, None) == rf.RF_INVALID_ARG
    # No samples: nothing to enqueue, success without a device.
    assert L.rf_fir(0, None, 8, None, 0, 3, 0, None) == rf.RF_OK


@pytest.mark.parametrize("count", [0, 9])
def test_batch_count_outside_1_to_8_is_a_limit(count: int) -> None:
    with pytest.raises(rf.RfError) as e:
        rf.fir_batched_device(0, [8] * count, [8] * count, [10] * count, 8, 3)
    assert e.value.status == rf.RF_LIMIT


def test_channelize_arguments() -> None:
    L = rf.lib()
    assert L.rf_channelize(None, 8, 8, 8, 100, 0, 3, None) == rf.RF_INVALID_ARG
    assert L.rf_channelize(None, 8, 8, 8, 100, 4, 0, None) == rf.RF_INVALID_ARG
    assert L.rf_channelize(None, None, None, None, 0, 4, 3, None) == rf.RF_OK
    assert [rf.channelize_frames(n, m) for n, m in [(0, 4), (1, 4), (8, 4), (9, 4), (7, 1)]] == \
        [0, 1, 2, 3, 7]


def test_xcorr_limits() -> None:
    L = rf.lib()
    h = ctypes.c_void_p()
    assert L.rf_xcorr_fft_create(0, 5, 0, ctypes.byref(h)) == rf.RF_INVALID_ARG
    # 2^24 + 1 lags need a 2^25-point FFT: above the supported length.
    assert L.rf_xcorr_fft_create(1 << 23, (1 << 23) + 2, 1, ctypes.byref(h)) == rf.RF_LIMIT
    assert "2^24" in rf.last_error()
    assert L.rf_xcorr_direct(0, 8, 8, 8, 0, 5, None) == rf.RF_INVALID_ARG
    assert L.rf_xcorr_direct(7, 8, 8, 8, 5, 5, None) == rf.RF_INVALID_ARG


def test_fft_plan_arguments() -> None:
    L = rf.lib()
    h = ctypes.c_void_p()
    assert L.rf_fft_plan_create(0, 1, ctypes.byref(h)) == rf.RF_INVALID_ARG
    assert L.rf_fft_plan_create(1 << 25, 1, ctypes.byref(h)) == rf.RF_LIMIT
    assert L.rf_fft_exec(None, 8, 8, -1, None) == rf.RF_INVALID_ARG


def test_peak_arguments() -> None:
    assert rf.lib().rf_peak(8, 0, 8, None) == rf.RF_INVALID_ARG


def test_delay_search_arguments() -> None:
    L = rf.lib()
    h = ctypes.c_void_p()
    assert L.rf_delay_search_create(4, 4000, ctypes.byref(h)) == rf.RF_LIMIT      # L = 2^13
    assert L.rf_delay_search_create(4, 600000, ctypes.byref(h)) == rf.RF_LIMIT    # L = 2^21
    assert L.rf_delay_search_create(0, 16384, ctypes.byref(h)) == rf.RF_INVALID_ARG
    assert L.rf_delay_mlp_create(0, 1, ctypes.byref(h)) == rf.RF_INVALID_ARG
    assert L.rf_synth_pairs(1, 0, 0, 16384, 8, 8, None, None) == rf.RF_INVALID_ARG
    assert L.rf_synth_pairs(1, 0, 4, 16384, None, 8, None, None) == rf.RF_INVALID_ARG
    assert L.rf_delay_search_run(None, 8, 8, None, 0.9, 8, None) == rf.RF_INVALID_ARG
    assert L.rf_delay_mlp_train(None, 1, 16, 16, 16384, 1e-3, None, None) == rf.RF_INVALID_ARG


def test_channelize_mfma_and_fir_tune_arguments() -> None:
    L = rf.lib()
    assert L.rf_channelize_mfma(8, 8, 8, 4096, 32, 1024, None) == rf.RF_LIMIT
    assert L.rf_channelize_mfma(8, 8, 8, 4096, 64, 1025, None) == rf.RF_LIMIT
    assert L.rf_channelize_mfma(8, 8, 8, 4096, 0, 1024, None) == rf.RF_INVALID_ARG
    assert L.rf_fir_tune(8, 8, 8, 1024, 63, 1 << 3, None, None) == rf.RF_INVALID_ARG
    cfg = rf.FirConfig(3, 256, 5, 0.0)
    assert L.rf_fir_run_config(ctypes.byref(cfg), 8, 8, 8, 1024, 63, None) == rf.RF_INVALID_ARG
    # the overlap-save transform length: only 0, 512 or 4096, only for fft_ola, and only for
    # the tap counts that length takes (1..449 and 1..2049), all refused before any launch
    ola = rf.fir_impl_index("fft_ola")
    for points, impl, num_taps, status in [(1000, ola, 63, rf.RF_INVALID_ARG),
                                           (512, 3, 63, rf.RF_INVALID_ARG),
                                           (512, ola, 450, rf.RF_LIMIT),
                                           (4096, ola, 2050, rf.RF_LIMIT)]:
        cfg = rf.FirConfig(impl, 256, 0, 0.0, points)
        assert L.rf_fir_run_config(ctypes.byref(cfg), 8, 8, 8, 1024, num_taps, None) == status


def test_spectrum_magnitude_arguments() -> None:
    assert rf.lib().rf_spectrum_magnitude(None, 8, 8, None) == rf.RF_INVALID_ARG


def test_fir_plan_arguments() -> None:
    # Checked before any device call: NULL handle, unknown variant, bad block, NULL taps,
    # constant-cache tap limit; a NULL plan to run or destroy.
    L = rf.lib()
    p = ctypes.c_void_p()
    assert L.rf_fir_plan_create(4, 8, 1025, 0, None, None) == rf.RF_INVALID_ARG
    assert L.rf_fir_plan_create(9, 8, 1025, 0, None, ctypes.byref(p)) == rf.RF_INVALID_ARG
    assert L.rf_fir_plan_create(0, 8, 1025, 33, None, ctypes.byref(p)) == rf.RF_INVALID_ARG
    assert L.rf_fir_plan_create(0, None, 1025, 0, None, ctypes.byref(p)) == rf.RF_INVALID_ARG
    assert L.rf_fir_plan_create(3, 8, 8193, 0, None, ctypes.byref(p)) == rf.RF_LIMIT
    assert p.value is None
    assert L.rf_fir_plan_run(None, 8, 8, 16, None) == rf.RF_INVALID_ARG
    assert L.rf_fir_plan_destroy(None) == rf.RF_OK


def test_graph_arguments() -> None:
    L = rf.lib()
    g = ctypes.c_void_p()
    assert L.rf_graph_begin_capture(None) == rf.RF_INVALID_ARG
    assert L.rf_graph_end_capture(None, ctypes.byref(g)) == rf.RF_INVALID_ARG
    assert L.rf_graph_launch(None, None) == rf.RF_INVALID_ARG
    assert L.rf_graph_destroy(None) == rf.RF_OK
