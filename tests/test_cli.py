"""The rfgpu command line (gpu-rf-pipeline_amd/rfgpu), as the reference's tests drive its own
binary (reference tests/test_pipeline.py:219-357,444-454): usage errors exit 2 and runtime
errors exit 1 before any device work (these run without a GPU); the bench report, the
streaming runner and DOA run on the GPU.
"""
from __future__ import annotations

import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from _common import FIXTURES, PACKAGE_DIR, cf32, meta, oracle, read_cf32, read_ci16_sigmf, taps

RFGPU = PACKAGE_DIR / "rfgpu"
FIR_IMPLS = ["naive", "tiled", "tiled_const", "register", "fft_ola", "mfma_f16x3", "mfma_bf16x3"]
DIRECT = ["naive", "tiled", "tiled_const", "register"]


@pytest.fixture(scope="module", autouse=True)
def built() -> None:
    if not RFGPU.exists():
        subprocess.run(["make", "-s", "-C", str(PACKAGE_DIR), "-j8", "rfgpu"], check=True)


def run(*args: object, timeout: int = 120) -> subprocess.CompletedProcess[str]:
    return subprocess.run([str(RFGPU), *map(str, args)], capture_output=True, text=True,
                          timeout=timeout)


def ok(*args: object) -> subprocess.CompletedProcess[str]:
    proc = run(*args)
    assert proc.returncode == 0, f"rfgpu {' '.join(map(str, args))}\n{proc.stderr}"
    return proc


def fx(name: str, ext: str = "cf32") -> Path:
    return FIXTURES / f"{name}.{ext}"


# --- usage and runtime errors (no GPU) -----------------------------------------------------

def test_unknown_impl_is_a_usage_error(tmp_path: Path) -> None:
    proc = run("fir", "--in", fx("tone"), "--taps", fx("lowpass", "f32"), "--out",
               tmp_path / "y.cf32", "--impl", "no_such_impl")
    assert proc.returncode == 2 and "no_such_impl" in proc.stderr
    assert not (tmp_path / "y.cf32").exists()


@pytest.mark.parametrize("block", ["0", "48", "2048", "abc"])
def test_bad_block_size_is_a_usage_error(tmp_path: Path, block: str) -> None:
    proc = run("fir", "--in", fx("tone"), "--taps", fx("lowpass", "f32"), "--out",
               tmp_path / "y.cf32", "--block", block)
    assert proc.returncode == 2 and "--block" in proc.stderr


def test_missing_input_file_is_a_runtime_error(tmp_path: Path) -> None:
    proc = run("spectrum", "--in", tmp_path / "missing.cf32", "--out", tmp_path / "s.cf32")
    assert proc.returncode == 1 and "missing.cf32" in proc.stderr


def test_truncated_cf32_is_rejected(tmp_path: Path) -> None:
    np.zeros(3, dtype="<f4").tofile(tmp_path / "bad.cf32")
    proc = run("spectrum", "--in", tmp_path / "bad.cf32", "--out", tmp_path / "s.cf32")
    assert proc.returncode == 1 and "bad.cf32" in proc.stderr


@pytest.mark.parametrize("meta_doc,message", [
    ({"core:datatype": "cf64_le"}, "cf64_le"),
    ({"core:datatype": "ci16_le", "core:num_channels": 2}, "channels"),
])
def test_unsupported_sigmf_is_a_runtime_error(tmp_path: Path, meta_doc: dict,
                                              message: str) -> None:
    (tmp_path / "rec.sigmf-meta").write_text(json.dumps({"global": meta_doc}))
    (tmp_path / "rec.sigmf-data").write_bytes(bytes(16))
    proc = run("magnitude", "--in", tmp_path / "rec.sigmf-meta", "--out", tmp_path / "m.f32")
    assert proc.returncode == 1 and message in proc.stderr


def test_bad_channel_count_is_a_usage_error(tmp_path: Path) -> None:
    proc = run("channelize", "--in", fx("tone"), "--taps", fx("channelizer_proto", "f32"),
               "--channels", "0", "--out", tmp_path / "ch.cf32")
    assert proc.returncode == 2


@pytest.mark.parametrize("args", [["--chunk", "0"], ["--depth", "0"], ["--memory", "disk"]])
def test_bad_stream_options_are_usage_errors(tmp_path: Path, args: list[str]) -> None:
    proc = run("stream", "--in", fx("tone"), "--taps", fx("lowpass", "f32"), "--out",
               tmp_path / "y.cf32", *args)
    assert proc.returncode == 2


def test_unknown_subcommand_and_warmup_without_bench(tmp_path: Path) -> None:
    assert run("frobnicate").returncode == 2
    proc = run("spectrum", "--in", fx("tone"), "--out", tmp_path / "s.cf32", "--warmup", "2")
    assert proc.returncode == 2


# --- GPU -------------------------------------------------------------------------------------

@pytest.mark.gpu
def test_info_lists_the_variant_tables() -> None:
    info = json.loads(ok("info").stdout)
    assert info["impls"]["fir"] == FIR_IMPLS
    assert info["impls"]["xcorr"] == ["naive", "tiled", "fft"]


@pytest.mark.gpu
def test_bench_reports_each_phase_separately(tmp_path: Path) -> None:
    proc = ok("spectrum", "--in", fx("tone"), "--out", tmp_path / "s.cf32", "--bench", 10,
              "--warmup", 2)
    report = json.loads(proc.stdout)
    assert report["op"] == "spectrum" and report["runs"] == 10 and report["warmup"] == 2
    for phase in ("h2d_ms", "kernel_ms", "d2h_ms"):
        assert report[phase]["mean"] > 0 and report[phase]["std"] >= 0


@pytest.mark.gpu
def test_bench_does_not_change_the_output(tmp_path: Path) -> None:
    ok("spectrum", "--in", fx("tone_odd"), "--out", tmp_path / "once.cf32")
    ok("spectrum", "--in", fx("tone_odd"), "--out", tmp_path / "bench.cf32", "--bench", 10)
    assert (tmp_path / "once.cf32").read_bytes() == (tmp_path / "bench.cf32").read_bytes()


@pytest.mark.gpu
@pytest.mark.parametrize("impl", FIR_IMPLS)
def test_cli_fir_matches_oracle(tmp_path: Path, impl: str) -> None:
    ok("fir", "--in", fx("tone_odd"), "--taps", fx("lowpass_long", "f32"), "--out",
       tmp_path / "y.cf32", "--impl", impl)
    tol = 2e-4 if impl == "mfma_bf16x3" else 1e-5
    np.testing.assert_allclose(read_cf32(tmp_path / "y.cf32"),
                               oracle.fir(cf32("tone_odd"), taps("lowpass_long")),
                               rtol=tol, atol=tol)


@pytest.mark.gpu
@pytest.mark.parametrize("suffix", [".sigmf-meta", ".sigmf-data"])
def test_sigmf_ci16_reads_as_scaled_cf32(tmp_path: Path, suffix: str) -> None:
    ok("fir", "--in", FIXTURES / f"tone_ci16{suffix}", "--taps", fx("identity", "f32"),
       "--out", tmp_path / "y.cf32")
    np.testing.assert_array_equal(read_cf32(tmp_path / "y.cf32"), read_ci16_sigmf("tone_ci16"))


@pytest.mark.gpu
@pytest.mark.parametrize("impl", FIR_IMPLS)
@pytest.mark.parametrize("chunk,depth", [(1000, 1), (1000, 3), (4096, 2), (100000, 3)])
@pytest.mark.parametrize("memory", ["copy", "registered", "zero-copy"])
def test_stream_matches_fir(tmp_path: Path, impl: str, chunk: int, depth: int,
                            memory: str) -> None:
    ok("fir", "--in", fx("tone"), "--taps", fx("lowpass_long", "f32"), "--out",
       tmp_path / "fir.cf32", "--impl", impl)
    ok("stream", "--in", fx("tone"), "--taps", fx("lowpass_long", "f32"), "--out",
       tmp_path / "stream.cf32", "--impl", impl, "--chunk", chunk, "--depth", depth,
       "--memory", memory)
    one = (tmp_path / "fir.cf32").read_bytes()
    streamed = (tmp_path / "stream.cf32").read_bytes()
    if impl in DIRECT:
        # Each output sums the same samples in the same order: bit-identical.
        assert one == streamed
    else:
        # Transform / GEMM blocks sit relative to each chunk: equal to rounding (BF16: to its
        # stated precision).
        tol = 2e-4 if impl == "mfma_bf16x3" else 1e-5
        np.testing.assert_allclose(np.frombuffer(streamed, np.complex64),
                                   np.frombuffer(one, np.complex64), rtol=tol, atol=tol)


@pytest.mark.gpu
@pytest.mark.parametrize("chunk,depth", [(1000, 1), (4096, 2), (100000, 3)])
@pytest.mark.parametrize("memory", ["copy", "zero-copy"])
def test_stream_matches_fir_short_filter(tmp_path: Path, chunk: int, depth: int,
                                         memory: str) -> None:
    # fft_ola with the 63-tap filter runs the 512-point overlap-save in both the one-shot and the
    # streaming runs (segments sit relative to each chunk: equal to rounding), and the register
    # kernel's streamed output is bit-identical to its one-shot output.
    for impl in ("fft_ola", "register"):
        ok("fir", "--in", fx("tone"), "--taps", fx("lowpass", "f32"), "--out",
           tmp_path / "fir.cf32", "--impl", impl)
        ok("stream", "--in", fx("tone"), "--taps", fx("lowpass", "f32"), "--out",
           tmp_path / "stream.cf32", "--impl", impl, "--chunk", chunk, "--depth", depth,
           "--memory", memory)
        one = np.frombuffer((tmp_path / "fir.cf32").read_bytes(), np.complex64)
        streamed = np.frombuffer((tmp_path / "stream.cf32").read_bytes(), np.complex64)
        if impl == "register":
            assert one.tobytes() == streamed.tobytes()
        else:
            np.testing.assert_allclose(streamed, one, rtol=1e-5, atol=1e-5)
            np.testing.assert_allclose(streamed, oracle.fir(cf32("tone"), taps("lowpass")),
                                       rtol=1e-5, atol=1e-5)


@pytest.mark.gpu
def test_stream_of_short_input(tmp_path: Path) -> None:
    ok("stream", "--in", fx("short"), "--taps", fx("lowpass", "f32"), "--out",
       tmp_path / "y.cf32", "--chunk", 3)
    np.testing.assert_allclose(read_cf32(tmp_path / "y.cf32"),
                               oracle.fir(cf32("short"), taps("lowpass")), rtol=1e-5, atol=1e-5)


@pytest.mark.gpu
def test_stream_bench_reports_throughput_against_the_sample_rate(tmp_path: Path) -> None:
    proc = ok("stream", "--in", fx("tone"), "--taps", fx("lowpass", "f32"), "--out",
              tmp_path / "y.cf32", "--chunk", 8192, "--bench", 10)
    report = json.loads(proc.stdout)
    assert report["op"] == "stream" and report["chunks"] == 8 and report["runs"] == 10
    assert report["sample_rate"] == meta("tone")["sample_rate"]
    assert report["msps"] > 0
    assert report["realtime_factor"] == pytest.approx(
        report["msps"] * 1e6 / meta("tone")["sample_rate"], rel=1e-3)


@pytest.mark.gpu
def test_cli_channelize_and_gcc_phat(tmp_path: Path) -> None:
    ok("channelize", "--in", fx("tone"), "--taps", fx("channelizer_proto", "f32"), "--channels",
       16, "--out", tmp_path / "z.cf32")
    np.testing.assert_allclose(read_cf32(tmp_path / "z.cf32"),
                               oracle.channelize(cf32("tone"), taps("channelizer_proto"), 16).ravel(),
                               rtol=1e-5, atol=1e-4)
    ok("gcc_phat", "--a", fx("ch0"), "--b", fx("ch1"), "--out", tmp_path / "g.cf32")
    np.testing.assert_allclose(read_cf32(tmp_path / "g.cf32"),
                               oracle.gcc_phat(cf32("ch0"), cf32("ch1")), atol=1e-4)


@pytest.mark.gpu
def test_doa_recovers_the_array_angle() -> None:
    m = meta("array0")
    channels = ",".join(str(fx(f"array{i}")) for i in range(m["channels"]))
    report = json.loads(ok("doa", "--in", channels, "--spacing", m["spacing_m"], "--speed",
                           m["speed_m_s"]).stdout)
    assert report["tdoa_samples"] == [2, 4, 6]
    assert report["angle_deg"] == pytest.approx(m["angle_deg"], abs=1e-3)
    chans = [cf32(f"array{i}") for i in range(4)]
    assert report["angle_deg"] == pytest.approx(
        oracle.doa(chans, m["spacing_m"], m["speed_m_s"], m["sample_rate"]), abs=1e-6)
