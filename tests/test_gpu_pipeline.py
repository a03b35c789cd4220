"""GPU results against the float64 oracle, through the C ABI (librfhip.so).

Mirrors the reference's tests/test_pipeline.py: the same fixtures, the same tolerances
(stated next to each operation), every --impl of the variant tables, exact results where the
reference demands them (identity taps, zero inputs). Full-size workloads are checked through
properties that do not need the oracle at full size (sampled outputs, round trips).
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from _common import cf32, load_rfhip, meta, oracle, read_ci16_sigmf, taps

pytestmark = pytest.mark.gpu
rf = load_rfhip()

FIR_IMPLS = [i.name for i in rf.fir_impls()]
XCORR_IMPLS = [i.name for i in rf.xcorr_impls()]
DIRECT_FIR = ["naive", "tiled", "tiled_const", "register"]
# Split-precision forms and their stated per-output error bound
#   |err[n]| <= rel * sum_k |h[k]| |x[n-k]|  +  floor * max|h| * sum_k |x[n-k]|
# rel: BF16 hi/lo keeps ~16 significant bits of each operand and drops lo*lo (2^-15); FP16
# hi/lo keeps ~22 bits (2^-20). floor: taps below max|h| 2^-24 fall into FP16 subnormals after
# the power-of-two tap scaling (BF16 has FP32's range: no floor).
SPLIT_BOUND = {"mfma_bf16x3": (2.0 ** -15, 0.0), "mfma_f16x3": (2.0 ** -20, 2.0 ** -24)}
REDUCED_BOUND = {"mfma_bf16x3": 2.0 ** -15}


def split_bound(impl: str, x: np.ndarray, h: np.ndarray) -> np.ndarray:
    rel, floor = SPLIT_BOUND[impl]
    ax = np.abs(x.astype(np.complex128))
    ah = np.abs(h.astype(np.float64))
    bound = rel * np.abs(oracle.fir(ax, ah))
    if floor:
        bound += floor * ah.max() * np.abs(oracle.fir(ax, np.ones_like(ah)))
    return bound

# Tolerances of the reference's tests/test_pipeline.py.
SPECTRUM_RTOL, SPECTRUM_ATOL = 1e-5, 2e-3      # :62-66
MAGNITUDE_RTOL, MAGNITUDE_ATOL = 1e-6, 1e-6    # :90-92
FIR_RTOL, FIR_ATOL = 1e-5, 1e-5                # :113-117
XCORR_RTOL, XCORR_ATOL = 1e-5, 5e-3            # :170-177
CHANNELIZE_RTOL, CHANNELIZE_ATOL = 1e-5, 1e-4  # :360-364
GCC_ATOL = 1e-4                                # :408-411


def kaiser_lowpass(num_taps: int, cutoff: float, beta: float) -> np.ndarray:
    """Kaiser-windowed sinc, cutoff in cycles per sample, unit DC gain, float32."""
    n = np.arange(num_taps) - (num_taps - 1) / 2
    h = 2 * cutoff * np.sinc(2 * cutoff * n) * np.kaiser(num_taps, beta)
    return (h / h.sum()).astype(np.float32)


def noise(seed: int, n: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)


def assert_fir_close(impl: str, got: np.ndarray, x: np.ndarray, h: np.ndarray,
                     want: np.ndarray | None = None) -> None:
    """FP32-equivalent variants: the reference's rtol = atol = 1e-5. Reduced-precision forms:
    |error[n]| <= bound * sum_k |h[k]| |x[n-k]| + 1e-7, checked output by output."""
    want = oracle.fir(x, h) if want is None else want
    assert got.shape == want.shape
    if impl in REDUCED_BOUND:
        scale = np.abs(oracle.fir(np.abs(x.astype(np.complex128)), np.abs(h.astype(np.float64))))
        err = np.abs(got.astype(np.complex128) - want)
        assert np.all(err <= REDUCED_BOUND[impl] * scale + 1e-7), float(np.max(err - REDUCED_BOUND[impl] * scale))
    else:
        np.testing.assert_allclose(got, want, rtol=FIR_RTOL, atol=FIR_ATOL)


# --- spectrum --------------------------------------------------------------------------------

@pytest.mark.parametrize("nfft", [1024, 256])
@pytest.mark.parametrize("name", ["tone", "tone_odd", "short"])
def test_spectrum_matches_oracle(name: str, nfft: int) -> None:
    x = cf32(name)
    got = rf.spectrum(x, nfft)
    want = oracle.spectrum(x, nfft).ravel()
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=SPECTRUM_RTOL, atol=SPECTRUM_ATOL)


@pytest.mark.parametrize("nfft", [1, 2, 4, 8, 16, 64, 100, 1000, 4096, 8192, 16384, 65536])
def test_spectrum_any_length_matches_oracle(nfft: int) -> None:
    # Powers of two up to 4096 in one workgroup, four-step above, exact DFT otherwise.
    x = cf32("tone")
    got = rf.spectrum(x, nfft)
    want = oracle.spectrum(x, nfft).ravel()
    scale = max(1.0, np.sqrt(nfft / 1024))
    np.testing.assert_allclose(got, want, rtol=SPECTRUM_RTOL, atol=SPECTRUM_ATOL * scale)


@pytest.mark.parametrize("nfft", [1, 2, 4, 8, 48, 64, 100, 1024, 4096, 8192, 16384, 65536,
                                  1 << 18])
def test_spectrum_magnitude_fusion(nfft: int) -> None:
    """One-pass |spectrum| equals spectrum then magnitude bit for bit, and the oracle's |X|
    within the spectrum tolerances, for every plan kind: 1/2/4-point transforms, exact DFT,
    one-workgroup transforms and the four-step transform (whose row pass stores |X|)."""
    x = cf32("tone")
    fused = rf.spectrum_magnitude(x, nfft)
    two_pass = rf.magnitude(rf.spectrum(x, nfft))
    assert fused.tobytes() == two_pass.tobytes()
    want = np.abs(oracle.spectrum(x, nfft).ravel())
    scale = max(1.0, np.sqrt(nfft / 1024))
    np.testing.assert_allclose(fused, want, rtol=SPECTRUM_RTOL, atol=SPECTRUM_ATOL * scale)


@pytest.mark.parametrize("log2n", list(range(13, 25)))
def test_four_step_every_column_length(log2n: int) -> None:
    """The four-step transform at every column length it has (n2 = 8 .. 4096 rows, one
    compile-time column transform each) against numpy's float64 FFT: rms error relative to the
    spectrum's rms within 2e-6 (float32 rounding over log2 n passes; a wrong factor or index
    gives errors of order one), and at 2^23 and 2^24 points the inverse transform back to the
    input."""
    n = 1 << log2n
    x = noise(log2n, n)
    got = rf.spectrum(x, n)
    want = np.fft.fft(x.astype(np.complex128))
    rel = np.sqrt(np.mean(np.abs(got - want) ** 2) / np.mean(np.abs(want) ** 2))
    assert rel < 2e-6, rel
    if log2n >= 23:
        plan = rf.FftPlan(n, 1)
        d = rf.DeviceBuffer.from_array(x)
        plan.exec(d.ptr, d.ptr, rf.RF_FFT_FORWARD)
        plan.exec(d.ptr, d.ptr, rf.RF_FFT_INVERSE)
        back = d.download() / n
        rf.synchronize()
        plan.destroy()
        d.free()
        np.testing.assert_allclose(back, x, atol=2e-5 * log2n)


def test_fft_inverse_round_trip() -> None:
    for nfft, batch in [(4096, 3), (1 << 15, 2), (1 << 18, 1), (48, 5)]:
        x = noise(nfft, nfft * batch)
        plan = rf.FftPlan(nfft, batch)
        d = rf.DeviceBuffer.from_array(x)
        plan.exec(d.ptr, d.ptr, rf.RF_FFT_FORWARD)
        plan.exec(d.ptr, d.ptr, rf.RF_FFT_INVERSE)
        back = d.download() / nfft
        rf.synchronize()
        plan.destroy()
        np.testing.assert_allclose(back, x, atol=2e-5 * np.log2(nfft + 1))


def test_spectrum_of_zeros_is_exactly_zero() -> None:
    got = rf.spectrum(cf32("zeros"), 1024)
    assert got.shape == (4096,) and not np.any(got)


# --- magnitude -----------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["tone", "tone_odd", "short", "ch0"])
def test_magnitude_matches_oracle(name: str) -> None:
    x = cf32(name)
    np.testing.assert_allclose(rf.magnitude(x), oracle.magnitude(x), rtol=MAGNITUDE_RTOL,
                               atol=MAGNITUDE_ATOL)


def test_magnitude_of_zeros_is_exactly_zero() -> None:
    got = rf.magnitude(cf32("zeros"))
    assert got.shape == (4096,) and not np.any(got)


def test_magnitude_unaligned_and_odd_lengths() -> None:
    x = noise(7, 1001)
    d = rf.DeviceBuffer.from_array(x)
    m = rf.DeviceBuffer(1001, np.float32)
    rf.magnitude_device(d.offset(1), m.ptr, 999)
    got = m.download(999)
    np.testing.assert_allclose(got, np.abs(x[1:1000].astype(np.complex128)), rtol=1e-6, atol=1e-6)


# --- FIR -----------------------------------------------------------------------------------

@pytest.mark.parametrize("impl", FIR_IMPLS)
@pytest.mark.parametrize("name,taps_name", [
    ("tone", "lowpass"), ("tone", "lowpass_long"), ("tone_odd", "lowpass"),
    ("tone_odd", "lowpass_long"), ("short", "lowpass"), ("short", "lowpass_long"),
])
def test_fir_matches_oracle(impl: str, name: str, taps_name: str) -> None:
    x, h = cf32(name), taps(taps_name)
    assert_fir_close(impl, rf.fir(x, h, impl), x, h)


@pytest.mark.parametrize("impl", FIR_IMPLS)
@pytest.mark.parametrize("block", [32, 96, 1024])
def test_fir_block_size_does_not_change_the_result(impl: str, block: int) -> None:
    x, h = cf32("tone_odd"), taps("lowpass_long")
    assert_fir_close(impl, rf.fir(x, h, impl, block), x, h)


@pytest.mark.parametrize("impl", FIR_IMPLS)
@pytest.mark.parametrize("name", ["tone", "tone_odd", "short"])
def test_fir_identity_returns_the_input_exactly(impl: str, name: str) -> None:
    x = cf32(name)
    np.testing.assert_array_equal(rf.fir(x, taps("identity"), impl), x)


@pytest.mark.parametrize("impl", FIR_IMPLS)
@pytest.mark.parametrize("taps_name", ["lowpass", "lowpass_long", "identity"])
def test_fir_of_zeros_is_exactly_zero(impl: str, taps_name: str) -> None:
    got = rf.fir(cf32("zeros"), taps(taps_name), impl)
    assert got.shape == (4096,) and not np.any(got)


def test_direct_fir_variants_agree_bit_for_bit() -> None:
    # Every direct variant sums fma(h[k], x[n-k]) in k order from 0.
    x, h = cf32("tone_odd"), taps("lowpass_long")
    results = [rf.fir(x, h, impl) for impl in DIRECT_FIR]
    for r in results[1:]:
        np.testing.assert_array_equal(r, results[0])


@pytest.mark.parametrize("impl", FIR_IMPLS)
@pytest.mark.parametrize("taps_name", ["lowpass", "lowpass_long"])
def test_fir_batched_matches_oracle(impl: str, taps_name: str) -> None:
    # Every channel of one batched launch against the float64 oracle, and bit-equal to the
    # single-channel call of the same variant (same kernels, same arithmetic).
    h = taps(taps_name)
    lens = [65536, 10007, 7, 4096, 1, 3000, 70001, 12]
    xs = [noise(100 + i, n) for i, n in enumerate(lens)]
    d_x = [rf.DeviceBuffer.from_array(x) for x in xs]
    d_y = [rf.DeviceBuffer(n) for n in lens]
    d_h = rf.DeviceBuffer.from_array(h)
    idx = rf.fir_impl_index(impl)
    rf.fir_batched_device(idx, [b.ptr for b in d_x], [b.ptr for b in d_y], lens, d_h.ptr, h.size)
    for x, y in zip(xs, d_y):
        got = y.download()
        assert_fir_close(impl, got, x, h)
        np.testing.assert_array_equal(got, rf.fir(x, h, impl))


# Which kernel each variant runs: overlap-save covers 17..225 taps with a 512-point transform and
# 226..2049 with a 4096-point one, the matrix-core GEMMs 66..2049 taps, the register kernel the
# rest (a single tap is then exact, and very short filters are cheaper direct).
@pytest.mark.parametrize("num_taps,ola,mfma", [
    (1, "fir_register_kernel", "fir_register_kernel"),
    (16, "fir_register_kernel", "fir_register_kernel"),
    (17, "fir_ola512_kernel", "fir_register_kernel"),
    (63, "fir_ola512_kernel", "fir_register_kernel"),
    (65, "fir_ola512_kernel", "fir_register_kernel"),
    (66, "fir_ola512_kernel", "fir_mfma_kernel"),
    (225, "fir_ola512_kernel", "fir_mfma_kernel"),
    (226, "fir_ola_kernel", "fir_mfma_kernel"),
    (1025, "fir_ola_kernel", "fir_mfma_kernel"),
    (2049, "fir_ola_kernel", "fir_mfma_kernel"),
    (2050, "fir_register_kernel", "fir_register_kernel"),
])
def test_fir_variant_kernel_by_tap_count(num_taps: int, ola: str, mfma: str) -> None:
    assert rf.fir_kernel("fft_ola", num_taps) == ola
    assert rf.fir_kernel("mfma_f16x3", num_taps) == mfma
    assert rf.fir_kernel("mfma_bf16x3", num_taps) == mfma
    assert rf.fir_kernel("register", num_taps) == "fir_register_kernel"
    assert rf.fir_kernel("naive", num_taps) == "fir_naive_kernel"


TRANSFORM_FIR = ["fft_ola", "mfma_f16x3", "mfma_bf16x3"]


@pytest.mark.parametrize("impl", TRANSFORM_FIR)
@pytest.mark.parametrize("num_taps", [66, 226, 1025, 2049])
def test_fir_transform_variants_at_their_tap_counts(impl: str, num_taps: int) -> None:
    # The matrix-core variants at both ends of their range (66..2049 taps) and at the long-FIR
    # size, fft_ola at the same counts (its 512-point form at 66, the 4096-point form from 226):
    # against the oracle, and different in the last bits from the direct register kernel (so
    # the variant's own arithmetic ran, not the direct fallback).
    h = kaiser_lowpass(num_taps, 0.1, 8.0)
    x = noise(40 + num_taps, 30011)
    got = rf.fir(x, h, impl)
    assert_fir_close(impl, got, x, h)
    assert not np.array_equal(got, rf.fir(x, h, "register"))
    # zeros in, exactly zero out, and no NaN from the transforms or the split scaling
    z = rf.fir(np.zeros(20000, np.complex64), h, impl)
    assert not np.any(z)


@pytest.mark.parametrize("num_taps", [17, 33, 63, 65, 129, 225])
@pytest.mark.parametrize("n", [1, 449, 512, 30011])
def test_fir_short_overlap_save(num_taps: int, n: int) -> None:
    # fft_ola's 512-point form (one wavefront per segment): both ends of its tap range, captures
    # shorter than one segment, exactly one, and many; against the oracle, and different in the
    # last bits from the register kernel (so the transform ran) wherever the outputs are not
    # trivially exact. Zeros give exactly zero.
    h = kaiser_lowpass(num_taps, 0.2, 8.0)
    x = noise(60 + num_taps + n, n)
    got = rf.fir(x, h, "fft_ola")
    assert_fir_close("fft_ola", got, x, h)
    if n > 64:
        assert not np.array_equal(got, rf.fir(x, h, "register"))
    z = rf.fir(np.zeros(max(n, 2000), np.complex64), h, "fft_ola")
    assert not np.any(z)


@pytest.mark.parametrize("impl", TRANSFORM_FIR)
@pytest.mark.parametrize("num_taps,delay", [(129, 100), (1025, 777), (2049, 2048)])
def test_fir_transform_variants_delayed_impulse(impl: str, num_taps: int, delay: int) -> None:
    # h = unit impulse at `delay`: y[n] = x[n - delay] (zero for n < delay). The transform and
    # GEMM forms round (FFT: ~1e-7 of the segment's energy; FP16 split: 2^-20 relative; BF16
    # split: 2^-15), so the check is to those bounds, sample by sample; a shifted or mirrored
    # index would miss by O(1).
    h = np.zeros(num_taps, np.float32)
    h[delay] = 1.0
    x = noise(50 + delay, 20011)
    got = rf.fir(x, h, impl).astype(np.complex128)
    want = np.concatenate([np.zeros(delay), x[:-delay].astype(np.complex128)])
    tol = {"fft_ola": 2e-6, "mfma_f16x3": 2.0 ** -19, "mfma_bf16x3": 2.0 ** -14}[impl]
    err = np.abs(got - want)
    assert np.all(err <= tol * (np.abs(want) + 1.0)), float(err.max())


@pytest.mark.parametrize("num_taps,delay", [(17, 16), (40, 3), (65, 64), (65, 0), (225, 224),
                                            (160, 97)])
def test_fir_short_overlap_save_delayed_impulse(num_taps: int, delay: int) -> None:
    # The 512-point form with h = unit impulse at `delay`: y[n] = x[n - delay], to the
    # transform's rounding, sample by sample (a shifted or mirrored index would miss by O(1)),
    # across segment borders (20011 samples, 288..496 outputs per segment).
    h = np.zeros(num_taps, np.float32)
    h[delay] = 1.0
    x = noise(70 + delay, 20011)
    got = rf.fir(x, h, "fft_ola").astype(np.complex128)
    want = np.concatenate([np.zeros(delay), x[:x.size - delay].astype(np.complex128)])
    err = np.abs(got - want)
    assert np.all(err <= 2e-6 * (np.abs(want) + 1.0)), float(err.max())


@pytest.mark.parametrize("impl", ["tiled_const", "register"])
def test_fir_const_tap_limit(impl: str) -> None:
    h = np.zeros(8193, dtype=np.float32)
    with pytest.raises(rf.RfError) as e:
        rf.fir(noise(1, 100), h, impl)
    assert e.value.status == rf.RF_LIMIT


def test_fir_tiled_lds_limit_is_reported() -> None:
    # (T - 1 + B) * 8 bytes beyond a workgroup's LDS
    h = np.zeros(40000, dtype=np.float32)
    h[0] = 1
    with pytest.raises(rf.RfError) as e:
        rf.fir(noise(1, 100), h, "tiled")
    assert e.value.status == rf.RF_LIMIT


@pytest.mark.parametrize("impl", FIR_IMPLS)
def test_fir_sigmf_ci16_input(impl: str) -> None:
    x = read_ci16_sigmf("tone_ci16")
    np.testing.assert_array_equal(rf.fir(x, taps("identity"), impl), x)
    assert_fir_close(impl, rf.fir(x, taps("lowpass"), impl), x, taps("lowpass"))


@pytest.mark.parametrize("impl", FIR_IMPLS)
def test_fir_long_random_filter(impl: str) -> None:
    # 2049 taps: the longest overlap-save filter; random signs stress cancellation.
    h = (np.random.default_rng(5).standard_normal(2049) / 40).astype(np.float32)
    x = noise(6, 20000)
    assert_fir_close(impl, rf.fir(x, h, impl), x, h)


@pytest.mark.parametrize("impl", FIR_IMPLS)
def test_fir_plan_gives_the_bits_of_fir(impl: str) -> None:
    # The plan holds its own copy of the taps (the caller's buffer is zeroed after create) and
    # serves inputs of different lengths, on the default and on a created stream.
    cases = [(cf32("tone_odd"), taps("lowpass_long")), (cf32("short"), taps("lowpass")),
             (noise(21, 20011), (np.random.default_rng(22).standard_normal(300) / 20)
              .astype(np.float32))]
    stream = rf.Stream()
    for x, h in cases:
        d_h = rf.DeviceBuffer.from_array(h)
        plan = rf.FirPlan(impl, d_h.ptr, h.size)
        d_h.zero()
        for n, s in [(x.size, None), (x.size // 3 + 1, stream)]:
            d_x = rf.DeviceBuffer.from_array(x[:n])
            d_y = rf.DeviceBuffer(n)
            plan.run(d_x.ptr, d_y.ptr, n, s)
            np.testing.assert_array_equal(d_y.download(stream=s), rf.fir(x[:n], h, impl))
        plan.destroy()
    stream.destroy()


def test_fir_calls_replayed_from_a_graph() -> None:
    # A captured graph of direct-variant rf_fir calls and prepared overlap-save runs, replayed
    # three times with eager work and synchronisations in between: every replay gives the eager
    # bits.
    x1, x2 = noise(31, 30011), noise(32, 5000)
    h_short, h_long = taps("lowpass"), taps("lowpass_long")
    stream = rf.Stream()
    d_x1, d_x2 = rf.DeviceBuffer.from_array(x1), rf.DeviceBuffer.from_array(x2)
    d_hs, d_hl = rf.DeviceBuffer.from_array(h_short), rf.DeviceBuffer.from_array(h_long)
    outs = [rf.DeviceBuffer(x1.size), rf.DeviceBuffer(x2.size), rf.DeviceBuffer(x1.size)]
    plan = rf.FirPlan("fft_ola", d_hl.ptr, h_long.size)
    reg = rf.fir_impl_index("register")

    def record():
        rf.fir_device(reg, d_x1.ptr, d_hs.ptr, outs[0].ptr, x1.size, h_short.size, 0, stream.handle)
        rf.fir_device(reg, d_x2.ptr, d_hs.ptr, outs[1].ptr, x2.size, h_short.size, 0, stream.handle)
        plan.run(d_x1.ptr, outs[2].ptr, x1.size, stream.handle)

    graph = rf.Graph(stream, record)
    want = [rf.fir(x1, h_short, "register"), rf.fir(x2, h_short, "register"),
            rf.fir(x1, h_long, "fft_ola")]
    for rep in range(3):
        for o in outs:
            o.zero(stream)
        graph.launch(stream)
        stream.synchronize()
        for o, w in zip(outs, want):
            np.testing.assert_array_equal(o.download(stream=stream), w)
        rf.fir(x2, h_long, "fft_ola")  # eager work between replays
    graph.destroy()
    plan.destroy()
    stream.destroy()


def test_fir_ola_one_shot_calls_on_two_streams() -> None:
    # A one-shot overlap-save call makes its tap spectrum in its stream's workspace block, then
    # filters. Calls queued on two streams with different filters, with no synchronisation
    # between them, each equal the prepared-filter result bit for bit (each stream has its own
    # block, and calls on one stream run in order).
    shapes = [(70000, 1025), (4096, 66), (123457, 2049), (9000, 513), (300000, 1500), (5000, 200)]
    rng = np.random.default_rng(7)
    streams = [rf.Stream(), rf.Stream()]
    ola = rf.fir_impl_index("fft_ola")
    cases = []
    for i, (n, num_taps) in enumerate(shapes):
        x = noise(100 + i, n)
        h = (rng.standard_normal(num_taps) / 20).astype(np.float32)
        cases.append((x, h, rf.DeviceBuffer.from_array(x), rf.DeviceBuffer.from_array(h),
                      rf.DeviceBuffer(n)))
    for i, (x, h, d_x, d_h, d_y) in enumerate(cases):
        rf.fir_device(ola, d_x.ptr, d_h.ptr, d_y.ptr, x.size, h.size, 0, streams[i % 2].handle)
    for s in streams:
        s.synchronize()
    s = streams[0]
    for x, h, d_x, d_h, d_y in cases:
        print(f"one-shot fft_ola n={x.size} taps={h.size}", flush=True)
        plan = rf.FirPlan("fft_ola", d_h.ptr, h.size)
        want = rf.DeviceBuffer(x.size)
        plan.run(d_x.ptr, want.ptr, x.size, s.handle)
        s.synchronize()
        np.testing.assert_array_equal(d_y.download(stream=s), want.download(stream=s))
        plan.destroy()
    for st in streams:
        st.destroy()


@pytest.mark.parametrize("num_taps", [66, 300, 777, 1500])
def test_fir_ola_uneven_halo_and_lengths(num_taps: int) -> None:
    # Overlap-save segments of 4096 keep 4097 - T outputs; a halo T - 1 that is not a multiple
    # of the 256-position blocks splits one block between discarded and kept positions, and the
    # first segment reads before x[0]. Lengths: shorter than the halo, one segment plus a few,
    # and many segments.
    h = (np.random.default_rng(num_taps).standard_normal(num_taps) / 20).astype(np.float32)
    lens = [1, num_taps - 1, 4097 - num_taps + 5, 20011]
    for i, n in enumerate(lens):
        x = noise(300 + i, n)
        assert_fir_close("fft_ola", rf.fir(x, h, "fft_ola"), x, h)
    xs = [noise(400 + i, n) for i, n in enumerate(lens)]
    d_x = [rf.DeviceBuffer.from_array(x) for x in xs]
    d_y = [rf.DeviceBuffer(n) for n in lens]
    d_h = rf.DeviceBuffer.from_array(h)
    rf.fir_batched_device(rf.fir_impl_index("fft_ola"), [b.ptr for b in d_x],
                          [b.ptr for b in d_y], lens, d_h.ptr, h.size)
    for x, y in zip(xs, d_y):
        assert_fir_close("fft_ola", y.download(), x, h)


@pytest.mark.parametrize("impl", ["mfma_f16x3", "mfma_bf16x3"])
def test_fir_wide_dynamic_range(impl: str) -> None:
    # Samples of 1e-6 in the first half and 1e4 in the second, taps from 1e-9 to 1e-1: the FP16
    # form scales taps and each workgroup's samples by powers of two. Outputs whose workgroup
    # saw only one of the two levels keep the per-output bound.
    rng = np.random.default_rng(9)
    n = 40000
    x = noise(10, n) * np.float32(1e-6)
    x[n // 2:] *= np.float32(1e10)
    h = (rng.standard_normal(300) * np.logspace(-9, -1, 300)).astype(np.float32)
    got = rf.fir(x, h, impl).astype(np.complex128)
    want = oracle.fir(x, h)
    bound = split_bound(impl, x, h)
    keep = np.r_[0:n // 2 - 8192, n // 2 + 8192:n]
    err = np.abs(got - want)[keep]
    assert np.all(err <= bound[keep] + 1e-35), float(np.max(err / bound[keep]))


def test_fir_config_plain_2e20_x_63() -> None:
    # The CPU-runnable workload: 2^20 samples, 63-tap Hamming sinc (cutoff 0.2 fs).
    n = 1 << 20
    t = np.arange(n)
    x = noise(0xC1, n) / np.sqrt(2) + sum(np.exp(2j * np.pi * f * t) for f in (0.05, 0.17, 0.31))
    x = x.astype(np.complex64)
    m = np.arange(63) - 31
    h = 0.4 * np.sinc(0.4 * m) * np.hamming(63)
    h = (h / h.sum()).astype(np.float32)
    want = oracle.fir(x, h)
    for impl in FIR_IMPLS:
        assert_fir_close(impl, rf.fir(x, h, impl), x, h, want)


def test_fir_autotuner_configs_give_identical_bits() -> None:
    """rf_fir_tune times the direct variants' tile shapes; every shape, and the tuned one, gives
    the same bits as rf_fir (same per-output summation order)."""
    rng = np.random.default_rng(11)
    n, t = 1 << 18, 255
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    h = rng.standard_normal(t).astype(np.float32)
    d_x = rf.DeviceBuffer.from_array(x)
    d_h = rf.DeviceBuffer.from_array(h)
    d_y = rf.DeviceBuffer(n)
    rf.fir_device(rf.fir_impl_index("register"), d_x.ptr, d_h.ptr, d_y.ptr, n, t)
    want = d_y.download()
    best = rf.fir_tune(d_x.ptr, d_h.ptr, d_y.ptr, n, t, ["register", "tiled", "naive"])
    assert best.impl in (rf.fir_impl_index(v) for v in ("register", "tiled", "naive"))
    assert best.ms > 0 and best.block in (64, 128, 256, 512)
    d_y.zero()
    rf.fir_run_config(best, d_x.ptr, d_h.ptr, d_y.ptr, n, t)
    assert d_y.download().tobytes() == want.tobytes()
    for block in (64, 512):
        for r in (4, 8, 16):
            cfg = rf.FirConfig(rf.fir_impl_index("register"), block, r, 0.0)
            d_y.zero()
            rf.fir_run_config(cfg, d_x.ptr, d_h.ptr, d_y.ptr, n, t)
            assert d_y.download().tobytes() == want.tobytes(), (block, r)
    with pytest.raises(rf.RfError):
        rf.fir_run_config(rf.FirConfig(rf.fir_impl_index("register"), 256, 5, 0.0),
                          d_x.ptr, d_h.ptr, d_y.ptr, n, t)


@pytest.mark.parametrize("num_taps", [17, 129, 300, 1025])
def test_fir_autotuner_picks_the_transform_length(num_taps: int) -> None:
    """fft_ola's tuner candidates are its 512- and 4096-point forms wherever the tap count
    allows them: the tuned configuration and each forced length match the oracle, and the length
    rf_fir picks for the tap count gives rf_fir's bits."""
    n = 100003
    x = noise(80 + num_taps, n)
    h = kaiser_lowpass(num_taps, 0.1, 8.0)
    d_x = rf.DeviceBuffer.from_array(x)
    d_h = rf.DeviceBuffer.from_array(h)
    d_y = rf.DeviceBuffer(n)
    want = oracle.fir(x, h)
    default = {"fir_ola512_kernel": 512, "fir_ola_kernel": 4096}[rf.fir_kernel("fft_ola", num_taps)]
    ola = rf.fir_impl_index("fft_ola")
    best = rf.fir_tune(d_x.ptr, d_h.ptr, d_y.ptr, n, num_taps, ["fft_ola"])
    assert best.impl == ola and best.ms > 0
    assert best.transform_points in ([512, 4096] if num_taps <= 449 else [4096])
    for points in sorted({best.transform_points, 512, 4096}):
        if points == 512 and num_taps > 449:
            continue
        d_y.zero()
        rf.fir_run_config(rf.FirConfig(ola, 256, 0, 0.0, points), d_x.ptr, d_h.ptr, d_y.ptr, n,
                          num_taps)
        got = d_y.download()
        np.testing.assert_allclose(got, want, rtol=FIR_RTOL, atol=FIR_ATOL)
        if points == default:
            assert got.tobytes() == rf.fir(x, h, "fft_ola").tobytes()
    for b in (d_x, d_h, d_y):
        b.free()


@pytest.mark.parametrize("impl", ["register", "fft_ola", "mfma_f16x3", "mfma_bf16x3"])
def test_fir_full_size_long_filter_sampled_outputs(impl: str) -> None:
    # 2^24 samples x 1025 Kaiser taps: 3000 outputs recomputed in float64 on the host.
    n = 1 << 24
    x = noise(0xC2, n) / np.float32(np.sqrt(2))
    h = kaiser_lowpass(1025, 0.1, 8.0)
    y = rf.fir(x, h, impl)
    rng = np.random.default_rng(3)
    idx = np.concatenate([np.arange(0, 1100), rng.integers(0, n, 1900), [n - 1]])
    k = np.arange(1025)
    src = idx[:, None] - k[None, :]
    xv = np.where(src >= 0, x[np.clip(src, 0, None)].astype(np.complex128), 0)
    want = xv @ h.astype(np.float64)
    if impl in REDUCED_BOUND:
        scale = np.abs(xv) @ np.abs(h.astype(np.float64))
        assert np.all(np.abs(y[idx] - want) <= REDUCED_BOUND[impl] * scale + 1e-7)
    else:
        np.testing.assert_allclose(y[idx], want, rtol=FIR_RTOL, atol=FIR_ATOL)


# --- cross-correlation ---------------------------------------------------------------------

@pytest.mark.parametrize("impl", XCORR_IMPLS)
@pytest.mark.parametrize("a,b", [("ch0", "ch1"), ("ch1", "ch0"), ("tone_odd", "short"),
                                 ("short", "tone_odd"), ("short", "short")])
def test_xcorr_matches_oracle(impl: str, a: str, b: str) -> None:
    xa, xb = cf32(a), cf32(b)
    got = rf.xcorr(xa, xb, impl)
    want = oracle.xcorr(xa, xb)
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=XCORR_RTOL, atol=XCORR_ATOL)


@pytest.mark.parametrize("impl", XCORR_IMPLS)
def test_xcorr_recovers_the_known_delay(impl: str) -> None:
    r = rf.xcorr(cf32("ch0"), cf32("ch1"), impl)
    lags = oracle.xcorr_lags(16384, 16384)
    assert lags[np.argmax(np.abs(r))] == meta("ch1")["true_delay_samples"]


@pytest.mark.parametrize("impl", XCORR_IMPLS)
def test_xcorr_of_zeros_is_exactly_zero(impl: str) -> None:
    got = rf.xcorr(cf32("zeros"), cf32("zeros"), impl)
    assert got.shape == (2 * 4096 - 1,) and not np.any(got)


@pytest.mark.parametrize("na,nb", [(1, 1), (1, 2), (2, 3), (3, 1), (5, 4), (100, 3)])
def test_xcorr_fft_tiny_lengths(na: int, nb: int) -> None:
    a, b = noise(na, na), noise(nb + 50, nb)
    np.testing.assert_allclose(rf.xcorr(a, b, "fft"), oracle.xcorr(a, b), rtol=XCORR_RTOL,
                               atol=1e-5)
    np.testing.assert_allclose(rf.gcc_phat(a, b), oracle.gcc_phat(a, b), atol=GCC_ATOL)


# --- GCC-PHAT, peak, DOA -------------------------------------------------------------------

@pytest.mark.parametrize("a,b", [("ch0", "ch1"), ("array0", "array3"), ("short", "tone_odd"),
                                 ("tone", "tone_odd"), ("tone", "tone")])
def test_gcc_phat_matches_oracle(a: str, b: str) -> None:
    xa, xb = cf32(a), cf32(b)
    got = rf.gcc_phat(xa, xb)
    want = oracle.gcc_phat(xa, xb)
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=0, atol=GCC_ATOL)


def test_gcc_phat_long_pair_matches_oracle() -> None:
    # L = 2^20: the four-step path with 256 rows of 4096.
    a = noise(41, 300000)
    b = np.concatenate([np.zeros(123, np.complex64), a[:-123]]) + 0.3 * noise(42, 300000)
    got = rf.gcc_phat(a, b.astype(np.complex64))
    want = oracle.gcc_phat(a, b)
    np.testing.assert_allclose(got, want, rtol=0, atol=GCC_ATOL)
    assert np.argmax(np.abs(got)) - (a.size - 1) == 123


def test_gcc_phat_peaks_at_the_known_delay() -> None:
    r = rf.gcc_phat(cf32("ch0"), cf32("ch1"))
    lags = oracle.xcorr_lags(16384, 16384)
    assert lags[np.argmax(np.abs(r))] == meta("ch1")["true_delay_samples"]
    peak = np.max(np.abs(r))
    assert peak > 20 * np.sort(np.abs(r))[-2]


def test_gcc_phat_of_zeros_is_exactly_zero() -> None:
    got = rf.gcc_phat(cf32("zeros"), cf32("zeros"))
    assert got.shape == (2 * 4096 - 1,) and not np.any(got)


def test_device_peak_search_takes_the_first_maximum() -> None:
    r = noise(8, 100000)
    r[777] = 50 + 0j
    r[90000] = 0 + 50j  # same magnitude, later index
    d = rf.DeviceBuffer.from_array(r)
    res = rf.peak(d.ptr, r.size)
    assert res.index == 777 and res.magnitude == pytest.approx(50.0)
    z = rf.DeviceBuffer.from_array(np.zeros(5000, np.complex64))
    res = rf.peak(z.ptr, 5000)
    assert res.index == 0 and res.magnitude == 0.0


def test_device_peak_interpolates_a_sampled_parabola() -> None:
    k = np.arange(2000, dtype=np.float64)
    offset = 0.3
    r = (100.0 - (k - 1000 - offset) ** 2).clip(0).astype(np.complex64)
    d = rf.DeviceBuffer.from_array(r)
    res = rf.peak(d.ptr, r.size)
    assert res.index == 1000
    assert res.frac == pytest.approx(offset, abs=1e-5)


def test_doa_recovers_the_array_angle() -> None:
    m = meta("array0")
    chans = [cf32(f"array{i}") for i in range(m["channels"])]
    report = rf.doa(chans, m["spacing_m"], m["speed_m_s"], m["sample_rate"])
    assert report["tdoa_samples"] == [2, 4, 6]
    assert report["angle_deg"] == pytest.approx(m["angle_deg"], abs=1e-3)
    ref = oracle.doa(chans, m["spacing_m"], m["speed_m_s"], m["sample_rate"])
    assert report["angle_deg"] == pytest.approx(ref, abs=1e-6)


def test_doa_batch_lags_match_the_oracle() -> None:
    # DOA's lags come from one batched GCC-PHAT over all channel pairs; pairs with at most 2048
    # output lags are too short for the batch and take the per-pair path. Every lag equals the
    # oracle's GCC-PHAT peak lag and the per-pair GPU path's.
    rng = np.random.default_rng(61)
    src = noise(62, 20000)
    ref = src[1000:2000]                                      # 1000 samples
    chans = [ref]
    for n, d in [(900, 7), (1500, -33), (6000, 120), (1049, 0)]:
        c = src[1000 - d:1000 - d + n] + 0.3 * noise(int(rng.integers(1 << 30)), n)
        chans.append(c.astype(np.complex64))
    lags = rf.tdoa_batch(chans)
    for c, lag in zip(chans[1:], lags):
        idx, _ = oracle.peak(oracle.gcc_phat(ref, c))
        assert lag == idx - (ref.size - 1)
        assert lag == rf.tdoa(ref, c)
    assert lags[:3] == [7, -33, 120]


# --- channelizer ---------------------------------------------------------------------------

@pytest.mark.parametrize("use_plan", [False, True])
@pytest.mark.parametrize("name,taps_name,channels", [
    ("tone", "channelizer_proto", 16), ("tone_odd", "channelizer_proto", 16),
    ("tone", "lowpass_long", 64), ("short", "lowpass", 4), ("tone_odd", "identity", 1),
    ("tone", "lowpass_long", 8), ("tone_odd", "lowpass", 6), ("tone", "lowpass_long", 1024),
    ("tone", "channelizer_proto", 2), ("tone_odd", "lowpass_long", 12),
])
def test_channelize_matches_oracle(name: str, taps_name: str, channels: int,
                                   use_plan: bool) -> None:
    x, h = cf32(name), taps(taps_name)
    got = rf.channelize(x, h, channels, use_plan=use_plan)
    want = oracle.channelize(x, h, channels).ravel()
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=CHANNELIZE_RTOL, atol=CHANNELIZE_ATOL)


def test_channelize_puts_a_tone_in_its_channel() -> None:
    n = np.arange(16 * 512)
    x = np.exp(2j * np.pi * 3 * n / 16).astype(np.complex64)
    z = rf.channelize(x, taps("channelizer_proto"), 16).reshape(-1, 16)[16:]
    power = np.mean(np.abs(z) ** 2, axis=0)
    assert np.argmax(power) == 3
    assert power[3] > 1000 * np.max(np.delete(power, 3))
    np.testing.assert_allclose(np.abs(z[:, 3]), 1.0, atol=0.02)


def test_channelize_config_shape_64_channels_1024_taps() -> None:
    # The channelizer workload's filterbank (M = 64, 16 taps per branch) on 2^20 samples.
    x = noise(0xC3, 1 << 20)
    h = kaiser_lowpass(1024, 1 / 128, 8.0)
    got = rf.channelize(x, h, 64)
    want = oracle.channelize(x, h, 64).ravel()
    np.testing.assert_allclose(got, want, rtol=CHANNELIZE_RTOL, atol=CHANNELIZE_ATOL)


# --- channelizer, DFT on the matrix cores -----------------------------------------------------

def channelize_term_sum(x: np.ndarray, h: np.ndarray, channels: int) -> np.ndarray:
    """S[t] = sum_l |h[l]| |x[tM - l]|: the scale of frame t's terms (float64)."""
    conv = np.convolve(np.abs(x.astype(np.complex128)), np.abs(h.astype(np.float64)))
    frames = -(-x.size // channels)
    return conv[np.arange(frames) * channels]


@pytest.mark.parametrize("name,taps_name", [("tone", "channelizer_proto"), ("tone_odd", "lowpass"),
                                            ("short", "identity"), ("ch0", "channelizer_proto")])
def test_channelize_mfma_matches_oracle(name: str, taps_name: str) -> None:
    x, h = cf32(name), taps(taps_name)
    got = rf.channelize_mfma(x, h, 64).reshape(-1, 64)
    want = oracle.channelize(x, h, 64)
    assert got.shape == want.shape
    bound = 2.0 ** -15 * channelize_term_sum(x, h, 64)[:, None]
    assert np.all(np.abs(got - want) <= bound + 1e-30), float(np.max(np.abs(got - want) / (bound + 1e-30)))


def test_channelize_mfma_config_shape_and_dynamic_range() -> None:
    """2^20 samples of the C3 kind (64 channels, 1024-tap Kaiser prototype), QPSK carriers over
    a 40 dB power range: per-output error within the stated bound, also for weak frames."""
    rng = np.random.default_rng(0xC3)
    n = 1 << 20
    t = np.arange(n)
    x = np.zeros(n, np.complex128)
    for k, p_db in zip(rng.choice(64, 8, replace=False), rng.uniform(-40, 0, 8)):
        sym = (rng.choice([-1, 1], n // 128 + 1) + 1j * rng.choice([-1, 1], n // 128 + 1)) / np.sqrt(2)
        x += 10 ** (p_db / 20) * np.repeat(sym, 128)[:n] * np.exp(2j * np.pi * k * t / 64)
    x += 1e-2 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    x = x.astype(np.complex64)
    m = np.arange(1024) - 511.5
    h = np.sinc(m / 128) * np.kaiser(1024, 8.0)
    h = (h / h.sum()).astype(np.float32)
    got = rf.channelize_mfma(x, h, 64).reshape(-1, 64)
    frames = np.sort(rng.choice(got.shape[0], 64, replace=False))
    xs = x.astype(np.complex128)
    for f in frames:
        l = np.arange(1024)
        idx = f * 64 - l
        ok = idx >= 0
        terms = np.zeros(1024, np.complex128)
        terms[ok] = h[ok] * xs[idx[ok]]
        want = np.array([np.sum(terms * np.exp(2j * np.pi * k * l / 64)) for k in range(64)])
        bound = 2.0 ** -15 * np.sum(np.abs(terms))
        assert np.all(np.abs(got[f] - want) <= bound + 1e-30), f
    fused = rf.channelize(x, h, 64).reshape(-1, 64)
    np.testing.assert_allclose(got, fused, rtol=0, atol=2e-4 * np.max(np.abs(fused)))


@pytest.mark.parametrize("n,num_taps", [(1000, 64), (4097, 300), (70001, 1000), (65536, 1024),
                                        (300000, 17)])
def test_channelize_64_branch_shapes(n: int, num_taps: int) -> None:
    # M = 64 with 1..16 taps per branch: the fixed-shape staging of the fused kernel (first
    # workgroup reaching before x[0], last one past x[n-1], inputs shorter than a workgroup) and
    # the matrix-core kernel over many frame blocks in XCD-major order.
    x = noise(n, n)
    h = (np.random.default_rng(num_taps).standard_normal(num_taps) / 10).astype(np.float32)
    want = oracle.channelize(x, h, 64)
    got = rf.channelize(x, h, 64).reshape(-1, 64)
    np.testing.assert_allclose(got, want, rtol=CHANNELIZE_RTOL, atol=CHANNELIZE_ATOL)
    got_mfma = rf.channelize_mfma(x, h, 64).reshape(-1, 64)
    bound = 2.0 ** -15 * channelize_term_sum(x, h, 64)[:, None]
    assert np.all(np.abs(got_mfma - want) <= bound + 1e-30)


def test_channelize_mfma_limits() -> None:
    x, h = cf32("tone"), taps("lowpass_long")
    with pytest.raises(rf.RfError):
        rf.channelize_mfma(x, h, 32)                  # only 64 channels
    with pytest.raises(rf.RfError):
        rf.channelize_mfma(x, np.ones(1025, np.float32), 64)   # more than 16 taps per branch
