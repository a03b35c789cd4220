"""Batched GCC-PHAT (rf_gcc_batch_*) against the float64 oracle and the single-pair path.

Every pair's full output is compared with the oracle (tolerance of the reference's GCC-PHAT
tests, atol 1e-4, tests/test_pipeline.py:408-411 of the reference); the fused peak must be the
first maximum of the GPU's own output and the oracle's peak lag (bit-exact, with a stated tie
rule); graph replays must reproduce the eager run exactly.
"""
from __future__ import annotations

import numpy as np
import pytest

from _common import load_rfhip, oracle

pytestmark = pytest.mark.gpu
rf = load_rfhip()

GCC_ATOL = 1e-4


def delayed_pair(seed: int, na: int, nb: int, delay: int, noise: float):
    # b[n] = a's source delayed by `delay`: b[n] = src[n + D - delay], a[n] = src[n + D].
    rng = np.random.default_rng(seed)
    total = max(na, nb) + 2 * abs(delay) + 8
    src = rng.standard_normal(total) + 1j * rng.standard_normal(total)
    a = src[abs(delay): abs(delay) + na]
    b = src[abs(delay) - delay: abs(delay) - delay + nb]
    a = a + noise * (rng.standard_normal(na) + 1j * rng.standard_normal(na))
    b = b + noise * (rng.standard_normal(nb) + 1j * rng.standard_normal(nb))
    return a.astype(np.complex64), b.astype(np.complex64)


LENGTHS = [(1500, 1200, 17), (2100, 2000, 13), (4096, 4096, -77), (5000, 3000, 150), (9000, 9001, 0),
           (20000, 16000, -200), (33000, 33000, 199), (70000, 65536, 5), (131072, 131072, -1),
           (300000, 200000, 123), (1 << 20, 1 << 20, -150), (3000, 70000, 40), (50000, 2500, -33),
           # FFT length 2^15 (the single-workgroup path): both channels at most 2^14, one channel
           # longer than 2^14 (a, then b), and the shortest and longest lag counts of that length;
           # 12000 + 4385 - 1 = 2^14 lags stays on the four-step path
           (1 << 14, 1 << 14, 77), (20000, 5000, -120), (3000, 29000, 33), (12000, 4386, 5),
           (32000, 769, 10), (12000, 4385, -9)]
# (1500, 1200): FFT length 4096, whose forward column tile (512 columns of 8) fills half of
# the forward pass's 1024 lanes


@pytest.fixture(scope="module")
def batch():
    pairs = [delayed_pair(i, na, nb, d, 0.3) for i, (na, nb, d) in enumerate(LENGTHS)]
    d_a = [rf.DeviceBuffer.from_array(a) for a, _ in pairs]
    d_b = [rf.DeviceBuffer.from_array(b) for _, b in pairs]
    d_r = [rf.DeviceBuffer(a.size + b.size - 1) for a, b in pairs]
    h = rf.GccBatch([a.size for a, _ in pairs], [b.size for _, b in pairs])
    h.bind([x.ptr for x in d_a], [x.ptr for x in d_b], [x.ptr for x in d_r])
    res = rf.DeviceBuffer(16 * len(pairs), np.uint8)
    h.run(res.ptr)
    peaks = rf.read_peaks(res, len(pairs))
    outs = [r.download() for r in d_r]
    return pairs, h, d_a, d_b, d_r, res, peaks, outs


def test_batched_outputs_match_the_oracle(batch) -> None:
    pairs, _, _, _, _, _, _, outs = batch
    for (a, b), got in zip(pairs, outs):
        np.testing.assert_allclose(got, oracle.gcc_phat(a, b), rtol=0, atol=GCC_ATOL)


def test_fused_peak_is_the_first_maximum_and_the_oracle_lag(batch) -> None:
    pairs, _, _, _, _, _, peaks, outs = batch
    for (a, b), (na, nb, delay), pk, got in zip(pairs, LENGTHS, peaks, outs):
        mags = np.abs(got.astype(np.complex128))
        assert pk.index == int(np.argmax(np.abs(got)))
        ref_idx, ref_mag = oracle.peak(oracle.gcc_phat(a, b))
        # Tie rule: a different index is accepted only if the oracle's float64 magnitudes
        # of the two candidates differ by less than the float32 error bound.
        if pk.index != ref_idx:
            ref = np.abs(oracle.gcc_phat(a, b))
            assert abs(ref[pk.index] - ref[ref_idx]) < 2 * GCC_ATOL
        assert pk.index - (na - 1) == delay
        assert pk.magnitude == pytest.approx(mags[pk.index], rel=1e-6)


def test_longest_fft_lengths_match_the_oracle() -> None:
    """Pairs whose FFT lengths are 2^22, 2^23 and 2^24 points (the batch's limit): column
    transforms of 1024, 2048 and 4096 rows, which the pairs above do not reach. Full outputs
    against the oracle, fused peak at the oracle's lag and the planted delay."""
    shapes = [((1 << 21), (1 << 21) - 5, 123), (3_000_000, 2_500_000, -77),
              (6_000_000, 5_000_000, 45)]
    pairs = [delayed_pair(100 + i, na, nb, d, 0.3) for i, (na, nb, d) in enumerate(shapes)]
    d_a = [rf.DeviceBuffer.from_array(a) for a, _ in pairs]
    d_b = [rf.DeviceBuffer.from_array(b) for _, b in pairs]
    d_r = [rf.DeviceBuffer(a.size + b.size - 1) for a, b in pairs]
    h = rf.GccBatch([a.size for a, _ in pairs], [b.size for _, b in pairs])
    h.bind([x.ptr for x in d_a], [x.ptr for x in d_b], [x.ptr for x in d_r])
    res = rf.DeviceBuffer(16 * len(pairs), np.uint8)
    print("gcc batch: FFT lengths 2^22, 2^23, 2^24", flush=True)
    h.run(res.ptr)
    peaks = rf.read_peaks(res, len(pairs))
    for (a, b), (na, nb, delay), pk, d in zip(pairs, shapes, peaks, d_r):
        got = d.download()
        want = oracle.gcc_phat(a, b)
        np.testing.assert_allclose(got, want, rtol=0, atol=GCC_ATOL)
        ref_idx, _ = oracle.peak(want)
        assert pk.index == ref_idx == int(np.argmax(np.abs(got)))
        assert pk.index - (na - 1) == delay
    for x in d_a + d_b + d_r + [res]:
        x.free()
    h.destroy()


def test_single_pair_path_agrees(batch) -> None:
    pairs, _, _, _, _, _, peaks, outs = batch
    for (a, b), pk, got in list(zip(pairs, peaks, outs))[:6]:
        single = rf.gcc_phat(a, b)
        np.testing.assert_allclose(got, single, rtol=0, atol=GCC_ATOL)
        assert int(np.argmax(np.abs(single))) == pk.index


def test_graph_replays_match_the_eager_run(batch) -> None:
    pairs, h, d_a, d_b, d_r, res, peaks, outs = batch
    want = res.download().tobytes()
    g_res = rf.DeviceBuffer(16 * len(pairs), np.uint8)
    for replay in range(4):
        g_res.zero()
        rf.synchronize()
        h.run(g_res.ptr, graph=True)
        rf.synchronize()
        assert g_res.download().tobytes() == want, f"replay {replay}"
        # eager work between replays
        h.run(res.ptr)
        rf.synchronize()
        assert res.download().tobytes() == want


def test_peaks_only_mode_gives_the_same_peaks(batch) -> None:
    pairs, h, d_a, d_b, d_r, res, peaks, outs = batch
    h2 = rf.GccBatch([a.size for a, _ in pairs], [b.size for _, b in pairs])
    h2.bind([x.ptr for x in d_a], [x.ptr for x in d_b])
    r2 = rf.DeviceBuffer(16 * len(pairs), np.uint8)
    h2.run(r2.ptr)
    assert [(p.index, p.magnitude) for p in rf.read_peaks(r2, len(pairs))] == \
        [(p.index, p.magnitude) for p in peaks]
    h2.destroy()


def test_fractional_delay_estimate() -> None:
    # A pure fractional delay of a white full-band signal: b = a shifted by d in frequency.
    n = 1 << 14
    rng = np.random.default_rng(77)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    truths = [10.3, -20.45, 7.0, 3.71, -0.25]
    pairs = []
    for d in truths:
        f = np.fft.fftfreq(n)
        b = np.fft.ifft(np.fft.fft(a) * np.exp(-2j * np.pi * f * d))
        pairs.append((a.astype(np.complex64), b.astype(np.complex64)))
    d_a = [rf.DeviceBuffer.from_array(x) for x, _ in pairs]
    d_b = [rf.DeviceBuffer.from_array(y) for _, y in pairs]
    h = rf.GccBatch([n] * len(pairs), [n] * len(pairs))
    h.bind([x.ptr for x in d_a], [x.ptr for x in d_b])
    res = rf.DeviceBuffer(16 * len(pairs), np.uint8)
    h.run(res.ptr)
    for d, pk in zip(truths, rf.read_peaks(res, len(pairs))):
        est = pk.index - (n - 1) + pk.frac
        assert est == pytest.approx(d, abs=0.1), (d, est)
    h.destroy()


def test_autotuned_grid_gives_the_same_results(batch) -> None:
    pairs, h, d_a, d_b, d_r, res, peaks, outs = batch
    want = res.download().tobytes()
    best = h.tune(res.ptr)
    assert best in (1, 2, 3, 4, 6, 8)
    h.run(res.ptr)
    rf.synchronize()
    assert res.download().tobytes() == want


def test_short_pairs_are_refused() -> None:
    with pytest.raises(rf.RfError) as e:
        rf.GccBatch([1000, 5000], [1000, 5000])
    assert e.value.status == rf.RF_LIMIT


def test_group_schedules_give_identical_results(batch) -> None:
    """The pairs run in groups sized to a byte budget; every schedule gives the same bits."""
    pairs, h, _, _, d_r, _, peaks, outs = batch
    want = [(p.index, p.magnitude, p.frac) for p in peaks]
    for budget, groups_at_least in ((0, 1), (1 << 20, 2), (96 << 20, 1)):
        h.set_group_bytes(budget)
        assert h.group_bytes()[0] == budget
        if budget == 0:
            assert h.group_bytes()[1] == 1
        assert h.group_bytes()[1] >= groups_at_least
        res = rf.DeviceBuffer(16 * len(pairs), np.uint8)
        h.run(res.ptr)
        got = rf.read_peaks(res, len(pairs))
        assert [(p.index, p.magnitude, p.frac) for p in got] == want
        for r, o in zip(d_r, outs):
            np.testing.assert_array_equal(r.download(), o)


def test_one_sample_channels_of_fft_length_2e15() -> None:
    """A channel of one sample beside one of 16385..20000 (FFT length 2^15, the single-workgroup
    path): every output against the oracle and the oracle's peak (same tie rule as above)."""
    rng = np.random.default_rng(15)
    shapes = [(1, 20000), (17000, 1), (16385, 1), (1, 16385)]
    pairs = []
    for na, nb in shapes:
        a = (rng.standard_normal(na) + 1j * rng.standard_normal(na)).astype(np.complex64)
        b = (rng.standard_normal(nb) + 1j * rng.standard_normal(nb)).astype(np.complex64)
        pairs.append((a, b))
    d_a = [rf.DeviceBuffer.from_array(a) for a, _ in pairs]
    d_b = [rf.DeviceBuffer.from_array(b) for _, b in pairs]
    d_r = [rf.DeviceBuffer(a.size + b.size - 1) for a, b in pairs]
    h = rf.GccBatch([a.size for a, _ in pairs], [b.size for _, b in pairs])
    h.bind([x.ptr for x in d_a], [x.ptr for x in d_b], [x.ptr for x in d_r])
    res = rf.DeviceBuffer(16 * len(pairs), np.uint8)
    h.run(res.ptr)
    for (a, b), pk, r in zip(pairs, rf.read_peaks(res, len(pairs)), d_r):
        want = oracle.gcc_phat(a, b)
        got = r.download()
        np.testing.assert_allclose(got, want, rtol=0, atol=GCC_ATOL)
        ref_idx, _ = oracle.peak(want)
        if pk.index != ref_idx:
            ref = np.abs(want)
            assert abs(ref[pk.index] - ref[ref_idx]) < 2 * GCC_ATOL
        assert pk.magnitude == pytest.approx(float(np.abs(got[pk.index])), rel=1e-6)
    h.destroy()
