"""The benchmark configurations at their full sizes, against the float64 oracle.

  C3  polyphase channelizer, 2^24 samples (40 QPSK carriers), M = 64, 1024-tap Kaiser
      prototype: the first frames and 96 random frames of the fused FP32 kernel and of the
      split-FP16 MFMA form against float64 sums, at the reference's channelizer tolerance
      (rtol 1e-5, atol 1e-4; tests/test_pipeline.py:360-364 of the reference) and the MFMA
      form's stated per-output bound;
  C4  batched GCC-PHAT, all 256 uneven pairs (2^12..2^20 samples, fractional delays), eager
      and replayed from the hipGraph: every peak index equal to the oracle's (the fixture
      tests/golden/gcc_batch_c4.json, written by make_gcc_batch_c4.py; no near-ties: the
      smallest peak margin there is 0.80) and every sub-sample offset within 1e-3 of the
      oracle's float64 interpolation of its own |r|;
  C5  learned delay search, 4096 pairs of 2^14 samples, network trained on 2^20 pairs: every
      lag equal to the exact batched GCC-PHAT's, and sampled pairs equal to the oracle's exact
      GCC-PHAT peak (oracle/delay.py); and, on the 2048 hardest-SNR pairs, what the network
      gains over the exact method without it: the same lags with at most 5 % of the pairs
      taking the full inverse transform.

Inputs come from bench.py's generators (seeded numpy) and from rf_synth_pairs (Philox on the
GPU, restated in oracle/delay.py).
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np
import pytest

from _common import ROOT, load_rfhip, oracle

pytestmark = pytest.mark.gpu
rf = load_rfhip()
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402  (the workload generators)

FIXTURE = ROOT / "tests" / "golden" / "gcc_batch_c4.json"


# --- C3 --------------------------------------------------------------------------------------

def channelizer_sums(x: np.ndarray, h: np.ndarray, m: int, frames: np.ndarray) -> tuple:
    """float64 z[f][k] = sum_l h[l] x[f M - l] e^{+2 pi i k l / M} and sum_l |h[l] x[f M - l]|
    for the given frames (tools/reference.py:128-146 of the reference)."""
    xs = x.astype(np.complex128)
    l = np.arange(h.size)
    dft = np.exp(2j * np.pi * np.outer(l, np.arange(m)) / m)
    want, scale = [], []
    for f in frames:
        idx = f * m - l
        ok = idx >= 0
        terms = np.zeros(h.size, np.complex128)
        terms[ok] = h[ok].astype(np.float64) * xs[idx[ok]]
        want.append(terms @ dft)
        scale.append(np.sum(np.abs(terms)))
    return np.array(want), np.array(scale)


def test_c3_channelizer_full_size() -> None:
    n, m, t = 1 << 24, 64, 1024
    x = bench.channelize_input(n)
    h = bench.kaiser_lowpass(t, 1.0 / 128, 8.0)
    frames = rf.channelize_frames(n, m)
    d_x, d_h = rf.DeviceBuffer.from_array(x), rf.DeviceBuffer.from_array(h)
    d_z = rf.DeviceBuffer(frames * m)
    check = np.concatenate([np.arange(20), np.random.default_rng(9).choice(frames, 96, False)])
    want, scale = channelizer_sums(x, h, m, check)
    print(f"C3 fused: {n} samples, {frames} frames", flush=True)
    rf.channelize_device(None, d_x.ptr, d_h.ptr, d_z.ptr, n, m, t)
    z = d_z.download().reshape(frames, m)
    np.testing.assert_allclose(z[check], want, rtol=1e-5, atol=1e-4)
    print("C3 split-FP16 MFMA", flush=True)
    d_z.zero()
    rf.channelize_mfma_device(d_x.ptr, d_h.ptr, d_z.ptr, n, m, t)
    z = d_z.download().reshape(frames, m)
    err = np.abs(z[check].astype(np.complex128) - want)
    assert np.all(err <= 2.0 ** -15 * scale[:, None] + 1e-7), float(err.max())
    for b in (d_x, d_h, d_z):
        b.free()


# --- C4 --------------------------------------------------------------------------------------

def test_c4_all_256_pairs_eager_and_graph() -> None:
    pairs, _ = bench.gcc_batch_input()
    want = json.loads(FIXTURE.read_text())
    na = [int(a.size) for a, _ in pairs]
    nb = [int(b.size) for _, b in pairs]
    assert na == want["num_a"] and nb == want["num_b"]
    d_a = [rf.DeviceBuffer.from_array(a) for a, _ in pairs]
    d_b = [rf.DeviceBuffer.from_array(b) for _, b in pairs]
    h = rf.GccBatch(na, nb)
    h.bind([x.ptr for x in d_a], [x.ptr for x in d_b])
    eager = rf.DeviceBuffer(16 * len(pairs), np.uint8)
    print(f"C4 eager: {len(pairs)} pairs, {sum(na) + sum(nb)} samples", flush=True)
    h.run(eager.ptr)
    rf.synchronize()
    peaks = rf.read_peaks(eager, len(pairs))
    got_index = [p.index for p in peaks]
    assert got_index == want["index"]
    frac_err = np.abs(np.array([p.frac for p in peaks]) - np.array(want["frac"]))
    assert frac_err.max() < 1e-3, float(frac_err.max())
    graph = rf.DeviceBuffer(16 * len(pairs), np.uint8)
    for replay in range(3):
        print(f"C4 graph replay {replay}", flush=True)
        graph.zero()
        rf.synchronize()
        h.run(graph.ptr, graph=True)
        rf.synchronize()
        assert graph.download().tobytes() == eager.download().tobytes()
    h.destroy()
    for b in [eager, graph, *d_a, *d_b]:
        b.free()


# --- C5 --------------------------------------------------------------------------------------

sys.path.insert(0, str(ROOT / "oracle"))
import delay  # noqa: E402  (the oracle's restatement of the delay search)


@pytest.fixture(scope="module")
def trained():
    """The C5 network: 2^20 synthetic pairs of 2^14 samples (seed 0xC5A), batch 1024, Adam
    lr 1e-3, trained on the GPU by the library's own forward and backward kernels."""
    stream = rf.Stream()
    mlp = rf.DelayMlp(1024, seed=0xC5A)
    print("C5 training on 2^20 pairs", flush=True)
    losses = mlp.train(0xC5A, 1 << 20, 1024, 1 << 14, 1e-3, stream)
    yield mlp, losses, stream
    mlp.destroy()
    stream.destroy()


def test_c5_learned_search_4096_pairs(trained) -> None:
    mlp, losses, stream = trained
    count, n = 4096, 1 << 14
    assert losses[-8:].mean() < 0.1 * losses[:8].mean()
    a, b = rf.DeviceBuffer(count * n), rf.DeviceBuffer(count * n)
    rf.synth_pairs(0xC5, 0, count, n, a.ptr, b.ptr, None, stream)
    search = rf.DelaySearch(count, n)
    res = rf.DeviceBuffer(32 * count, np.uint8)
    print("C5 search", flush=True)
    search.run(a.ptr, b.ptr, mlp, 0.9, res.ptr, stream)
    stream.synchronize()
    found = rf.read_delay_results(res, count)
    exact = rf.GccBatch([n] * count, [n] * count)
    exact.bind([a.offset(k * n) for k in range(count)], [b.offset(k * n) for k in range(count)])
    pres = rf.DeviceBuffer(16 * count, np.uint8)
    exact.run(pres.ptr, stream.handle)
    stream.synchronize()
    peaks = rf.read_peaks(pres, count)
    assert [r.index for r in found] == [p.index for p in peaks]
    fallback = sum(r.path == 1 for r in found)
    assert fallback < count // 20
    ha = a.download(16 * n).reshape(16, n)
    hb = b.download(16 * n).reshape(16, n)
    for k in range(16):
        w = delay.phat_spectrum(ha[k], hb[k])
        assert found[k].index == int(np.argmax(np.abs(np.fft.ifft(w)[:2 * n - 1])))
    exact.destroy()
    search.destroy()
    for x in (a, b, res, pres):
        x.free()


def pair_snr_db(seed: int, first: int, count: int) -> np.ndarray:
    """SNR (dB) of pairs first .. first + count - 1 of rf_synth_pairs(seed): the second
    uniform of the pair's Philox block, as oracle/delay.py pair_params restates it."""
    ctr = np.zeros((count, 4), np.uint32)
    ctr[:, 1] = np.arange(first, first + count, dtype=np.uint32)
    x = delay.philox4x32_10(ctr, seed & 0xFFFFFFFF, seed >> 32)
    return np.float32(-10.0) + np.float32(30.0) * delay.uniform24(x[:, 1]).astype(np.float32)


def test_learned_search_on_the_hardest_pairs(trained) -> None:
    # What the trained network gains over the method without it (the exact GCC-PHAT with its
    # full inverse transform and peak search on every pair), on the hardest pairs of the C5
    # distribution: SNR <= -7 dB, 2048 pairs picked from the first 2^15 of seed 0xC5. Every lag
    # equals the exact batched GCC-PHAT's, while at most 5 % of the pairs need the full inverse
    # transform (low confidence, or a window peak at the edge or below the PHAT noise floor).
    # The coarse-proposal ablation is run beside it for the record.
    mlp, _, stream = trained
    n = 1 << 14
    pick = np.where(pair_snr_db(0xC5, 0, 1 << 15) <= -7.0)[0][:2048]
    count = int(pick.size)
    assert count == 2048
    a, b = rf.DeviceBuffer(count * n), rf.DeviceBuffer(count * n)
    for k, p in enumerate(pick):
        rf.synth_pairs(0xC5, int(p), 1, n, a.offset(k * n), b.offset(k * n), None, stream)
    exact = rf.GccBatch([n] * count, [n] * count)
    exact.bind([a.offset(k * n) for k in range(count)], [b.offset(k * n) for k in range(count)])
    pres = rf.DeviceBuffer(16 * count, np.uint8)
    exact.run(pres.ptr, stream.handle)
    stream.synchronize()
    want = [p.index for p in rf.read_peaks(pres, count)]
    search = rf.DelaySearch(count, n)
    res = rf.DeviceBuffer(32 * count, np.uint8)
    outcome = {}
    for name, net, threshold in [("network", mlp, 0.9), ("coarse", None, 0.0)]:
        search.run(a.ptr, b.ptr, net, threshold, res.ptr, stream)
        stream.synchronize()
        got = rf.read_delay_results(res, count)
        outcome[name] = (sum(g.index != w for g, w in zip(got, want)),
                         sum(g.path == 1 for g in got))
    print(f"wrong lags / full inverse transforms: {outcome}", flush=True)
    assert outcome["network"][0] == 0
    assert outcome["network"][1] <= count // 20
    exact.destroy()
    search.destroy()
    for x in (a, b, res, pres):
        x.free()
