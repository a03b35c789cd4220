"""Learned delay search (rf_synth_pairs, rf_delay_mlp_*, rf_delay_search_*) against the float64
oracle (oracle/delay.py) and against the exact batched GCC-PHAT (rf_gcc_batch) on the same
pairs. Tolerances are written per test; integer results (peak index, proposal, delays) are
compared exactly.
"""
from __future__ import annotations

import numpy as np
import pytest

from _common import load_rfhip, oracle  # noqa: F401  (puts oracle/ on sys.path)
import delay as D

pytestmark = pytest.mark.gpu
rf = load_rfhip()

N = 1 << 14                 # samples per channel, L = 2^15
SEED_EVAL = 0xC5
SEED_TRAIN = 0xC5A


def synth(count: int, first: int = 0, seed: int = SEED_EVAL):
    a = rf.DeviceBuffer(count * N)
    b = rf.DeviceBuffer(count * N)
    d = rf.DeviceBuffer(count, np.float32)
    rf.synth_pairs(seed, first, count, N, a.ptr, b.ptr, d.ptr)
    return a, b, d


def exact_peaks(a, b, count: int):
    """rf_gcc_batch (the full inverse transform) on the same device pairs."""
    h = rf.GccBatch([N] * count, [N] * count)
    h.bind([a.offset(k * N) for k in range(count)], [b.offset(k * N) for k in range(count)])
    res = rf.DeviceBuffer(16 * count, np.uint8)
    h.run(res.ptr)
    return rf.read_peaks(res, count)


def search(a, b, count: int, mlp, threshold: float):
    s = rf.DelaySearch(count, N)
    res = rf.DeviceBuffer(32 * count, np.uint8)
    s.run(a.ptr, b.ptr, mlp, threshold, res.ptr)
    return rf.read_delay_results(res, count)


# --- generator and features --------------------------------------------------------------

def test_synthetic_pairs_match_the_oracle() -> None:
    count, first = 3, 5
    a, b, d = synth(count, first)
    ga = a.download().reshape(count, N)
    gb = b.download().reshape(count, N)
    ea, eb, ed = D.synth_pairs(SEED_EVAL, first, count, N)
    np.testing.assert_array_equal(d.download(), ed.astype(np.float32))
    # float32 Box-Muller, AR(1) recursion and 65-tap filter against float64
    np.testing.assert_allclose(ga, ea, rtol=0, atol=2e-4)
    np.testing.assert_allclose(gb, eb, rtol=0, atol=2e-4)


def test_features_match_the_oracle() -> None:
    count = 8
    a, b, _ = synth(count, 40)
    s = rf.DelaySearch(count, N)
    feats = rf.DeviceBuffer(count * 256, np.float32)
    coarse = rf.DeviceBuffer(count, np.int32)
    s.features(a.ptr, b.ptr, feats.ptr, coarse.ptr)
    gf = feats.download().reshape(count, 256)
    gc = coarse.download()
    ha = a.download().reshape(count, N)
    hb = b.download().reshape(count, N)
    for k in range(count):
        f, c = D.features(ha[k], hb[k])
        np.testing.assert_allclose(gf[k], f, rtol=0, atol=2e-4)
        assert gc[k] == c


# --- network ----------------------------------------------------------------------------

def random_params(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, 0.06, D.PARAM_COUNT).astype(np.float32)


def test_mlp_forward_matches_the_oracle() -> None:
    batch = 96                                   # not a multiple of the 32-row tile
    params = random_params(1)
    x = np.random.default_rng(2).uniform(0, 1, (batch, 256)).astype(np.float32)
    mlp = rf.DelayMlp(128)
    mlp.set_params(params)
    dx = rf.DeviceBuffer.from_array(x.ravel())
    probs = rf.DeviceBuffer(batch * 64, np.float32)
    mlp.forward(dx.ptr, batch, probs.ptr)
    got = probs.download().reshape(batch, 64)
    want = D.mlp_forward(params.astype(np.float64), x.astype(np.float64), round_bf16=True)["p"]
    # BF16 operands on both sides; a rare one-ulp BF16 flip of an FP32 activation moves a
    # probability by well under 2e-3
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-3)
    np.testing.assert_allclose(got.sum(axis=1), 1.0, atol=1e-5)


def test_mlp_gradients_match_the_oracle() -> None:
    batch = 64
    params = random_params(3)
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 1, (batch, 256)).astype(np.float32)
    labels = rng.integers(0, 64, batch).astype(np.int32)
    mlp = rf.DelayMlp(batch)
    mlp.set_params(params)
    dx = rf.DeviceBuffer.from_array(x.ravel())
    dl = rf.DeviceBuffer.from_array(labels)
    grads = rf.DeviceBuffer(D.PARAM_COUNT, np.float32)
    loss = rf.DeviceBuffer(1, np.float32)
    mlp.gradients(dx.ptr, dl.ptr, batch, grads.ptr, loss.ptr)
    g = grads.download().astype(np.float64)
    p64 = params.astype(np.float64)
    want = D.mlp_backward(p64, x.astype(np.float64), labels, round_bf16=True)
    assert float(loss.download()[0]) == pytest.approx(
        D.mlp_loss(p64, x.astype(np.float64), labels, round_bf16=True), rel=1e-4)
    # per parameter block: relative error in the 2-norm (BF16 operands, FP32 accumulation)
    edges = [0, 65536, 65792, 131328, 131584, 147968, 148032]
    for lo, hi in zip(edges[:-1], edges[1:]):
        err = np.linalg.norm(g[lo:hi] - want[lo:hi]) / np.linalg.norm(want[lo:hi])
        assert err < 2e-2, (lo, hi, err)


def test_train_step_is_one_adam_step() -> None:
    batch = 64
    params = random_params(5)
    rng = np.random.default_rng(6)
    x = rng.uniform(0, 1, (batch, 256)).astype(np.float32)
    labels = rng.integers(0, 64, batch).astype(np.int32)
    mlp = rf.DelayMlp(batch)
    mlp.set_params(params)
    dx = rf.DeviceBuffer.from_array(x.ravel())
    dl = rf.DeviceBuffer.from_array(labels)
    grads = rf.DeviceBuffer(D.PARAM_COUNT, np.float32)
    loss = rf.DeviceBuffer(1, np.float32)
    mlp.gradients(dx.ptr, dl.ptr, batch, grads.ptr, loss.ptr)
    g = grads.download().astype(np.float64)
    lr = 1e-3
    mlp.train_step(dx.ptr, dl.ptr, batch, lr, loss.ptr)
    want, _, _ = D.adam_step(params.astype(np.float64), g, np.zeros_like(g), np.zeros_like(g),
                             1, lr)
    np.testing.assert_allclose(mlp.get_params(), want, rtol=0, atol=lr * 1e-2)


# --- search ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def eval_pairs():
    count = 48
    a, b, d = synth(count, 1000)
    return a, b, d.download(), exact_peaks(a, b, count), count


def test_full_path_equals_the_exact_batch(eval_pairs) -> None:
    a, b, _, exact, count = eval_pairs
    res = search(a, b, count, None, 2.0)             # threshold > 1: every pair falls back
    assert all(r.path == 1 for r in res)
    for r, e in zip(res, exact):
        # same transforms as rf_gcc_batch, W stored between the forward and inverse row FFTs
        assert r.index == e.index
        assert r.magnitude == pytest.approx(e.magnitude, rel=1e-6)
        assert r.frac == pytest.approx(e.frac, abs=1e-6)


def test_window_refine_equals_the_exact_peak(eval_pairs) -> None:
    a, b, delays, exact, count = eval_pairs
    res = search(a, b, count, None, 0.0)             # coarse proposal, always refine
    refined = [k for k, r in enumerate(res) if r.path == 0]
    assert len(refined) >= count - 2
    for k in refined:
        assert res[k].index == exact[k].index, k
        assert res[k].magnitude == pytest.approx(exact[k].magnitude, rel=1e-4, abs=1e-6)
        assert res[k].frac == pytest.approx(exact[k].frac, abs=2e-3)
        # the coarse bin holds the exact peak; the label rounds the true delay
        assert res[k].proposal == (res[k].index - (N - 1) + 512) >> 4
        assert abs(res[k].proposal - D.label_of(delays[k])) <= 1


def test_refine_against_the_oracle_window() -> None:
    count = 4
    a, b, d = synth(count, 77)
    res = search(a, b, count, None, 0.0)
    ha = a.download().reshape(count, N)
    hb = b.download().reshape(count, N)
    for k in range(count):
        out = D.refine(D.phat_spectrum(ha[k], hb[k]), N, res[k].proposal)
        assert out is not None and res[k].path == 0
        assert res[k].index == out[0]
        assert res[k].magnitude == pytest.approx(out[1], rel=1e-4)
        assert res[k].frac == pytest.approx(out[2], abs=2e-3)


def test_trained_network_proposals(eval_pairs) -> None:
    a, b, delays, exact, count = eval_pairs
    mlp = rf.DelayMlp(512, seed=SEED_TRAIN)
    losses = mlp.train(SEED_TRAIN, 1 << 16, 512, N, 1e-3)
    assert np.all(np.isfinite(losses))
    assert losses[-4:].mean() < 0.5 * losses[:4].mean()
    res = search(a, b, count, mlp, 0.9)
    agree = sum(r.index == e.index for r, e in zip(res, exact))
    assert agree >= count - 1
    for r, e in zip(res, exact):
        if r.path == 1:
            assert r.index == e.index                # the fallback is the exact path
    near = sum(abs(r.proposal - D.label_of(d)) <= 1 for r, d in zip(res, delays))
    assert near >= count - 2


def test_bad_arguments() -> None:
    with pytest.raises(rf.RfError):
        rf.DelaySearch(4, 4000)                      # L = 2^13: below the supported range
    with pytest.raises(rf.RfError):
        rf.DelayMlp(0)
    mlp = rf.DelayMlp(8)
    x = rf.DeviceBuffer(16 * 256, np.float32)
    p = rf.DeviceBuffer(16 * 64, np.float32)
    with pytest.raises(rf.RfError):
        mlp.forward(x.ptr, 16, p.ptr)                # above max_batch
