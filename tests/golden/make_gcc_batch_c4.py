"""Writes gcc_batch_c4.json: the float64 oracle's GCC-PHAT peak for each of the 256 pairs of the
batched GCC-PHAT workload (bench.py gcc_batch_input(): seed 0xC4, lengths log-uniform in
[2^12, 2^20], delays U[-200, 200] with a fractional part, SNR U[0, 20] dB).

Per pair: the index of the first maximum of |r| (lag = index - (num_a - 1)); the sub-sample
offset the batched GCC-PHAT reports, f = m1 / (m0 + m1) towards the larger neighbour m1 of the
peak m0 (exact for the sinc-shaped peak of a pure fractional delay), here on the oracle's
float64 |r|; the relative margin of the peak over the largest |r| outside index - 2 ..
index + 2 (how far the pair is from a tie); the true delay.

Run from the repository root: python tests/golden/make_gcc_batch_c4.py  (about a minute, 8
processes). The oracle is the checker; nothing here runs on the GPU.
"""
from __future__ import annotations

import json
import sys
from multiprocessing import Pool
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

import bench  # noqa: E402  (the workload generator)
import oracle  # noqa: E402  (the float64 oracle)

OUT = Path(__file__).resolve().parent / "gcc_batch_c4.json"


def neighbour_offset(mag: np.ndarray, idx: int) -> float:
    if idx == 0 or idx == mag.size - 1:
        return 0.0
    m0, ml, mh = mag[idx], mag[idx - 1], mag[idx + 1]
    f = mh / (m0 + mh) if mh >= ml else -ml / (m0 + ml)
    return float(min(0.5, max(-0.5, f)))


def one(pair):
    a, b = pair
    r = oracle.gcc_phat(a, b)
    idx, _ = oracle.peak(r)
    mag = np.abs(r)
    rest = mag.copy()
    rest[max(idx - 2, 0):idx + 3] = 0.0
    return idx, neighbour_offset(mag, idx), float((mag[idx] - rest.max()) / mag[idx])


def main() -> None:
    pairs, delays = bench.gcc_batch_input()
    with Pool(8) as pool:
        res = pool.map(one, pairs, chunksize=1)
    doc = {"workload": "bench.py gcc_batch_input(), 256 pairs, seed 0xC4",
           "num_a": [int(a.size) for a, _ in pairs], "num_b": [int(b.size) for _, b in pairs],
           "index": [r[0] for r in res], "frac": [r[1] for r in res],
           "margin": [r[2] for r in res], "delay": [float(d) for d in delays]}
    OUT.write_text(json.dumps(doc))
    print(f"wrote {OUT} ({len(res)} pairs)")


if __name__ == "__main__":
    main()
