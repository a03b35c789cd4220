"""Shared helpers of the test suite: fixture I/O, and loaders for the library binding and the
float64 oracle."""
from __future__ import annotations

import importlib.util
import json
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
FIXTURES = ROOT / "tests" / "golden" / "fixtures"
PACKAGE_DIR = ROOT / "gpu-rf-pipeline_amd"
LIB = PACKAGE_DIR / "librfhip.so"

sys.path.insert(0, str(ROOT / "oracle"))
import oracle  # noqa: E402  (the float64 oracle)


def load_rfhip():
    """Imports the package directory (its name has dashes) as module `rfhip`, building the
    library first only if nothing is built."""
    if "rfhip" in sys.modules:
        return sys.modules["rfhip"]
    if not LIB.exists():
        subprocess.run(["make", "-s", "-C", str(PACKAGE_DIR), "-j8", "librfhip.so"], check=True)
    spec = importlib.util.spec_from_file_location("rfhip", PACKAGE_DIR / "__init__.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rfhip"] = mod
    spec.loader.exec_module(mod)
    return mod


def read_cf32(path: Path) -> np.ndarray:
    """Interleaved little-endian float32 as complex64 (the bits the GPU sees)."""
    iq = np.fromfile(path, dtype="<f4")
    if iq.size % 2:
        raise ValueError(f"{path}: odd number of float32 values")
    return iq.view(np.complex64).copy()


def read_f32(path: Path) -> np.ndarray:
    return np.fromfile(path, dtype="<f4").astype(np.float32)


def cf32(name: str) -> np.ndarray:
    return read_cf32(FIXTURES / f"{name}.cf32")


def taps(name: str) -> np.ndarray:
    return read_f32(FIXTURES / f"{name}.f32")


def meta(name: str) -> dict:
    return json.loads((FIXTURES / f"{name}.json").read_text())


def read_ci16_sigmf(name: str) -> np.ndarray:
    raw = np.fromfile(FIXTURES / f"{name}.sigmf-data", dtype="<i2").astype(np.float32) / 32768
    return (raw[0::2] + 1j * raw[1::2]).astype(np.complex64)
