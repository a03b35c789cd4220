"""The constant tables compiled into the library (csrc/tables.hpp, csrc/table_gen.cpp): their
values against float64 numpy (exp(-2 pi i m / n) rounded once to float32; the FP16 hi/lo DFT
split), and their presence in the gfx950 code object with the sizes the kernels index (a table
missing from the code object makes the HIP runtime abort the whole process at load).
"""
from __future__ import annotations

import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from _common import PACKAGE_DIR

BUILD = PACKAGE_DIR / "_build"
LLVM = Path("/opt/rocm/llvm/bin")


def tables() -> dict[str, np.ndarray]:
    path = BUILD / "tables_data.inc"
    if not path.exists():
        subprocess.run(["make", "-s", "-C", str(PACKAGE_DIR), "_build/tables_data.inc"],
                       check=True)
    text = path.read_text()
    out = {}
    for m in re.finditer(r"unsigned (int|short) (\w+)\[(\d+)\] = \{([^}]*)\}", text):
        kind, name, size, body = m.groups()
        words = np.array([int(w.strip().rstrip("u"), 16) for w in body.split(",")],
                         dtype=np.uint32 if kind == "int" else np.uint16)
        assert words.size == int(size)
        out[name] = words
    return out


def as_complex(words: np.ndarray) -> np.ndarray:
    return words.view(np.float32).astype(np.float64).view(np.complex128)


def rounded(z: np.ndarray) -> np.ndarray:
    return (z.real.astype(np.float32) + 1j * z.imag.astype(np.float32)).astype(np.complex128)


def w(num: np.ndarray, n: float) -> np.ndarray:
    return np.exp(-2j * np.pi * (num % n) / n)


def close_to_rounded(got: np.ndarray, want: np.ndarray) -> None:
    # float64 sin/cos of two libraries may differ in the last double bit, which moves the
    # float32 rounding only at a tie: allow one float32 ulp, and require most entries exact.
    np.testing.assert_allclose(got.real, rounded(want).real, rtol=0, atol=6e-8)
    np.testing.assert_allclose(got.imag, rounded(want).imag, rtol=0, atol=6e-8)
    assert np.mean(got == rounded(want)) > 0.999


def test_twiddle_4096() -> None:
    m = np.arange(4096)
    close_to_rounded(as_complex(tables()["rf_table_tw4096"]), w(m, 4096))


def test_stage_table_4096() -> None:
    got = as_complex(tables()["rf_table_stage4096"])
    want = []
    for span in (16, 256):
        for k in range(span):
            want.append(w(np.arange(16) * k, 16 * span))
    close_to_rounded(got, np.concatenate(want))


def test_table_16384() -> None:
    got = as_complex(tables()["rf_table_16384"])
    assert got.size == 64 + 64 * 17 + 128 + 128 + 1024 + 32
    r = np.arange(16)
    span4 = np.concatenate([w(r * k, 64) for k in range(4)])
    span64 = np.concatenate([np.append(w(r * k, 1024), 0) for k in range(64)])
    want = np.concatenate([span4, span64, w(np.arange(128), 128), w(np.arange(128), 16384),
                           w(np.arange(1024), 32768), w(np.arange(32), 64)])
    close_to_rounded(got, want)


def test_dft64_split() -> None:
    words = tables()["rf_table_dft64"]
    half = words.size // 2
    hi = words[:half].view(np.float16).astype(np.float64)
    lo = words[half:].view(np.float16).astype(np.float64)
    idx = np.arange(half)
    q, l = idx % 8, (idx // 8) % 64
    s, wave = (idx // 512) % 8, idx // 4096
    kk = 16 * s + 8 * (l // 32) + q          # 2r + d
    n = 32 * wave + l % 32                   # 2k + e
    r, d, k, e = kk >> 1, kk & 1, n >> 1, n & 1
    theta = 2 * np.pi * ((k * r) % 64) / 64
    want = np.where(e == 0, np.where(d == 0, np.cos(theta), -np.sin(theta)),
                    np.where(d == 0, np.sin(theta), np.cos(theta)))
    # hi + lo carries about 22 bits: the split error is below 2^-21 of an entry of size <= 1
    assert np.max(np.abs(hi + lo - want)) < 2.0 ** -21
    assert np.max(np.abs(hi - want)) < 2.0 ** -11


def device_symbols(obj: Path, tmp_path: Path) -> str:
    """llvm-readelf --syms of the gfx950 code object bundled into one host object file."""
    if not obj.exists():
        pytest.skip("library not built in-tree")
    fatbin = tmp_path / "fatbin.bin"
    dev = tmp_path / "dev.o"
    subprocess.run([str(LLVM / "llvm-objcopy"), f"--dump-section=.hip_fatbin={fatbin}", str(obj),
                    str(tmp_path / "host.o")], check=True)
    subprocess.run([str(LLVM / "clang-offload-bundler"), "--unbundle", "--type=o",
                    f"--input={fatbin}", "--targets=hipv4-amdgcn-amd-amdhsa--gfx950",
                    f"--output={dev}"], check=True)
    return subprocess.run([str(LLVM / "llvm-readelf"), "--syms", str(dev)], check=True,
                          capture_output=True, text=True).stdout


TOOLS = pytest.mark.skipif(shutil.which(str(LLVM / "llvm-readelf")) is None,
                           reason="no ROCm llvm tools")


@TOOLS
def test_tables_are_in_the_gfx950_code_object(tmp_path: Path) -> None:
    syms = device_symbols(BUILD / "twiddle.o", tmp_path)
    sizes = {"rf_table_tw4096": 4096 * 8, "rf_table_stage4096": 4352 * 8,
             "rf_table_16384": 2464 * 8, "rf_table_dft64": 2 * 16384 * 2}
    for name, size in sizes.items():
        m = re.search(rf"\s(\d+) OBJECT\s+GLOBAL\s+\w+\s+\d+ {name}\n", syms)
        assert m is not None, f"{name} missing from the device code object"
        assert int(m.group(1)) == size


@TOOLS
def test_noalloc_scratch_is_in_the_gfx950_code_object(tmp_path: Path) -> None:
    # rf_fir_noalloc's file-scope scratch array (72 KiB), which the runtime registers by name
    syms = device_symbols(BUILD / "fir.o", tmp_path)
    m = re.search(r"\s(\d+) OBJECT\s+GLOBAL\s+\w+\s+\d+ _ZL7scratch\n", syms)
    assert m is not None and int(m.group(1)) == 72 * 1024
