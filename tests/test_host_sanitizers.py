"""Host parsers under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

sigflowc (its lexer, parser and passes read untrusted program text) and the sample-file
readers of host/io.cpp (SigMF metadata is JSON from outside) are built with g++
-fsanitize=address,undefined into a temporary directory and run over every golden and example
program with every --emit mode, and over mutated programs and SigMF recordings: every run must
end in output or a diagnostic (exit 0, 1 or 2; the readers' errors are exceptions), with no
sanitizer report.
"""
from __future__ import annotations

import random
import subprocess
from pathlib import Path

import pytest

from _common import FIXTURES, PACKAGE_DIR, ROOT

SANITIZE = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
            "-fno-omit-frame-pointer", "-O1", "-g", "-std=c++17"]
PROGRAMS = (sorted((ROOT / "tests" / "golden" / "sigflow").rglob("*.sf"))
            + sorted((PACKAGE_DIR / "sigflow" / "examples").glob("*.sf")))
EMITS = ["tokens", "ast", "check", "ir", "schedule", "liveness", "json", "hip"]
ENV = {"ASAN_OPTIONS": "detect_leaks=0", "UBSAN_OPTIONS": "print_stacktrace=1",
       "PATH": "/usr/bin:/bin"}


def build(tmp: Path, name: str, sources: list[Path], includes: list[Path]) -> Path:
    exe = tmp / name
    cmd = ["g++", *SANITIZE, *[f"-I{i}" for i in includes], *map(str, sources), "-o", str(exe)]
    proc = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if proc.returncode != 0:
        pytest.skip(f"g++ cannot build with sanitizers here: {proc.stderr[-400:]}")
    return exe


def clean(proc: subprocess.CompletedProcess, allowed: tuple[int, ...]) -> bool:
    reported = "AddressSanitizer" in proc.stderr or "runtime error" in proc.stderr
    return not reported and proc.returncode in allowed


@pytest.fixture(scope="module")
def tools(tmp_path_factory) -> dict[str, Path]:
    tmp = tmp_path_factory.mktemp("sanitized")
    sigflow = PACKAGE_DIR / "sigflow"
    host = PACKAGE_DIR / "host"
    return {
        "sigflowc": build(tmp, "sigflowc", sorted(sigflow.glob("*.cpp")), [sigflow]),
        "reader": build(tmp, "reader", [host / "io.cpp", ROOT / "tests" / "host" / "sigmf_read_driver.cpp"],
                        [ROOT / "include", host]),
    }


def mutate(text: str, rng: random.Random, pieces: list[str]) -> str:
    for _ in range(rng.randint(1, 6)):
        pos = rng.randint(0, len(text))
        action = rng.random()
        if action < 0.4:
            text = text[:pos] + rng.choice(pieces) + text[pos:]
        elif action < 0.7 and text:
            text = text[:pos] + text[pos + rng.randint(1, 8):]
        else:
            a, b = sorted((rng.randint(0, len(text)), rng.randint(0, len(text))))
            text = text[:pos] + text[a:b] + text[pos:]
    return text


def test_sigflowc_every_program_every_emit(tools, tmp_path: Path) -> None:
    failures = []
    for program in PROGRAMS:
        for emit in EMITS:
            proc = subprocess.run([str(tools["sigflowc"]), str(program), f"--emit={emit}",
                                   f"--input-dir={ROOT / 'tests' / 'golden'}"],
                                  capture_output=True, text=True, timeout=60, cwd=program.parent,
                                  env=ENV)
            if not clean(proc, (0, 1, 2)):
                failures.append(f"{program.name} --emit={emit}: {proc.returncode} {proc.stderr[:300]}")
    assert not failures, "\n".join(failures[:5])


@pytest.mark.parametrize("seed", [1, 2])
def test_sigflowc_mutated_programs(tools, tmp_path: Path, seed: int) -> None:
    rng = random.Random(seed)
    texts = [p.read_text() for p in PROGRAMS]
    pieces = list("abxyz_09().,=|>#\"\n \t-+e") + ["fir", "fft", "emit", "capture", "xcorr",
                                                  "gcc_phat", "channelize", "magnitude", "n=",
                                                  "taps=", "|>", "->", "1e999", "-1", "9" * 30]
    case = tmp_path / "case.sf"
    for _ in range(150):
        case.write_text(mutate(rng.choice(texts), rng, pieces))
        for emit in ("ir", "hip"):
            proc = subprocess.run([str(tools["sigflowc"]), str(case), f"--emit={emit}"],
                                  capture_output=True, text=True, timeout=60, env=ENV)
            assert clean(proc, (0, 1, 2)), f"{case.read_text()!r}\n{proc.stderr[:500]}"


def test_sigmf_reader_mutated_recordings(tools, tmp_path: Path) -> None:
    rng = random.Random(3)
    meta = (FIXTURES / "tone_ci16.sigmf-meta").read_text()
    data = (FIXTURES / "tone_ci16.sigmf-data").read_bytes()[:4096]
    pieces = ['"core:datatype"', '"cf32_le"', '"ci16_le"', '"ci8"', '"ri16_le"',
              '"core:sample_rate"', ":", ",", "{", "}", "[", "]", '"', "\\", "1e308", "-1",
              "null", "\n", '"\\u0000"', "9" * 40]
    targets = []
    for i in range(200):
        stem = tmp_path / f"r{i}"
        (tmp_path / f"r{i}.sigmf-meta").write_text(mutate(meta, rng, pieces))
        (tmp_path / f"r{i}.sigmf-data").write_bytes(data[:rng.choice([0, 1, 3, 7, 8, 100, 4096])])
        targets += [f"{stem}.sigmf-meta", f"{stem}.sigmf-data"]
    for nbytes in (0, 1, 7, 8, 9, 16):
        raw = tmp_path / f"raw{nbytes}.cf32"
        raw.write_bytes(b"\x01" * nbytes)
        targets.append(str(raw))
    for start in range(0, len(targets), 100):
        proc = subprocess.run([str(tools["reader"]), *targets[start:start + 100]],
                              capture_output=True, text=True, timeout=120, env=ENV)
        assert clean(proc, (0,)), proc.stderr[:800]
    proc = subprocess.run([str(tools["reader"]), str(FIXTURES / "tone_ci16.sigmf-meta")],
                          capture_output=True, text=True, timeout=60, env=ENV)
    assert clean(proc, (0,)) and not proc.stdout.startswith("error"), proc.stdout + proc.stderr
