"""Compiles every program in sigflow/examples/ with sigflowc for each kernel / pass variant and
builds it with build_program.sh, into sigflow/_programs/EXAMPLE.VARIANT (plus the generated
EXAMPLE.VARIANT.cpp). The end-to-end GPU tests run these binaries, so they are built here, on
the build machine, before any GPU run.

    python gpu-rf-pipeline_amd/sigflow/build_examples.py

Needs the package built first (make -C gpu-rf-pipeline_amd: librfhip.so, libsigflow_rt.a,
sigflowc).
"""
from __future__ import annotations

import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
PACKAGE = HERE.parent
SIGFLOWC = PACKAGE / "sigflowc"
EXAMPLES = sorted((HERE / "examples").glob("*.sf"))
OUT = HERE / "_programs"

# Variant name -> sigflowc flags. With reuse, a later value overwrites a dead value's slot, so a
# liveness mistake that frees a slot too early shows as a wrong output; --no-fuse runs a fusion
# group's FIRs (and fft |> magnitude) as separate calls; every FIR and xcorr variant appears.
VARIANTS: dict[str, list[str]] = {
    "default": [],
    "no_reuse": ["--no-reuse"],
    "no_fuse": ["--no-fuse"],
    "tiled": ["--fir-impl=tiled", "--xcorr-impl=tiled"],
    "const_fft": ["--fir-impl=tiled_const", "--xcorr-impl=fft"],
    "register_fft": ["--fir-impl=register", "--xcorr-impl=fft"],
    "register_no_fuse": ["--fir-impl=register", "--xcorr-impl=fft", "--no-fuse"],
    "ola_fft": ["--fir-impl=fft_ola", "--xcorr-impl=fft"],
}


def program_path(example: str, variant: str) -> Path:
    return OUT / f"{example}.{variant}"


def build_one(example: Path, variant: str) -> str:
    exe = program_path(example.stem, variant)
    source = exe.with_name(exe.name + ".cpp")
    subprocess.run([str(SIGFLOWC), str(example), *VARIANTS[variant], "-o", str(source)],
                   check=True, capture_output=True, text=True)
    subprocess.run([str(HERE / "build_program.sh"), str(source), str(exe)], check=True,
                   capture_output=True, text=True)
    return exe.name


def main() -> int:
    if not SIGFLOWC.exists():
        print(f"{SIGFLOWC} is not built (make -C {PACKAGE})", file=sys.stderr)
        return 2
    OUT.mkdir(exist_ok=True)
    jobs = [(e, v) for e in EXAMPLES for v in VARIANTS]
    with ThreadPoolExecutor(max_workers=8) as pool:
        for future in [pool.submit(build_one, e, v) for e, v in jobs]:
            try:
                future.result()
            except subprocess.CalledProcessError as err:
                print(f"failed: {' '.join(map(str, err.cmd))}\n{err.stdout}{err.stderr}",
                      file=sys.stderr)
                return 1
    print(f"built {len(jobs)} programs in {OUT}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
