"""SigFlow end to end on the GPU (reference: compiler/tests/e2e/test_examples.py).

Every program in gpu-rf-pipeline_amd/sigflow/examples/ was compiled by sigflowc with each
variant of build_examples.VARIANTS (kernel choices, --no-reuse, --no-fuse) and built against
librfhip.so before this run. Each runs here in a temporary directory next to a `fixtures` link;
its IR (--emit=json) is evaluated node by node with the float64 oracle, and every emitted file
is compared with the oracle's value of the node it emits, with the reference's end-to-end
tolerance for that node's operation. With slot reuse a later value overwrites a dead value's
buffer, so a liveness mistake that frees a slot too early shows up as a wrong output.
"""
from __future__ import annotations

import importlib.util
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from _common import FIXTURES, PACKAGE_DIR, oracle, read_cf32, read_f32

pytestmark = pytest.mark.gpu

SIGFLOW = PACKAGE_DIR / "sigflow"
SIGFLOWC = PACKAGE_DIR / "sigflowc"
_spec = importlib.util.spec_from_file_location("build_examples", SIGFLOW / "build_examples.py")
build_examples = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(build_examples)

EXAMPLES = [e.stem for e in build_examples.EXAMPLES]
VARIANTS = list(build_examples.VARIANTS)

# Tolerance of an emitted value by the operation that produced it: the reference's end-to-end
# table (compiler/tests/e2e/test_examples.py:48-54); fft frames and channelizer outputs, which
# its examples never emit directly, take its operator tests' spectrum and channelizer bounds
# (tests/test_pipeline.py:62-66, 360-364).
TOLERANCES: dict[str, tuple[float, float]] = {
    "capture": (0.0, 0.0),
    "fir": (1e-5, 1e-5),
    "magnitude": (1e-5, 1e-3),
    "xcorr": (1e-5, 5e-3),
    "gcc_phat": (0.0, 1e-4),
    "fft": (1e-5, 2e-3),
    "channelize": (1e-5, 1e-4),
}


def evaluate(graph: dict, workdir: Path) -> dict[int, np.ndarray]:
    """The oracle's value of every live value node, in id order (inputs have smaller ids)."""
    inputs: dict[int, list[int]] = {}
    for e in sorted(graph["edges"], key=lambda e: (e["to"], e["input"])):
        inputs.setdefault(e["to"], []).append(e["from"])
    values: dict[int, np.ndarray] = {}
    for node in sorted(graph["nodes"], key=lambda n: n["id"]):
        if not node["live"] or node["op"] == "emit":
            continue
        args = [values[i] for i in inputs.get(node["id"], [])]
        op = node["op"]
        if op == "capture":
            value = read_cf32(workdir / node["path"]).astype(np.complex128)
        elif op == "fir":
            value = oracle.fir(args[0], read_f32(workdir / node["taps"]))
        elif op == "fft":
            value = oracle.spectrum(args[0], node["n"]).ravel()
        elif op == "magnitude":
            value = oracle.magnitude(args[0])
        elif op == "xcorr":
            value = oracle.xcorr(args[0], args[1])
        elif op == "channelize":
            value = oracle.channelize(args[0], read_f32(workdir / node["taps"]),
                                      node["channels"]).ravel()
        elif op == "gcc_phat":
            value = oracle.gcc_phat(args[0], args[1])
        else:
            raise AssertionError(f"unknown op {op}")
        values[node["id"]] = value
    return values


def run_program(example: str, variant: str, workdir: Path, *args: str) -> tuple[dict, dict]:
    """Runs the prebuilt program in `workdir`; returns (its IR as JSON, its report)."""
    exe = build_examples.program_path(example, variant)
    assert exe.exists(), f"{exe} is not built (python {SIGFLOW / 'build_examples.py'})"
    workdir.mkdir(parents=True, exist_ok=True)
    (workdir / "fixtures").symlink_to(FIXTURES, target_is_directory=True)
    source = SIGFLOW / "examples" / f"{example}.sf"
    graph = json.loads(subprocess.run([str(SIGFLOWC), str(source), "--emit=json",
                                       *build_examples.VARIANTS[variant]],
                                      check=True, capture_output=True, text=True).stdout)
    print(f"run {example}.{variant} {' '.join(args)}", flush=True)
    proc = subprocess.run([str(exe), *args], cwd=workdir, capture_output=True, text=True,
                          timeout=120)
    assert proc.returncode == 0, f"{exe.name}: {proc.stdout}\n{proc.stderr}"
    return graph, json.loads(proc.stdout.strip().splitlines()[-1])


def check_outputs(graph: dict, workdir: Path, label: str) -> None:
    nodes = {n["id"]: n for n in graph["nodes"]}
    sources = {e["to"]: e["from"] for e in graph["edges"] if nodes[e["to"]]["op"] == "emit"}
    values = evaluate(graph, workdir)
    emits = [n for n in graph["nodes"] if n["op"] == "emit"]
    assert emits
    for emit in emits:
        source = nodes[sources[emit["id"]]]
        path = workdir / emit["path"]
        real = source["type"].startswith("frames<f32")
        got = read_f32(path) if real else read_cf32(path)
        want = values[source["id"]]
        rtol, atol = TOLERANCES[source["op"]]
        assert got.shape == want.shape, f"{label}: {emit['path']} shape"
        np.testing.assert_allclose(got, want, rtol=rtol, atol=atol,
                                   err_msg=f"{label}: {emit['path']}")


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("example", EXAMPLES)
def test_example_matches_oracle(tmp_path: Path, example: str, variant: str) -> None:
    graph, report = run_program(example, variant, tmp_path)
    check_outputs(graph, tmp_path, f"{example}.{variant}")
    assert report["runs"] == 1 and report["calls"] >= 1
    fused_groups = {n["group"] for n in graph["nodes"] if n.get("group") is not None}
    assert report["calls_by_kind"].get("fir_batched", 0) == len(fused_groups)


@pytest.mark.parametrize("example", EXAMPLES)
def test_example_as_graph_replays(tmp_path: Path, example: str) -> None:
    """The compute phase recorded once as a hipGraph and replayed: four timed replays after a
    warm-up, uploads and downloads between them; the files hold the last replay's values."""
    graph, report = run_program(example, "register_fft", tmp_path, "--graph", "--bench", "4",
                                "--warmup", "1")
    check_outputs(graph, tmp_path, f"{example} graph")
    assert report["graph"] is True and report["runs"] == 4 and report["kernel_ms"] > 0


def test_fusion_saves_calls(tmp_path: Path) -> None:
    """two_channel: one batched call for the FIR pair and one spectrum-magnitude call instead of
    two FIR calls, a spectrum and a magnitude."""
    _, fused = run_program("two_channel", "register_fft", tmp_path / "a")
    _, plain = run_program("two_channel", "register_no_fuse", tmp_path / "b")
    assert fused["calls_by_kind"] == {"fir_batched": 1, "spectrum_magnitude": 1, "xcorr_fft": 1}
    assert plain["calls_by_kind"] == {"fir": 2, "spectrum": 1, "magnitude": 1, "xcorr_fft": 1}
    assert fused["fused"] is True and plain["fused"] is False


def test_delay_and_direction_peaks(tmp_path: Path) -> None:
    """The delay example's xcorr peaks at the ch1 fixture's true delay; the direction example's
    GCC-PHAT peaks at each receiver's delay behind receiver 0."""
    run_program("delay", "default", tmp_path / "d")
    r = read_cf32(tmp_path / "d" / "out" / "delay_xcorr.cf32")
    na = read_cf32(FIXTURES / "ch0.cf32").size
    lag = int(np.argmax(np.abs(r))) - (na - 1)
    assert lag == json.loads((FIXTURES / "ch1.json").read_text())["true_delay_samples"]
    run_program("direction", "default", tmp_path / "g")
    n0 = read_cf32(FIXTURES / "array0.cf32").size
    for k in (1, 2, 3):
        r = read_cf32(tmp_path / "g" / "out" / f"tdoa_{k}.cf32")
        delay = json.loads((FIXTURES / f"array{k}.json").read_text())["delay_samples"]
        assert int(np.argmax(np.abs(r))) - (n0 - 1) == delay
