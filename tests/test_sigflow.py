"""The SigFlow compiler (sigflowc) on the CPU.

- The reference compiler's golden cases (tests/golden/sigflow/reference, copied byte for byte
  from its compiler/tests/golden/): lexer, parser and semantic diagnostics, IR, dead-node
  elimination, fusion, schedule, liveness and JSON, each compared exactly.
- This build's HIP-backend cases (tests/golden/sigflow/hip): the emitted programs, compared
  exactly with files regenerated on purpose (SIGFLOW_UPDATE_GOLDEN=1 rewrites them), and
  compiled against host/device_ops.hpp and host/pipeline_runtime.hpp.
- Invariants of the passes on random programs (the schedule respects every edge, fusion groups
  share taps and never depend on each other, two values share a slot only when their launches
  do not overlap, value sizes follow the fixtures), and robustness: mutated programs end in
  output or diagnostics, never in a crash (the reference fuzzes its compiler for the same).
"""
from __future__ import annotations

import hashlib
import json
import os
import random
import re
import subprocess
from pathlib import Path

import pytest

from _common import FIXTURES, PACKAGE_DIR, ROOT

SIGFLOWC = PACKAGE_DIR / "sigflowc"
GOLDEN = ROOT / "tests" / "golden" / "sigflow"
REFERENCE_CASES = sorted((GOLDEN / "reference").glob("*.sf"))
HIP_CASES = sorted((GOLDEN / "hip").glob("*.sf"))
EXAMPLES = sorted((PACKAGE_DIR / "sigflow" / "examples").glob("*.sf"))
UPDATE = os.environ.get("SIGFLOW_UPDATE_GOLDEN") == "1"


@pytest.fixture(scope="module", autouse=True)
def sigflowc_built() -> None:
    sources = list((PACKAGE_DIR / "sigflow").glob("*.[ch]pp"))
    if not SIGFLOWC.exists() or max(p.stat().st_mtime for p in sources) > SIGFLOWC.stat().st_mtime:
        subprocess.run(["make", "-s", "-C", str(PACKAGE_DIR), "sigflowc"], check=True)


def sigflowc(*args: object, cwd: Path | None = None) -> subprocess.CompletedProcess:
    return subprocess.run([str(SIGFLOWC), *map(str, args)], cwd=cwd, capture_output=True,
                          text=True, timeout=60)


def case_output(case: Path) -> str:
    """Runs a golden case the way its harness does: from the case's directory, with the flags of
    its first line; stderr (diagnostics come first) then stdout, and `[exit N]` on failure."""
    first = case.read_text().split("\n", 1)[0]
    args = first[len("# args:"):].split() if first.startswith("# args:") else []
    # the reference's fixture root, as seen from its golden directory
    args = [f"--input-dir={ROOT / 'tests' / 'golden'}" if a == "--input-dir=../../../tests" else a
            for a in args]
    proc = sigflowc(*args, case.name, cwd=case.parent)
    text = proc.stderr + proc.stdout
    if proc.returncode != 0:
        text += f"[exit {proc.returncode}]\n"
    return text


def test_reference_cases_are_unchanged() -> None:
    listed = dict(reversed(line.split()) for line in (GOLDEN / "PROVENANCE.md").read_text().splitlines()
                  if re.fullmatch(r"[0-9a-f]{64}  \S+", line))
    files = sorted(p.name for p in (GOLDEN / "reference").iterdir())
    assert sorted(listed) == files
    for name, digest in listed.items():
        assert hashlib.sha256((GOLDEN / "reference" / name).read_bytes()).hexdigest() == digest, name


@pytest.mark.parametrize("case", REFERENCE_CASES, ids=[c.stem for c in REFERENCE_CASES])
def test_reference_golden(case: Path) -> None:
    assert case_output(case) == case.with_suffix(".expected").read_text()


@pytest.mark.parametrize("case", HIP_CASES, ids=[c.stem for c in HIP_CASES])
def test_hip_golden(case: Path) -> None:
    expected = case.with_suffix(".expected")
    text = case_output(case)
    if UPDATE:
        expected.write_text(text)
    assert text == expected.read_text()


PROGRAM_CASES = [c for c in HIP_CASES if c.stem not in ("hip_kernels_unknown", "hip_bad_mode")]


@pytest.mark.parametrize("case", PROGRAM_CASES, ids=[c.stem for c in PROGRAM_CASES])
def test_hip_program_compiles(case: Path) -> None:
    source = case.with_suffix(".expected")
    proc = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                           f"-I{ROOT / 'include'}", f"-I{PACKAGE_DIR / 'host'}", "-x", "c++",
                           str(source)], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stderr


@pytest.mark.parametrize("example", EXAMPLES, ids=[e.stem for e in EXAMPLES])
def test_examples_check_and_match_fixture_rates(example: Path) -> None:
    proc = sigflowc(example, "--emit=json")
    assert proc.returncode == 0 and proc.stderr == "", proc.stderr
    graph = json.loads(proc.stdout)
    for node in graph["nodes"]:
        if node["op"] == "capture":
            meta = json.loads((FIXTURES / Path(node["path"]).with_suffix(".json").name).read_text())
            assert float(node["rate"]) == meta["sample_rate"]


def test_usage_errors() -> None:
    assert sigflowc().returncode == 2
    proc = sigflowc("--emit=hip")
    assert proc.returncode == 2 and "no input file" in proc.stderr
    proc = sigflowc(EXAMPLES[0], "--bogus")
    assert proc.returncode == 2 and "unknown option '--bogus'" in proc.stderr
    proc = sigflowc("missing.sf")
    assert proc.returncode == 2 and "cannot read 'missing.sf'" in proc.stderr


# --- invariants on random programs ---------------------------------------------------------

CAPTURES = ["ch0", "ch1", "tone", "tone_odd", "short", "zeros"]
TAPS = ["lowpass", "lowpass_long", "identity"]


def random_program(rng: random.Random) -> str:
    """A random valid program: captures, then FIR / FFT / magnitude / xcorr / channelize /
    gcc_phat over earlier stream values, some values emitted, statements shuffled (names may be
    used before their binding)."""
    lines, streams, frames, values = [], [], [], []
    for k in range(rng.randint(1, 4)):
        name = f"c{k}"
        lines.append(f'{name} = capture("fixtures/{rng.choice(CAPTURES)}.cf32", rate=1e6)')
        streams.append(name)
    for k in range(rng.randint(1, 10)):
        name = f"v{k}"
        op = rng.choice(["fir", "fir", "fir", "fft", "xcorr", "gcc", "chan", "mag"])
        if op == "fir":
            lines.append(f'{name} = {rng.choice(streams)} |> fir(taps="fixtures/{rng.choice(TAPS)}.f32")')
            streams.append(name)
        elif op == "fft":
            lines.append(f"{name} = {rng.choice(streams)} |> fft(n={rng.choice([8, 64, 1024, 8192])})")
            frames.append(name)
        elif op == "chan":
            lines.append(f'{name} = {rng.choice(streams)} |> channelize('
                         f'taps="fixtures/channelizer_proto.f32", channels={rng.choice([4, 16, 64])})')
            frames.append(name)
        elif op == "mag" and frames:
            lines.append(f"{name} = {rng.choice(frames)} |> magnitude")
            values.append((name, ".f32"))
            continue
        elif op in ("xcorr", "gcc"):
            fn = "xcorr" if op == "xcorr" else "gcc_phat"
            lines.append(f"{name} = {fn}({rng.choice(streams)}, {rng.choice(streams)})")
            streams.append(name)
        else:
            continue
        values.append((name, ".cf32"))
    emits = rng.sample(values, rng.randint(1, len(values))) if values else [(streams[0], ".cf32")]
    for name, ext in emits:
        lines.append(f'emit {name} -> "out/{name}{ext}"')
    rng.shuffle(lines)
    return "\n".join(lines) + "\n"


def launches(graph: dict) -> dict[int, int]:
    """Launch index of every scheduled node: a fusion group's consecutive members share one."""
    nodes = {n["id"]: n for n in graph["nodes"]}
    out, launch, last_group = {}, -1, None
    for node_id in graph["schedule"]:
        group = nodes[node_id].get("group")
        if group is None or group != last_group:
            launch += 1
        out[node_id] = launch
        last_group = group
    return out


def check_plan(graph: dict) -> None:
    nodes = {n["id"]: n for n in graph["nodes"]}
    step = {node_id: k for k, node_id in enumerate(graph["schedule"])}
    assert sorted(step) == sorted(n["id"] for n in graph["nodes"] if n["live"])
    for e in graph["edges"]:
        if nodes[e["to"]]["live"]:
            assert step[e["from"]] < step[e["to"]], e
    users: dict[int, list[int]] = {}
    for e in graph["edges"]:
        users.setdefault(e["from"], []).append(e["to"])

    def reaches(a: int, b: int) -> bool:
        todo, seen = [a], set()
        while todo:
            v = todo.pop()
            if v == b:
                return True
            if v not in seen:
                seen.add(v)
                todo.extend(users.get(v, []))
        return False

    groups: dict[int, list[int]] = {}
    for n in graph["nodes"]:
        if n.get("group") is not None:
            groups.setdefault(n["group"], []).append(n["id"])
    for members in groups.values():
        assert 2 <= len(members) <= 8
        assert len({nodes[m]["taps"] for m in members}) == 1
        assert all(not reaches(a, b) for a in members for b in members if a != b)
        steps = sorted(step[m] for m in members)
        assert steps == list(range(steps[0], steps[0] + len(steps)))  # consecutive
    launch = launches(graph)
    by_slot: dict[int, list[dict]] = {}
    for n in graph["nodes"]:
        if n["live"] and n["op"] != "emit":
            lo, hi = n["live_range"]
            assert lo == step[n["id"]]
            readers = [step[u] for u in users.get(n["id"], []) if nodes[u]["live"]]
            assert hi == max(readers, default=lo)
            by_slot.setdefault(n["slot"], []).append(n)
    for values in by_slot.values():
        values.sort(key=lambda n: n["live_range"][0])
        for older, newer in zip(values, values[1:]):
            # the newer value is written only after the launch of the older one's last reader
            assert launch[graph["schedule"][older["live_range"][1]]] < \
                launch[graph["schedule"][newer["live_range"][0]]]


def expected_bytes(graph: dict) -> dict[int, int]:
    nodes = {n["id"]: n for n in graph["nodes"]}
    inputs: dict[int, list[int]] = {}
    for e in sorted(graph["edges"], key=lambda e: (e["to"], e["input"])):
        inputs.setdefault(e["to"], []).append(e["from"])
    samples: dict[int, int] = {}
    for node_id in sorted(nodes):
        n = nodes[node_id]
        ins = [samples[i] for i in inputs.get(node_id, [])]
        if n["op"] == "capture":
            samples[node_id] = (FIXTURES / Path(n["path"]).name).stat().st_size // 8
        elif n["op"] in ("fir", "magnitude"):
            samples[node_id] = ins[0]
        elif n["op"] in ("fft", "channelize"):
            m = n.get("n", n.get("channels"))
            samples[node_id] = -(-ins[0] // m) * m
        elif n["op"] in ("xcorr", "gcc_phat"):
            samples[node_id] = ins[0] + ins[1] - 1
    return {i: samples[i] * (4 if nodes[i]["type"].startswith("frames<f32") else 8)
            for i in samples}


@pytest.mark.parametrize("seed", range(40))
def test_random_program_plans(tmp_path: Path, seed: int) -> None:
    rng = random.Random(seed)
    source = tmp_path / f"p{seed}.sf"
    source.write_text(random_program(rng))
    for flags in ([], ["--no-fuse"], ["--no-reuse"], ["--no-dce"]):
        proc = sigflowc(source, "--emit=json", f"--input-dir={FIXTURES.parent}", *flags)
        assert proc.returncode == 0, source.read_text() + proc.stderr
        graph = json.loads(proc.stdout)
        check_plan(graph)
        sizes = expected_bytes(graph)
        for n in graph["nodes"]:
            if n["live"] and n["op"] != "emit":
                assert n["bytes"] == sizes[n["id"]], n
        if "--no-reuse" in flags:
            assert graph["slots"] == sum(1 for n in graph["nodes"] if n["live"] and n["op"] != "emit")
        hip = sigflowc(source, "--emit=hip", *flags)
        assert hip.returncode == 0 and "int main" in hip.stdout
        check_hip_slots(hip.stdout, graph)


def check_hip_slots(program: str, graph: dict) -> None:
    """The HIP backend's own slot plan (after fft |> magnitude fusion), from the launch ranges
    it writes beside each rt.need: a slot passes to a new value only after the launch of the old
    value's last reader, and emitted values (copied back after the compute phase) keep their
    slots to the last launch."""
    ranges = {}
    for slot, node, first, last in re.findall(
            r"rt\.need\((\d+), n\d+ \* sizeof\(\w+\)\);  // %(\d+), launches (\d+)\.\.(\d+)",
            program):
        ranges[int(node)] = (int(slot), int(first), int(last))
    assert ranges
    final = max(last for _, _, last in ranges.values())
    nodes = {n["id"]: n for n in graph["nodes"]}
    for e in graph["edges"]:
        if nodes[e["to"]]["op"] == "emit" and nodes[e["from"]]["op"] != "capture":
            assert ranges[e["from"]][2] == final, e
    by_slot: dict[int, list[tuple[int, int]]] = {}
    for slot, first, last in ranges.values():
        by_slot.setdefault(slot, []).append((first, last))
    for spans in by_slot.values():
        spans.sort()
        for (_, old_last), (new_first, _) in zip(spans, spans[1:]):
            assert old_last < new_first, spans


def mutate(text: str, rng: random.Random) -> str:
    alphabet = list('abcxyz_019.e+-|>()=,"#\n \t$') + ["emit", "fir", "|>", "->", "capture("]
    chars = list(text)
    for _ in range(rng.randint(1, 6)):
        k = rng.randrange(len(chars) + 1)
        action = rng.random()
        if action < 0.4 and chars:
            del chars[min(k, len(chars) - 1)]
        elif action < 0.8:
            chars.insert(k, rng.choice(alphabet))
        elif chars:
            j = rng.randrange(len(chars))
            chars[min(k, len(chars) - 1)], chars[j] = chars[j], chars[min(k, len(chars) - 1)]
    return "".join(chars)


def test_mutated_programs_never_crash(tmp_path: Path) -> None:
    rng = random.Random(0x5F)
    seeds = [e.read_text() for e in EXAMPLES] + [c.read_text() for c in REFERENCE_CASES]
    modes = ["--emit=tokens", "--emit=ast", "--emit=check", "--emit=json", "--emit=hip"]
    source = tmp_path / "m.sf"
    outcomes = {0: 0, 1: 0}
    for k in range(300):
        source.write_bytes(mutate(rng.choice(seeds), rng).encode("utf-8", "replace"))
        proc = sigflowc(source.name, rng.choice(modes), cwd=tmp_path)
        assert proc.returncode in (0, 1), (source.read_text(), proc.returncode, proc.stderr)
        if proc.returncode == 1:
            assert re.search(r"^m\.sf:\d+:\d+: error: ", proc.stderr, re.M), proc.stderr
        outcomes[proc.returncode] += 1
    assert outcomes[0] > 20 and outcomes[1] > 20


def test_limits_end_in_diagnostics(tmp_path: Path) -> None:
    """Inputs that would exhaust the stack or run long: one error, exit 1, quickly."""
    fir = 'fir(taps="h.f32")'
    cases = {
        "nest": "a = " + "(" * 300 + 'capture("x.cf32", rate=1e6)' + ")" * 300
                + '\nemit a -> "o.cf32"\n',
        "pipe": 'a = capture("x.cf32", rate=1e6)' + f" |> {fir}" * 300 + '\nemit a -> "o.cf32"\n',
        # used before bound, so resolving the emit walks the whole chain at once
        "chain": "\n".join(['emit a1499 -> "o.cf32"']
                           + [f"a{i} = a{i - 1} |> {fir}" for i in range(1499, 0, -1)]
                           + ['a0 = capture("x.cf32", rate=1e6)']) + "\n",
        "operations": "\n".join(['a0 = capture("x.cf32", rate=1e6)']
                                + [f"a{i} = a0 |> {fir}" for i in range(1, 10050)]
                                + ['emit a1 -> "o.cf32"']) + "\n",
    }
    messages = {"nest": "expression nested too deeply", "pipe": "expression nested too deeply",
                "chain": "bindings nested too deeply", "operations": "more than 10000 operations"}
    for name, text in cases.items():
        (tmp_path / f"{name}.sf").write_text(text)
        proc = sigflowc(f"{name}.sf", "--emit=hip", cwd=tmp_path)
        assert proc.returncode == 1, name
        assert proc.stderr.count(": error: ") == 1 and messages[name] in proc.stderr, proc.stderr[:300]


def test_large_programs_compile(tmp_path: Path) -> None:
    """Near the limits the passes stay fast: 9000 FIRs of one capture, 1000 of them emitted
    (125 fusion groups of 8 after dead-node elimination), and a chain of 900 filters."""
    fir = 'fir(taps="h.f32")'
    wide = (['a0 = capture("x.cf32", rate=1e6)'] + [f"a{i} = a0 |> {fir}" for i in range(1, 9000)]
            + [f'emit a{i} -> "o{i}.cf32"' for i in range(1, 9000, 9)])
    chain = (['a0 = capture("x.cf32", rate=1e6)'] + [f"a{i} = a{i - 1} |> {fir}" for i in range(1, 900)]
             + ['emit a899 -> "o.cf32"'])
    for name, lines in [("wide", wide), ("chain", chain)]:
        (tmp_path / f"{name}.sf").write_text("\n".join(lines) + "\n")
        proc = sigflowc(f"{name}.sf", "--emit=json", cwd=tmp_path)
        assert proc.returncode == 0, proc.stderr[:300]
        graph = json.loads(proc.stdout)
        check_plan(graph)
        groups = {n["group"] for n in graph["nodes"] if n.get("group") is not None}
        assert len(groups) == (125 if name == "wide" else 0)
