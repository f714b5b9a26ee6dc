"""The mm_1d / mm_2d executables and the runner on the GPU: reference CLI,
validation, BENCHMARK_RESULT contract and exit codes (1D-row-block/mmb.cu:
927-967, 1036; 2D-block-matrix/mmb.cu:377-401, 439).

With no MGX_* knob set, the result line must be the reference's format string
exactly (no extra keys); a planted bad element of C (MGX_TEST_PERTURB) must
turn validation to FAIL and the exit code to 2 (1D) or 1 (2D)."""
import json
import os
import re
import subprocess
import sys
from pathlib import Path

import pytest

from mgxbench_amd.runner import parse_result
from tests.gpu_util import announce

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "mgxbench_amd" / "bin"

# printf("BENCHMARK_RESULT decomposition=1D library=%s size=%d ranks=%d seconds=%.9f valid=%d
# loops=%d\n") (1D-row-block/mmb.cu:964) and its 2D twin with "loops=1 grid=%dx%d"
# (2D-block-matrix/mmb.cu:398), as whole-line patterns.
LINE_1D = re.compile(r"^BENCHMARK_RESULT decomposition=1D library=(\S+) size=(-?\d+) "
                     r"ranks=(-?\d+) seconds=(-?\d+\.\d{9}) valid=(-?\d+) loops=(-?\d+)$")
LINE_2D = re.compile(r"^BENCHMARK_RESULT decomposition=2D library=(\S+) size=(-?\d+) "
                     r"ranks=(-?\d+) seconds=(-?\d+\.\d{9}) valid=(-?\d+) loops=1 "
                     r"grid=(-?\d+)x(-?\d+)$")
MGX_KNOBS = ("MGX_FORM", "MGX_SLICES", "MGX_INIT", "MGX_SHARE_RANKS", "MGX_SHARE_GRID",
             "MGX_SPLIT_K", "MGX_SUMMA_GRAPH", "MGX_HOST_PINNED", "MGX_SUMMA_SERIALIZE",
             "MGX_RESULT_EXTRAS", "MGX_TEST_PERTURB")


def _result_lines(out):
    return [l for l in out.splitlines() if l.startswith("BENCHMARK_RESULT")]


def _run(exe, args, env=None, timeout=120):
    e = {k: v for k, v in os.environ.items() if k not in MGX_KNOBS}
    e.update(env or {})
    announce(f"{exe} {' '.join(args)} {env or ''}")
    proc = subprocess.run([str(BIN / exe), *args], capture_output=True, text=True, env=e,
                          timeout=timeout)
    return proc, proc.stdout + proc.stderr


@pytest.mark.parametrize("lib", ["MMM", "NNN", "SSS", "YYY", "XXX", "CCC", "WWW", "SNS", "MYX"])
def test_mm_1d_codes(lib):
    proc, out = _run("mm_1d", ["auto", "1024", lib],
                     {"MGXBENCH_INNER_LOOPS": "2", "MM_PIPELINE_K_CHUNK": "300"})
    assert proc.returncode == 0, out
    rec = parse_result(out)
    assert rec["decomposition"] == "1D" and rec["library"] == lib and rec["size"] == "1024"
    assert rec["ranks"] == "1" and rec["valid"] == "1" and rec["loops"] == "2"
    assert float(rec["seconds"]) > 0
    assert "NUMERICAL VALIDATION" in out and "Result             : PASS" in out
    lines = _result_lines(proc.stdout)
    assert len(lines) == 1 and LINE_1D.match(lines[0]), lines


def test_mm_1d_int8_and_uniform():
    for env in ({"MGX_FORM": "int8", "MGX_SLICES": "8"},
                {"MGX_INIT": "uniform:5"},
                {"MGX_FORM": "int8", "MGX_INIT": "uniform:5"},
                {"MGX_FORM": "int8", "MGX_INIT": "uniform:5", "MGX_SHARE_RANKS": "4"},
                {"MGX_FORM": "mod", "MGX_INIT": "uniform:5"},
                {"MGX_FORM": "mod", "MGX_SLICES": "16", "MGX_INIT": "uniform:5",
                 "MGX_SHARE_RANKS": "4"},
                {"MGX_FORM": "mod", "MGX_INIT": "uniform:5", "MM_PIPELINE_K_CHUNK": "256",
                 "MGX_SPLIT_K": "300"}):
        proc, out = _run("mm_1d", ["0", "1024", "SSS"], {"MGXBENCH_INNER_LOOPS": "1", **env})
        assert proc.returncode == 0, out
        assert parse_result(out)["valid"] == "1"


def test_mm_1d_rejects_bad_cli():
    proc, out = _run("mm_1d", ["auto", "1024", "QQQ"])
    assert proc.returncode != 0 and "Invalid communication string" in out
    proc, out = _run("mm_1d", ["auto", "1024"])
    assert proc.returncode != 0


@pytest.mark.parametrize("lib", ["MMM", "YYY", "CCC", "XXX", "NNN", "WWW", "SSS"])
def test_mm_2d_codes(lib):
    proc, out = _run("mm_2d", ["auto", "1000", lib], {"MGXBENCH_SUMMA_BK": "256"})
    assert proc.returncode == 0, out
    rec = parse_result(out)
    assert rec["decomposition"] == "2D" and rec["grid"] == "1x1" and rec["loops"] == "1"
    assert rec["valid"] == "1"
    lines = _result_lines(proc.stdout)
    assert len(lines) == 1 and LINE_2D.match(lines[0]), lines
    # the extra keys appear only on request
    proc, out = _run("mm_2d", ["auto", "1000", lib],
                     {"MGXBENCH_SUMMA_BK": "256", "MGX_RESULT_EXTRAS": "1"})
    assert proc.returncode == 0, out
    rec = parse_result(out)
    assert rec["panels"] == "4" and rec["form"] == "fp64"     # 256,256,256,232


def test_mm_2d_variants():
    for env in ({"MGX_FORM": "int8"}, {"MGX_SUMMA_GRAPH": "1"}, {"MGX_HOST_PINNED": "1"},
                {"MGX_INIT": "uniform:2", "MGX_SHARE_GRID": "2x2"},
                {"MGX_INIT": "uniform:2", "MGX_FORM": "int8", "MGX_SUMMA_GRAPH": "1"},
                {"MGX_INIT": "uniform:2", "MGX_FORM": "mod", "MGX_SUMMA_GRAPH": "1"},
                {"MGX_FORM": "mod", "MGX_SHARE_GRID": "2x2"},
                {"MGX_INIT": "uniform:2", "MGX_FORM": "mod", "MGX_SUMMA_BK": "128",
                 "MGX_SPLIT_K": "300"}):
        proc, out = _run("mm_2d", ["0", "1024", "NNN"], env)
        assert proc.returncode == 0, out
        assert parse_result(out)["valid"] == "1"


@pytest.mark.parametrize("perturb,count", [
    ("5:2048.5", 1),          # |err| 0.5: over both tolerances
    ("0:2048.0000000015", 0),  # |err| 1.5e-9 > abs_tol but relative 7.3e-13 <= rel_tol
    ("9:2048.00000001", 1),    # |err| 1e-8: relative 4.9e-12, over both
    ("1048575:inf", 1),       # Inf: counted, max_abs = inf
    ("77:nan", 0),            # NaN compares false in the reference rule: not counted
])
def test_mm_1d_planted_element(perturb, count):
    # N = 1024: expected C = 2048, abs_tol 1e-9, rel_tol 1e-12 (rel bound 2.048e-9)
    proc, out = _run("mm_1d", ["auto", "1024", "NNN"],
                     {"MGXBENCH_INNER_LOOPS": "1", "MGX_TEST_PERTURB": perturb})
    lines = _result_lines(proc.stdout)
    assert len(lines) == 1 and LINE_1D.match(lines[0]), lines
    rec = parse_result(out)
    assert f"Invalid elements   : {count}" in out, out
    if count:
        assert proc.returncode == 2 and rec["valid"] == "0"
        assert "Result             : FAIL" in out
    else:
        assert proc.returncode == 0 and rec["valid"] == "1"
    if "inf" in perturb:
        assert "Max absolute error : inf" in out


@pytest.mark.parametrize("perturb,bad", [
    ("12345:2048.5", True), ("0:nan", True), ("999999:-inf", True),
    ("3:2048.000000002", False),   # |err| 2e-9 <= 1e-9 + 1e-12 * 2048 = 3.048e-9
])
def test_mm_2d_planted_element(perturb, bad):
    proc, out = _run("mm_2d", ["auto", "1024", "NNN"], {"MGX_TEST_PERTURB": perturb})
    lines = _result_lines(proc.stdout)
    assert len(lines) == 1 and LINE_2D.match(lines[0]), lines
    rec = parse_result(out)
    if bad:
        idx = perturb.split(":")[0]
        assert proc.returncode == 1 and rec["valid"] == "0", out
        assert f"2D validation failed at local index={idx}" in out
    else:
        assert proc.returncode == 0 and rec["valid"] == "1", out


def test_perturbation_outside_c_is_refused():
    proc, out = _run("mm_1d", ["auto", "64", "NNN"], {"MGX_TEST_PERTURB": "4096:1"})
    assert proc.returncode != 0 and "MGX_TEST_PERTURB" in out
    proc, out = _run("mm_2d", ["auto", "64", "NNN"], {"MGX_TEST_PERTURB": "x"})
    assert proc.returncode != 0 and "MGX_TEST_PERTURB" in out


def test_mm_2d_rejects_bad_grid():
    proc, out = _run("mm_2d", ["auto", "1024", "NNN"], {"MGXBENCH_GRID_ROWS": "2"})
    assert proc.returncode != 0 and "invalid process grid" in out


def test_runner_drives_executables(tmp_path):
    announce("runner end to end")
    cmd = [sys.executable, "-m", "mgxbench_amd.runner", "--decomposition", "2d", "--size", "1024",
           "--ranks", "1", "--warmup", "1", "--repeats", "2", "--libraries", "MMM", "NNN",
           "--output", str(tmp_path)]
    proc = subprocess.run(cmd, capture_output=True, text=True, cwd=ROOT, timeout=300)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    doc = json.loads(next(tmp_path.rglob("benchmark.json")).read_text())
    assert [r["library"] for r in doc["results"]] == ["MMM", "NNN"]
    assert all(len(r["samples_s"]) == 2 for r in doc["results"])
