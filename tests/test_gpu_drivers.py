"""The mm_1d / mm_2d executables and the runner on the GPU: reference CLI,
validation, BENCHMARK_RESULT contract and exit codes (1D-row-block/mmb.cu:
927-967, 2D-block-matrix/mmb.cu:377-401)."""
import json
import os
import subprocess
import sys
from pathlib import Path

import pytest

from mgxbench_amd.runner import parse_result
from tests.gpu_util import announce

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "mgxbench_amd" / "bin"


def _run(exe, args, env=None, timeout=120):
    e = os.environ.copy()
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
    assert rec["valid"] == "1" and rec["panels"] == "4"     # 256,256,256,232


def test_mm_2d_variants():
    for env in ({"MGX_FORM": "int8"}, {"MGX_SUMMA_GRAPH": "1"}, {"MGX_HOST_PINNED": "0"},
                {"MGX_INIT": "uniform:2", "MGX_SHARE_GRID": "2x2"},
                {"MGX_INIT": "uniform:2", "MGX_FORM": "int8", "MGX_SUMMA_GRAPH": "1"},
                {"MGX_INIT": "uniform:2", "MGX_FORM": "mod", "MGX_SUMMA_GRAPH": "1"},
                {"MGX_FORM": "mod", "MGX_SHARE_GRID": "2x2"},
                {"MGX_INIT": "uniform:2", "MGX_FORM": "mod", "MGX_SUMMA_BK": "128",
                 "MGX_SPLIT_K": "300"}):
        proc, out = _run("mm_2d", ["0", "1024", "NNN"], env)
        assert proc.returncode == 0, out
        assert parse_result(out)["valid"] == "1"


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
