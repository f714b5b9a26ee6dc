"""Runner, launcher shim, grid and statistics (CPU only, fake launchers).

Mirrors the reference's own runner tests (tests/test_workflow.py:20-99,
tests/test_rectangular_summa.py:11-21): a fake launcher prints a
BENCHMARK_RESULT record, and the runner's acceptance rules, statistics and
JSON / CSV outputs are checked against the reference schema.
"""
import csv
import json
import subprocess
import sys
from pathlib import Path

import pytest

from mgxbench_amd import experiment, grid, launch, runner

ROOT = Path(__file__).resolve().parents[1]


# ---- record parsing and statistics ----------------------------------------------------

def test_parser_rejects_duplicate_records():
    with pytest.raises(ValueError):
        runner.parse_result("BENCHMARK_RESULT valid=1\nBENCHMARK_RESULT valid=1")
    with pytest.raises(ValueError):
        runner.parse_result("no record")
    rec = runner.parse_result("x\nBENCHMARK_RESULT decomposition=2D library=NNN size=8 grid=1x1\n")
    assert rec == {"decomposition": "2D", "library": "NNN", "size": "8", "grid": "1x1"}


def test_statistics_and_history(tmp_path):
    assert experiment.sample_statistics([1, 1, 1])["ci95_low_s"] == 1
    assert experiment.sample_statistics([1])["ci95_low_s"] is None
    a, _ = experiment.new_run(tmp_path)
    b, _ = experiment.new_run(tmp_path)
    assert a != b


def test_bootstrap_bounds_closed_form():
    # Two samples: every resample mean is one of {a, (a+b)/2, b}; the bounds of
    # 2000 resamples must be a and b, and the CV is the population CV.
    st = experiment.sample_statistics([1.0, 3.0])
    assert st["ci95_low_s"] == 1.0 and st["ci95_high_s"] == 3.0
    assert st["cv_pct"] == pytest.approx(50.0)


def test_percentile_linear_interpolation():
    assert runner.percentile([1.0, 2.0, 3.0, 4.0], 0.95) == pytest.approx(3.85)
    assert runner.percentile([5.0], 0.95) == 5.0


# ---- grid ------------------------------------------------------------------------------

@pytest.mark.parametrize("ranks,g", [(1, (1, 1)), (4, (2, 2)), (8, (2, 4)), (12, (3, 4)), (16, (4, 4))])
def test_auto_grid(ranks, g):
    assert grid.resolve_grid(ranks) == g


@pytest.mark.parametrize("ranks,rows,cols", [(8, 2, 4), (8, 4, 2), (16, 4, 4), (16, 2, 8)])
def test_explicit_grid(ranks, rows, cols):
    assert grid.resolve_grid(ranks, rows, cols) == (rows, cols)


@pytest.mark.parametrize("ranks,rows,cols", [(8, 3, 4), (16, 2, 4), (8, 0, 8)])
def test_invalid_grid(ranks, rows, cols):
    with pytest.raises(ValueError):
        grid.resolve_grid(ranks, rows, cols)


# ---- launcher shim ---------------------------------------------------------------------

def test_launch_segments():
    segs = launch.split_segments(["-np", "1", "a", "x", ":", "-np", "1", "b"])
    assert [launch.parse_segment(s) for s in segs] == [(1, ["a", "x"]), (1, ["b"])]


def test_launch_runs_one_rank_and_refuses_more(tmp_path):
    prog = tmp_path / "p.py"
    prog.write_text("import sys; print('hi', sys.argv[1:]); sys.exit(7)")
    rc = launch.main(["-np", "1", sys.executable, str(prog), "auto", "16", "NNN"])
    assert rc == 7
    assert launch.main(["-np", "1", "x", ":", "-np", "1", "y"]) == 2
    assert launch.main(["-np", "2", "x"]) == 2
    assert launch.main([]) == 2


def test_launch_help(capsys):
    assert launch.main(["--help"]) == 0
    assert "-np 1" in capsys.readouterr().out


# ---- end-to-end runner with a simulated launcher ------------------------------------

def _fake_launcher(tmp_path, body):
    launcher = tmp_path / "launcher.py"
    launcher.write_text(body)
    build = tmp_path / "build"
    build.mkdir(exist_ok=True)
    for exe in ("mm_1d", "mm_2d"):
        (build / exe).touch()
    return launcher, build


def _cmd(tmp_path, launcher, build, *extra):
    return [sys.executable, "-m", "mgxbench_amd.runner", "--decomposition", "1d", "--size", "16",
            "--ranks", "1", "--warmup", "1", "--repeats", "2", "--libraries", "MMM", "NNN",
            "--launcher", sys.executable, "--launcher-args", str(launcher), "--build-dir",
            str(build), "--output", str(tmp_path / "out"), *extra]


def test_runner_with_simulated_launcher(tmp_path):
    launcher, build = _fake_launcher(
        tmp_path,
        "import sys\nn=sys.argv[-2];lib=sys.argv[-1]\n"
        "print(f'BENCHMARK_RESULT decomposition=1D library={lib} size={n} ranks=1 "
        "seconds=0.5 valid=1 loops=1')\n")
    cmd = _cmd(tmp_path, launcher, build)
    for _ in range(2):
        proc = subprocess.run(cmd, capture_output=True, text=True, cwd=ROOT)
        assert proc.returncode == 0, proc.stderr
    files = sorted((tmp_path / "out").rglob("benchmark.json"))
    assert len(files) == 2
    doc = json.loads(files[0].read_text())
    assert doc["schema_version"] == 2
    assert doc["results"][0]["samples_s"] == [0.5, 0.5]
    assert doc["results"][1]["speedup_vs_MMM"] == 1
    assert doc["results"][0]["gflops"] == pytest.approx(2 * 16 ** 3 / 0.5 / 1e9)
    for key in ("run_id", "label", "environment", "application", "measurement", "note",
                "bandwidth", "timestamp_utc", "benchmark_name", "benchmark_full_name",
                "command_args", "results", "failures"):
        assert key in doc
    with open(files[0].parent / "benchmark.csv") as f:
        rows = list(csv.DictReader(f))
    assert list(rows[0].keys()) == runner.CSV_COLUMNS
    assert rows[0]["library"] == "MMM"


@pytest.mark.parametrize("record,why", [
    ("size=16 ranks=1 seconds=0.5 valid=0 loops=1", "valid=0"),
    ("size=32 ranks=1 seconds=0.5 valid=1 loops=1", "size mismatch"),
    ("size=16 ranks=1 seconds=0.5 valid=1 loops=3", "loops mismatch"),
    ("size=16 ranks=1 seconds=-1 valid=1 loops=1", "bad time"),
])
def test_runner_rejects_bad_records(tmp_path, record, why):
    launcher, build = _fake_launcher(
        tmp_path,
        "import sys\nlib=sys.argv[-1]\n"
        f"print(f'BENCHMARK_RESULT decomposition=1D library={{lib}} {record}')\n")
    proc = subprocess.run(_cmd(tmp_path, launcher, build), capture_output=True, text=True, cwd=ROOT)
    assert proc.returncode == 2, why
    doc = json.loads(next((tmp_path / "out").rglob("benchmark.json")).read_text())
    assert doc["results"] == [] and len(doc["failures"]) == 2


def test_runner_timeout_and_retry(tmp_path):
    launcher, build = _fake_launcher(tmp_path, "import time\ntime.sleep(5)\n")
    cmd = _cmd(tmp_path, launcher, build, "--timeout", "1", "--retries", "1")
    proc = subprocess.run(cmd, capture_output=True, text=True, cwd=ROOT)
    assert proc.returncode == 2
    assert "retry 1/1" in proc.stderr


def test_runner_argument_rules(tmp_path):
    base = [sys.executable, "-m", "mgxbench_amd.runner", "--size", "16", "--ranks", "3"]
    bad = [
        ["--decomposition", "1d"],                                   # 16 % 3
        ["--decomposition", "2d", "--inner-loops", "2"],
        ["--decomposition", "1d", "--grid-rows", "1"],
        ["--decomposition", "1d", "--libraries", "AUTO", "MMM"],
        ["--decomposition", "1d", "--auto-feedback"],
    ]
    for extra in bad:
        proc = subprocess.run(base + extra, capture_output=True, text=True, cwd=ROOT)
        assert proc.returncode == 2, extra


def test_child_environment_keys():
    ap = runner.build_parser()
    args = ap.parse_args(["--decomposition", "2d", "--size", "16", "--ranks", "4",
                          "--form", "int8", "--slices", "7", "--init", "uniform:3"])
    env = runner.child_environment(args, (2, 2))
    assert env["MGXBENCH_GRID_ROWS"] == "2" and env["MGXBENCH_GRID_COLS"] == "2"
    assert env["MGXBENCH_SUMMA_BK"] == "256" and env["MGXBENCH_INNER_LOOPS"] == "1"
    assert env["MGX_FORM"] == "int8" and env["MGX_SLICES"] == "7" and env["MGX_INIT"] == "uniform:3"
    cmd = runner.launch_command(args, Path("/x/mm_2d"), "NNN")
    assert cmd.count(":") == 3
    assert cmd[-6:] == ["-np", "1", "/x/mm_2d", "auto", "16", "NNN"]


def test_environment_probe_without_rocm_smi(monkeypatch):
    monkeypatch.setattr(experiment, "probe", lambda cmd: None)
    assert experiment._gpu_inventory() is None
    env = experiment.collect_environment()
    assert len(env["id"]) == 16 and "library_versions" in env


def test_gpu_inventory_parses_rocm_smi_csv(monkeypatch):
    csv_text = ("device,PCI Bus,Card Series,Card Model,Card Vendor,Card SKU,GFX Version\n"
                "card0,0000:05:00.0,AMD Instinct MI355X,0x75a3,AMD,MI355X,gfx950\n")
    monkeypatch.setattr(experiment, "probe", lambda cmd: csv_text)
    inv = experiment._gpu_inventory()
    assert inv[0]["name"] == "AMD Instinct MI355X" and inv[0]["gfx_version"] == "gfx950"
    assert inv[0]["index"] == "0"
