"""Campaign sizes (mgxbench_amd.campaign, the run_campaign.py counterpart) and
the offline report (mgxbench_amd.report, the plot_benchmark.py counterpart),
on the CPU: the weak-scaling size rule against hand-computed cases, pooling
and scaling rules on synthetic reports, and a campaign driving the runner
over a stand-in executable end to end."""
import csv
import json
import sys
from pathlib import Path

import pytest

from mgxbench_amd import campaign, report


def test_weak_scaling_sizes():
    assert campaign.weak_scaling_size(4096, 1) == (4096.0, 4096)
    ideal, n = campaign.weak_scaling_size(4096, 2)
    assert ideal == pytest.approx(4096 * 2 ** (1 / 3)) and n == 5160
    assert campaign.weak_scaling_size(4096, 8)[1] == 8192
    # 2D: the multiple is lcm(alignment, lcm(grid rows, grid cols))
    assert campaign.weak_scaling_size(4096, 6, "2d", 4)[1] % 12 == 0
    with pytest.raises(ValueError):
        campaign.weak_scaling_size(0, 2)


def _points(*args, **kw):
    return [pt.as_tuple() for pt in campaign.plan(*args, **kw)]


def test_plan_rules():
    logs = []
    cfg = _points(["1d"], [4096, 4098], [1, 4], "sweep", 4096, log=logs.append)
    assert cfg == [("1d", 4096, 1), ("1d", 4096, 4)]
    assert any("4098" in s for s in logs)
    weak = _points(["1d", "2d"], [], [1, 2, 4, 8], "weak", 4096)
    for d, n, p in weak:
        assert abs(n ** 3 / p / 4096 ** 3 - 1) <= 0.025
    with pytest.raises(ValueError):                      # rounding moves N^3/p by 12.5 %
        _points(["1d"], [], [3], "weak", 8)
    logs.clear()
    # fixed 2 grid rows: p = 4 -> 2x2 (multiple of 4), p = 6 -> 2x3 (multiple of 12:
    # 4096 is not), p = 3 has no grid with 2 rows
    cfg = _points(["2d"], [4096], [4, 6, 3], "sweep", 4096, rows=2, log=logs.append)
    assert cfg == [("2d", 4096, 4)]
    assert any("p=3" in s for s in logs) and any("p=6" in s for s in logs)


def _doc(path: Path, label, env, rows):
    path.mkdir(parents=True, exist_ok=True)
    (path / "benchmark.json").write_text(json.dumps(
        {"label": label, "environment": {"id": env}, "results": rows}))


def _row(lib, n, p, samples, d="2d", validated=True):
    import statistics
    return {"decomposition": d, "library": lib, "size": n, "ranks": p,
            "mean_s": statistics.mean(samples), "stddev_s": statistics.pstdev(samples),
            "validated": validated, "samples_s": samples, "original_inner_loops": 1}


def test_report_pooling_speedup_and_scaling(tmp_path):
    src = tmp_path / "runs"
    # two runs of the same configuration pool their samples
    _doc(src / "a", "c1", "E", [_row("MMM", 4096, 1, [2.0, 2.2]), _row("NNN", 4096, 1, [1.0])])
    _doc(src / "b", "c1", "E", [_row("NNN", 4096, 1, [1.2]), _row("NNN", 4096, 2, [0.6]),
                                _row("NNN", 4096, 4, [0.4]), _row("MMM", 8192, 1, [16.0]),
                                _row("SSS", 4096, 1, [9.0], validated=False)])
    # weak scaling: N^3 / p constant
    _doc(src / "w", "weak:w1", "E", [_row("NNN", 4096, 1, [1.0]), _row("NNN", 5160, 2, [1.1])])
    rows = report.load_results([src])
    assert len(rows) == 8          # the unvalidated row is dropped
    data = report.consolidate(rows)
    nnn = next(r for r in data if r["library"] == "NNN" and r["ranks"] == 1
               and r["label"] == "c1")
    assert nnn["samples_s"] == [1.0, 1.2] and nnn["mean_s"] == pytest.approx(1.1)
    assert nnn["run_count"] == 2
    assert nnn["speedup_vs_MMM"] == pytest.approx(2.1 / 1.1)
    out = report.generate(rows, tmp_path / "rep")
    assert out.name == "report.html" and out.exists()
    figs = {p.name for p in (tmp_path / "rep").glob("*.png")}
    assert any(f.endswith("_2d_N4096_strong_scaling.png") for f in figs)
    assert any(f.endswith("_2d_NNN_weak_scaling.png") for f in figs)
    assert any(f.endswith("_2d_p1_size_scalability.png") for f in figs)
    assert any(f.endswith("_N4096_p1_speedup.png") for f in figs)
    summary = list(csv.DictReader((tmp_path / "rep" / "summary.csv").open()))
    assert len(summary) == len(data)
    assert "Only numerically validated" in (tmp_path / "rep" / "report.md").read_text()


def test_report_rejects_unpoolable_and_invalid(tmp_path):
    for i in range(2):
        d = tmp_path / "csv" / f"r{i}"
        d.mkdir(parents=True)
        with (d / "benchmark.csv").open("w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=["decomposition", "library", "size", "ranks",
                                              "mean_s", "stddev_s", "validated",
                                              "original_inner_loops", "label",
                                              "environment_id"])
            w.writeheader()
            w.writerow({"decomposition": "1d", "library": "MMM", "size": 64, "ranks": 1,
                        "mean_s": 1.0, "stddev_s": 0.0, "validated": "True",
                        "original_inner_loops": 1, "label": "x", "environment_id": "E"})
    rows = report.load_results([tmp_path / "csv"])
    with pytest.raises(ValueError):
        report.consolidate(rows)
    bad = tmp_path / "bad"
    _doc(bad / "r", "x", "E", [_row("MMM", 64, 1, [1.0]) | {"mean_s": -1.0}])
    with pytest.raises(SystemExit) as exc:
        report.main(["--input", str(bad), "--output", str(tmp_path / "o")])
    assert exc.value.code == 2
    with pytest.raises(SystemExit) as exc:
        report.main(["--input", str(tmp_path / "nothing")])
    assert exc.value.code == 2


def test_campaign_end_to_end_through_the_runner(tmp_path, monkeypatch):
    monkeypatch.chdir(Path(__file__).resolve().parents[1])
    build = tmp_path / "bin"
    build.mkdir()
    exe = build / "mm_1d"
    exe.write_text("#!" + sys.executable + "\n" + (
        "import sys\n"
        "n, lib = int(sys.argv[2]), sys.argv[3]\n"
        "rate = {'MMM': 4e9, 'NNN': 8e9}[lib]\n"
        "print('BENCHMARK_RESULT decomposition=1D library=%s size=%d ranks=1 seconds=%.9f "
        "valid=1 loops=1' % (lib, n, n ** 3 / rate))\n"))
    exe.chmod(0o755)
    out = tmp_path / "camp"
    rc = campaign.main(["--decompositions", "1d", "--sizes", "256", "512", "--ranks", "1",
                        "--output", str(out), "--label", "t", "--",
                        "--libraries", "MMM", "NNN", "--repeats", "2", "--warmup", "0",
                        "--build-dir", str(build)])
    assert rc == 0
    manifest = list(csv.DictReader((out / "t-size-manifest.csv").open()))
    assert [(m["decomposition"], m["N_actual"], m["ranks"]) for m in manifest] == \
        [("1d", "256", "1"), ("1d", "512", "1")]
    summary = list(csv.DictReader((out / "report" / "summary.csv").open()))
    assert len(summary) == 4
    nnn = next(r for r in summary if r["library"] == "NNN" and r["size"] == "512")
    assert float(nnn["mean_s"]) == pytest.approx(512 ** 3 / 8e9, rel=1e-6)
    assert float(nnn["speedup_vs_MMM"]) == pytest.approx(2.0, rel=1e-6)
    with pytest.raises(SystemExit):
        campaign.main(["--mode", "strong", "--sizes", "256", "512"])
    with pytest.raises(SystemExit):
        campaign.main(["--sizes", "256", "--", "--size", "512"])
