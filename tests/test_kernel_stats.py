"""rocprofv3 kernel-summary reader, plots and MFMA join (mgxbench_amd.kernel_stats),
following the cases of the reference's tests/test_nsys_stats.py: preamble and
byte-order mark, quoted names, invalid rows discarded and counted, top
selection not renormalised, missing columns / empty / invalid tables
rejected, export through a (simulated) profiler tool with a logged
transcript and no unrelated fallback, CLI exit codes.  The rocprofv3 and
counter fixtures under tests/golden/rocprof were written on an MI355X."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

from mgxbench_amd import kernel_stats as ks

ROOT = Path(__file__).resolve().parents[1]
GOLD = ROOT / "tests" / "golden" / "rocprof"


def test_preamble_bom_quoted_names_and_invalid_rows(tmp_path):
    path = tmp_path / "kernels.csv"
    path.write_text('﻿rocprofv3 message\n\nTime (%),Total Time (ns),Name\n60%,600,"gemm<double, double>"\n'
                    "40,400,rcclKernel\nnan,0,bad\n-1,0,negative\ninf,0,infinite\n")
    df, name, percent, discarded = ks.read_summary(path)
    assert discarded == 3
    assert df[name].tolist() == ["gemm<double, double>", "rcclKernel"]
    assert df[percent].tolist() == [60, 40]


def test_rocprofv3_kernel_stats_layout():
    df, name, percent, discarded = ks.read_summary(GOLD / "bench_kernel_stats.csv")
    assert (name, percent, discarded) == ("Name", "Percentage", 0)
    assert len(df) == 19
    assert df[percent].sum() == pytest.approx(100.0, abs=0.05)
    assert df[name].iloc[0].startswith("void (anonymous namespace)::dgemm_mfma<128, 128, false>")


def test_rocpd_summary_layout():
    # written by `rocpd summary -i run_results.db --format csv` on an MI355X
    df, name, percent, discarded = ks.read_summary(GOLD / "rocpd_kernels_summary.csv")
    assert (name, percent, discarded) == ("Name", "Percent (Inc)", 0)
    assert df[percent].sum() == pytest.approx(100.0, abs=1e-6)
    assert df[name].iloc[0].startswith("void (anonymous namespace)::dgemm_mfma<128, 128, false>")


def test_table_ends_at_blank_line_or_column_change(tmp_path):
    path = tmp_path / "k.csv"
    path.write_text("Name,Percentage\na,60\nb,40\n\nc,5\n")
    assert ks.read_summary(path)[0]["Name"].tolist() == ["a", "b"]
    path.write_text("Name,Percentage\na,60\nb,40\nfooter line with,three,fields\n")
    assert ks.read_summary(path)[0]["Name"].tolist() == ["a", "b"]
    path.write_text("Name,Percentage\nx,y,z\n")
    with pytest.raises(ValueError):
        ks.read_summary(path)
    path.write_text("Name,Percentage\na,60\nName,Percentage\n")
    with pytest.raises(ValueError):
        ks.read_summary(path)


def test_missing_columns_empty_and_invalid_data(tmp_path):
    path = tmp_path / "bad.csv"
    for text in ("Name,Duration\nx,2\n", "Time (%),Name\n", "Time (%),Name\nno,x\n",
                 "Percentage,Name\n50,\n"):
        path.write_text(text)
        with pytest.raises(ValueError):
            ks.read_summary(path)


def test_plot_top_names_preserved_without_renormalizing(tmp_path):
    path = tmp_path / "kernels.csv"
    path.write_text("Time(%),Kernel Name\n70,same\n20,same\n10,other\n")
    outputs, _ = ks.plot_summary(path, top=2)
    assert all(x.stat().st_size > 1000 for x in outputs)
    data = json.loads((tmp_path / "kernels_plot.json").read_text())
    assert data["displayed_percent_sum"] == 90
    assert data["plotted_rows"] == 2 and data["valid_rows"] == 3


def test_mfma_counters_joined_per_kernel(tmp_path):
    mf = ks.mfma_table(GOLD / "pmc_counter_collection.csv")
    F = 2 * 8192 * 8192 * 16384
    dg = next(v for k, v in mf.items() if "dgemm_mfma" in k)
    mod = next(v for k, v in mf.items() if "int8_product_kernel<256, 2, 4, true" in k)
    sl = next(v for k, v in mf.items() if "int8_product_kernel<256, 2, 4, false" in k)
    assert dg["mfma_type"] == "fp64" and dg["mfma_ops_per_dispatch"] == F
    assert mod["mfma_type"] == "int8" and mod["mfma_ops_per_dispatch"] == 15 * F
    assert sl["mfma_ops_per_dispatch"] == 28 * F
    for v in (dg, mod, sl):
        assert 0.2 < v["mfma_frac"] < 1.0
        assert v["mfma_rate"] == pytest.approx(v["mfma_ops_per_dispatch"] / (v["dispatch_ms"] * 1e-3))
    assert not any("row_absmax" in k for k in mf)      # no matrix-core work
    src = tmp_path / "bench_kernel_stats.csv"
    src.write_bytes((GOLD / "bench_kernel_stats.csv").read_bytes())
    ks.plot_summary(src, top=5, counters=GOLD / "pmc_counter_collection.csv")
    meta = json.loads((tmp_path / "bench_kernel_stats_plot.json").read_text())
    assert len(meta["mfma"]) == 3
    norm = (tmp_path / "bench_kernel_stats_normalized.csv").read_text()
    assert "mfma_frac" in norm.splitlines()[0]


def test_export_from_a_rocprofv3_directory(tmp_path):
    run = tmp_path / "prof" / "host" / "1234"
    run.mkdir(parents=True)
    (run / "bench_kernel_stats.csv").write_bytes((GOLD / "bench_kernel_stats.csv").read_bytes())
    dest, kind = ks.export_stats(tmp_path / "prof", tmp_path / "out")
    assert kind == "kernel_stats" and dest.exists() and dest.name == "prof_kernel_stats.csv"
    (run / "other_kernel_stats.csv").write_text("Name,Percentage\na,100\n")
    with pytest.raises(ValueError):
        ks.export_stats(tmp_path / "prof", tmp_path / "out")
    (tmp_path / "empty").mkdir()
    with pytest.raises(ValueError):
        ks.export_stats(tmp_path / "empty", tmp_path / "o2")
    (tmp_path / "notes.txt").write_text("x")
    with pytest.raises(ValueError):
        ks.export_stats(tmp_path / "notes.txt", tmp_path / "o3")


def _fake_rocpd(path: Path, body: str) -> Path:
    path.write_text("#!" + sys.executable + "\n" + body)
    path.chmod(0o755)
    return path


def test_export_from_a_database_with_simulated_rocpd(tmp_path):
    db = tmp_path / "run.db"
    db.touch()
    fake = _fake_rocpd(tmp_path / "rocpd", (
        "import sys\nfrom pathlib import Path\na = sys.argv\n"
        "assert a[1] == 'summary' and '--format' in a and a[a.index('--format') + 1] == 'csv'\n"
        "out = Path(a[a.index('--output-path') + 1])\n"
        "stem = a[a.index('--output-file') + 1]\n"
        "(out / (stem + '_kernels_summary.csv')).write_text("
        "'Name,Calls,Duration (Nsec),Percent (Inc)\\nk1,2,30,75\\nk2,1,10,25\\n')\n"
        "(out / (stem + '_hip_api_summary.csv')).write_text('Name,Percent (Inc)\\nhipLaunchKernel,100\\n')\n"))
    dest, kind = ks.export_stats(db, tmp_path / "out", rocpd=str(fake))
    assert kind == "rocpd_summary" and dest.exists()
    assert ks.read_summary(dest)[0]["Name"].tolist() == ["k1", "k2"]
    assert "COMMAND:" in (tmp_path / "out" / "run_rocpd_summary.log").read_text()
    cli = subprocess.run([sys.executable, "-m", "mgxbench_amd.kernel_stats", str(db), "--rocpd",
                          str(fake), "--output-dir", str(tmp_path / "cli")],
                         capture_output=True, text=True, cwd=ROOT)
    assert cli.returncode == 0, cli.stderr
    assert (tmp_path / "cli" / "run_kernel_summary_plot.png").exists()


def test_rocpd_failure_does_not_try_a_fallback(tmp_path):
    db = tmp_path / "run.db"
    db.touch()
    fake = _fake_rocpd(tmp_path / "rocpd",
                       "import sys\nprint('Permission denied', file=sys.stderr)\nsys.exit(1)\n")
    with pytest.raises(ValueError):
        ks.export_stats(db, tmp_path / "out", rocpd=str(fake))
    log = (tmp_path / "out" / "run_rocpd_summary.log").read_text()
    assert log.count("COMMAND:") == 1 and "Permission denied" in log


def test_cli_csv_and_input_failures(tmp_path):
    path = tmp_path / "kernels.csv"
    path.write_text("Percentage,Name\n100,gemm\n")
    cli = [sys.executable, "-m", "mgxbench_amd.kernel_stats"]
    res = subprocess.run(cli + [str(path), "--output-dir", str(tmp_path / "plots")],
                         capture_output=True, text=True, cwd=ROOT)
    assert res.returncode == 0, res.stderr
    assert (tmp_path / "plots" / "kernels_plot.png").exists()
    for extra in ([str(path), "--top", "0"], [str(tmp_path / "missing.csv")],
                  [str(tmp_path / "kernels.txt")]):
        if extra[0].endswith(".txt"):
            Path(extra[0]).write_text("x")
        res = subprocess.run(cli + extra, capture_output=True, text=True, cwd=ROOT)
        assert res.returncode == 2, (extra, res.stderr)
