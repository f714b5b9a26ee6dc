"""MGXBench runner for the MI355X build: the benchmark.py counterpart.

Same command line (benchmark.py:53-79), same per-run launch contract
(`launcher launcher_args -np 1 exe <auto|rank> N LIB [: ...]`,
benchmark.py:127-130), same acceptance rules for the single BENCHMARK_RESULT
record (benchmark.py:45-50, 147-156) and the same outputs: runs/<run_id>/
benchmark.json (schema_version 2) and benchmark.csv with the reference's
columns (benchmark.py:166-194).  Defaults differ where the platform does:
the launcher is `python -m mgxbench_amd.launch` (one rank, no MPI) and the
executables are read from mgxbench_amd/bin.  New options --form, --slices and
--init travel to the executables as MGX_FORM, MGX_SLICES and MGX_INIT.

    python -m mgxbench_amd.runner --decomposition 1d --size 4096 --ranks 1 --libraries NNN
"""
from __future__ import annotations

import argparse
import csv
import datetime
import json
import math
import os
import re
import shlex
import statistics
import subprocess
import sys
from pathlib import Path

from .experiment import collect_environment, new_run, sample_statistics

PKG = Path(__file__).resolve().parent

LIBRARY_NAMES = {
    "MMM": "MPI-SYNC",
    "YYY": "MPI-CHUNKED-ASYNC",
    "CCC": "CUDA-AWARE-SYNC",
    "XXX": "CUDA-AWARE-ASYNC",
    "NNN": "NCCL-ASYNC",
    "WWW": "NVSHMEM-CHUNKED-SYNC",
    "SSS": "NVSHMEM-ASYNC",
}

CSV_COLUMNS = ["run_id", "label", "environment_id", "cv_pct", "ci95_low_s", "ci95_high_s",
               "decomposition", "library", "size", "ranks", "mean_s", "median_s", "p95_s",
               "stddev_s", "gflops", "speedup_vs_MMM", "validated", "original_inner_loops",
               "seconds_semantics"]

_RECORD = re.compile(r"BENCHMARK_RESULT\s+([^\r\n]+)")


def parse_result(output: str) -> dict:
    """key=value fields of the one BENCHMARK_RESULT record in `output`."""
    found = _RECORD.findall(output)
    if len(found) != 1:
        raise ValueError(f"Expected exactly one BENCHMARK_RESULT record; got {len(found)}")
    return dict(field.split("=", 1) for field in found[0].split())


def percentile(sorted_values: list[float], q: float) -> float:
    """Linear interpolation between closest ranks (benchmark.py:169-171)."""
    pos = (len(sorted_values) - 1) * q
    lo = int(pos)
    hi = min(lo + 1, len(sorted_values) - 1)
    return sorted_values[lo] + (sorted_values[hi] - sorted_values[lo]) * (pos - lo)


def build_parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(description="MGXBench - Multi-GPU Execution Benchmark (MI355X build)")
    ap.add_argument("--decomposition", choices=["1d", "2d"], required=True)
    ap.add_argument("--size", type=int, required=True)
    ap.add_argument("--ranks", type=int, required=True)
    ap.add_argument("--libraries", nargs="+", default=list(LIBRARY_NAMES))
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--warmup-mode", choices=["external", "internal"], default="external")
    ap.add_argument("--retries", type=int, default=0)
    ap.add_argument("--repeats", type=int, default=5)
    ap.add_argument("--launcher", default=f"{shlex.quote(sys.executable)} -m mgxbench_amd.launch")
    ap.add_argument("--launcher-args", default="")
    ap.add_argument("--build-dir", default=str(PKG / "bin"))
    ap.add_argument("--output", default="results")
    ap.add_argument("--timeout", type=int, default=600)
    ap.add_argument("--plot", action="store_true")
    ap.add_argument("--inner-loops", type=int, default=1)
    ap.add_argument("--environment")
    ap.add_argument("--label", default="default")
    ap.add_argument("--device-map", choices=["local", "explicit"], default="local")
    ap.add_argument("--grid-rows", type=int, default=None)
    ap.add_argument("--grid-cols", type=int, default=None)
    ap.add_argument("--summa-bk", type=int, default=256)
    ap.add_argument("--prediction-model", choices=["extrapolation", "ml", "auto", "mlp"],
                    default="extrapolation")
    ap.add_argument("--auto-model", default=None)
    ap.add_argument("--auto-validate", action="store_true")
    ap.add_argument("--auto-feedback", action="store_true")
    ap.add_argument("--form", choices=["fp64", "int8", "bf16", "mod"], default=None,
                    help="product form of the executables (MGX_FORM)")
    ap.add_argument("--slices", type=int, default=None, help="INT8 slices (MGX_SLICES)")
    ap.add_argument("--init", default=None, help="const | uniform:SEED (MGX_INIT)")
    return ap


def validate_args(ap: argparse.ArgumentParser, args) -> tuple[int, int] | None:
    """Reference argument rules (benchmark.py:80-107); returns the 2D grid."""
    if args.auto_feedback and not args.auto_validate:
        ap.error("--auto-feedback requires --auto-validate")
    if "AUTO" in args.libraries and args.libraries != ["AUTO"]:
        ap.error("AUTO must be the only value of --libraries")
    if args.retries < 0:
        ap.error("retries must be nonnegative")
    if args.warmup_mode == "internal" and args.decomposition != "1d":
        ap.error("internal warmup currently supports 1D only")
    if args.inner_loops < 1 or (args.decomposition == "2d" and args.inner_loops != 1):
        ap.error("inner-loops must be positive; 2D supports 1")
    if args.libraries != ["AUTO"] and (any(l not in LIBRARY_NAMES for l in args.libraries)
                                       or len(set(args.libraries)) != len(args.libraries)):
        ap.error("select distinct supported library codes")
    if args.size <= 0 or args.ranks <= 0 or args.repeats <= 0 or args.warmup < 0:
        ap.error("invalid size/ranks/repeats/warmup")
    if args.decomposition == "1d" and args.size % args.ranks:
        ap.error("1D requires size divisible by ranks")
    if args.summa_bk <= 0:
        ap.error("summa-bk must be positive")
    grid = None
    if args.decomposition == "2d":
        from .grid import resolve_grid
        try:
            grid = resolve_grid(args.ranks, args.grid_rows, args.grid_cols)
        except ValueError as exc:
            ap.error(str(exc))
        if args.size % grid[0] or args.size % grid[1]:
            ap.error("2D requires N divisible by grid rows and columns")
    elif args.grid_rows is not None or args.grid_cols is not None:
        ap.error("--grid-rows/--grid-cols are only supported with 2D")
    return grid


def child_environment(args, grid) -> dict:
    env = os.environ.copy()
    if args.decomposition == "2d":
        env["MGXBENCH_GRID_ROWS"] = str(grid[0])
        env["MGXBENCH_GRID_COLS"] = str(grid[1])
        env["MGXBENCH_SUMMA_BK"] = str(args.summa_bk)
    env["MGXBENCH_INNER_LOOPS"] = str(args.inner_loops)
    env["MGXBENCH_INTERNAL_WARMUP"] = str(args.warmup if args.warmup_mode == "internal" else 0)
    if args.form:
        env["MGX_FORM"] = args.form
    if args.slices is not None:
        env["MGX_SLICES"] = str(args.slices)
    if args.init:
        env["MGX_INIT"] = args.init
    return env


def launch_command(args, exe: Path, lib: str) -> list[str]:
    cmd = shlex.split(args.launcher) + shlex.split(args.launcher_args)
    for rank in range(args.ranks):
        if rank:
            cmd.append(":")
        device = "auto" if args.device_map == "local" else str(rank)
        cmd += ["-np", "1", str(exe), device, str(args.size), lib]
    return cmd


def check_record(record: dict, returncode: int, args, lib: str) -> float:
    """Acceptance rules of one measured launch; returns its seconds."""
    if returncode != 0 or record.get("valid") != "1":
        raise RuntimeError(f"validation/exit failed: exit={returncode}, valid={record.get('valid')}")
    if int(record["size"]) != args.size or int(record["ranks"]) != args.ranks or record["library"] != lib:
        raise RuntimeError("metadata mismatch")
    if (record.get("decomposition", "").lower() != args.decomposition
            or int(record.get("loops", 0)) != args.inner_loops):
        raise RuntimeError("decomposition/inner-loop metadata mismatch")
    seconds = float(record["seconds"])
    if not math.isfinite(seconds) or seconds <= 0:
        raise RuntimeError("invalid time")
    return seconds


def run_library(args, exe: Path, lib: str, env: dict, dest: Path, failures: list) -> dict | None:
    samples, runs = [], []
    first = -args.warmup if args.warmup_mode == "external" else 0
    for rep in range(first, args.repeats):
        cmd = launch_command(args, exe, lib)
        stage = f"warmup {rep + args.warmup + 1}/{args.warmup}" if rep < 0 else f"run {rep + 1}/{args.repeats}"
        print(f"[{lib:3}] {LIBRARY_NAMES.get(lib, lib):<22} {stage:<15}", flush=True)
        output = ""
        try:
            proc = None
            for attempt in range(args.retries + 1):
                try:
                    proc = subprocess.run(cmd, capture_output=True, text=True, timeout=args.timeout,
                                          env=env)
                except subprocess.TimeoutExpired:
                    if attempt == args.retries:
                        raise
                    print(f"  timeout; retry {attempt + 1}/{args.retries}", file=sys.stderr)
                    continue
                output = proc.stdout + "\n" + proc.stderr
                if proc.returncode != 0 and "BENCHMARK_RESULT" not in output and attempt < args.retries:
                    print(f"  launcher failure; retry {attempt + 1}/{args.retries}", file=sys.stderr)
                    continue
                break
            record = parse_result(output)
            (dest / f"{lib}_{rep}.log").write_text(output)
            seconds = check_record(record, proc.returncode, args, lib)
            if rep >= 0:
                samples.append(seconds)
                runs.append({"iteration": rep + 1, "seconds": seconds, "validated": True})
        except Exception as exc:  # a failed launch ends this library's series
            log = dest / f"{args.decomposition}_{lib}_failed_{rep}.log"
            log.write_text((" ".join(cmd) + "\n" + output + "\n" + str(exc))[-100000:])
            failures.append({"library": lib, "iteration": rep, "error": str(exc), "log": str(log)})
            print("ERROR:", exc, file=sys.stderr)
            break
    if len(samples) != args.repeats:
        return None
    ordered = sorted(samples)
    mean = statistics.mean(samples)
    return {
        "decomposition": args.decomposition, "library": lib, "size": args.size,
        "ranks": args.ranks, "mean_s": mean, "median_s": statistics.median(samples),
        "p95_s": percentile(ordered, 0.95), "stddev_s": statistics.pstdev(samples),
        "gflops": 2.0 * args.size ** 3 / (mean * 1e9), "speedup_vs_MMM": None,
        "validated": True, "samples_s": samples, "runs": runs,
        "original_inner_loops": args.inner_loops,
        "seconds_semantics": "mean_seconds_per_multiplication",
        **sample_statistics(samples),
    }


def print_summary(results, failures, args, dest: Path) -> None:
    width = 91
    print("\n" + "=" * width)
    print("MGXBench | Multi-GPU Execution Benchmark (MI355X)".center(width))
    print("=" * width)
    print(f"Workload: Distributed Matrix Multiplication | Decomposition: {args.decomposition.upper()}")
    print(f"Matrix: {args.size} x {args.size} | ranks / GPUs: {args.ranks} | "
          f"Warmups: {args.warmup} | Repeats: {args.repeats}")
    print("-" * width)
    print(f'{"Rank":>4}  {"Code":<4}  {"Communication backend":<22} {"Mean (s)":>10} '
          f'{"GFLOP/s":>12} {"Speedup":>9} {"Status":>8}')
    print("-" * width)
    ranked = sorted(results, key=lambda r: r["mean_s"])
    for i, r in enumerate(ranked, 1):
        sp = f"{r['speedup_vs_MMM']:.2f}x" if r["speedup_vs_MMM"] is not None else "N/A"
        print(f"{i:>4}  {r['library']:<4}  {LIBRARY_NAMES.get(r['library'], 'Unknown'):<22} "
              f"{r['mean_s']:>10.6f} {r['gflops']:>12.2f} {sp:>9} {'PASS':>8}")
    if not results:
        print("No successful measurements.")
    print("-" * width)
    if ranked:
        b = ranked[0]
        print(f"Fastest backend: {b['library']} ({LIBRARY_NAMES.get(b['library'], 'Unknown')}) | "
              f"{b['mean_s']:.6f} s | {b['gflops']:.2f} GFLOP/s")
    print(f"Validated: {len(results)}/{len(args.libraries)} | Failed: {len(failures)}")
    print("Metric: executable-reported region time (steady clock around the timed loop).")
    print("GFLOP/s: theoretical 2*N^3 / time; not a hardware counter.")
    print(f"Results: {dest / 'benchmark.csv'} | {dest / 'benchmark.json'}")
    print("=" * width)


def write_outputs(dest: Path, run_id: str, args, environment: dict, results, failures) -> None:
    base = next((r["mean_s"] for r in results if r["library"] == "MMM"), None)
    for r in results:
        if base:
            r["speedup_vs_MMM"] = base / r["mean_s"]
        r.update(run_id=run_id, label=args.label, environment_id=environment["id"])
    doc = {
        "run_id": run_id, "label": args.label, "environment": environment, "schema_version": 2,
        "application": "distributed-multi-gpu-mm-domain-decomposition",
        "measurement": "executable-reported computation/communication region; not process startup",
        "note": ("Independent measured process launches; warmups excluded. 1D inner loops "
                 "configurable; 2D one iteration. seconds is mean time per multiplication."),
        "bandwidth": "not reported: actual communication bytes vary with backend and decomposition",
        "timestamp_utc": datetime.datetime.now(datetime.timezone.utc).isoformat(),
        "benchmark_name": "MGXBench", "benchmark_full_name": "Multi-GPU Execution Benchmark",
        "command_args": vars(args), "results": results, "failures": failures,
    }
    (dest / "benchmark.json").write_text(json.dumps(doc, indent=2))
    with (dest / "benchmark.csv").open("w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=CSV_COLUMNS)
        w.writeheader()
        w.writerows({k: r[k] for k in CSV_COLUMNS} for r in results)


def main(argv: list[str] | None = None) -> int:
    ap = build_parser()
    args = ap.parse_args(argv)
    grid = validate_args(ap, args)
    if args.libraries == ["AUTO"]:
        from .predictive import auto
        return auto(args)
    for lib in args.libraries:
        if len(lib) != 3 or any(c not in "MYCXNSW" for c in lib):
            ap.error("invalid three-character library code: " + lib)
    exe = Path(args.build_dir).resolve() / ("mm_1d" if args.decomposition == "1d" else "mm_2d")
    if not exe.is_file():
        ap.error(f"missing executable: {exe}; build it with `python -m mgxbench_amd.build`")
    environment = collect_environment(args.environment, args.build_dir)
    dest, run_id = new_run(args.output)
    (dest / "environment.json").write_text(json.dumps(environment, indent=2, ensure_ascii=False) + "\n")
    if environment.get("gpu_topology"):
        (dest / "gpu-topology.txt").write_text(environment["gpu_topology"] + "\n")
    env = child_environment(args, grid)
    print("\nMGXBench - Multi-GPU Execution Benchmark")
    print(f"Running {args.decomposition.upper()} matrix multiplication, N={args.size}, "
          f"GPUs={args.ranks}\n", flush=True)
    results, failures = [], []
    for lib in args.libraries:
        row = run_library(args, exe, lib, env, dest, failures)
        if row is not None:
            results.append(row)
    write_outputs(dest, run_id, args, environment, results, failures)
    print_summary(results, failures, args, dest)
    if args.plot and results:
        from .report import plot_results
        try:
            plot_results(results, dest / "plots")
        except Exception as exc:  # measurements are saved either way
            print(f"Measurements saved; visualization failed: {exc}", file=sys.stderr)
            return 3
    return 2 if failures or not results else 0


if __name__ == "__main__":
    sys.exit(main())
