"""Benchmark runner of the MI355X build (the reference's benchmark.py role).

Contract kept with the reference (benchmark.py:52-79 for the flags,
:127-130 for the launch line, :147-156 for what a valid record is,
:166-194 for the outputs):

* the same command-line flags and argument rules;
* one launch per measurement, `launcher launcher-args -np 1 exe <auto|rank>
  N CODE [: -np 1 ...]` (one segment per rank);
* a launch counts only if it exits 0 and prints exactly one
  BENCHMARK_RESULT record with valid=1, the requested size, ranks, code,
  decomposition and inner-loop count, and a positive finite time;
* results/<run_id>/benchmark.json (schema_version 2) and benchmark.csv with
  the reference's keys and columns.

Platform defaults: the launcher is `python -m mgxbench_amd.launch` (one rank,
no MPI) and the executables are mgxbench_amd/bin/mm_1d and mm_2d.  The build's
own knobs --form, --slices and --init reach the executables as MGX_FORM,
MGX_SLICES and MGX_INIT.

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
from dataclasses import dataclass, field
from pathlib import Path

from .experiment import collect_environment, new_run, sample_statistics

PKG = Path(__file__).resolve().parent

# library code -> communication backend it stands for in the reference
LIBRARY_NAMES = {
    "MMM": "MPI-SYNC",
    "YYY": "MPI-CHUNKED-ASYNC",
    "CCC": "CUDA-AWARE-SYNC",
    "XXX": "CUDA-AWARE-ASYNC",
    "NNN": "NCCL-ASYNC",
    "WWW": "NVSHMEM-CHUNKED-SYNC",
    "SSS": "NVSHMEM-ASYNC",
}

# benchmark.csv columns, in the reference's order
CSV_COLUMNS = ["run_id", "label", "environment_id", "cv_pct", "ci95_low_s", "ci95_high_s",
               "decomposition", "library", "size", "ranks", "mean_s", "median_s", "p95_s",
               "stddev_s", "gflops", "speedup_vs_MMM", "validated", "original_inner_loops",
               "seconds_semantics"]

RESULT_TAG = "BENCHMARK_RESULT"
_TAGGED_LINE = re.compile(RESULT_TAG + r"\s+([^\r\n]+)")


def parse_result(output: str) -> dict:
    """The key=value fields of the single BENCHMARK_RESULT line in `output`;
    ValueError when there is none or more than one."""
    lines = _TAGGED_LINE.findall(output)
    if len(lines) != 1:
        raise ValueError(f"{len(lines)} {RESULT_TAG} lines in the output (one is required)")
    fields = {}
    for token in lines[0].split():
        key, _, value = token.partition("=")
        fields[key] = value
    return fields


def percentile(sorted_values: list[float], q: float) -> float:
    """q-quantile of ascending samples, interpolating linearly between the
    two order statistics around position (n - 1) q."""
    position = (len(sorted_values) - 1) * q
    below = math.floor(position)
    above = min(below + 1, len(sorted_values) - 1)
    weight = position - below
    return (1.0 - weight) * sorted_values[below] + weight * sorted_values[above]


# ---- command line ------------------------------------------------------------------------

def build_parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(
        description="MGXBench runner on MI355X: times the mm_1d / mm_2d executables")
    add = ap.add_argument
    add("--decomposition", choices=["1d", "2d"], required=True)
    add("--size", type=int, required=True, help="global matrix size N")
    add("--ranks", type=int, required=True, help="ranks (GPUs) of the run")
    add("--libraries", nargs="+", default=list(LIBRARY_NAMES),
        help="library codes to time, or AUTO for the autotuner's choice")
    add("--warmup", type=int, default=2, help="untimed launches (or in-process loops)")
    add("--warmup-mode", choices=["external", "internal"], default="external",
        help="internal: warm up inside the executable (1D only)")
    add("--retries", type=int, default=0,
        help="extra attempts after a timeout or a launcher error without a record")
    add("--repeats", type=int, default=5, help="timed launches per library")
    add("--launcher", default=f"{shlex.quote(sys.executable)} -m mgxbench_amd.launch")
    add("--launcher-args", default="")
    add("--build-dir", default=str(PKG / "bin"), help="directory holding mm_1d and mm_2d")
    add("--output", default="results")
    add("--timeout", type=int, default=600, help="seconds per launch")
    add("--plot", action="store_true", help="write the report figures after measuring")
    add("--inner-loops", type=int, default=1, help="products per launch (1D; 2D runs one)")
    add("--environment", help="JSON file of environment metadata shared by a campaign")
    add("--label", default="default")
    add("--device-map", choices=["local", "explicit"], default="local")
    add("--grid-rows", type=int, default=None, help="2D process-grid rows")
    add("--grid-cols", type=int, default=None, help="2D process-grid columns")
    add("--summa-bk", type=int, default=256, help="2D panel width")
    add("--prediction-model", choices=["extrapolation", "ml", "auto", "mlp"],
        default="extrapolation", help="runtime model used by --libraries AUTO")
    add("--auto-model", default=None, help="calibration file for --libraries AUTO")
    add("--auto-validate", action="store_true", help="time the AUTO choice")
    add("--auto-feedback", action="store_true",
        help="add the timed AUTO choice to its calibration (needs --auto-validate)")
    add("--form", choices=["fp64", "int8", "bf16", "mod"], default=None,
        help="product form of the executables (MGX_FORM)")
    add("--slices", type=int, default=None, help="slices or moduli (MGX_SLICES)")
    add("--init", default=None, help="const | uniform:SEED (MGX_INIT)")
    return ap


def _argument_problem(args):
    """The first rule the arguments break (the reference's rules,
    benchmark.py:80-107, in this build's words), or None."""
    libs = args.libraries
    auto = "AUTO" in libs
    rules = [
        (lambda: args.auto_feedback and not args.auto_validate,
         "--auto-feedback only makes sense with --auto-validate"),
        (lambda: auto and libs != ["AUTO"], "--libraries AUTO cannot be combined with other codes"),
        (lambda: args.retries < 0, "--retries cannot be negative"),
        (lambda: args.warmup_mode == "internal" and args.decomposition != "1d",
         "--warmup-mode internal exists for 1D only"),
        (lambda: args.inner_loops < 1, "--inner-loops must be at least 1"),
        (lambda: args.decomposition == "2d" and args.inner_loops != 1,
         "2D runs exactly one product per launch (--inner-loops 1)"),
        (lambda: not auto and (len(set(libs)) != len(libs) or not set(libs) <= set(LIBRARY_NAMES)),
         "--libraries takes distinct codes out of " + " ".join(LIBRARY_NAMES)),
        (lambda: min(args.size, args.ranks, args.repeats) <= 0 or args.warmup < 0,
         "--size, --ranks and --repeats must be positive and --warmup non-negative"),
        (lambda: args.decomposition == "1d" and args.size % args.ranks != 0,
         "1D splits the rows evenly: --size must be a multiple of --ranks"),
        (lambda: args.summa_bk <= 0, "--summa-bk must be positive"),
        (lambda: args.decomposition == "1d" and (args.grid_rows, args.grid_cols) != (None, None),
         "--grid-rows / --grid-cols apply to 2D only"),
    ]
    # checked in order, so a later rule may assume the earlier ones hold
    return next((msg for broken, msg in rules if broken()), None)


def validate_args(ap: argparse.ArgumentParser, args) -> tuple[int, int] | None:
    """Applies the argument rules (parser error on the first broken one) and
    returns the 2D process grid (None for 1D)."""
    problem = _argument_problem(args)
    if problem:
        ap.error(problem)
    if args.decomposition != "2d":
        return None
    from .grid import resolve_grid
    try:
        rows, cols = resolve_grid(args.ranks, args.grid_rows, args.grid_cols)
    except ValueError as exc:
        ap.error(str(exc))
    if args.size % rows or args.size % cols:
        ap.error(f"2D blocks must be whole: N={args.size} is not a multiple of the "
                 f"{rows}x{cols} grid")
    return rows, cols


def child_environment(args, grid) -> dict:
    """Environment of every launch: the reference's MGXBENCH_* variables plus
    this build's MGX_* knobs."""
    env = dict(os.environ)
    settings = {
        "MGXBENCH_INNER_LOOPS": args.inner_loops,
        "MGXBENCH_INTERNAL_WARMUP": args.warmup if args.warmup_mode == "internal" else 0,
    }
    if args.decomposition == "2d":
        settings.update(MGXBENCH_GRID_ROWS=grid[0], MGXBENCH_GRID_COLS=grid[1],
                        MGXBENCH_SUMMA_BK=args.summa_bk)
    knobs = {"MGX_FORM": args.form, "MGX_SLICES": args.slices, "MGX_INIT": args.init}
    settings.update({k: v for k, v in knobs.items() if v is not None and v != ""})
    env.update({k: str(v) for k, v in settings.items()})
    return env


def launch_command(args, exe: Path, lib: str) -> list[str]:
    """The MPMD launch line: one `-np 1 exe device N CODE` segment per rank."""
    head = shlex.split(args.launcher) + shlex.split(args.launcher_args)
    segments = []
    for rank in range(args.ranks):
        device = "auto" if args.device_map == "local" else str(rank)
        segments.append(["-np", "1", str(exe), device, str(args.size), lib])
    line = list(head)
    for i, seg in enumerate(segments):
        line += ([":"] if i else []) + seg
    return line


# ---- one library's series of launches ---------------------------------------------------

class RejectedLaunch(RuntimeError):
    """A launch whose output does not count as a measurement."""


def check_record(record: dict, returncode: int, args, lib: str) -> float:
    """The launch's time in seconds, or RejectedLaunch with the reason."""
    if returncode != 0 or record.get("valid") != "1":
        raise RejectedLaunch(f"not a validated product: exit status {returncode}, "
                             f"valid={record.get('valid')}")
    asked = (args.size, args.ranks, lib)
    got = (int(record.get("size", -1)), int(record.get("ranks", -1)), record.get("library"))
    if got != asked:
        raise RejectedLaunch(f"record is for size/ranks/library {got}, the run asked for {asked}")
    shape = (record.get("decomposition", "").lower(), int(record.get("loops", 0)))
    if shape != (args.decomposition, args.inner_loops):
        raise RejectedLaunch(f"record is for decomposition/loops {shape}, the run asked for "
                             f"{(args.decomposition, args.inner_loops)}")
    seconds = float(record["seconds"])
    if not (math.isfinite(seconds) and seconds > 0):
        raise RejectedLaunch(f"record time {record['seconds']} is not a positive finite number")
    return seconds


def _run_with_retries(cmd, args, env):
    """Runs one launch; timeouts and launcher errors without any record are
    tried again up to --retries times.  Returns (returncode, output)."""
    for attempt in range(args.retries + 1):
        last = attempt == args.retries
        try:
            proc = subprocess.run(cmd, capture_output=True, text=True, timeout=args.timeout,
                                  env=env)
        except subprocess.TimeoutExpired:
            if last:
                raise
            print(f"  no result within {args.timeout} s; retry {attempt + 1}/{args.retries}",
                  file=sys.stderr)
            continue
        output = proc.stdout + "\n" + proc.stderr
        if proc.returncode != 0 and RESULT_TAG not in output and not last:
            print(f"  launcher exited {proc.returncode} without a record; "
                  f"retry {attempt + 1}/{args.retries}", file=sys.stderr)
            continue
        return proc.returncode, output
    raise AssertionError("unreachable")


@dataclass
class Series:
    """The launches of one library code: warmups first, then timed repeats;
    the first rejected launch ends the series."""
    lib: str
    samples: list = field(default_factory=list)
    runs: list = field(default_factory=list)
    failure: dict | None = None

    def complete(self, repeats: int) -> bool:
        return self.failure is None and len(self.samples) == repeats

    def summary(self, args) -> dict:
        mean = statistics.mean(self.samples)
        row = {
            "decomposition": args.decomposition, "library": self.lib, "size": args.size,
            "ranks": args.ranks, "mean_s": mean, "median_s": statistics.median(self.samples),
            "p95_s": percentile(sorted(self.samples), 0.95),
            "stddev_s": statistics.pstdev(self.samples),
            # one dense N x N product is 2 N^3 flop, whatever the decomposition
            "gflops": 2.0 * args.size ** 3 / (mean * 1e9),
            "speedup_vs_MMM": None, "validated": True,
            "samples_s": list(self.samples), "runs": list(self.runs),
            "original_inner_loops": args.inner_loops,
            "seconds_semantics": "mean_seconds_per_multiplication",
        }
        row.update(sample_statistics(self.samples))
        return row


def measure_library(args, exe: Path, lib: str, env: dict, dest: Path) -> Series:
    series = Series(lib)
    indices = range(-args.warmup if args.warmup_mode == "external" else 0, args.repeats)
    for index in indices:
        timed = index >= 0
        what = f"run {index + 1}/{args.repeats}" if timed else \
            f"warmup {index + args.warmup + 1}/{args.warmup}"
        print(f"[{lib:3}] {LIBRARY_NAMES.get(lib, lib):<22} {what:<15}", flush=True)
        cmd = launch_command(args, exe, lib)
        output = ""
        try:
            returncode, output = _run_with_retries(cmd, args, env)
            record = parse_result(output)
            (dest / f"{lib}_{index}.log").write_text(output)
            seconds = check_record(record, returncode, args, lib)
        except (RejectedLaunch, ValueError, KeyError, subprocess.TimeoutExpired, OSError) as exc:
            log = dest / f"{args.decomposition}_{lib}_failed_{index}.log"
            log.write_text((" ".join(cmd) + "\n" + output + "\n" + str(exc))[-100000:])
            series.failure = {"library": lib, "iteration": index, "error": str(exc),
                              "log": str(log)}
            print(f"ERROR: {lib}: {exc}", file=sys.stderr)
            break
        if timed:
            series.samples.append(seconds)
            series.runs.append({"iteration": index + 1, "seconds": seconds, "validated": True})
    return series


# ---- outputs --------------------------------------------------------------------------------

def write_outputs(dest: Path, run_id: str, args, environment: dict, results, failures) -> None:
    reference = {r["library"]: r["mean_s"] for r in results}.get("MMM")
    for r in results:
        if reference:
            r["speedup_vs_MMM"] = reference / r["mean_s"]
        r["run_id"], r["label"], r["environment_id"] = run_id, args.label, environment["id"]
    doc = {
        "run_id": run_id, "label": args.label, "environment": environment, "schema_version": 2,
        "application": "distributed-multi-gpu-mm-domain-decomposition",
        "measurement": ("time the executable reports for its timed loop (steady clock); "
                        "process start-up excluded"),
        "note": ("each sample is a separate launch and warmup launches are not recorded; "
                 "seconds is the mean time of one product (1D may loop several, 2D runs one)"),
        "bandwidth": "not reported: bytes exchanged depend on the backend and the decomposition",
        "timestamp_utc": datetime.datetime.now(datetime.timezone.utc).isoformat(),
        "benchmark_name": "MGXBench", "benchmark_full_name": "Multi-GPU Execution Benchmark",
        "command_args": vars(args), "results": results, "failures": failures,
    }
    (dest / "benchmark.json").write_text(json.dumps(doc, indent=2))
    with (dest / "benchmark.csv").open("w", newline="") as f:
        writer = csv.DictWriter(f, fieldnames=CSV_COLUMNS, extrasaction="ignore")
        writer.writeheader()
        writer.writerows(results)


def print_summary(results, failures, args, dest: Path) -> None:
    rule = "-" * 88
    print(f"\n{rule}\nMGXBench on MI355X | {args.decomposition.upper()} product, N = {args.size}, "
          f"{args.ranks} rank(s), {args.warmup} warmup(s), {args.repeats} timed launch(es)\n{rule}")
    if not results:
        print("no library produced a complete series")
    ordered = sorted(results, key=lambda r: r["mean_s"])
    for place, r in enumerate(ordered, 1):
        rel = "-" if r["speedup_vs_MMM"] is None else f"{r['speedup_vs_MMM']:.2f}x MMM"
        print(f"{place:>3}. {r['library']} {LIBRARY_NAMES.get(r['library'], '?'):<22} "
              f"mean {r['mean_s']:.6f} s  {r['gflops']:10.2f} GFLOP/s  {rel}")
    print(rule)
    print(f"complete series: {len(results)} of {len(args.libraries)}; failed: {len(failures)}")
    print("seconds: executable-reported timed region; GFLOP/s: 2 N^3 / seconds")
    print(f"written: {dest / 'benchmark.json'} and {dest / 'benchmark.csv'}\n{rule}")


# ---- entry point ----------------------------------------------------------------------------

def main(argv: list[str] | None = None) -> int:
    ap = build_parser()
    args = ap.parse_args(argv)
    grid = validate_args(ap, args)
    if args.libraries == ["AUTO"]:
        from .predictive import run_auto
        return run_auto(args)
    bad = [lib for lib in args.libraries if not re.fullmatch("[MYCXNSW]{3}", lib)]
    if bad:
        ap.error(f"not a three-letter library code: {bad[0]}")
    exe = Path(args.build_dir).resolve() / f"mm_{args.decomposition}"
    if not exe.is_file():
        ap.error(f"{exe} does not exist; build the executables with "
                 "`python -m mgxbench_amd.build`")
    environment = collect_environment(args.environment, args.build_dir)
    dest, run_id = new_run(args.output)
    (dest / "environment.json").write_text(json.dumps(environment, indent=2, ensure_ascii=False)
                                           + "\n")
    if environment.get("gpu_topology"):
        (dest / "gpu-topology.txt").write_text(environment["gpu_topology"] + "\n")
    env = child_environment(args, grid)
    print(f"\nMGXBench on MI355X: {args.decomposition.upper()} product, N={args.size}, "
          f"GPUs={args.ranks}\n", flush=True)
    results, failures = [], []
    for lib in args.libraries:
        series = measure_library(args, exe, lib, env, dest)
        if series.failure:
            failures.append(series.failure)
        if series.complete(args.repeats):
            results.append(series.summary(args))
    write_outputs(dest, run_id, args, environment, results, failures)
    print_summary(results, failures, args, dest)
    if args.plot and results:
        from .report import plot_results
        try:
            plot_results(results, dest / "plots")
        except Exception as exc:  # the measurements are written already
            print(f"measurements written; the figures failed: {exc}", file=sys.stderr)
            return 3
    return 0 if results and not failures else 2


if __name__ == "__main__":
    sys.exit(main())
