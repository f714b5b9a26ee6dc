"""Campaigns: size sweeps, strong and weak scaling over the runner.

Role of the reference's configuration/run_campaign.py.  A campaign is a
list of points (decomposition, N, ranks); each point is one `python -m
mgxbench_amd.runner` run into the campaign directory, with the arguments
after `--` passed through.  Before running, the campaign writes
<label>-size-manifest.csv (requested and actual N per point, and the work per
GPU relative to the base size for weak scaling); after the last point it
writes campaign.json (every command with its exit status) and builds the
report (`python -m mgxbench_amd.report`) over the directory.

Size rules kept from the reference (run_campaign.py:20-63):
  * N must be a multiple of step = lcm(--size-alignment, d), d = ranks in 1D,
    lcm(grid rows, grid cols) in 2D;
  * sweep / strong: the sizes given, a size that is no multiple of the step
    is skipped (strong takes exactly one size);
  * weak: N^3 / p held at base^3, N = base * p^(1/3) rounded to the nearest
    multiple of the step (at least one step); the campaign stops if that
    rounding moves N^3 / p more than 2.5 % away from base^3;
  * weak-scaling runs are labelled 'weak:<label>', which the report uses.

    python -m mgxbench_amd.campaign --decompositions 1d 2d --sizes 4096 8192 \
        --ranks 1 --mode sweep -- --libraries NNN SSS --repeats 3
"""
from __future__ import annotations

import argparse
import csv
import json
import math
import subprocess
import sys
from dataclasses import asdict, dataclass
from pathlib import Path

from .grid import resolve_grid

# runner arguments a campaign sets itself
OWNED_BY_CAMPAIGN = ("--size", "--ranks", "--decomposition", "--output", "--label",
                     "--grid-rows", "--grid-cols")
WEAK_TOLERANCE = 0.025
MANIFEST_FIELDS = ("mode", "decomposition", "ranks", "N_theoretical", "N_actual",
                   "work_per_gpu_ratio", "size_alignment", "grid_rows", "grid_cols")


def grid_divisor(decomposition: str, ranks: int, rows=None, cols=None) -> int:
    """What N must be divisible by for the decomposition's blocks."""
    if decomposition == "1d":
        return ranks
    if decomposition == "2d":
        r, c = resolve_grid(ranks, rows, cols)
        return math.lcm(r, c)
    raise ValueError(f"unknown decomposition {decomposition!r}")


def size_step(decomposition, ranks, alignment, rows=None, cols=None) -> int:
    return math.lcm(alignment, grid_divisor(decomposition, ranks, rows, cols))


def weak_scaling_size(base: int, ranks: int, decomposition: str = "1d", alignment: int = 4,
                      rows=None, cols=None) -> tuple[float, int]:
    """(base * ranks^(1/3), that value rounded half up to a multiple of the
    step, at least one step)."""
    if min(base, ranks, alignment) <= 0:
        raise ValueError("base size, ranks and alignment must all be positive")
    step = size_step(decomposition, ranks, alignment, rows, cols)
    ideal = base * ranks ** (1.0 / 3.0)
    multiples = max(1, math.floor(ideal / step + 0.5))
    return ideal, step * multiples


@dataclass(frozen=True)
class Point:
    decomposition: str
    size: int
    ranks: int
    ideal: float
    grid_rows: int | None = None
    grid_cols: int | None = None

    def as_tuple(self) -> tuple[str, int, int]:
        return self.decomposition, self.size, self.ranks


def plan(decompositions, sizes, ranks, mode, base, alignment=4, rows=None, cols=None,
         log=lambda s: print(s, file=sys.stderr)) -> list[Point]:
    """The campaign's points in run order (decomposition, then ranks, then N)."""
    points = []
    for d in decompositions:
        for p in ranks:
            if p <= 0:
                raise ValueError(f"rank counts must be positive (got {p})")
            try:
                step = size_step(d, p, alignment, rows, cols)
                grid = resolve_grid(p, rows, cols) if d == "2d" else (None, None)
            except ValueError as exc:
                log(f"skip {d} p={p}: no grid ({exc})")
                continue
            if mode == "weak":
                ideal, n = weak_scaling_size(base, p, d, alignment, rows, cols)
                drift = n ** 3 / p / base ** 3 - 1
                if abs(drift) > WEAK_TOLERANCE:
                    raise ValueError(f"{d} p={p}: N={n} puts N^3/p {100 * drift:+.1f}% off "
                                     f"{base}^3 (limit 2.5%); use a larger --base-size")
                points.append(Point(d, n, p, ideal, *grid))
                continue
            for n in sizes:
                if n <= 0:
                    raise ValueError(f"matrix sizes must be positive (got {n})")
                if n % step:
                    log(f"skip {d} N={n} p={p}: N is no multiple of {step}")
                    continue
                points.append(Point(d, n, p, float(n), *grid))
    if not points:
        raise ValueError("no point of the campaign fits its ranks, grids and sizes")
    return points


def write_manifest(path: Path, points: list[Point], mode: str, base: int, alignment: int) -> None:
    with path.open("w", newline="", encoding="utf-8") as f:
        out = csv.DictWriter(f, fieldnames=MANIFEST_FIELDS)
        out.writeheader()
        for pt in points:
            ratio = f"{(pt.size / base) ** 3 / pt.ranks:.6f}" if mode == "weak" else ""
            out.writerow({"mode": mode, "decomposition": pt.decomposition, "ranks": pt.ranks,
                          "N_theoretical": f"{pt.ideal:.6f}", "N_actual": pt.size,
                          "work_per_gpu_ratio": ratio, "size_alignment": alignment,
                          "grid_rows": "" if pt.grid_rows is None else pt.grid_rows,
                          "grid_cols": "" if pt.grid_cols is None else pt.grid_cols})


def runner_command(pt: Point, output: Path, label: str, passthrough: list[str]) -> list[str]:
    cmd = [sys.executable, "-m", "mgxbench_amd.runner", "--decomposition", pt.decomposition,
           "--size", str(pt.size), "--ranks", str(pt.ranks), "--output", str(output),
           "--label", label, *passthrough]
    if pt.grid_rows is not None:
        cmd += ["--grid-rows", str(pt.grid_rows), "--grid-cols", str(pt.grid_cols)]
    return cmd


def build_parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(
        description="MGXBench campaign: runner points over sizes and rank counts, then a report",
        epilog="arguments after -- go to every runner call")
    ap.add_argument("--decompositions", nargs="+", choices=["1d", "2d"], default=["1d", "2d"])
    ap.add_argument("--sizes", nargs="+", type=int, default=[2048, 4096, 8192, 16384, 32768])
    ap.add_argument("--ranks", nargs="+", type=int, default=[1])
    ap.add_argument("--mode", choices=["sweep", "strong", "weak"], default="sweep")
    ap.add_argument("--base-size", type=int, default=4096, help="weak scaling: N at one rank")
    ap.add_argument("--size-alignment", type=int, default=4,
                    help="N is a multiple of this and of the grid's divisor")
    ap.add_argument("--grid-rows", type=int, default=None, help="fixed 2D grid rows")
    ap.add_argument("--grid-cols", type=int, default=None, help="fixed 2D grid columns")
    ap.add_argument("--output", default="results/campaign")
    ap.add_argument("--label", default="campaign")
    return ap


def main(argv=None) -> int:
    ap = build_parser()
    args, passthrough = ap.parse_known_args(argv)
    passthrough = [x for x in passthrough if x != "--"]
    flags = {x.split("=", 1)[0] for x in passthrough if x.startswith("--")}
    if "--dry-run" in flags:
        ap.error("a campaign always runs its points; --dry-run does not exist")
    clash = sorted(flags & set(OWNED_BY_CAMPAIGN))
    if clash:
        ap.error(f"{', '.join(clash)} is set by the campaign itself; remove it after --")
    if args.base_size <= 0 or args.size_alignment <= 0 or min(args.sizes) <= 0:
        ap.error("--sizes, --base-size and --size-alignment must be positive")
    if args.mode == "strong" and len(args.sizes) != 1:
        ap.error("strong scaling holds one N fixed: give exactly one --sizes value")
    try:
        points = plan(args.decompositions, args.sizes, args.ranks, args.mode, args.base_size,
                      args.size_alignment, args.grid_rows, args.grid_cols)
    except ValueError as exc:
        ap.error(str(exc))
    out = Path(args.output)
    out.mkdir(parents=True, exist_ok=True)
    manifest = out / f"{args.label}-size-manifest.csv"
    write_manifest(manifest, points, args.mode, args.base_size, args.size_alignment)
    print(f"campaign: {len(points)} point(s); manifest {manifest}", flush=True)
    label = f"weak:{args.label}" if args.mode == "weak" else args.label
    journal = []
    for i, pt in enumerate(points, 1):
        cmd = runner_command(pt, out, label, passthrough)
        print(f"[{i}/{len(points)}] {subprocess.list2cmdline(cmd)}", flush=True)
        rc = subprocess.run(cmd).returncode
        journal.append({**asdict(pt), "command": cmd, "exit_status": rc})
        if rc:
            print(f"[{i}/{len(points)}] runner exited {rc}", file=sys.stderr, flush=True)
    report_cmd = [sys.executable, "-m", "mgxbench_amd.report", "--input", str(out),
                  "--output", str(out / "report")]
    report_rc = subprocess.run(report_cmd).returncode
    (out / "campaign.json").write_text(json.dumps(
        {"mode": args.mode, "label": label, "points": journal,
         "report": {"command": report_cmd, "exit_status": report_rc}}, indent=2) + "\n")
    failed = sum(1 for j in journal if j["exit_status"]) + (1 if report_rc else 0)
    if failed:
        print(f"campaign: {failed} step(s) failed", file=sys.stderr)
    return 2 if failed else 0


if __name__ == "__main__":
    sys.exit(main())
