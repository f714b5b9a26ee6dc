"""Size sweeps and strong / weak scaling campaigns: the counterpart of
configuration/run_campaign.py.

Each configuration (decomposition, N, ranks) is one `python -m
mgxbench_amd.runner` invocation; extra arguments are forwarded to it.  The
campaign writes a size manifest first (theoretical and actual N, the work
per GPU ratio for weak scaling) and ends with the report of
`python -m mgxbench_amd.report` over its output directory.

Size rules (run_campaign.py:20-63): 'sweep' and 'strong' use the given sizes
and skip those not divisible by the ranks (1D) or both grid dimensions
(2D); 'weak' keeps N^3 / p constant: N = base * p^(1/3) rounded to the
nearest multiple of lcm(size alignment, grid divisor), and refuses a
rounding that moves N^3 / p by more than 2.5 %.

    python -m mgxbench_amd.campaign --decompositions 1d 2d --sizes 4096 8192 \
        --ranks 1 --mode sweep -- --libraries NNN SSS --repeats 3
"""
from __future__ import annotations

import argparse
import csv
import math
import subprocess
import sys
from pathlib import Path

from .grid import resolve_grid

FORWARD_FORBIDDEN = ("--size", "--ranks", "--decomposition", "--output", "--label",
                     "--grid-rows", "--grid-cols")


def _lcm(a: int, b: int) -> int:
    return abs(a * b) // math.gcd(a, b)


def grid_divisor(decomposition: str, ranks: int, rows=None, cols=None) -> int:
    if decomposition == "1d":
        return ranks
    if decomposition == "2d":
        r, c = resolve_grid(ranks, rows, cols)
        return _lcm(r, c)
    raise ValueError("Invalid decomposition")


def weak_scaling_size(base: int, ranks: int, decomposition: str = "1d", alignment: int = 4,
                      rows=None, cols=None) -> tuple[float, int]:
    """(theoretical N = base * ranks^(1/3), N rounded to the nearest multiple
    of lcm(alignment, grid divisor), at least one multiple)."""
    if base <= 0 or ranks <= 0 or alignment <= 0:
        raise ValueError("Base size, ranks and alignment must be positive")
    step = _lcm(alignment, grid_divisor(decomposition, ranks, rows, cols))
    ideal = base * ranks ** (1.0 / 3.0)
    return ideal, max(step, step * math.floor(ideal / step + 0.5))


def configurations(decompositions, sizes, ranks, mode, base, alignment=4, rows=None, cols=None,
                   log=lambda s: print(s, file=sys.stderr)) -> list[tuple[str, int, int]]:
    out = []
    for d in decompositions:
        for p in ranks:
            if p <= 0:
                raise ValueError("Ranks must be positive")
            try:
                step = _lcm(alignment, grid_divisor(d, p, rows, cols))
            except ValueError as exc:
                log(f"Skipping {d} p={p}: {exc}")
                continue
            if mode == "weak":
                _, n = weak_scaling_size(base, p, d, alignment, rows, cols)
                if abs(n ** 3 / p / base ** 3 - 1) > 0.025:
                    raise ValueError("Weak-scaling rounding exceeds 2.5%; choose a larger base size")
                out.append((d, n, p))
                continue
            for n in sizes:
                if n <= 0:
                    raise ValueError("Matrix size must be positive")
                if n % step:
                    log(f"Skipping {d} N={n}, p={p}: divisibility")
                    continue
                out.append((d, n, p))
    if not out:
        raise ValueError("No compatible configurations")
    return out


def write_manifest(path: Path, configs, args) -> None:
    with path.open("w", newline="", encoding="utf-8") as f:
        w = csv.DictWriter(f, fieldnames=["mode", "decomposition", "ranks", "N_theoretical",
                                          "N_actual", "work_per_gpu_ratio", "size_alignment",
                                          "grid_rows", "grid_cols"])
        w.writeheader()
        for d, n, p in configs:
            ideal = args.base_size * p ** (1.0 / 3.0) if args.mode == "weak" else n
            grid = resolve_grid(p, args.grid_rows, args.grid_cols) if d == "2d" else ("", "")
            w.writerow({"mode": args.mode, "decomposition": d, "ranks": p,
                        "N_theoretical": f"{ideal:.6f}", "N_actual": n,
                        "work_per_gpu_ratio": (f"{(n / args.base_size) ** 3 / p:.6f}"
                                               if args.mode == "weak" else ""),
                        "size_alignment": args.size_alignment,
                        "grid_rows": grid[0], "grid_cols": grid[1]})


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--decompositions", nargs="+", choices=["1d", "2d"], default=["1d", "2d"])
    ap.add_argument("--sizes", nargs="+", type=int, default=[2048, 4096, 8192, 16384, 32768])
    ap.add_argument("--ranks", nargs="+", type=int, default=[1])
    ap.add_argument("--mode", choices=["sweep", "strong", "weak"], default="sweep")
    ap.add_argument("--base-size", type=int, default=4096)
    ap.add_argument("--size-alignment", type=int, default=4)
    ap.add_argument("--grid-rows", type=int, default=None)
    ap.add_argument("--grid-cols", type=int, default=None)
    ap.add_argument("--output", default="results/campaign")
    ap.add_argument("--label", default="campaign")
    args, extra = ap.parse_known_args(argv)
    extra = [x for x in extra if x != "--"]
    if any(x.split("=")[0] == "--dry-run" for x in extra):
        ap.error("--dry-run is not supported: campaigns run their benchmarks")
    if any(x.split("=")[0] in FORWARD_FORBIDDEN for x in extra):
        ap.error("Conflicting forwarded benchmark argument")
    if args.base_size <= 0 or args.size_alignment <= 0 or any(n <= 0 for n in args.sizes):
        ap.error("Sizes and size alignment must be positive")
    if args.mode == "strong" and len(args.sizes) != 1:
        ap.error("Strong scaling requires exactly one --sizes value")
    try:
        configs = configurations(args.decompositions, args.sizes, args.ranks, args.mode,
                                 args.base_size, args.size_alignment, args.grid_rows,
                                 args.grid_cols)
    except ValueError as exc:
        ap.error(str(exc))
    out = Path(args.output)
    out.mkdir(parents=True, exist_ok=True)
    manifest = out / f"{args.label}-size-manifest.csv"
    write_manifest(manifest, configs, args)
    print(f"Size manifest: {manifest}", flush=True)
    label = ("weak:" if args.mode == "weak" else "") + args.label
    failures = 0
    for d, n, p in configs:
        cmd = [sys.executable, "-m", "mgxbench_amd.runner", "--decomposition", d, "--size", str(n),
               "--ranks", str(p), "--output", str(out), "--label", label, *extra]
        if d == "2d":
            r, c = resolve_grid(p, args.grid_rows, args.grid_cols)
            cmd += ["--grid-rows", str(r), "--grid-cols", str(c)]
        print(subprocess.list2cmdline(cmd), flush=True)
        if subprocess.run(cmd).returncode:
            failures += 1
    rep = subprocess.run([sys.executable, "-m", "mgxbench_amd.report", "--input", str(out),
                          "--output", str(out / "report")])
    if rep.returncode:
        failures += 1
    return 2 if failures else 0


if __name__ == "__main__":
    sys.exit(main())
