"""Minimal --plot output of the runner: mean time and GFLOP/s per library code.

The reference's campaign/report tooling (configuration/plot_benchmark.py) is
not part of this build; this writes one PNG per run so `--plot` keeps working.
"""
from __future__ import annotations

from pathlib import Path


def plot_results(results: list[dict], out_dir: Path) -> Path:
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    out_dir.mkdir(parents=True, exist_ok=True)
    libs = [r["library"] for r in results]
    fig, (ax1, ax2) = plt.subplots(1, 2, figsize=(9, 3.5))
    ax1.bar(libs, [r["mean_s"] for r in results],
            yerr=[r["stddev_s"] for r in results], color="tab:red")
    ax1.set_ylabel("mean time per multiplication (s)")
    ax2.bar(libs, [r["gflops"] for r in results], color="tab:blue")
    ax2.set_ylabel("GFLOP/s (2N^3 / t)")
    r0 = results[0]
    fig.suptitle(f"{r0['decomposition'].upper()} N={r0['size']} ranks={r0['ranks']}")
    fig.tight_layout()
    path = out_dir / f"{r0['decomposition']}_N{r0['size']}_p{r0['ranks']}.png"
    fig.savefig(path)
    plt.close(fig)
    return path
