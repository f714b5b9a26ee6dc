"""Benchmark reports: the runner's --plot figure and the offline report over
benchmark.json / benchmark.csv files (counterpart of
configuration/plot_benchmark.py).

    python -m mgxbench_amd.report --input results/campaign [more dirs or files]
        [--output results/comparison] [--size N] [--ranks P] [--environment-id ID] [--label L]

Rules kept from the reference report (plot_benchmark.py:24-199): only
validated rows are read; measurements are pooled only within one
environment id, label and inner-loop count (raw launch samples are
concatenated, CSV summaries without samples cannot be pooled); GFLOP/s is
2 N^3 / mean time; speedups are relative to MMM of the same configuration;
strong scaling needs a one-GPU point; weak scaling is drawn only for labels
starting with 'weak:' whose work per GPU N^3 / p varies by at most 5 %.
Outputs: per (N, ranks) execution time / throughput / speedup bars and
launch-time distributions, workload growth over N, strong and weak scaling
curves (PNG and PDF), summary.csv, report.md and a self-contained
report.html.
"""
from __future__ import annotations

import argparse
import csv
import hashlib
import html
import json
import math
import statistics
import sys
from pathlib import Path

from .experiment import sample_statistics
from .runner import LIBRARY_NAMES

NOTES = [
    "Only numerically validated measurements are plotted.",
    "Bars show means and population standard deviations; CI95 is a seeded percentile bootstrap "
    "of independent launches (2,000 resamples).",
    "GFLOP/s is effective application throughput, 2 N^3 / elapsed time, not a counter reading.",
    "Comparisons require the same environment, label and internal-loop count; repeated matching "
    "runs pool their raw launch samples.",
    "Strong scaling fixes N; weak scaling for dense products holds N^3 / p about constant.",
]


def _plt():
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    return plt


def plot_results(results: list[dict], out_dir: Path) -> Path:
    """The runner's --plot: mean time and GFLOP/s per library code of one run."""
    plt = _plt()
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


# ---- offline report --------------------------------------------------------------------------

def _anon_env(path: Path) -> str:
    return "unknown-" + hashlib.sha256(str(path.parent.resolve()).encode()).hexdigest()[:8]


def load_results(inputs) -> list[dict]:
    """Validated result rows of benchmark.json files (and of benchmark.csv
    files without a JSON twin) under the given files / directories."""
    rows, seen = [], set()
    for raw in inputs:
        p = Path(raw)
        if p.is_dir():
            files = sorted(p.rglob("benchmark.json"))
            files += [f for f in sorted(p.rglob("benchmark.csv"))
                      if not f.with_suffix(".json").exists()]
        else:
            files = [p]
        for f in files:
            if f.resolve() in seen:
                continue
            seen.add(f.resolve())
            if f.suffix == ".json":
                doc = json.loads(f.read_text())
                if not isinstance(doc, dict) or "results" not in doc:
                    raise ValueError(f"{f}: expected a benchmark JSON object")
                entries = doc["results"]
                env = (doc.get("environment") or {}).get("id") or _anon_env(f)
                label = doc.get("label", "legacy")
            else:
                with f.open(newline="") as s:
                    entries = list(csv.DictReader(s))
                env, label = _anon_env(f), "legacy"
            for e in entries:
                r = dict(e)
                if str(r.get("validated", "")).lower() not in ("true", "1"):
                    continue
                r["mean_s"] = float(r["mean_s"])
                r["stddev_s"] = float(r.get("stddev_s") or 0.0)
                for k in ("size", "ranks", "original_inner_loops"):
                    r[k] = int(r.get(k) or 1)
                if (r["size"] <= 0 or r["ranks"] <= 0 or r["original_inner_loops"] <= 0
                        or not math.isfinite(r["mean_s"]) or r["mean_s"] <= 0
                        or not math.isfinite(r["stddev_s"]) or r["stddev_s"] < 0):
                    raise ValueError(f"{f}: invalid numeric measurement")
                samples = r.get("samples_s") or []
                if isinstance(samples, str):
                    samples = json.loads(samples) if samples.strip() else []
                if any(not math.isfinite(float(x)) or float(x) <= 0 for x in samples):
                    raise ValueError(f"{f}: invalid samples")
                r["samples_s"] = [float(x) for x in samples]
                r["decomposition"] = str(r["decomposition"]).lower()
                r["environment_id"] = r.get("environment_id") or env
                r["label"] = r.get("label") or label
                r["run_id"] = r.get("run_id") or f.parent.name
                r["source"] = str(f)
                r["gflops"] = 2 * r["size"] ** 3 / (r["mean_s"] * 1e9)
                rows.append(r)
    if not rows:
        raise ValueError("No validated measurements found")
    return rows


IDENTITY = ("environment_id", "label", "original_inner_loops")


def consolidate(rows: list[dict]) -> list[dict]:
    """One row per (identity, decomposition, library, size, ranks), pooling
    the raw samples of repeated runs; speedup vs MMM of the same setting."""
    groups = {}
    for r in rows:
        k = tuple(r[x] for x in IDENTITY) + (r["decomposition"], r["library"], r["size"],
                                                r["ranks"])
        groups.setdefault(k, []).append(r)
    out = []
    for grp in groups.values():
        r = dict(grp[0])
        if len(grp) > 1 and not all(g["samples_s"] for g in grp):
            raise ValueError("Repeated CSV summaries cannot be pooled without raw samples; "
                             "select a single run")
        samples = [s for g in grp for s in g["samples_s"]]
        if samples:
            ordered = sorted(samples)
            r.update(mean_s=statistics.mean(samples), stddev_s=statistics.pstdev(samples),
                     median_s=statistics.median(samples),
                     p95_s=ordered[min(len(ordered) - 1, int(math.ceil(0.95 * len(ordered))) - 1)],
                     samples_s=samples, **sample_statistics(samples))
        r["gflops"] = 2 * r["size"] ** 3 / (r["mean_s"] * 1e9)
        r["run_count"] = len(grp)
        out.append(r)
    base = {tuple(r[x] for x in IDENTITY) + (r["decomposition"], r["size"], r["ranks"]): r["mean_s"]
            for r in out if r["library"] == "MMM"}
    for r in out:
        b = base.get(tuple(r[x] for x in IDENTITY) + (r["decomposition"], r["size"], r["ranks"]))
        r["speedup_vs_MMM"] = b / r["mean_s"] if b else float("nan")
    return out


def _order(libs) -> list[str]:
    known = [x for x in LIBRARY_NAMES if x in libs]
    return known + sorted(set(libs) - set(known))


def generate(rows: list[dict], output) -> Path:
    plt = _plt()
    out = Path(output)
    out.mkdir(parents=True, exist_ok=True)
    data = consolidate(rows)
    figures, notes = [], list(NOTES)

    def save(fig, name):
        fig.tight_layout()
        for ext in ("png", "pdf"):
            fig.savefig(out / f"{name}.{ext}", dpi=160, bbox_inches="tight")
        figures.append(name + ".png")
        plt.close(fig)

    families = {}
    for r in data:
        families.setdefault(tuple(r[x] for x in IDENTITY), []).append(r)
    for ident, fam in families.items():
        tag = hashlib.sha256(str(ident).encode()).hexdigest()[:10]
        by_np = {}
        for r in fam:
            by_np.setdefault((r["size"], r["ranks"]), []).append(r)
        for (size, ranks), grp in sorted(by_np.items()):
            libs = _order({r["library"] for r in grp})
            decomps = sorted({r["decomposition"] for r in grp})
            cell = {(r["library"], r["decomposition"]): r for r in grp}
            width = 0.8 / len(decomps)
            for metric in ("execution_time", "throughput", "speedup"):
                fig, ax = plt.subplots(figsize=(11, 5.5))
                drawn = False
                for j, d in enumerate(decomps):
                    xs, ys, errs = [], [], []
                    for i, lib in enumerate(libs):
                        r = cell.get((lib, d))
                        if r is None:
                            continue
                        if metric == "execution_time":
                            y, e = r["mean_s"], r["stddev_s"]
                        elif metric == "throughput":
                            y, e = r["gflops"], 0.0
                        else:
                            if not math.isfinite(r["speedup_vs_MMM"]):
                                continue
                            y, e = r["speedup_vs_MMM"], 0.0
                        xs.append(i + (j - (len(decomps) - 1) / 2) * width)
                        ys.append(y)
                        errs.append(e)
                    if not xs:
                        if metric == "speedup":
                            notes.append(f"{tag}_N{size}_p{ranks}/{d}: MPI-SYNC baseline absent; "
                                         "speedup omitted.")
                        continue
                    ax.bar(xs, ys, width, yerr=errs if metric == "execution_time" else None,
                           capsize=3, label=d.upper())
                    drawn = True
                if not drawn:
                    plt.close(fig)
                    continue
                if metric == "speedup":
                    ax.axhline(1, color="black", ls="--", lw=1)
                ax.set_xticks(range(len(libs)))
                ax.set_xticklabels([LIBRARY_NAMES.get(x, x) for x in libs], rotation=25,
                                   ha="right")
                ax.set_ylabel({"execution_time": "Execution time (s)",
                               "throughput": "Effective GFLOP/s",
                               "speedup": "Speedup relative to MPI-SYNC"}[metric])
                ax.set_title(f"{metric.replace('_', ' ').title()} | N={size}, GPUs={ranks}\n"
                             f"environment {ident[0]} | {ident[1]} | inner loops={ident[2]}")
                ax.legend()
                ax.grid(axis="y", alpha=0.25)
                save(fig, f"{tag}_N{size}_p{ranks}_{metric}")
            for d in decomps:
                avail = [cell[(lib, d)] for lib in libs if (lib, d) in cell and
                         cell[(lib, d)]["samples_s"]]
                if not avail:
                    continue
                fig, ax = plt.subplots(figsize=(11, 5.5))
                ax.boxplot([r["samples_s"] for r in avail], showmeans=True)
                ax.set_xticks(range(1, len(avail) + 1))
                ax.set_xticklabels([LIBRARY_NAMES.get(r["library"], r["library"]) for r in avail],
                                   rotation=25)
                ax.set_ylabel("Launch sample time (s)")
                ax.set_title(f"{d.upper()} time distribution | N={size}, GPUs={ranks}")
                save(fig, f"{tag}_N{size}_p{ranks}_{d}_time_distribution")
        by_dp = {}
        for r in fam:
            by_dp.setdefault((r["decomposition"], r["ranks"]), []).append(r)
        for (d, ranks), grp in sorted(by_dp.items()):
            if len({r["size"] for r in grp}) < 2:
                continue
            fig, ax = plt.subplots(figsize=(9, 5.5))
            for lib in _order({r["library"] for r in grp}):
                v = sorted((r for r in grp if r["library"] == lib), key=lambda r: r["size"])
                ax.plot([r["size"] for r in v], [r["mean_s"] for r in v], "o-",
                        label=LIBRARY_NAMES.get(lib, lib))
            ax.set(xlabel="Matrix dimension N", ylabel="Execution time (s)", xscale="log",
                   yscale="log", title=f"{d.upper()} workload growth | GPUs={ranks}")
            ax.legend(fontsize=8)
            ax.grid(alpha=0.25)
            save(fig, f"{tag}_{d}_p{ranks}_size_scalability")
        by_dn = {}
        for r in fam:
            by_dn.setdefault((r["decomposition"], r["size"]), []).append(r)
        for (d, size), grp in sorted(by_dn.items()):
            if len({r["ranks"] for r in grp}) < 2:
                continue
            fig, axes = plt.subplots(1, 2, figsize=(12, 5))
            drawn = False
            for lib in _order({r["library"] for r in grp}):
                v = sorted((r for r in grp if r["library"] == lib), key=lambda r: r["ranks"])
                one = [r for r in v if r["ranks"] == 1]
                if not one:
                    notes.append(f"{tag}/{d}/N{size}/{lib}: no one-GPU baseline; strong scaling "
                                 "omitted.")
                    continue
                sp = [one[0]["mean_s"] / r["mean_s"] for r in v]
                axes[0].plot([r["ranks"] for r in v], sp, "o-", label=lib)
                axes[1].plot([r["ranks"] for r in v], [100 * s / r["ranks"] for s, r in zip(sp, v)],
                             "o-", label=lib)
                drawn = True
            if not drawn:
                plt.close(fig)
                continue
            ps = sorted({r["ranks"] for r in grp})
            axes[0].plot(ps, ps, "k--", label="Ideal")
            axes[1].axhline(100, color="black", ls="--")
            for ax in axes:
                ax.set_xlabel("GPUs / ranks")
                ax.legend()
                ax.grid(alpha=0.25)
            axes[0].set_ylabel("Strong scaling speedup")
            axes[1].set_ylabel("Efficiency (%)")
            fig.suptitle(f"{d.upper()} strong scaling | N={size}")
            save(fig, f"{tag}_{d}_N{size}_strong_scaling")
        if str(ident[1]).startswith("weak:"):
            by_dl = {}
            for r in fam:
                by_dl.setdefault((r["decomposition"], r["library"]), []).append(r)
            for (d, lib), grp in sorted(by_dl.items()):
                per_p = {}
                for r in grp:
                    per_p.setdefault(r["ranks"], []).append(r)
                if max(len(x) for x in per_p.values()) != 1 or len(grp) < 2:
                    notes.append(f"{tag}/{d}/{lib}: ambiguous weak-scaling curve; one N per rank "
                                 "count required.")
                    continue
                v = sorted(grp, key=lambda r: r["ranks"])
                work = [r["size"] ** 3 / r["ranks"] for r in v]
                if max(work) / min(work) > 1.05:
                    notes.append(f"{tag}/{d}/{lib}: weak-scaling work per GPU varies by >5%; "
                                 "omitted.")
                    continue
                one = [r for r in v if r["ranks"] == 1]
                if not one:
                    continue
                fig, ax = plt.subplots(figsize=(8, 5))
                ax.plot([r["ranks"] for r in v], [100 * one[0]["mean_s"] / r["mean_s"] for r in v],
                        "o-")
                ax.axhline(100, color="black", ls="--")
                ax.set(xlabel="GPUs", ylabel="Weak scaling efficiency (%)",
                       title=f"{d.upper()} {lib} | N^3/p about constant")
                save(fig, f"{tag}_{d}_{lib}_weak_scaling")
    cols = ["environment_id", "label", "original_inner_loops", "decomposition", "library", "size",
            "ranks", "run_count", "mean_s", "median_s", "stddev_s", "p95_s", "gflops",
            "speedup_vs_MMM", "cv_pct", "ci95_low_s", "ci95_high_s"]
    with (out / "summary.csv").open("w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=cols, extrasaction="ignore")
        w.writeheader()
        for r in sorted(data, key=lambda r: (r["decomposition"], r["ranks"], r["size"],
                                              r["library"])):
            w.writerow({c: r.get(c) for c in cols})
    notes = list(dict.fromkeys(notes))
    (out / "report.md").write_text("# MGXBench performance report\n\n" +
                                   "\n".join("- " + n for n in notes) + "\n\n" +
                                   "\n\n".join(f"![{f}]({f})" for f in figures) + "\n")
    table = "".join("<tr>" + "".join(f"<td>{html.escape(str(r.get(c)))}</td>" for c in cols) +
                    "</tr>" for r in data)
    head = "".join(f"<th>{html.escape(c)}</th>" for c in cols)
    figs = "".join(f'<figure><img loading="lazy" src="{html.escape(f)}"><figcaption>'
                   f'{html.escape(f)}</figcaption><a href="{html.escape(f[:-4])}.pdf">PDF</a>'
                   "</figure>" for f in figures)
    (out / "report.html").write_text(
        '<!doctype html><meta charset="utf-8"><title>MGXBench performance report</title>'
        "<style>body{font:15px system-ui;max-width:1250px;margin:30px auto;padding:20px}"
        "img{width:100%}table{border-collapse:collapse;font-size:12px}"
        "td,th{border:1px solid #ddd;padding:5px}</style><h1>MGXBench performance report</h1>"
        "<ul>" + "".join(f"<li>{html.escape(n)}</li>" for n in notes) + "</ul>"
        f"<div style='overflow:auto'><table><tr>{head}</tr>{table}</table></div>" + figs)
    return out / "report.html"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Offline performance report from runner outputs")
    ap.add_argument("--input", nargs="+", required=True)
    ap.add_argument("--output", default="results/comparison")
    ap.add_argument("--size", type=int)
    ap.add_argument("--ranks", type=int)
    ap.add_argument("--environment-id")
    ap.add_argument("--label")
    a = ap.parse_args(argv)
    try:
        rows = load_results(a.input)
        for k in ("size", "ranks", "environment_id", "label"):
            v = getattr(a, k)
            if v is not None:
                rows = [r for r in rows if r[k] == v]
        if not rows:
            raise ValueError("No measurements match the selected filters")
        print(generate(rows, a.output))
    except (ValueError, KeyError, OSError, json.JSONDecodeError) as exc:
        ap.exit(2, f"Cannot generate report: {exc}\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
