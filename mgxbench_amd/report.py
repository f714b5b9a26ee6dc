"""Benchmark reports: the runner's --plot chart and the offline report.

Role of the reference's configuration/plot_benchmark.py.

    python -m mgxbench_amd.report --input results/campaign [more dirs or files]
        [--output results/comparison] [--size N] [--ranks P] [--environment-id ID] [--label L]

Rules kept from the reference report (plot_benchmark.py:24-199):
  * only rows with validated = true are read; a row with a non-positive or
    non-finite time, size, rank count or loop count is an error;
  * runs are comparable only with the same environment id, label and
    inner-loop count; repeated runs of one configuration pool their raw
    launch samples, and CSV summaries (no samples) cannot be pooled;
  * GFLOP/s = 2 N^3 / mean time; speedup is against MMM of the same
    configuration;
  * strong scaling is drawn only against a one-GPU point; weak scaling only
    for labels 'weak:...', one N per rank count, with N^3 / p within 5 %.
Outputs (PNG and PDF per figure): per (N, ranks) an overview (mean time and
GFLOP/s per library code) and a speedup chart, launch-time distributions,
time over N per rank count, strong and weak scaling; summary.csv, report.md
and a self-contained report.html.
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

IDENTITY = ("environment_id", "label", "original_inner_loops")
SUMMARY_COLUMNS = ["environment_id", "label", "original_inner_loops", "decomposition", "library",
                   "size", "ranks", "run_count", "mean_s", "median_s", "stddev_s", "p95_s",
                   "gflops", "speedup_vs_MMM", "cv_pct", "ci95_low_s", "ci95_high_s"]
READING_NOTES = [
    "Only numerically validated measurements are used.",
    "Times are the executables' own timed regions; bars give the mean with the population "
    "standard deviation, and CI95 is a seeded 2,000-resample percentile bootstrap over launches.",
    "GFLOP/s means 2 N^3 / mean time: delivered application rate, not a hardware counter.",
    "Only runs with the same environment id, label and inner-loop count are compared; repeated "
    "runs of one configuration pool their launch samples.",
    "Strong scaling holds N fixed; weak scaling holds N^3 / p roughly fixed.",
]


def _pyplot():
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    return plt


def _display(code: str) -> str:
    return f"{code} {LIBRARY_NAMES[code]}" if code in LIBRARY_NAMES else code


def plot_results(results: list[dict], out_dir: Path) -> Path:
    """The runner's --plot: GFLOP/s per library code of one run, each bar
    labelled with its mean time."""
    plt = _pyplot()
    out_dir.mkdir(parents=True, exist_ok=True)
    ordered = sorted(results, key=lambda r: r["gflops"])
    fig, ax = plt.subplots(figsize=(9, 0.55 * len(ordered) + 1.8))
    bars = ax.barh([_display(r["library"]) for r in ordered], [r["gflops"] for r in ordered],
                   color="#C8102E")
    ax.bar_label(bars, labels=[f"{r['mean_s']:.4g} s" for r in ordered], padding=3, fontsize=8)
    first = results[0]
    ax.set_xlabel("GFLOP/s (2 N^3 / mean time)")
    ax.set_title(f"{first['decomposition'].upper()} N={first['size']}, {first['ranks']} rank(s)")
    ax.margins(x=0.15)
    fig.tight_layout()
    path = out_dir / f"{first['decomposition']}_N{first['size']}_p{first['ranks']}.png"
    fig.savefig(path, dpi=150)
    plt.close(fig)
    return path


# ---- reading ------------------------------------------------------------------------------

def _anonymous_environment(path: Path) -> str:
    return "unknown-" + hashlib.sha256(str(path.parent.resolve()).encode()).hexdigest()[:8]


def _source_files(inputs) -> list[Path]:
    """benchmark.json files, and benchmark.csv files without a JSON twin,
    under the inputs (files are taken as given); each file once."""
    files = []
    for item in map(Path, inputs):
        if item.is_dir():
            found = sorted(item.rglob("benchmark.json"))
            found += [c for c in sorted(item.rglob("benchmark.csv"))
                      if not c.with_suffix(".json").exists()]
        else:
            found = [item]
        files.extend(found)
    unique = {}
    for f in files:
        unique.setdefault(f.resolve(), f)
    return list(unique.values())


def _entries(path: Path):
    """(result entries, environment id, label) of one report file."""
    if path.suffix == ".json":
        doc = json.loads(path.read_text())
        if not isinstance(doc, dict) or not isinstance(doc.get("results"), list):
            raise ValueError(f"{path}: not a runner report (no results list)")
        env = (doc.get("environment") or {}).get("id") or _anonymous_environment(path)
        return doc["results"], env, doc.get("label", "legacy")
    with path.open(newline="") as f:
        return list(csv.DictReader(f)), _anonymous_environment(path), "legacy"


def _positive_finite(value) -> bool:
    return math.isfinite(value) and value > 0


def _clean(entry: dict, path: Path, env: str, label: str) -> dict | None:
    """The report's form of one validated entry; None when not validated."""
    if str(entry.get("validated", "")).strip().lower() not in ("true", "1"):
        return None
    row = dict(entry)
    try:
        row["mean_s"] = float(row["mean_s"])
        row["stddev_s"] = float(row.get("stddev_s") or 0.0)
        for k in ("size", "ranks", "original_inner_loops"):
            row[k] = int(row.get(k) or 1)
        samples = row.get("samples_s") or []
        if isinstance(samples, str):
            samples = json.loads(samples) if samples.strip() else []
        samples = [float(s) for s in samples]
    except (KeyError, TypeError, ValueError) as exc:
        raise ValueError(f"{path}: unreadable measurement ({exc})") from None
    if not (min(row["size"], row["ranks"], row["original_inner_loops"]) > 0
            and _positive_finite(row["mean_s"])
            and math.isfinite(row["stddev_s"]) and row["stddev_s"] >= 0):
        raise ValueError(f"{path}: {row.get('library')} N={row['size']} holds a time, size, "
                         "rank or loop count that is not positive and finite")
    if not all(_positive_finite(s) for s in samples):
        raise ValueError(f"{path}: {row.get('library')} N={row['size']} has a launch sample "
                         "that is not positive and finite")
    row.update(samples_s=samples, decomposition=str(row["decomposition"]).lower(),
               environment_id=row.get("environment_id") or env, label=row.get("label") or label,
               run_id=row.get("run_id") or path.parent.name, source=str(path),
               gflops=2 * row["size"] ** 3 / (row["mean_s"] * 1e9))
    return row


def load_results(inputs) -> list[dict]:
    rows = []
    for path in _source_files(inputs):
        entries, env, label = _entries(path)
        rows.extend(r for r in (_clean(e, path, env, label) for e in entries) if r is not None)
    if not rows:
        raise ValueError("the inputs hold no validated measurement")
    return rows


# ---- pooling --------------------------------------------------------------------------------

def _identity(row) -> tuple:
    return tuple(row[k] for k in IDENTITY)


def _setting(row) -> tuple:
    return _identity(row) + (row["decomposition"], row["size"], row["ranks"])


def _pool(runs: list[dict]) -> dict:
    if len(runs) > 1 and not all(r["samples_s"] for r in runs):
        raise ValueError(f"{runs[0]['library']} N={runs[0]['size']} p={runs[0]['ranks']} appears "
                         f"in {len(runs)} runs, some without launch samples (CSV summaries); "
                         "report one of them")
    merged = dict(runs[0])
    samples = [s for r in runs for s in r["samples_s"]]
    if samples:
        ranked = sorted(samples)
        p95_index = min(len(ranked), math.ceil(0.95 * len(ranked))) - 1
        merged.update(samples_s=samples, mean_s=statistics.mean(samples),
                      stddev_s=statistics.pstdev(samples), median_s=statistics.median(samples),
                      p95_s=ranked[p95_index], **sample_statistics(samples))
    merged["gflops"] = 2 * merged["size"] ** 3 / (merged["mean_s"] * 1e9)
    merged["run_count"] = len(runs)
    return merged


def consolidate(rows: list[dict]) -> list[dict]:
    """One row per comparable configuration and library code, repeated runs
    pooled; speedup_vs_MMM is NaN where MMM of the same setting is absent."""
    runs: dict[tuple, list[dict]] = {}
    for r in rows:
        runs.setdefault(_setting(r) + (r["library"],), []).append(r)
    pooled = [_pool(group) for group in runs.values()]
    mmm = {_setting(r): r["mean_s"] for r in pooled if r["library"] == "MMM"}
    for r in pooled:
        base = mmm.get(_setting(r))
        r["speedup_vs_MMM"] = base / r["mean_s"] if base else float("nan")
    return pooled


def _codes(rows) -> list[str]:
    present = {r["library"] for r in rows}
    return [c for c in LIBRARY_NAMES if c in present] + sorted(present - set(LIBRARY_NAMES))


def _groups(rows, *keys) -> dict:
    out: dict[tuple, list[dict]] = {}
    for r in rows:
        out.setdefault(tuple(r[k] for k in keys), []).append(r)
    return dict(sorted(out.items(), key=lambda kv: str(kv[0])))


# ---- figures --------------------------------------------------------------------------------

class FigureBook:
    """Saves figures as <tag>_<name>.png/.pdf and keeps the list and the
    notes on what could not be drawn."""

    def __init__(self, out: Path):
        self.out, self.plt = out, _pyplot()
        self.figures: list[tuple[str, str]] = []   # (section, png name)
        self.notes: list[str] = []

    def save(self, fig, name: str, section: str) -> None:
        fig.tight_layout()
        for ext in ("png", "pdf"):
            fig.savefig(self.out / f"{name}.{ext}", dpi=160, bbox_inches="tight")
        self.figures.append((section, f"{name}.png"))
        self.plt.close(fig)

    def per_point(self, tag, ident, size, ranks, rows) -> None:
        """Overview (time and GFLOP/s) and speedup of every code at one (N, p)."""
        codes = _codes(rows)
        decomps = sorted({r["decomposition"] for r in rows})
        cell = {(r["library"], r["decomposition"]): r for r in rows}
        width = 0.8 / len(decomps)
        offsets = {d: (j - (len(decomps) - 1) / 2) * width for j, d in enumerate(decomps)}
        fig, (t_ax, g_ax) = self.plt.subplots(1, 2, figsize=(14, 5.2))
        for d in decomps:
            have = [(i, cell[(c, d)]) for i, c in enumerate(codes) if (c, d) in cell]
            xs = [i + offsets[d] for i, _ in have]
            t_ax.bar(xs, [r["mean_s"] for _, r in have], width,
                     yerr=[r["stddev_s"] for _, r in have], capsize=3, label=d.upper())
            g_ax.bar(xs, [r["gflops"] for _, r in have], width, label=d.upper())
        for ax, ylabel in ((t_ax, "mean time (s), +/- std"), (g_ax, "GFLOP/s")):
            ax.set_xticks(range(len(codes)))
            ax.set_xticklabels([_display(c) for c in codes], rotation=30, ha="right",
                               fontsize=8)
            ax.set_ylabel(ylabel)
            ax.grid(axis="y", alpha=0.3)
            ax.legend(fontsize=8)
        fig.suptitle(f"N={size}, {ranks} GPU(s) | environment {ident[0]}, label {ident[1]}, "
                     f"inner loops {ident[2]}")
        self.save(fig, f"{tag}_N{size}_p{ranks}_overview", "per size and rank count")

        speed = [(d, [(i, cell[(c, d)]["speedup_vs_MMM"]) for i, c in enumerate(codes)
                      if (c, d) in cell and math.isfinite(cell[(c, d)]["speedup_vs_MMM"])])
                 for d in decomps]
        for d, pts in speed:
            if not pts:
                self.notes.append(f"N={size} p={ranks} {d.upper()} ({tag}): no MMM run, so no "
                                  "speedup chart")
        if not any(pts for _, pts in speed):
            return
        fig, ax = self.plt.subplots(figsize=(11, 5))
        for d, pts in speed:
            if pts:
                ax.bar([i + offsets[d] for i, _ in pts], [s for _, s in pts], width,
                       label=d.upper())
        ax.axhline(1.0, color="black", lw=1, ls=":")
        ax.set_xticks(range(len(codes)))
        ax.set_xticklabels([_display(c) for c in codes], rotation=30, ha="right", fontsize=8)
        ax.set_ylabel("speedup over MMM (MPI-SYNC)")
        ax.set_title(f"N={size}, {ranks} GPU(s)")
        ax.legend(fontsize=8)
        self.save(fig, f"{tag}_N{size}_p{ranks}_speedup", "per size and rank count")

    def distributions(self, tag, size, ranks, rows) -> None:
        for (d,), grp in _groups(rows, "decomposition").items():
            have = [r for c in _codes(grp) for r in grp if r["library"] == c and r["samples_s"]]
            if not have:
                continue
            fig, ax = self.plt.subplots(figsize=(11, 5))
            ax.boxplot([r["samples_s"] for r in have], showmeans=True)
            ax.set_xticks(range(1, len(have) + 1))
            ax.set_xticklabels([_display(r["library"]) for r in have], rotation=30, ha="right",
                               fontsize=8)
            ax.set_ylabel("launch time (s)")
            ax.set_title(f"{d.upper()} launch times, N={size}, {ranks} GPU(s)")
            self.save(fig, f"{tag}_N{size}_p{ranks}_{d}_time_distribution", "launch times")

    def growth(self, tag, family) -> None:
        for (d, ranks), grp in _groups(family, "decomposition", "ranks").items():
            if len({r["size"] for r in grp}) < 2:
                continue
            fig, ax = self.plt.subplots(figsize=(9, 5.5))
            for code in _codes(grp):
                line = sorted((r for r in grp if r["library"] == code), key=lambda r: r["size"])
                ax.plot([r["size"] for r in line], [r["mean_s"] for r in line], "o-",
                        label=_display(code))
            ax.set_xscale("log", base=2)
            ax.set_yscale("log")
            ax.set_xlabel("N")
            ax.set_ylabel("mean time (s)")
            ax.set_title(f"{d.upper()} time over N, {ranks} GPU(s)")
            ax.legend(fontsize=8)
            ax.grid(alpha=0.3, which="both")
            self.save(fig, f"{tag}_{d}_p{ranks}_size_scalability", "growth with N")

    def strong(self, tag, family) -> None:
        for (d, size), grp in _groups(family, "decomposition", "size").items():
            counts = sorted({r["ranks"] for r in grp})
            if len(counts) < 2:
                continue
            curves = []
            for code in _codes(grp):
                line = sorted((r for r in grp if r["library"] == code), key=lambda r: r["ranks"])
                single = next((r for r in line if r["ranks"] == 1), None)
                if single is None:
                    self.notes.append(f"{d.upper()} N={size} {code} ({tag}): no one-GPU run, so "
                                      "no strong-scaling curve")
                    continue
                curves.append((code, [(r["ranks"], single["mean_s"] / r["mean_s"]) for r in line]))
            if not curves:
                continue
            fig, (s_ax, e_ax) = self.plt.subplots(1, 2, figsize=(12, 5))
            for code, pts in curves:
                s_ax.plot([p for p, _ in pts], [s for _, s in pts], "o-", label=code)
                e_ax.plot([p for p, _ in pts], [100 * s / p for p, s in pts], "o-", label=code)
            s_ax.plot(counts, counts, "k:", label="linear")
            e_ax.axhline(100, color="black", ls=":")
            s_ax.set_ylabel("speedup over one GPU")
            e_ax.set_ylabel("parallel efficiency (%)")
            for ax in (s_ax, e_ax):
                ax.set_xlabel("GPUs")
                ax.grid(alpha=0.3)
                ax.legend(fontsize=8)
            fig.suptitle(f"{d.upper()} strong scaling at N={size}")
            self.save(fig, f"{tag}_{d}_N{size}_strong_scaling", "scaling")

    def weak(self, tag, family) -> None:
        for (d, code), grp in _groups(family, "decomposition", "library").items():
            per_count = _groups(grp, "ranks")
            if len(grp) < 2 or any(len(v) != 1 for v in per_count.values()):
                self.notes.append(f"{d.upper()} {code} ({tag}): weak scaling needs one N per "
                                  "rank count; not drawn")
                continue
            line = sorted(grp, key=lambda r: r["ranks"])
            work = [r["size"] ** 3 / r["ranks"] for r in line]
            if max(work) > 1.05 * min(work):
                self.notes.append(f"{d.upper()} {code} ({tag}): N^3/p differs by more than 5% "
                                  "across rank counts; not drawn")
                continue
            single = next((r for r in line if r["ranks"] == 1), None)
            if single is None:
                continue
            fig, ax = self.plt.subplots(figsize=(8, 5))
            ax.plot([r["ranks"] for r in line], [100 * single["mean_s"] / r["mean_s"]
                                                 for r in line], "o-")
            for r in line:
                ax.annotate(f"N={r['size']}", (r["ranks"], 100 * single["mean_s"] / r["mean_s"]),
                            textcoords="offset points", xytext=(4, 6), fontsize=8)
            ax.axhline(100, color="black", ls=":")
            ax.set_xlabel("GPUs")
            ax.set_ylabel("weak-scaling efficiency (%)")
            ax.set_title(f"{d.upper()} {_display(code)}: N^3/p held about constant")
            self.save(fig, f"{tag}_{d}_{code}_weak_scaling", "scaling")


# ---- documents ------------------------------------------------------------------------------

def _write_summary(path: Path, data: list[dict]) -> None:
    with path.open("w", newline="") as f:
        out = csv.DictWriter(f, fieldnames=SUMMARY_COLUMNS, extrasaction="ignore")
        out.writeheader()
        for r in sorted(data, key=lambda r: (r["decomposition"], r["ranks"], r["size"],
                                              r["library"])):
            out.writerow({c: r.get(c) for c in SUMMARY_COLUMNS})


def _write_markdown(path: Path, notes: list[str], figures) -> None:
    parts = ["# MGXBench performance report", "", *[f"- {n}" for n in notes], ""]
    for section in dict.fromkeys(s for s, _ in figures):
        parts += [f"## {section}", ""]
        parts += [f"![{name}]({name})" for s, name in figures if s == section]
        parts.append("")
    path.write_text("\n".join(parts))


def _write_html(path: Path, notes: list[str], figures, data: list[dict]) -> None:
    esc = html.escape
    head = "".join(f"<th>{esc(c)}</th>" for c in SUMMARY_COLUMNS)
    body = "".join("<tr>" + "".join(f"<td>{esc(str(r.get(c)))}</td>" for c in SUMMARY_COLUMNS)
                   + "</tr>" for r in data)
    sections = []
    for section in dict.fromkeys(s for s, _ in figures):
        cards = "".join(
            f'<figure><img loading="lazy" src="{esc(name)}" alt="{esc(name)}">'
            f'<figcaption>{esc(name)} &middot; <a href="{esc(name[:-4])}.pdf">PDF</a>'
            "</figcaption></figure>" for s, name in figures if s == section)
        sections.append(f"<h2>{esc(section)}</h2><div class='grid'>{cards}</div>")
    path.write_text(
        "<!doctype html><html><head><meta charset='utf-8'>"
        "<title>MGXBench performance report</title><style>"
        "body{font:14px sans-serif;margin:24px auto;max-width:1300px;color:#222}"
        ".grid{display:grid;grid-template-columns:repeat(auto-fill,minmax(560px,1fr));gap:16px}"
        "figure{margin:0}img{width:100%;border:1px solid #ccc}"
        "table{border-collapse:collapse;font-size:12px}td,th{border:1px solid #ccc;padding:3px 6px}"
        "th{background:#eee}</style></head><body><h1>MGXBench performance report</h1><ul>"
        + "".join(f"<li>{esc(n)}</li>" for n in notes)
        + f"</ul><h2>Summary</h2><div style='overflow-x:auto'><table><tr>{head}</tr>{body}"
        "</table></div>" + "".join(sections) + "</body></html>")


def generate(rows: list[dict], output) -> Path:
    """Figures, summary.csv, report.md and report.html under `output`;
    returns the HTML path."""
    out = Path(output)
    out.mkdir(parents=True, exist_ok=True)
    data = consolidate(rows)
    book = FigureBook(out)
    for ident, family in _groups(data, *IDENTITY).items():
        tag = hashlib.sha256(str(ident).encode()).hexdigest()[:10]
        for (size, ranks), rows_at in _groups(family, "size", "ranks").items():
            book.per_point(tag, ident, size, ranks, rows_at)
            book.distributions(tag, size, ranks, rows_at)
        book.growth(tag, family)
        book.strong(tag, family)
        if str(ident[1]).startswith("weak:"):
            book.weak(tag, family)
    notes = list(dict.fromkeys(READING_NOTES + book.notes))
    _write_summary(out / "summary.csv", data)
    _write_markdown(out / "report.md", notes, book.figures)
    _write_html(out / "report.html", notes, book.figures, data)
    return out / "report.html"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="offline MGXBench report over runner outputs")
    ap.add_argument("--input", nargs="+", required=True,
                    help="run directories or benchmark.json / benchmark.csv files")
    ap.add_argument("--output", default="results/comparison")
    ap.add_argument("--size", type=int, help="keep only this N")
    ap.add_argument("--ranks", type=int, help="keep only this rank count")
    ap.add_argument("--environment-id", help="keep only this environment")
    ap.add_argument("--label", help="keep only this label")
    args = ap.parse_args(argv)
    try:
        rows = load_results(args.input)
        wanted = {k: getattr(args, k) for k in ("size", "ranks", "environment_id", "label")
                  if getattr(args, k) is not None}
        rows = [r for r in rows if all(r[k] == v for k, v in wanted.items())]
        if not rows:
            raise ValueError(f"no validated measurement matches {wanted}")
        print(generate(rows, args.output))
    except (ValueError, KeyError, OSError) as exc:
        ap.exit(2, f"report: {exc}\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
