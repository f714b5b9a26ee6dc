"""Kernel-time summaries of rocprofv3 captures, joined with MFMA counters.

Role of the reference's configuration/plot_nsys_stats.py (which turns an
Nsight Systems kernel summary into a top-N share chart) for MI355X captures.
Accepted inputs:

  * a rocprofv3 output directory (`--kernel-trace --stats --output-format
    csv`) holding exactly one `*_kernel_stats.csv` (Name, Calls,
    TotalDurationNs, AverageNs, Percentage, ...);
  * a rocpd database (`.db`): `rocpd summary -i DB --format csv` is run and
    its one kernel table (Name, ..., Percent (Inc), ...) is taken;
  * a summary CSV in the rocprofv3, rocpd or Nsight ("Time (%)") layout.

Table rules kept from the reference (plot_nsys_stats.py:29-56): console
lines and a byte-order mark before the header are skipped; the header is the
first row naming a kernel column and a percentage column; the table stops at
a blank row or a row of another width once it has records (before that, a
row of another width is an error); a second header is an error; rows with an
empty name or a share that is not a finite number in [0, 100] are dropped
and counted; shares are plotted as reported, never rescaled.

Added here: `--counters` takes the counter_collection.csv of a `rocprofv3
--pmc SQ_INSTS_VALU_MFMA_MOPS_*` run and adds each kernel's matrix-core
operations per dispatch (one MOPS count is 512 operations), the rate over
that run's dispatch time and the fraction of the MI355X dense peak of the
data type it used; the chart then gets a second panel with those fractions.

    python -m mgxbench_amd.kernel_stats INPUT [--top 15] [--output-dir DIR]
                                              [--counters CSV] [--rocpd rocpd]
"""
from __future__ import annotations

import argparse
import csv
import json
import math
import shutil
import subprocess
import tempfile
import textwrap
from dataclasses import dataclass, field
from pathlib import Path

# MI355X dense matrix-core peak (operations / s) for each MOPS counter
MFMA_PEAKS = {
    "SQ_INSTS_VALU_MFMA_MOPS_F64": ("fp64", 78.6e12),
    "SQ_INSTS_VALU_MFMA_MOPS_F32": ("fp32", 157.3e12),
    "SQ_INSTS_VALU_MFMA_MOPS_F16": ("fp16", 2.5e15),
    "SQ_INSTS_VALU_MFMA_MOPS_BF16": ("bf16", 2.5e15),
    "SQ_INSTS_VALU_MFMA_MOPS_F8": ("fp8", 5.0e15),
    "SQ_INSTS_VALU_MFMA_MOPS_I8": ("int8", 5.0e15),
}
MOPS_UNIT = 512

NOTE = ("Shares are the profiler's own, not rescaled to the kernels shown. They are shares of "
        "summed kernel time, which counts concurrent kernels twice: not wall time, not overlap.")


def _is_name_column(label: str) -> bool:
    return label.lower() in ("name", "kernel name", "kernel_name")


def _is_share_column(label: str) -> bool:
    low = label.lower()
    # Nsight "Time (%)", rocprofv3 "Percentage", rocpd "Percent (Inc)"
    return ("time" in low and "%" in low) or low.startswith("percent")


def _header_roles(cells) -> tuple[int, int] | None:
    """(name index, share index) when `cells` is a summary header."""
    labels = [str(c).strip() for c in cells]
    name = next((i for i, c in enumerate(labels) if _is_name_column(c)), None)
    share = next((i for i, c in enumerate(labels) if _is_share_column(c)), None)
    return None if name is None or share is None else (name, share)


def _share(text: str) -> float | None:
    """A percentage cell as a number in [0, 100], or None."""
    try:
        value = float(text.replace("%", "").strip())
    except ValueError:
        return None
    return value if math.isfinite(value) and 0.0 <= value <= 100.0 else None


@dataclass
class SummaryTable:
    """One kernel-summary table: its header, the kept rows (cells as read,
    share cell replaced by its number) and how many rows were dropped."""
    header: list[str]
    name_col: str
    share_col: str
    rows: list[list] = field(default_factory=list)
    dropped: int = 0

    def frame(self):
        import pandas as pd
        return pd.DataFrame(self.rows, columns=self.header)


def scan_table(path) -> SummaryTable:
    """Reads the one kernel summary in a CSV file by the rules above."""
    text = Path(path).read_text(encoding="utf-8-sig")
    table = None
    body = []
    for cells in csv.reader(text.splitlines()):
        if table is None:
            roles = _header_roles(cells)
            if roles:
                header = [c.strip() for c in cells]
                table = SummaryTable(header, header[roles[0]], header[roles[1]])
            continue
        blank = not any(c.strip() for c in cells)
        if blank or len(cells) != len(table.header):
            if body:
                break
            if blank:
                continue
            raise ValueError(f"{path}: a row of {len(cells)} cells under a header of "
                             f"{len(table.header)}")
        if _header_roles(cells):
            raise ValueError(f"{path}: a second summary header; give one table per file")
        body.append(cells)
    if table is None:
        raise ValueError(f"{path}: no header with a kernel-name column (Name) and a share column "
                         "(Percentage, Percent (Inc) or Time (%)); export the summary as CSV")
    if not body:
        raise ValueError(f"{path}: the summary lists no kernels; was any GPU work captured?")
    ni, si = table.header.index(table.name_col), table.header.index(table.share_col)
    for cells in body:
        value = _share(cells[si])
        if value is None or not cells[ni].strip():
            table.dropped += 1
            continue
        row = list(cells)
        row[si] = value
        table.rows.append(row)
    if not table.rows:
        raise ValueError(f"{path}: no row has a kernel name and a share in [0, 100]")
    return table


def read_summary(path):
    """(DataFrame of the kept rows, name column, share column, dropped rows)."""
    t = scan_table(path)
    return t.frame(), t.name_col, t.share_col, t.dropped


# ---- from a capture to a summary CSV --------------------------------------------------------

def _from_directory(source: Path, output: Path) -> Path:
    found = sorted(source.rglob("*_kernel_stats.csv"))
    if len(found) != 1:
        raise ValueError(f"{source}: {len(found)} *_kernel_stats.csv files; expected exactly one "
                         "(a rocprofv3 --stats --output-format csv directory)")
    scan_table(found[0])
    dest = output / f"{source.name}_kernel_stats.csv"
    shutil.copyfile(found[0], dest)
    return dest


def _from_database(source: Path, output: Path, rocpd: str, timeout: int) -> Path:
    tool = shutil.which(rocpd)
    if tool is None:
        raise ValueError(f"{rocpd!r} is not on PATH; export the summary CSV elsewhere and pass it")
    log = output / f"{source.stem}_rocpd_summary.log"
    transcript = []
    try:
        with tempfile.TemporaryDirectory(prefix="mgx-rocpd-") as scratch:
            cmd = [tool, "summary", "-i", str(source.resolve()), "--format", "csv",
                   "--output-path", scratch, "--output-file", source.stem]
            transcript.append("COMMAND: " + subprocess.list2cmdline(cmd))
            try:
                done = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
            except subprocess.TimeoutExpired:
                raise ValueError(f"rocpd summary ran past {timeout} s (log {log})") from None
            transcript += [f"EXIT: {done.returncode}", done.stdout, done.stderr]
            if done.returncode != 0:
                raise ValueError(f"rocpd summary exited {done.returncode} (log {log})")
            tables = [p for p in Path(scratch).rglob("*.csv") if "kernel" in p.name.lower()]
            if len(tables) != 1:
                raise ValueError(f"rocpd summary wrote {len(tables)} kernel tables (log {log})")
            scan_table(tables[0])
            dest = output / f"{source.stem}_kernel_summary.csv"
            shutil.copyfile(tables[0], dest)
            return dest
    finally:
        log.write_text("\n".join(transcript) + "\n")


def export_stats(source, output, rocpd="rocpd", timeout=600):
    """Installs the kernel summary of a capture under `output` after checking
    it; returns (CSV path, 'kernel_stats' | 'rocpd_summary').  A failing
    profiler tool is an error: no other route is tried."""
    source, output = Path(source), Path(output)
    output.mkdir(parents=True, exist_ok=True)
    if source.is_dir():
        return _from_directory(source, output), "kernel_stats"
    if source.suffix.lower() == ".db":
        return _from_database(source, output, rocpd, timeout), "rocpd_summary"
    raise ValueError(f"{source}: give a summary .csv, a rocprofv3 output directory or a rocpd .db")


# ---- MFMA counters --------------------------------------------------------------------------

def mfma_table(counter_csv) -> dict:
    """Per kernel name (kernels with matrix-core work only): dispatches, the
    dominant MFMA data type, its operations per dispatch, the mean dispatch
    time of the counter run, the achieved rate and its fraction of peak."""
    durations: dict[str, dict] = {}
    ops: dict[str, dict] = {}
    with open(counter_csv, newline="", encoding="utf-8-sig") as f:
        for rec in csv.DictReader(f):
            counter = rec.get("Counter_Name", "")
            if counter not in MFMA_PEAKS:
                continue
            kernel = rec["Kernel_Name"]
            span = int(rec["End_Timestamp"]) - int(rec["Start_Timestamp"])
            durations.setdefault(kernel, {})[rec.get("Dispatch_Id")] = span
            per = ops.setdefault(kernel, {})
            per[counter] = per.get(counter, 0.0) + float(rec["Counter_Value"]) * MOPS_UNIT
    joined = {}
    for kernel, spans in durations.items():
        nonzero = {c: v for c, v in ops[kernel].items() if v > 0}
        mean_s = sum(spans.values()) / len(spans) * 1e-9
        if not nonzero or mean_s <= 0:
            continue
        counter = max(nonzero, key=nonzero.get)
        dtype, peak = MFMA_PEAKS[counter]
        per_dispatch = nonzero[counter] / len(spans)
        joined[kernel] = {"dispatches": len(spans), "mfma_type": dtype,
                          "mfma_ops_per_dispatch": per_dispatch, "dispatch_ms": mean_s * 1e3,
                          "mfma_rate": per_dispatch / mean_s, "mfma_peak": peak,
                          "mfma_frac": per_dispatch / mean_s / peak}
    return joined


# ---- chart ----------------------------------------------------------------------------------

def _draw(shown: list[dict], title: str, with_mfma: bool):
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    labels = ["\n".join(textwrap.wrap(r["name"], width=72, break_long_words=True)) for r in shown]
    lines = sum(max(1, lab.count("\n") + 1) for lab in labels)
    fig, axes = plt.subplots(1, 2 if with_mfma else 1, sharey=True,
                             figsize=(17 if with_mfma else 14, max(5.0, 0.36 * lines + 2.2)),
                             gridspec_kw={"width_ratios": [3, 1]} if with_mfma else None)
    share_ax = axes[0] if with_mfma else axes
    ypos = list(range(len(shown)))
    bars = share_ax.barh(ypos, [r["share"] for r in shown], color="#C8102E")
    share_ax.bar_label(bars, labels=[f"{r['share']:.2f}%" for r in shown], padding=3)
    share_ax.set_yticks(ypos)
    share_ax.set_yticklabels(labels, fontsize=8.5)
    share_ax.set_xlabel("share of summed kernel time (%)")
    share_ax.margins(x=0.12)
    share_ax.set_title(title)
    if with_mfma:
        frac_ax = axes[1]
        fracs = [100 * r["mfma_frac"] if r.get("mfma_frac") is not None else 0.0 for r in shown]
        fb = frac_ax.barh(ypos, fracs, color="#5B6770")
        frac_ax.bar_label(fb, labels=[f"{r['mfma_type']} {f:.1f}%" if r.get("mfma_frac")
                                      is not None else "" for r, f in zip(shown, fracs)],
                          padding=3, fontsize=8)
        frac_ax.set_xlim(0, 130)
        frac_ax.set_xlabel("MFMA rate, % of dense peak")
    for ax in (axes if with_mfma else [axes]):
        ax.grid(axis="x", linestyle=":", alpha=0.4)
        ax.set_axisbelow(True)
    fig.text(0.5, 0.01, "Summed kernel time: concurrent kernels count twice; not wall time.",
             ha="center", fontsize=8.5)
    fig.tight_layout(rect=(0, 0.035, 1, 1))
    return fig, plt


def plot_summary(path, top=15, output_dir=None, report=None, counters=None):
    """Writes <stem>_plot.png/.pdf (the `top` largest shares), the kept rows
    as <stem>_normalized.csv (with MFMA columns when counters are given) and
    <stem>_plot.json; returns (chart paths, dropped rows)."""
    if top < 1:
        raise ValueError("--top must be at least 1")
    path = Path(path)
    table = scan_table(path)
    ni = table.header.index(table.name_col)
    si = table.header.index(table.share_col)
    mfma = mfma_table(counters) if counters else {}
    records = []
    for row in table.rows:
        rec = {"name": str(row[ni]), "share": float(row[si])}
        rec.update({k: v for k, v in mfma.get(rec["name"].strip(), {}).items()
                    if k in ("mfma_type", "mfma_ops_per_dispatch", "mfma_rate", "mfma_frac")})
        records.append(rec)
    # the largest shares, drawn smallest at the bottom; ties keep file order
    order = sorted(range(len(records)), key=lambda i: (-records[i]["share"], i))[:top]
    shown = [records[i] for i in sorted(order, key=lambda i: (records[i]["share"], -i))]

    out_dir = Path(output_dir) if output_dir else path.parent
    out_dir.mkdir(parents=True, exist_ok=True)
    fig, plt = _draw(shown, f"rocprofv3 kernel time - {path.stem}", bool(mfma))
    charts = []
    for ext in ("png", "pdf"):
        dest = out_dir / f"{path.stem}_plot.{ext}"
        fig.savefig(dest, dpi=300, bbox_inches="tight")
        charts.append(dest)
    plt.close(fig)

    frame = table.frame()
    if mfma:
        for col in ("mfma_type", "mfma_ops_per_dispatch", "mfma_rate", "mfma_frac"):
            frame[col] = [r.get(col) for r in records]
    frame.to_csv(out_dir / f"{path.stem}_normalized.csv", index=False)
    meta = {"input": str(path.resolve()), "report": report, "valid_rows": len(records),
            "discarded_rows": table.dropped, "plotted_rows": len(shown),
            "displayed_percent_sum": float(sum(r["share"] for r in shown)), "note": NOTE}
    if mfma:
        meta["counters"] = str(Path(counters).resolve())
        meta["mfma"] = mfma
    (out_dir / f"{path.stem}_plot.json").write_text(json.dumps(meta, indent=2) + "\n")
    return charts, table.dropped


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(
        description="top-N kernel-time chart of a rocprofv3 capture, with MFMA rates")
    ap.add_argument("input", help="summary CSV, rocprofv3 output directory or rocpd .db")
    ap.add_argument("--top", type=int, default=15, help="kernels to draw")
    ap.add_argument("--output-dir", type=Path, help="default: beside the input")
    ap.add_argument("--counters", type=Path, default=None,
                    help="counter_collection.csv of a rocprofv3 --pmc MFMA MOPS run")
    ap.add_argument("--rocpd", default="rocpd", help="rocpd tool, for .db inputs")
    ap.add_argument("--timeout", type=int, default=600, help="seconds allowed to rocpd")
    args = ap.parse_args(argv)
    if args.top < 1 or args.timeout < 1:
        ap.error("--top and --timeout must be at least 1")
    source = Path(args.input)
    if not source.exists():
        ap.error(f"{source} does not exist")
    if args.counters is not None and not args.counters.is_file():
        ap.error(f"{args.counters} does not exist")
    kind = None
    try:
        if source.is_dir() or source.suffix.lower() == ".db":
            source, kind = export_stats(source, args.output_dir or source.parent / "stats",
                                        args.rocpd, args.timeout)
            print(f"summary CSV: {source}")
        elif source.suffix.lower() != ".csv":
            raise ValueError(f"{source}: give a summary .csv, a rocprofv3 output directory or "
                             "a rocpd .db")
        charts, dropped = plot_summary(source, args.top, args.output_dir, kind, args.counters)
    except (ValueError, OSError) as exc:
        ap.exit(2, f"kernel_stats: {exc}\n")
    if dropped:
        print(f"{dropped} row(s) without a kernel name or a valid share were left out")
    for chart in charts:
        print(f"chart: {chart}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
