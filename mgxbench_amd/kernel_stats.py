"""rocprofv3 kernel summaries with MFMA counters: the counterpart of
configuration/plot_nsys_stats.py for MI355X runs.

The reference exports an Nsight Systems kernel summary (`nsys stats --report
gpukernsum | cuda_gpu_kern_sum --format csv`) and plots each kernel's share of
the accumulated kernel time (plot_nsys_stats.py:59-96, 99-126).  Here the
summary comes from rocprofv3:

  * a `rocprofv3 --kernel-trace --stats --output-format csv` directory holds
    one `<name>_kernel_stats.csv` (columns Name, Calls, TotalDurationNs,
    AverageNs, Percentage, ...);
  * a rocpd database (`.db`, `rocprofv3 --output-format rocpd`) is
    summarised by `rocpd summary -i <db> --format csv`, whose
    `<name>_kernels_summary.csv` (Name, Calls, Duration (Nsec), ...,
    Percent (Inc), ...) is taken;
  * a CSV in either the rocprofv3 or the Nsight layout can be given directly.

The CSV reader keeps the reference's tolerance rules (plot_nsys_stats.py:
29-56): leading console lines and a byte-order mark are skipped until a
header with a kernel-name column and a percentage column; the table ends at
the first blank line or a row with a different column count after at least
one record; a second header in the body is an error; rows whose percentage
is not a finite number in [0, 100] or whose name is empty are discarded and
counted.  Percentages are kept as reported (the top selection is not
renormalised).

New here: with `--counters <pmc_counter_collection.csv>` (a `rocprofv3 --pmc
SQ_INSTS_VALU_MFMA_MOPS_F64 SQ_INSTS_VALU_MFMA_MOPS_I8 ...` run) each kernel
gets its matrix-core work per dispatch (MOPS counters count 512 operations
each), its achieved rate over the dispatch duration of that run and the
fraction of the MI355X dense peak for the data type.

    python -m mgxbench_amd.kernel_stats INPUT [--top 15] [--output-dir DIR]
                                              [--counters CSV] [--rocpd rocpd]
"""
from __future__ import annotations

import argparse
import csv
import io
import json
import math
import shutil
import subprocess
import tempfile
import textwrap
from pathlib import Path

NAME_COLUMNS = ("name", "kernel name", "kernel_name")

# MI355X dense matrix-core peaks (operations per second) per MOPS counter
MFMA_PEAKS = {
    "SQ_INSTS_VALU_MFMA_MOPS_F64": ("fp64", 78.6e12),
    "SQ_INSTS_VALU_MFMA_MOPS_F32": ("fp32", 157.3e12),
    "SQ_INSTS_VALU_MFMA_MOPS_F16": ("fp16", 2.5e15),
    "SQ_INSTS_VALU_MFMA_MOPS_BF16": ("bf16", 2.5e15),
    "SQ_INSTS_VALU_MFMA_MOPS_F8": ("fp8", 5.0e15),
    "SQ_INSTS_VALU_MFMA_MOPS_I8": ("int8", 5.0e15),
}
MOPS_UNIT = 512

NOTE = ("Percentages as reported by the profiler; the top selection is not renormalised. "
        "Kernel sums may include concurrent work: this is neither wall time nor overlap.")


def columns(header):
    """(names, kernel-name column, percentage column) of a candidate header row."""
    names = [str(c).strip() for c in header]
    name = next((c for c in names if c.lower() in NAME_COLUMNS), None)
    # Nsight "Time (%)", rocprofv3 "Percentage", rocpd summary "Percent (Inc)"
    percent = next((c for c in names if ("time" in c.lower() and "%" in c)
                    or c.lower().startswith("percent")), None)
    return names, name, percent


def read_summary(path):
    """One kernel-summary table from a CSV file -> (DataFrame, name column,
    percentage column, discarded row count)."""
    import pandas as pd
    lines = Path(path).read_text(encoding="utf-8-sig").splitlines()
    for index, line in enumerate(lines):
        names, name, percent = columns(next(csv.reader([line]), []))
        if name and percent:
            break
    else:
        raise ValueError('no kernel-name ("Name") and percentage ("Percentage" or "Time (%)") '
                         "columns found; export the kernel summary as CSV")
    records = []
    for record in csv.reader(io.StringIO("\n".join(lines[index + 1:]))):
        if not record or not any(x.strip() for x in record):
            if records:
                break
            continue
        if len(record) != len(names):
            if records:
                break
            raise ValueError("CSV row with an unexpected number of columns")
        if columns(record)[1] and columns(record)[2]:
            raise ValueError("one kernel summary per file")
        records.append(record)
    df = pd.DataFrame(records, columns=names)
    if df.empty:
        raise ValueError("summary with no kernels; check that the capture holds GPU activity")
    numeric = pd.to_numeric(df[percent].astype(str).str.replace("%", "", regex=False).str.strip(),
                            errors="coerce")
    valid = numeric.notna() & numeric.map(lambda x: math.isfinite(x) and 0 <= x <= 100)
    valid &= df[name].astype(str).str.strip().ne("")
    discarded = int((~valid).sum())
    df = df.loc[valid].copy()
    df[percent] = numeric.loc[valid]
    if df.empty:
        raise ValueError("the CSV holds no valid kernel percentages")
    return df, name, percent, discarded


def export_stats(source, output, rocpd="rocpd", timeout=600):
    """Kernel-summary CSV from a rocprofv3 capture -> (installed CSV, kind).

    A directory must hold exactly one *_kernel_stats.csv; a .db is summarised
    by `rocpd summary --format csv`.  The table is validated before it is
    installed; the tool transcript goes to <output>/<stem>_rocpd_summary.log.
    """
    source = Path(source)
    output = Path(output)
    output.mkdir(parents=True, exist_ok=True)
    if source.is_dir():
        found = sorted(source.rglob("*_kernel_stats.csv"))
        if len(found) != 1:
            raise ValueError(f"{source}: expected one *_kernel_stats.csv, found {len(found)}")
        read_summary(found[0])
        dest = output / f"{source.name}_kernel_stats.csv"
        shutil.copyfile(found[0], dest)
        return dest, "kernel_stats"
    if source.suffix.lower() != ".db":
        raise ValueError("use a .csv summary, a rocprofv3 output directory or a rocpd .db")
    exe = shutil.which(rocpd)
    if not exe:
        raise ValueError(f"executable {rocpd!r} not found; give an exported CSV instead")
    log = output / f"{source.stem}_rocpd_summary.log"
    transcript = []
    try:
        with tempfile.TemporaryDirectory(prefix="mgx-rocpd-") as tmp:
            cmd = [exe, "summary", "-i", str(source.resolve()), "--format", "csv",
                   "--output-path", tmp, "--output-file", source.stem]
            transcript.append("COMMAND: " + subprocess.list2cmdline(cmd))
            try:
                proc = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
            except subprocess.TimeoutExpired as exc:
                raise ValueError(f"rocpd summary exceeded {timeout}s; see {log}") from exc
            transcript.extend([proc.stdout, proc.stderr])
            if proc.returncode != 0:
                raise ValueError(f"rocpd summary failed (exit {proc.returncode}); see {log}")
            kernels = [p for p in Path(tmp).rglob("*.csv") if "kernel" in p.name.lower()]
            if len(kernels) != 1:
                raise ValueError(f"rocpd summary wrote {len(kernels)} kernel tables; see {log}")
            read_summary(kernels[0])
            dest = output / f"{source.stem}_kernel_summary.csv"
            shutil.copyfile(kernels[0], dest)
            return dest, "rocpd_summary"
    finally:
        log.write_text("\n".join(transcript))


def mfma_table(counter_csv) -> dict:
    """Per kernel name: dispatches, matrix-core operations per dispatch, mean
    dispatch duration and the fraction of the data type's peak, from a
    rocprofv3 --pmc counter collection CSV."""
    per = {}
    with open(counter_csv, newline="", encoding="utf-8-sig") as f:
        for r in csv.DictReader(f):
            cname = r.get("Counter_Name", "")
            if cname not in MFMA_PEAKS:
                continue
            k = per.setdefault(r["Kernel_Name"], {"dispatch": {}, "ops": {}})
            did = r.get("Dispatch_Id")
            k["dispatch"][did] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            k["ops"][cname] = k["ops"].get(cname, 0.0) + float(r["Counter_Value"]) * MOPS_UNIT
    out = {}
    for name, k in per.items():
        n = len(k["dispatch"])
        dur = sum(k["dispatch"].values()) / n * 1e-9 if n else 0.0
        busy = {c: v for c, v in k["ops"].items() if v > 0}
        if not busy or dur <= 0:
            continue
        cname = max(busy, key=busy.get)
        dtype, peak = MFMA_PEAKS[cname]
        ops = busy[cname] / n
        out[name] = {"dispatches": n, "mfma_type": dtype, "mfma_ops_per_dispatch": ops,
                     "dispatch_ms": dur * 1e3, "mfma_rate": ops / dur, "mfma_peak": peak,
                     "mfma_frac": ops / dur / peak}
    return out


def plot_summary(path, top=15, output_dir=None, report=None, counters=None):
    """Top-N kernel-share bar chart (PNG and PDF), normalised CSV and a JSON
    note; with counters, MFMA columns and bar annotations."""
    if top < 1:
        raise ValueError("--top must be positive")
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    path = Path(path)
    df, name, percent, discarded = read_summary(path)
    mfma = mfma_table(counters) if counters else {}
    if mfma:
        for col in ("mfma_type", "mfma_ops_per_dispatch", "mfma_rate", "mfma_frac"):
            df[col] = [mfma.get(str(x).strip(), {}).get(col) for x in df[name]]
    selected = df.nlargest(top, percent).sort_values(percent)
    output = Path(output_dir) if output_dir else path.parent
    output.mkdir(parents=True, exist_ok=True)
    labels = ["\n".join(textwrap.wrap(str(x), width=80, break_long_words=True))
              for x in selected[name]]
    height = max(5, sum(max(1, lab.count("\n") + 1) for lab in labels) * 0.35 + 2)
    fig, ax = plt.subplots(figsize=(14, height))
    pos = range(len(selected))
    bars = ax.barh(pos, selected[percent], color="#C8102E")
    ax.set_yticks(list(pos))
    ax.set_yticklabels(labels, fontsize=9)
    texts = []
    for _, row in selected.iterrows():
        t = f"{row[percent]:.2f}%"
        frac = row.get("mfma_frac") if mfma else None
        if frac is not None and not (isinstance(frac, float) and math.isnan(frac)):
            t += f"  MFMA {row['mfma_type']} {100 * frac:.1f}% of peak"
        texts.append(t)
    ax.bar_label(bars, labels=texts, padding=4)
    ax.set_xlabel("share of the accumulated kernel time (%)")
    ax.set_ylabel("kernel")
    ax.set_title(f"rocprofv3 - {path.stem}")
    ax.grid(axis="x", linestyle="--", alpha=0.3)
    ax.set_axisbelow(True)
    ax.margins(x=0.3 if mfma else 0.15)
    fig.text(0.5, 0.01, "Accumulated kernel time: neither total application time nor overlap.",
             ha="center", fontsize=9)
    fig.tight_layout(rect=(0, 0.035, 1, 1))
    outputs = []
    for ext in ("png", "pdf"):
        target = output / f"{path.stem}_plot.{ext}"
        fig.savefig(target, dpi=300, bbox_inches="tight")
        outputs.append(target)
    plt.close(fig)
    df.to_csv(output / f"{path.stem}_normalized.csv", index=False)
    meta = {"input": str(path.resolve()), "report": report, "valid_rows": len(df),
            "discarded_rows": discarded, "plotted_rows": len(selected),
            "displayed_percent_sum": float(selected[percent].sum()), "note": NOTE}
    if mfma:
        meta["counters"] = str(Path(counters).resolve())
        meta["mfma"] = mfma
    (output / f"{path.stem}_plot.json").write_text(json.dumps(meta, indent=2))
    return outputs, discarded


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("input", help="kernel summary CSV, rocprofv3 output directory or rocpd .db")
    ap.add_argument("--top", type=int, default=15)
    ap.add_argument("--output-dir", type=Path)
    ap.add_argument("--counters", type=Path, default=None,
                    help="rocprofv3 --pmc counter collection CSV with MFMA MOPS counters")
    ap.add_argument("--rocpd", default="rocpd", help="rocpd executable (for .db inputs)")
    ap.add_argument("--timeout", type=int, default=600)
    args = ap.parse_args(argv)
    if args.top < 1 or args.timeout < 1:
        ap.error("--top and --timeout must be positive")
    source = Path(args.input)
    if not source.exists():
        ap.error(f"not found: {source}")
    if args.counters is not None and not args.counters.is_file():
        ap.error(f"not found: {args.counters}")
    report = None
    try:
        if source.is_dir() or source.suffix.lower() == ".db":
            source, report = export_stats(source, args.output_dir or source.parent / "stats",
                                          args.rocpd, args.timeout)
            print(f"CSV exported: {source}")
        elif source.suffix.lower() != ".csv":
            raise ValueError("use a .csv summary, a rocprofv3 output directory or a rocpd .db")
        outputs, discarded = plot_summary(source, args.top, args.output_dir, report, args.counters)
        if discarded:
            print(f"Warning: {discarded} rows without a valid name or percentage were discarded.")
        for out in outputs:
            print(f"plot saved: {out}")
    except (ValueError, OSError) as exc:
        ap.exit(2, f"Error: {exc}\n")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
