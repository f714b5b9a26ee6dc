"""Calibrated configuration prediction: the autotuning/predictive.py counterpart.

The reference's AUTO path (autotuning/predictive.py; benchmark.py:83-85)
calibrates every candidate configuration -- communication library code,
and for 2D the process grid and SUMMA panel width BK -- by short measured
benchmark runs over several sizes, stores the medians as a schema-1
calibration model, and at a requested size ranks the candidates by predicted
time.  This module keeps that workflow, its CLI (calibrate / recover /
auto), the schema-1 model format and the prediction methods:

  extrapolation  piecewise log-log interpolation between the bracketing
                 calibrated sizes, exponent clipped to [0, 4]
                 (predictive.py:98-127);
  ml             ridge regression on (x^2, x^3), x = N / max N, with
                 population-std standardisation and a free intercept
                 (predictive.py:140-158; solved here in closed form);
                 with >= 4 sizes the least-squares fit of the same
                 features is the 'extrapolation' method's first choice;
  auto           the method with the lower median hold-out-largest-size
                 error over the candidates (predictive.py:213-233);
  mlp            new: one 3-layer width-64 FP32 network over all candidates
                 (features: log2 N, library one-hot, log2 grid, log2 BK),
                 trained on the calibration medians by libmgx's GPU kernels.

MI355X differences: the architecture fingerprint comes from rocm-smi, the
calibration runs drive `python -m mgxbench_amd.runner` (the benchmark.py
counterpart, same report schema) with the one-rank launcher stand-in, and
models live in mgxbench_amd/models/.

    python -m mgxbench_amd.predictive calibrate --decomposition 2d --ranks 1 --sizes 2048 4096 8192
    python -m mgxbench_amd.predictive auto --decomposition 2d --ranks 1 --size 16384 [--prediction-model mlp]
"""
from __future__ import annotations

import argparse
import datetime
import hashlib
import json
import math
import os
import shlex
import shutil
import statistics
import subprocess
import sys
from pathlib import Path

from .runner import LIBRARY_NAMES

LIBRARIES = tuple(LIBRARY_NAMES)
PKG = Path(__file__).resolve().parent
SCHEMA = 1
METHODS = ("extrapolation", "ml", "auto", "mlp")


# ---- architecture fingerprint -------------------------------------------------------------

def _rocm_smi(args: list[str]) -> str:
    exe = shutil.which("rocm-smi")
    if not exe:
        raise ValueError("Cannot identify GPU architecture (rocm-smi unavailable)")
    try:
        res = subprocess.run([exe, *args], capture_output=True, text=True, timeout=15)
    except (OSError, subprocess.TimeoutExpired) as exc:
        raise ValueError(f"Cannot identify GPU architecture ({exc})") from None
    if res.returncode != 0:
        raise ValueError("Cannot identify GPU architecture (rocm-smi failed)")
    return res.stdout


def gpu_identity() -> tuple[list[str], list[str]]:
    """(GPU model names, driver versions) reported by rocm-smi."""
    import csv
    import io
    text = _rocm_smi(["--showproductname", "--csv"])
    rows = list(csv.DictReader(io.StringIO("\n".join(l for l in text.splitlines() if "," in l))))

    def ident(r):
        # series when reported, else the PCI model id; with the gfx target
        parts = [(r.get(k) or "").strip() for k in ("Card Series", "Card Model", "GFX Version")]
        parts = [x for x in parts if x and x.upper() != "N/A"]
        return " ".join(parts[:1] + parts[-1:]) if len(parts) > 1 else "".join(parts)

    names = sorted({ident(r) for r in rows
                    if (r.get("device") or "").lower().startswith("card")} - {""})
    if not names:
        raise ValueError("No GPU detected")
    drv = _rocm_smi(["--showdriverversion", "--csv"])
    versions = sorted({line.split(",")[-1].strip().strip('"') for line in drv.splitlines()[1:]
                       if "," in line} - {""})
    return names, versions


def fingerprint(root, decomposition: str, ranks: int, build_dir) -> dict:
    """Stable compatibility key (GPU models, ranks, decomposition, executable
    hash, node count) plus diagnostic fields; `id` hashes the stable part."""
    exe = Path(root) / build_dir / ("mm_2d" if decomposition == "2d" else "mm_1d")
    if not exe.is_file():
        raise ValueError(f"Executable not found: {exe}; build with `python -m mgxbench_amd.build`")
    names, versions = gpu_identity()
    stable = {"gpu_models": names, "ranks": ranks, "decomposition": decomposition,
              "binary_sha256": hashlib.sha256(exe.read_bytes()).hexdigest(),
              "slurm_nodes": os.environ.get("SLURM_NNODES", "1")}
    sig = dict(stable)
    sig["driver_versions"] = versions
    sig["visible_devices"] = os.environ.get("HIP_VISIBLE_DEVICES",
                                            os.environ.get("ROCR_VISIBLE_DEVICES", ""))
    sig["gpus_per_node"] = os.environ.get("SLURM_GPUS_ON_NODE", "")
    sig["id"] = hashlib.sha256(json.dumps(stable, sort_keys=True).encode()).hexdigest()[:16]
    return sig


def compatible_signature(stored: dict, actual: dict) -> bool:
    def models(x):
        return sorted({str(v).split(",")[0].strip() for v in x})
    return (models(stored.get("gpu_models", [])) == models(actual.get("gpu_models", []))
            and all(str(stored.get(k)) == str(actual.get(k))
                    for k in ("ranks", "decomposition", "binary_sha256", "slurm_nodes")))


def model_filename(decomposition: str, ranks: int) -> str:
    return f"calibration-{decomposition}-{ranks}gpu.json"


def resolve_model(decomposition: str, ranks: int, supplied=None, root: Path | None = None) -> Path:
    if supplied:
        p = Path(supplied)
        return p if p.is_absolute() or root is None else root / p
    models = PKG / "models"
    named = models / model_filename(decomposition, ranks)
    legacy = models / "calibration.json"
    return named if named.exists() or not legacy.exists() else legacy


# ---- candidates ------------------------------------------------------------------------------

def factor_grids(ranks: int) -> list[tuple[int, int]]:
    return [(r, ranks // r) for r in range(1, ranks + 1) if ranks % r == 0]


def candidate_configs(decomposition, ranks, sizes, libraries, bks, grids=None) -> list[dict]:
    if decomposition == "1d":
        return [{"library": lib, "grid_rows": None, "grid_cols": None, "summa_bk": None}
                for lib in libraries]
    out = []
    for r, c in (grids if grids is not None else factor_grids(ranks)):
        if r * c != ranks or any(n % r or n % c for n in sizes):
            continue
        out += [{"library": lib, "grid_rows": r, "grid_cols": c, "summa_bk": bk}
                for bk in bks if bk > 0 for lib in libraries]
    return out


def key(cfg: dict) -> tuple:
    return (cfg["library"], cfg.get("grid_rows"), cfg.get("grid_cols"), cfg.get("summa_bk"))


# ---- prediction methods ----------------------------------------------------------------------

def points(samples) -> list[tuple[int, float]]:
    """(size, median seconds) per distinct valid size, ascending."""
    by = {}
    for s in samples:
        n, t = int(s["size"]), float(s["median_s"])
        if n > 0 and math.isfinite(t) and t > 0:
            by.setdefault(n, []).append(t)
    return sorted((n, statistics.median(ts)) for n, ts in by.items())


def estimate(samples, target: int) -> tuple[float, bool]:
    """Piecewise log-log interpolation at `target` (bracketing sizes, or the
    two end sizes when outside), exponent clipped to [0, 4]; returns
    (seconds, extrapolated)."""
    pts = points(samples)
    if len(pts) < 2:
        raise ValueError("At least two calibrated sizes required per candidate")
    exact = [t for n, t in pts if n == target]
    if exact:
        return exact[0], False
    if target < pts[0][0]:
        lo, hi = pts[0], pts[1]
    elif target > pts[-1][0]:
        lo, hi = pts[-2], pts[-1]
    else:
        j = next(i for i in range(len(pts) - 1) if pts[i][0] < target < pts[i + 1][0])
        lo, hi = pts[j], pts[j + 1]
    slope = math.log(hi[1] / lo[1]) / math.log(hi[0] / lo[0])
    slope = min(4.0, max(0.0, slope))
    return lo[1] * (target / lo[0]) ** slope, not (pts[0][0] <= target <= pts[-1][0])


def physical_fit(pts, target: int, method: str, alpha: float = 0.1) -> float:
    """Fit t(N) on the features (x^2, x^3), x = N / max calibrated N.
    'ml': ridge (alpha) on standardised features (population std) with an
    unpenalised intercept, i.e. centred features and targets and
    w = (Z'Z + alpha I)^-1 Z'(t - mean t); otherwise ordinary least squares
    with an intercept."""
    import numpy as np
    if len(pts) < 4:
        raise ValueError("Physical regression requires >=4 distinct sizes")
    n = np.array([p[0] for p in pts], dtype=float)
    t = np.array([p[1] for p in pts], dtype=float)
    ref = float(n.max())

    def feats(v):
        x = np.asarray(v, dtype=float) / ref
        return np.column_stack((x ** 2, x ** 3))

    X = feats(n)
    xt = feats([target])[0]
    if method == "ml":
        mu = X.mean(axis=0)
        sd = X.std(axis=0)
        sd[sd == 0] = 1.0
        Z = (X - mu) / sd
        w = np.linalg.solve(Z.T @ Z + alpha * np.eye(2), Z.T @ (t - t.mean()))
        pred = float(t.mean() + ((xt - mu) / sd) @ w)
    else:
        D = np.column_stack((np.ones(len(n)), X))
        coef, *_ = np.linalg.lstsq(D, t, rcond=None)
        pred = float(np.r_[1.0, xt] @ coef)
    if not math.isfinite(pred) or pred <= 0:
        raise ValueError("Nonpositive or nonfinite physical prediction")
    return pred


def mlp_features(size: int, cfg: dict) -> list[float]:
    onehot = [1.0 if cfg["library"] == lib else 0.0 for lib in LIBRARIES]
    gr, gc, bk = cfg.get("grid_rows") or 1, cfg.get("grid_cols") or 1, cfg.get("summa_bk") or 1
    return [math.log2(size)] + onehot + [math.log2(gr), math.log2(gc), math.log2(bk)]


def mlp_predictions(model: dict, size: int, cands: list[dict], epochs: int = 3000,
                    seed: int = 3) -> list[float]:
    """Seconds predicted by an MLP trained (on the GPU, libmgx kernels) on the
    log calibration medians of every candidate."""
    import ctypes
    import numpy as np
    from . import _lib
    X, y = [], []
    for c in model["candidates"]:
        for n, t in points(c["measurements"]):
            X.append(mlp_features(n, c))
            y.append(math.log(t))
    if len(y) < 4:
        raise ValueError("MLP needs at least four calibration points")
    X = np.array(X, dtype=np.float64)
    mean, std = X.mean(axis=0), X.std(axis=0)
    std[std == 0] = 1.0
    Xs = np.ascontiguousarray((X - mean) / std, dtype=np.float32)
    ys = np.ascontiguousarray(y, dtype=np.float32)
    D = Xs.shape[1]
    theta = np.empty(_lib.lib().mgx_mlp_param_count(D), dtype=np.float32)
    _lib.call("mgx_mlp_init", theta.ctypes.data, D, seed)
    hist = np.empty(epochs, dtype=np.float32)
    _lib.call("mgx_mlp_train", theta.ctypes.data, D, Xs.ctypes.data, ys.ctypes.data, len(ys),
              epochs, ctypes.c_float(3e-3), ctypes.c_float(1e-5), hist.ctypes.data)
    Q = np.ascontiguousarray((np.array([mlp_features(size, c) for c in cands]) - mean) / std,
                             dtype=np.float32)
    out = np.empty(len(cands), dtype=np.float32)
    _lib.call("mgx_mlp_predict", theta.ctypes.data, D, Q.ctypes.data, len(cands), out.ctypes.data)
    return [float(math.exp(v)) for v in out]


def predict_method(model: dict, size: int, ranks: int, decomposition: str, grid, method: str):
    """Candidates ranked by predicted seconds at `size` under one method."""
    if size <= 0 or ranks <= 0:
        raise ValueError("size and ranks must be positive")
    sig = model["signature"]
    if sig["ranks"] != ranks or sig["decomposition"] != decomposition:
        raise ValueError("Model is for a different rank count or decomposition")
    usable = []
    for c in model["candidates"]:
        if grid and (c["grid_rows"], c["grid_cols"]) != tuple(grid):
            continue
        if decomposition == "2d" and (size % c["grid_rows"] or size % c["grid_cols"]):
            continue
        if decomposition == "1d" and size % ranks:
            continue
        usable.append(c)
    out = []
    mlp = mlp_predictions(model, size, usable) if method == "mlp" and usable else None
    for idx, c in enumerate(usable):
        pts = points(c["measurements"])
        if len(pts) < 1:
            continue
        if mlp is not None:
            seconds, algorithm = mlp[idx], "mlp-gpu"
        else:
            seconds = None
            if len(pts) >= 4:
                try:
                    seconds = physical_fit(pts, size, method)
                    algorithm = "ridge-physical" if method == "ml" else "physical-polynomial"
                except (ValueError, ArithmeticError):
                    seconds = None
            if seconds is None:
                if method != "extrapolation":
                    continue
                try:
                    seconds, _ = estimate(c["measurements"], size)
                except ValueError:
                    continue
                algorithm = "piecewise-power-law-fallback"
        out.append({"library": c["library"], "grid_rows": c["grid_rows"],
                    "grid_cols": c["grid_cols"], "summa_bk": c["summa_bk"],
                    "predicted_s": seconds, "algorithm": algorithm,
                    "calibration_sizes": len(pts),
                    "extrapolated": not (pts[0][0] <= size <= pts[-1][0])})
    if not out:
        raise ValueError("No compatible candidates with sufficient valid calibration sizes")
    return sorted(out, key=lambda r: r["predicted_s"])


def select_predictions(model, size, ranks, decomposition, grid, method):
    """(ranking, method used).  'auto' picks, between the least-squares and
    the ridge fit, the one with the lower median relative error when each
    candidate's largest size is predicted from the others (>= 5 sizes)."""
    if method in ("extrapolation", "ml", "mlp"):
        return predict_method(model, size, ranks, decomposition, grid, method), method
    if method != "auto":
        raise ValueError(f"Unknown prediction model: {method}")
    errors = {"extrapolation": [], "ml": []}
    for c in model["candidates"]:
        pts = points(c["measurements"])
        if len(pts) < 5:
            continue
        for m in errors:
            try:
                p = physical_fit(pts[:-1], pts[-1][0], m)
            except ValueError:
                continue
            errors[m].append(abs(p - pts[-1][1]) / pts[-1][1])
    scored = sorted((statistics.median(e), m) for m, e in errors.items() if e)
    chosen = scored[0][1] if scored else "extrapolation"
    return predict_method(model, size, ranks, decomposition, grid, chosen), chosen


def load_model(path) -> dict:
    model = json.loads(Path(path).read_text())
    if model.get("schema") != SCHEMA:
        raise ValueError("Unsupported autotuning model schema")
    return model


def _write_model(model: dict, out: Path) -> None:
    out.parent.mkdir(parents=True, exist_ok=True)
    if out.exists():
        shutil.copy2(out, out.with_suffix(out.suffix + ".bak"))
    tmp = out.with_name(out.name + ".tmp")
    tmp.write_text(json.dumps(model, indent=2) + "\n")
    tmp.replace(out)


# ---- calibration -----------------------------------------------------------------------------

def _runner_command(args, cfg: dict, size: int, output: Path, label: str) -> list[str]:
    cmd = [sys.executable, "-m", "mgxbench_amd.runner", "--decomposition", args.decomposition,
           "--size", str(size), "--ranks", str(args.ranks), "--libraries", cfg["library"],
           "--warmup", str(args.warmup), "--repeats", str(args.repeats),
           "--build-dir", str(args.build_dir), "--output", str(output), "--label", label]
    if args.launcher:
        cmd += ["--launcher", args.launcher]
    if args.launcher_args:
        cmd += ["--launcher-args", args.launcher_args]
    if args.decomposition == "2d":
        cmd += ["--grid-rows", str(cfg["grid_rows"]), "--grid-cols", str(cfg["grid_cols"]),
                "--summa-bk", str(cfg["summa_bk"])]
    return cmd


def _report_matches(doc: dict, cfg: dict, size: int, args, label: str) -> bool:
    a = doc.get("command_args") or {}
    libs = a.get("libraries")
    if isinstance(libs, str):
        libs = [libs]
    if (a.get("label", doc.get("label")) != label or a.get("size") != size
            or a.get("ranks") != args.ranks or a.get("decomposition") != args.decomposition
            or libs != [cfg["library"]]):
        return False
    return args.decomposition != "2d" or all(a.get(k) == cfg[k]
                                              for k in ("grid_rows", "grid_cols", "summa_bk"))


def _validated_median(doc: dict, lib: str) -> float | None:
    for row in doc.get("results") or []:
        if not isinstance(row, dict) or row.get("library") != lib or row.get("validated") is not True:
            continue
        m = row.get("median_s")
        if isinstance(m, (int, float)) and not isinstance(m, bool) and math.isfinite(m) and m > 0:
            return float(m)
    return None


def _new_reports(runs_dir: Path, before: dict):
    for rep in runs_dir.glob("*/benchmark.json"):
        st = rep.stat()
        if before.get(rep.resolve()) == (st.st_mtime_ns, st.st_size):
            continue
        try:
            yield st.st_mtime_ns, rep, json.loads(rep.read_text())
        except (OSError, ValueError):
            continue


def calibrate(args) -> int:
    root = Path(args.project_root).resolve()
    sig = fingerprint(root, args.decomposition, args.ranks, args.build_dir)
    sizes = sorted(set(args.sizes))
    if len(sizes) < 2 or any(n <= 0 for n in sizes):
        raise ValueError("Supply at least two distinct positive calibration sizes")
    if args.decomposition == "1d" and any(n % args.ranks for n in sizes):
        raise ValueError("1D calibration sizes must be divisible by ranks")
    grids = [tuple(map(int, g.lower().split("x"))) for g in args.grids] if args.grids else None
    if args.grid_rows is not None:
        grids = [(args.grid_rows, args.grid_cols)]
    configs = candidate_configs(args.decomposition, args.ranks, sizes, args.libraries, args.bks,
                                grids)
    if not configs:
        raise ValueError("No valid calibration candidates")
    model = {"schema": SCHEMA, "signature": sig, "sizes": sizes, "candidates": [],
             "warning": "Empirical calibration; predictions are estimates, not measured timings"}
    out = resolve_model(args.decomposition, args.ranks, args.model, root)
    partial = out.with_name(out.stem + ".partial.json")
    out.parent.mkdir(parents=True, exist_ok=True)
    output = Path(args.output) if Path(args.output).is_absolute() else root / args.output
    runs_dir, logs = output / "runs", output / "calibration_logs"
    logs.mkdir(parents=True, exist_ok=True)
    stats = dict.fromkeys(("attempted", "valid_measurements", "failed_launches",
                           "missing_reports", "rejected_reports", "rejected_results"), 0)
    for idx, cfg in enumerate(configs, 1):
        item = {**cfg, "measurements": []}
        for n in sizes:
            stats["attempted"] += 1
            cmd = _runner_command(args, cfg, n, output, "auto-calibration")
            tag = (f"grid{cfg['grid_rows']}x{cfg['grid_cols']}_bk{cfg['summa_bk']}"
                   if args.decomposition == "2d" else "1d")
            name = f"{idx:03d}_{cfg['library']}_N{n}_{tag}"
            print(f"Calibrating {idx}/{len(configs)} {cfg} N={n}", flush=True)
            before = {p.resolve(): (p.stat().st_mtime_ns, p.stat().st_size)
                      for p in runs_dir.glob("*/benchmark.json")}
            try:
                done = subprocess.run(cmd, cwd=root, capture_output=True, text=True)
            except OSError as exc:
                stats["failed_launches"] += 1
                (logs / (name + ".stderr")).write_text(str(exc) + "\n")
                continue
            (logs / (name + ".stdout")).write_text(done.stdout)
            (logs / (name + ".stderr")).write_text(done.stderr)
            (logs / (name + ".command.txt")).write_text(shlex.join(cmd) + "\n")
            if done.returncode:
                stats["failed_launches"] += 1
                print(f"  FAILED (exit {done.returncode}); logs: {logs / name}", file=sys.stderr,
                      flush=True)
                continue
            found = [r for r in _new_reports(runs_dir, before)
                     if _report_matches(r[2], cfg, n, args, "auto-calibration")]
            if not found:
                stats["missing_reports"] += 1
                continue
            _, path, doc = max(found, key=lambda r: r[0])
            if doc.get("failures"):
                stats["rejected_reports"] += 1
                continue
            med = _validated_median(doc, cfg["library"])
            if med is None:
                stats["rejected_results"] += 1
                continue
            item["measurements"].append({"size": n, "median_s": med})
            stats["valid_measurements"] += 1
            print(f"  SUCCESS: N={n} median={med:.6f}s ({path})", flush=True)
        distinct = len({m["size"] for m in item["measurements"]})
        if distinct >= 2:
            model["candidates"].append(item)
        else:
            print(f"  REJECTED candidate: only {distinct} distinct valid sizes", file=sys.stderr)
        partial.write_text(json.dumps(model, indent=2) + "\n")
    print(f"Calibration summary: {stats}; eligible candidates="
          f"{len(model['candidates'])}/{len(configs)}", flush=True)
    if not model["candidates"]:
        raise ValueError(f"Calibration produced no candidate with two valid measurements; "
                         f"partial model: {partial}; logs: {logs}")
    _write_model(model, out)
    partial.unlink(missing_ok=True)
    print(f"Calibration saved: {out} ({len(model['candidates'])} candidates)")
    return 0


def recover_calibration(args) -> int:
    """Rebuild a calibration model from existing auto-calibration reports
    (requires --assume-compatible: old reports carry no GPU/binary proof)."""
    root = Path(args.project_root).resolve()
    src = Path(args.input) if Path(args.input).is_absolute() else root / args.input
    reports = sorted(src.glob("*/benchmark.json"))
    if not reports:
        raise ValueError(f"No benchmark.json files found under {src}")
    if not args.assume_compatible:
        raise ValueError("Cannot verify GPU/binary provenance of historical reports; review them "
                         "and rerun with --assume-compatible if they match this system")
    sig = fingerprint(root, args.decomposition, args.ranks, args.build_dir)
    grouped, reasons = {}, {}

    def reject(why):
        reasons[why] = reasons.get(why, 0) + 1

    for rep in reports:
        try:
            doc = json.loads(rep.read_text())
        except (OSError, ValueError) as exc:
            reject(f"unreadable JSON: {type(exc).__name__}")
            continue
        a = doc.get("command_args") or {}
        if not isinstance(a, dict) or a.get("label", doc.get("label")) != "auto-calibration":
            reject("not an auto-calibration report")
            continue
        try:
            size, ranks = int(a["size"]), int(a["ranks"])
        except (KeyError, TypeError, ValueError):
            reject("missing/invalid size or ranks")
            continue
        if ranks != args.ranks or a.get("decomposition") != args.decomposition:
            reject("different rank count or decomposition")
            continue
        if size <= 0 or (args.decomposition == "1d" and size % ranks):
            reject("invalid matrix size")
            continue
        libs = a.get("libraries")
        libs = [libs] if isinstance(libs, str) else libs
        if not isinstance(libs, list) or len(libs) != 1 or libs[0] not in LIBRARIES:
            reject("missing/ambiguous library")
            continue
        gr = gc = bk = None
        if args.decomposition == "2d":
            try:
                gr, gc, bk = (int(a[k]) for k in ("grid_rows", "grid_cols", "summa_bk"))
            except (KeyError, TypeError, ValueError):
                reject("missing/invalid grid or BK")
                continue
            if gr <= 0 or gc <= 0 or gr * gc != ranks or bk <= 0 or size % gr or size % gc:
                reject("incompatible grid/BK")
                continue
        if doc.get("failures"):
            reject("report has failures")
            continue
        med = _validated_median(doc, libs[0])
        if med is None:
            reject("no validated positive median_s")
            continue
        grouped.setdefault((libs[0], gr, gc, bk), {}).setdefault(size, []).append(med)
    cands = []
    for (lib, gr, gc, bk), by in sorted(grouped.items(), key=lambda kv: str(kv[0])):
        if len(by) < 2:
            reject(f"candidate {lib}/{gr}x{gc}/BK{bk} has only {len(by)} size(s)")
            continue
        cands.append({"library": lib, "grid_rows": gr, "grid_cols": gc, "summa_bk": bk,
                      "measurements": [{"size": n, "median_s": statistics.median(ts)}
                                       for n, ts in sorted(by.items())]})
    print(f"Reports scanned: {len(reports)}; eligible candidates: {len(cands)}", flush=True)
    for why, count in sorted(reasons.items()):
        print(f"  REJECTED {count}: {why}", flush=True)
    if not cands:
        raise ValueError("Recovery found no candidate with >=2 validated sizes")
    model = {"schema": SCHEMA, "signature": sig,
             "sizes": sorted({m["size"] for c in cands for m in c["measurements"]}),
             "candidates": cands,
             "warning": ("Recovered empirical calibration; predictions are estimates. Hardware "
                         "and binary compatibility was confirmed with --assume-compatible.")}
    out = resolve_model(args.decomposition, args.ranks, args.model, root)
    _write_model(model, out)
    print(f"Recovered model saved: {out}; candidates={len(cands)}", flush=True)
    return 0


# ---- AUTO ----------------------------------------------------------------------------------

def validate_and_feedback(args, best: dict, report: dict, model: dict, model_path: Path) -> int:
    """Measure the chosen configuration; with feedback, append the validated
    median to its candidate in the calibration model (previous file kept)."""
    output = Path(args.output) / "validation"
    cmd = _runner_command(args, best, args.size, output, "auto-validation")
    print("AUTO validation: " + shlex.join(cmd), flush=True)
    before = {p.resolve(): (p.stat().st_mtime_ns, p.stat().st_size)
              for p in (output / "runs").glob("*/benchmark.json")}
    done = subprocess.run(cmd)
    if done.returncode:
        report["validation"] = {"status": "FAILED", "exit_code": done.returncode}
        return 2
    found = [r for r in _new_reports(output / "runs", before)
             if _report_matches(r[2], best, args.size, args, "auto-validation")
             and not r[2].get("failures")]
    if not found:
        report["validation"] = {"status": "FAILED", "reason": "No matching benchmark report"}
        return 2
    _, path, doc = max(found, key=lambda r: r[0])
    measured = _validated_median(doc, best["library"])
    if measured is None:
        report["validation"] = {"status": "FAILED", "reason": "No validated result"}
        return 2
    pred = best["predicted_s"]
    report["validation"] = {"status": "PASS", "median_s": measured, "predicted_s": pred,
                            "absolute_error_s": abs(measured - pred),
                            "relative_error": abs(measured - pred) / measured,
                            "benchmark_report": str(path)}
    print(f"AUTO validation: predicted={pred:.6f}s measured={measured:.6f}s "
          f"error={100 * abs(measured - pred) / measured:.2f}%", flush=True)
    if not args.auto_feedback:
        return 0
    cand = next((c for c in model["candidates"] if key(c) == key(best)), None)
    if cand is None:
        report["feedback"] = {"status": "SKIPPED", "reason": "Candidate not found"}
        return 0
    cand["measurements"].append({"size": args.size, "median_s": measured, "source": str(path),
                                 "timestamp_utc": datetime.datetime.now(
                                     datetime.timezone.utc).isoformat()})
    model["sizes"] = sorted(set(model.get("sizes", [])) | {args.size})
    _write_model(model, model_path)
    report["feedback"] = {"status": "SAVED", "model": str(model_path),
                          "backup": str(model_path.with_suffix(model_path.suffix + ".bak"))}
    return 0


def auto(args) -> int:
    """--libraries AUTO: rank the calibrated candidates at the requested size."""
    path = resolve_model(args.decomposition, args.ranks, getattr(args, "auto_model", None),
                         Path.cwd())
    if not path.is_file():
        print(f"AUTO model missing ({path}). Run: python -m mgxbench_amd.predictive calibrate "
              "--help", file=sys.stderr)
        return 2
    try:
        model = load_model(path)
        actual = fingerprint(Path.cwd(), args.decomposition, args.ranks, args.build_dir)
        if not compatible_signature(model["signature"], actual):
            raise ValueError("Model does not match this GPU architecture/binary/allocation; "
                             "recalibrate")
        grid = (args.grid_rows, args.grid_cols) if args.grid_rows and args.grid_cols else None
        ranking, chosen = select_predictions(model, args.size, args.ranks, args.decomposition,
                                             grid, getattr(args, "prediction_model",
                                                           "extrapolation"))
    except (ValueError, ImportError, OSError) as exc:
        print(f"AUTO: {exc}", file=sys.stderr)
        return 2
    best = ranking[0]
    base = next((r for r in ranking if r["library"] == "MMM" and
                 (r["grid_rows"], r["grid_cols"], r["summa_bk"]) ==
                 (best["grid_rows"], best["grid_cols"], best["summa_bk"])), None)
    speedup = base["predicted_s"] / best["predicted_s"] if base else None
    report = {"size": args.size, "ranks": args.ranks, "decomposition": args.decomposition,
              "architecture_id": actual["id"], "best": best, "ranking": ranking,
              "speedup_vs_MPI_predicted": speedup, "prediction_model": chosen,
              "calibration_file": str(path), "status": "PREDICTED", "measured": False}
    status = 0
    if getattr(args, "auto_validate", False):
        status = validate_and_feedback(args, best, report, model, path)
        report["status"] = "VALIDATED" if status == 0 else "VALIDATION_FAILED"
    out = Path(args.output) / "autotuning_prediction.json"
    out.parent.mkdir(parents=True, exist_ok=True)
    out.write_text(json.dumps(report, indent=2) + "\n")
    bar = "=" * 48
    lines = [bar, "         MGXBench - AUTOTUNING REPORT", bar,
             f"Prediction Model  : {chosen}", "Configuration:",
             f"  Matrix Size       : {args.size}", f"  GPUs              : {args.ranks}",
             f"  Decomposition     : {args.decomposition.upper()}"]
    if args.decomposition == "2d":
        lines.append(f"  Grid              : {best['grid_rows']} x {best['grid_cols']}")
    lines += ["", "Best Predicted Configuration:",
              f"  Communication API : {LIBRARY_NAMES[best['library']]}",
              f"  Implementation    : {best['library']}",
              f"  SUMMA BK          : {best['summa_bk'] if args.decomposition == '2d' else 'N/A'}",
              f"  Predicted Time    : {best['predicted_s']:.6f} s",
              f"  Speedup vs MPI    : {f'{speedup:.2f}x' if speedup else 'N/A'}",
              f"  Improvement       : {f'{(1 - 1 / speedup) * 100:.2f}%' if speedup else 'N/A'}",
              f"  Extrapolation     : {'YES' if best['extrapolated'] else 'NO'}", "",
              f"Status: {report['status']}", f"Report: {out}", bar]
    print("\n" + "\n".join(lines))
    if getattr(args, "auto_validate", False):
        print("Validation status:", report.get("validation", {}).get("status", "UNKNOWN"))
        if report.get("feedback"):
            print("Feedback status:", report["feedback"]["status"])
    return status


def _check_grid_args(args, parser) -> None:
    if args.ranks <= 0:
        parser.error("--ranks must be a positive integer")
    rows, cols = getattr(args, "grid_rows", None), getattr(args, "grid_cols", None)
    if (rows is None) != (cols is None):
        parser.error("--grid-rows and --grid-cols must be supplied together")
    if rows is not None:
        if args.decomposition != "2d":
            parser.error("--grid-rows/--grid-cols require --decomposition 2d")
        if rows <= 0 or cols <= 0 or rows * cols != args.ranks:
            parser.error("--grid-rows * --grid-cols must equal --ranks (positive integers)")
        if args.command == "calibrate" and args.grids:
            parser.error("Use either --grids or --grid-rows/--grid-cols, not both")
        if args.command == "auto" and (args.size % rows or args.size % cols):
            parser.error("--size must be divisible by both grid dimensions")
    if args.command == "calibrate" and args.grids:
        for g in args.grids:
            try:
                r, c = map(int, g.lower().split("x"))
                if r <= 0 or c <= 0 or r * c != args.ranks:
                    raise ValueError
            except ValueError:
                parser.error(f"Invalid --grids entry {g!r}; expected RxC with R*C=--ranks")


def build_parser() -> argparse.ArgumentParser:
    default_launcher = f"{shlex.quote(sys.executable)} -m mgxbench_amd.launch"
    p = argparse.ArgumentParser(description="Calibrate MGXBench AUTO with short measured runs")
    sub = p.add_subparsers(dest="command", required=True)
    c = sub.add_parser("calibrate")
    c.add_argument("--project-root", default=".")
    c.add_argument("--decomposition", choices=("1d", "2d"), required=True)
    c.add_argument("--ranks", type=int, required=True)
    c.add_argument("--sizes", type=int, nargs="+", default=[2048, 4096, 8192])
    c.add_argument("--libraries", nargs="+", choices=LIBRARIES, default=list(LIBRARIES))
    c.add_argument("--bks", type=int, nargs="+", default=[128, 256])
    c.add_argument("--grids", nargs="+", default=None, help="e.g. 2x4 4x2")
    c.add_argument("--grid-rows", type=int, default=None)
    c.add_argument("--grid-cols", type=int, default=None)
    c.add_argument("--warmup", type=int, default=1)
    c.add_argument("--repeats", type=int, default=3)
    c.add_argument("--launcher", default=default_launcher)
    c.add_argument("--launcher-args", default="")
    c.add_argument("--build-dir", default=str(PKG / "bin"))
    c.add_argument("--output", default="results/auto_calibration")
    c.add_argument("--model", default=None)
    r = sub.add_parser("recover", help="Rebuild a model from existing calibration reports")
    r.add_argument("--project-root", default=".")
    r.add_argument("--decomposition", choices=("1d", "2d"), required=True)
    r.add_argument("--ranks", type=int, required=True)
    r.add_argument("--input", default="results/auto_calibration/runs")
    r.add_argument("--model", default=None)
    r.add_argument("--build-dir", default=str(PKG / "bin"))
    r.add_argument("--assume-compatible", action="store_true")
    a = sub.add_parser("auto", help="Predict the fastest calibrated configuration")
    a.add_argument("--decomposition", choices=("1d", "2d"), required=True)
    a.add_argument("--ranks", type=int, required=True)
    a.add_argument("--size", type=int, required=True)
    a.add_argument("--grid-rows", type=int, default=None)
    a.add_argument("--grid-cols", type=int, default=None)
    a.add_argument("--prediction-model", choices=METHODS, default="extrapolation")
    a.add_argument("--build-dir", default=str(PKG / "bin"))
    a.add_argument("--auto-model", default=None)
    a.add_argument("--auto-validate", action="store_true")
    a.add_argument("--auto-feedback", action="store_true")
    a.add_argument("--warmup", type=int, default=2)
    a.add_argument("--repeats", type=int, default=5)
    a.add_argument("--launcher", default=default_launcher)
    a.add_argument("--launcher-args", default="")
    a.add_argument("--timeout", type=int, default=600)
    a.add_argument("--output", default="results/auto_prediction")
    return p


def main(argv=None) -> int:
    p = build_parser()
    a = p.parse_args(argv)
    _check_grid_args(a, p)
    if a.command == "auto" and a.auto_feedback and not a.auto_validate:
        p.error("--auto-feedback requires --auto-validate")
    if a.command == "auto" and a.size <= 0:
        p.error("--size must be positive")
    try:
        if a.command == "calibrate":
            return calibrate(a)
        if a.command == "recover":
            return recover_calibration(a)
        return auto(a)
    except ValueError as exc:
        print(f"Error: {exc}", file=sys.stderr)
        return 2


if __name__ == "__main__":
    sys.exit(main())
