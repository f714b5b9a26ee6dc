"""Calibrated AUTO selection (the reference's autotuning/predictive.py role).

`--libraries AUTO` picks, for a requested matrix size, the configuration of a
run -- library code and, for 2D, the process grid and panel width -- from
short measured calibration runs.  Contract kept with the reference:

* the CLI: calibrate / recover / auto with the reference's flags
  (predictive.py:604-659), plus an `mlp` prediction method;
* the schema-1 calibration file {schema, signature, sizes, candidates:
  [{library, grid_rows, grid_cols, summa_bk, measurements: [{size,
  median_s}]}], warning} (predictive.py:281-298, 367) and the
  autotuning_prediction.json report keys;
* the runtime models (predictive.py:98-233):
    extrapolation  power law between the two calibrated sizes around the
                   target (the two end sizes outside the range), exponent
                   clipped to [0, 4]; with >= 4 sizes a least-squares fit of
                   t = c0 + c1 x^2 + c2 x^3 (x = N / largest calibrated N)
                   comes first;
    ml             ridge regression (alpha 0.1) on the standardised
                   (x^2, x^3) with a free intercept, solved in closed form
                   (scikit-learn's StandardScaler + Ridge give the same);
    auto           the one of the two fits with the lower median error when
                   every candidate's largest size is predicted from the rest;
    mlp            one 3-layer width-64 FP32 network over all candidates,
                   trained on the calibration medians by libmgx's GPU kernels.

On this platform the hardware fingerprint comes from rocm-smi, calibration
runs go through `python -m mgxbench_amd.runner` with the one-rank launcher,
and calibration files live in mgxbench_amd/models/.

    python -m mgxbench_amd.predictive calibrate --decomposition 2d --ranks 1 --sizes 2048 4096 8192
    python -m mgxbench_amd.predictive auto --decomposition 2d --ranks 1 --size 16384
"""
from __future__ import annotations

import argparse
import csv
import datetime
import hashlib
import io
import json
import math
import os
import shlex
import shutil
import statistics
import subprocess
import sys
from dataclasses import asdict, dataclass
from pathlib import Path

from .runner import LIBRARY_NAMES

LIBRARIES = tuple(LIBRARY_NAMES)
PKG = Path(__file__).resolve().parent
SCHEMA = 1
METHODS = ("extrapolation", "ml", "auto", "mlp")
CALIBRATION_LABEL = "auto-calibration"
VALIDATION_LABEL = "auto-validation"


def _now() -> str:
    return datetime.datetime.now(datetime.timezone.utc).isoformat()


# ---- hardware identity -------------------------------------------------------------------

def _rocm_smi(args: list[str]) -> str:
    """stdout of `rocm-smi args...`; ValueError when the tool is absent or fails."""
    tool = shutil.which("rocm-smi")
    if tool is None:
        raise ValueError("rocm-smi is not on PATH: the GPU model cannot be identified")
    try:
        done = subprocess.run([tool, *args], capture_output=True, text=True, timeout=15)
    except (OSError, subprocess.TimeoutExpired) as exc:
        raise ValueError(f"rocm-smi {' '.join(args)} did not answer: {exc}") from None
    if done.returncode != 0:
        raise ValueError(f"rocm-smi {' '.join(args)} exited {done.returncode}")
    return done.stdout


def _csv_rows(text: str) -> list[dict]:
    return list(csv.DictReader(io.StringIO("\n".join(l for l in text.splitlines() if "," in l))))


def gpu_identity() -> tuple[list[str], list[str]]:
    """(GPU model names, driver versions).  A model name is the card series
    when rocm-smi reports one, else its PCI model id, followed by the gfx
    target ("0x75a3 gfx950" on an MI355X, which reports no series)."""
    names = set()
    for row in _csv_rows(_rocm_smi(["--showproductname", "--csv"])):
        if not (row.get("device") or "").lower().startswith("card"):
            continue
        known = [(row.get(k) or "").strip() for k in ("Card Series", "Card Model", "GFX Version")]
        known = [v for v in known if v and v.upper() != "N/A"]
        if known:
            names.add(known[0] if len(known) == 1 else f"{known[0]} {known[-1]}")
    if not names:
        raise ValueError("rocm-smi lists no GPU")
    drivers = set()
    for line in _rocm_smi(["--showdriverversion", "--csv"]).splitlines()[1:]:
        if "," in line:
            value = line.rsplit(",", 1)[1].strip().strip('"')
            if value:
                drivers.add(value)
    return sorted(names), sorted(drivers)


STABLE_KEYS = ("ranks", "decomposition", "binary_sha256", "slurm_nodes")


def fingerprint(root, decomposition: str, ranks: int, build_dir) -> dict:
    """What a calibration is valid for: GPU models, rank count,
    decomposition, the executable's hash and the node count (plus
    informational fields); `id` hashes the part that must match."""
    exe = Path(root) / build_dir / f"mm_{decomposition}"
    if not exe.is_file():
        raise ValueError(f"{exe} does not exist; build it with `python -m mgxbench_amd.build`")
    models, drivers = gpu_identity()
    must_match = {"gpu_models": models, "ranks": ranks, "decomposition": decomposition,
                  "binary_sha256": hashlib.sha256(exe.read_bytes()).hexdigest(),
                  "slurm_nodes": os.environ.get("SLURM_NNODES", "1")}
    digest = hashlib.sha256(json.dumps(must_match, sort_keys=True).encode()).hexdigest()[:16]
    return {**must_match, "driver_versions": drivers,
            "visible_devices": os.environ.get("HIP_VISIBLE_DEVICES",
                                              os.environ.get("ROCR_VISIBLE_DEVICES", "")),
            "gpus_per_node": os.environ.get("SLURM_GPUS_ON_NODE", ""), "id": digest}


def compatible_signature(stored: dict, actual: dict) -> bool:
    def model_set(sig):
        return sorted({str(v).split(",")[0].strip() for v in sig.get("gpu_models", [])})
    same_models = model_set(stored) == model_set(actual)
    return same_models and all(str(stored.get(k)) == str(actual.get(k)) for k in STABLE_KEYS)


def model_filename(decomposition: str, ranks: int) -> str:
    return f"calibration-{decomposition}-{ranks}gpu.json"


def resolve_model(decomposition: str, ranks: int, supplied=None, root: Path | None = None) -> Path:
    """The calibration file: the one supplied (relative to `root`), else the
    per-decomposition / rank-count file in mgxbench_amd/models/ (or an older
    single calibration.json when only that exists)."""
    if supplied:
        path = Path(supplied)
        return path if path.is_absolute() or root is None else root / path
    per_shape = PKG / "models" / model_filename(decomposition, ranks)
    single = PKG / "models" / "calibration.json"
    return single if single.exists() and not per_shape.exists() else per_shape


# ---- candidates -----------------------------------------------------------------------------

@dataclass(frozen=True)
class Candidate:
    library: str
    grid_rows: int | None = None
    grid_cols: int | None = None
    summa_bk: int | None = None

    @classmethod
    def of(cls, d: dict) -> "Candidate":
        return cls(d["library"], d.get("grid_rows"), d.get("grid_cols"), d.get("summa_bk"))

    def divides(self, size: int) -> bool:
        return self.grid_rows is None or (size % self.grid_rows == 0 and size % self.grid_cols == 0)


def factor_grids(ranks: int) -> list[tuple[int, int]]:
    """Every rows x cols grid of `ranks` processes, rows ascending."""
    return [(r, ranks // r) for r in range(1, ranks + 1) if ranks % r == 0]


def candidate_configs(decomposition, ranks, sizes, libraries, bks, grids=None) -> list[dict]:
    """Configurations to calibrate: one per library in 1D; in 2D every
    library x positive panel width x grid whose blocks divide every size."""
    if decomposition == "1d":
        return [asdict(Candidate(lib)) for lib in libraries]
    configs = []
    for rows, cols in (factor_grids(ranks) if grids is None else grids):
        whole = all(n % rows == 0 and n % cols == 0 for n in sizes)
        if rows * cols != ranks or not whole:
            continue
        configs.extend(asdict(Candidate(lib, rows, cols, bk))
                       for bk in bks if bk > 0 for lib in libraries)
    return configs


def key(cfg: dict) -> tuple:
    return tuple(asdict(Candidate.of(cfg)).values())


# ---- runtime models ---------------------------------------------------------------------

def points(samples) -> list[tuple[int, float]]:
    """(size, median seconds) of every size with valid samples, ascending."""
    per_size: dict[int, list[float]] = {}
    for s in samples:
        size, seconds = int(s["size"]), float(s["median_s"])
        if size > 0 and seconds > 0 and math.isfinite(seconds):
            per_size.setdefault(size, []).append(seconds)
    return [(n, statistics.median(ts)) for n, ts in sorted(per_size.items())]


def estimate(samples, target: int) -> tuple[float, bool]:
    """Power-law estimate at `target` from the two calibrated sizes around
    it (the two end sizes when outside the calibrated range), the exponent
    clipped to [0, 4]; returns (seconds, outside the range)."""
    pts = points(samples)
    if len(pts) < 2:
        raise ValueError("a power-law estimate needs two calibrated sizes")
    sizes = [n for n, _ in pts]
    for n, t in pts:
        if n == target:
            return t, False
    # index of the lower point of the pair used
    below = sum(1 for n in sizes if n < target) - 1
    below = min(max(below, 0), len(pts) - 2)
    (n0, t0), (n1, t1) = pts[below], pts[below + 1]
    exponent = min(4.0, max(0.0, math.log(t1 / t0) / math.log(n1 / n0)))
    return t0 * (target / n0) ** exponent, not sizes[0] <= target <= sizes[-1]


def physical_fit(pts, target: int, method: str, alpha: float = 0.1) -> float:
    """t(target) from a fit of t = f(x^2, x^3), x = N / largest calibrated N.

    method 'ml': ridge with penalty `alpha` on the standardised features
    (population standard deviation) and an unpenalised intercept, i.e.
    w = (Z^T Z + alpha I)^-1 Z^T (t - mean t) on the centred, scaled features
    Z; anything else: ordinary least squares with an intercept."""
    import numpy as np
    if len(pts) < 4:
        raise ValueError("the x^2 / x^3 fits need four calibrated sizes")
    sizes = np.array([p[0] for p in pts], dtype=float)
    secs = np.array([p[1] for p in pts], dtype=float)
    scale = sizes.max()
    design = np.stack([(sizes / scale) ** 2, (sizes / scale) ** 3], axis=1)
    query = np.array([(target / scale) ** 2, (target / scale) ** 3])
    if method == "ml":
        centre = design.mean(axis=0)
        spread = design.std(axis=0)
        spread = np.where(spread > 0, spread, 1.0)
        z = (design - centre) / spread
        gram = z.T @ z + alpha * np.eye(z.shape[1])
        w = np.linalg.solve(gram, z.T @ (secs - secs.mean()))
        value = secs.mean() + float(((query - centre) / spread) @ w)
    else:
        full = np.hstack([np.ones((len(sizes), 1)), design])
        coef = np.linalg.lstsq(full, secs, rcond=None)[0]
        value = float(coef[0] + query @ coef[1:])
    if not (math.isfinite(value) and value > 0):
        raise ValueError(f"the {method} fit predicts {value} s at N={target}")
    return value


def mlp_features(size: int, cfg: dict) -> list[float]:
    lib = [float(cfg["library"] == code) for code in LIBRARIES]
    shape = [math.log2(cfg.get(k) or 1) for k in ("grid_rows", "grid_cols", "summa_bk")]
    return [math.log2(size)] + lib + shape


def mlp_predictions(model: dict, size: int, cands: list[dict], epochs: int = 3000,
                    seed: int = 3) -> list[float]:
    """Seconds at `size` for each of `cands` from one network trained (on the
    GPU, by libmgx's kernels) on the log medians of every calibrated point."""
    import ctypes
    import numpy as np
    from . import _lib
    rows = [(mlp_features(n, c), math.log(t)) for c in model["candidates"]
            for n, t in points(c["measurements"])]
    if len(rows) < 4:
        raise ValueError("the network needs at least four calibrated points")
    X = np.array([r[0] for r in rows])
    centre = X.mean(axis=0)
    spread = np.where(X.std(axis=0) > 0, X.std(axis=0), 1.0)
    train_x = np.ascontiguousarray((X - centre) / spread, dtype=np.float32)
    train_y = np.ascontiguousarray([r[1] for r in rows], dtype=np.float32)
    dim = train_x.shape[1]
    theta = np.empty(_lib.lib().mgx_mlp_param_count(dim), dtype=np.float32)
    losses = np.empty(epochs, dtype=np.float32)
    _lib.call("mgx_mlp_init", theta.ctypes.data, dim, seed)
    _lib.call("mgx_mlp_train", theta.ctypes.data, dim, train_x.ctypes.data, train_y.ctypes.data,
              len(train_y), epochs, ctypes.c_float(3e-3), ctypes.c_float(1e-5), losses.ctypes.data)
    query = np.ascontiguousarray((np.array([mlp_features(size, c) for c in cands]) - centre)
                                 / spread, dtype=np.float32)
    logs = np.empty(len(cands), dtype=np.float32)
    _lib.call("mgx_mlp_predict", theta.ctypes.data, dim, query.ctypes.data, len(cands),
              logs.ctypes.data)
    return [math.exp(float(v)) for v in logs]


def _usable(model: dict, size: int, ranks: int, decomposition: str, grid) -> list[dict]:
    sig = model["signature"]
    if (sig["ranks"], sig["decomposition"]) != (ranks, decomposition):
        raise ValueError(f"the calibration is for {sig['ranks']} rank(s), {sig['decomposition']}; "
                         f"asked for {ranks} rank(s), {decomposition}")
    out = []
    for c in model["candidates"]:
        cand = Candidate.of(c)
        if grid and (cand.grid_rows, cand.grid_cols) != tuple(grid):
            continue
        if decomposition == "2d" and not cand.divides(size):
            continue
        if decomposition == "1d" and size % ranks:
            continue
        out.append(c)
    return out


def _predict_one(c: dict, size: int, method: str, network_seconds):
    """(seconds, algorithm) for one candidate, or None when the method cannot
    predict it."""
    pts = points(c["measurements"])
    if not pts:
        return None
    if network_seconds is not None:
        return network_seconds, "mlp-gpu"
    if len(pts) >= 4:
        try:
            fitted = physical_fit(pts, size, method)
            return fitted, "ridge-physical" if method == "ml" else "physical-polynomial"
        except (ValueError, ArithmeticError):
            pass
    if method == "extrapolation":
        try:
            return estimate(c["measurements"], size)[0], "piecewise-power-law-fallback"
        except ValueError:
            return None
    return None


def predict_method(model: dict, size: int, ranks: int, decomposition: str, grid, method: str):
    """Candidates that `method` can predict at `size`, fastest first."""
    if size <= 0 or ranks <= 0:
        raise ValueError("size and ranks must be positive")
    usable = _usable(model, size, ranks, decomposition, grid)
    network = mlp_predictions(model, size, usable) if method == "mlp" and usable else None
    ranking = []
    for i, c in enumerate(usable):
        got = _predict_one(c, size, method, None if network is None else network[i])
        if got is None:
            continue
        pts = points(c["measurements"])
        ranking.append({**asdict(Candidate.of(c)), "predicted_s": got[0], "algorithm": got[1],
                        "calibration_sizes": len(pts),
                        "extrapolated": not pts[0][0] <= size <= pts[-1][0]})
    if not ranking:
        raise ValueError(f"method {method}: no calibrated candidate can be predicted at N={size}")
    ranking.sort(key=lambda r: r["predicted_s"])
    return ranking


def _holdout_error(model: dict, method: str) -> list[float]:
    """Relative errors of `method` predicting each candidate's largest
    calibrated size from its other sizes (candidates with >= 5 sizes)."""
    errors = []
    for c in model["candidates"]:
        pts = points(c["measurements"])
        if len(pts) < 5:
            continue
        *train, (n_out, t_out) = pts
        try:
            errors.append(abs(physical_fit(train, n_out, method) - t_out) / t_out)
        except ValueError:
            pass
    return errors


def select_predictions(model, size, ranks, decomposition, grid, method):
    """(ranking, method used); 'auto' resolves to whichever fit has the lower
    median hold-out error ('extrapolation' when neither can be scored)."""
    if method not in METHODS:
        raise ValueError(f"unknown prediction model {method!r}; one of {', '.join(METHODS)}")
    if method == "auto":
        scores = {m: _holdout_error(model, m) for m in ("extrapolation", "ml")}
        scored = [(statistics.median(e), m) for m, e in scores.items() if e]
        method = min(scored)[1] if scored else "extrapolation"
    return predict_method(model, size, ranks, decomposition, grid, method), method


# ---- calibration files ----------------------------------------------------------------------

def load_model(path) -> dict:
    model = json.loads(Path(path).read_text())
    if model.get("schema") != SCHEMA:
        raise ValueError(f"{path}: calibration schema {model.get('schema')!r}, expected {SCHEMA}")
    return model


def save_model(model: dict, path: Path) -> None:
    """Writes atomically; an existing file is kept as <name>.bak first."""
    path.parent.mkdir(parents=True, exist_ok=True)
    if path.exists():
        shutil.copy2(path, path.with_suffix(path.suffix + ".bak"))
    staging = path.with_name(path.name + ".tmp")
    staging.write_text(json.dumps(model, indent=2) + "\n")
    staging.replace(path)


def _median_of(doc: dict, lib: str) -> float | None:
    """The validated, positive, finite median_s of `lib` in a report."""
    for row in doc.get("results") or []:
        if not isinstance(row, dict) or row.get("library") != lib or row.get("validated") is not True:
            continue
        m = row.get("median_s")
        if type(m) in (int, float) and math.isfinite(m) and m > 0:
            return float(m)
    return None


class RunnerJob:
    """One runner invocation for one configuration and size, and the report
    it leaves under <output>/runs."""

    def __init__(self, args, cfg: dict, size: int, output: Path, label: str):
        self.args, self.cfg, self.size, self.output, self.label = args, cfg, size, output, label

    def command(self) -> list[str]:
        a, cfg = self.args, self.cfg
        cmd = [sys.executable, "-m", "mgxbench_amd.runner", "--decomposition", a.decomposition,
               "--size", str(self.size), "--ranks", str(a.ranks), "--libraries", cfg["library"],
               "--warmup", str(a.warmup), "--repeats", str(a.repeats),
               "--build-dir", str(a.build_dir), "--output", str(self.output),
               "--label", self.label]
        if a.launcher:
            cmd += ["--launcher", a.launcher]
        if a.launcher_args:
            cmd += ["--launcher-args", a.launcher_args]
        if a.decomposition == "2d":
            cmd += ["--grid-rows", str(cfg["grid_rows"]), "--grid-cols", str(cfg["grid_cols"]),
                    "--summa-bk", str(cfg["summa_bk"])]
        return cmd

    def _reports(self) -> dict:
        return {p.resolve(): (p.stat().st_mtime_ns, p.stat().st_size)
                for p in (self.output / "runs").glob("*/benchmark.json")}

    def _is_mine(self, doc: dict) -> bool:
        a, cfg = doc.get("command_args") or {}, self.cfg
        libs = a.get("libraries")
        libs = [libs] if isinstance(libs, str) else libs
        same = (a.get("label", doc.get("label")) == self.label and a.get("size") == self.size
                and a.get("ranks") == self.args.ranks
                and a.get("decomposition") == self.args.decomposition and libs == [cfg["library"]])
        if same and self.args.decomposition == "2d":
            same = all(a.get(k) == cfg[k] for k in ("grid_rows", "grid_cols", "summa_bk"))
        return same

    def run(self, cwd=None, capture=True):
        """Runs the runner; returns (completed process, newest matching report
        path or None, its document or None)."""
        seen = self._reports()
        done = subprocess.run(self.command(), cwd=cwd, capture_output=capture, text=True)
        newest = None
        for path, stamp in self._reports().items():
            if seen.get(path) == stamp:
                continue
            try:
                doc = json.loads(path.read_text())
            except (OSError, ValueError):
                continue
            if self._is_mine(doc) and (newest is None or stamp[0] > newest[0]):
                newest = (stamp[0], path, doc)
        return done, (newest[1] if newest else None), (newest[2] if newest else None)


def calibrate(args) -> int:
    root = Path(args.project_root).resolve()
    signature = fingerprint(root, args.decomposition, args.ranks, args.build_dir)
    sizes = sorted(set(args.sizes))
    if len(sizes) < 2 or min(sizes) <= 0:
        raise ValueError("calibration needs two or more distinct positive sizes")
    if args.decomposition == "1d" and any(n % args.ranks for n in sizes):
        raise ValueError("1D calibration sizes must be multiples of --ranks")
    grids = None
    if args.grids:
        grids = [tuple(int(v) for v in g.lower().split("x")) for g in args.grids]
    if args.grid_rows is not None:
        grids = [(args.grid_rows, args.grid_cols)]
    configs = candidate_configs(args.decomposition, args.ranks, sizes, args.libraries, args.bks,
                                grids)
    if not configs:
        raise ValueError("no configuration fits the ranks, grids and sizes given")
    out = resolve_model(args.decomposition, args.ranks, args.model, root)
    partial = out.with_name(out.stem + ".partial.json")
    output = Path(args.output) if Path(args.output).is_absolute() else root / args.output
    logs = output / "calibration_logs"
    logs.mkdir(parents=True, exist_ok=True)
    model = {"schema": SCHEMA, "signature": signature, "sizes": sizes, "candidates": [],
             "warning": ("predictions from short calibration runs are estimates; "
                         "validate a choice with --auto-validate")}
    tally = dict.fromkeys(("attempted", "valid_measurements", "failed_launches",
                           "missing_reports", "rejected_reports", "rejected_results"), 0)
    for index, cfg in enumerate(configs, 1):
        measured = []
        for size in sizes:
            tally["attempted"] += 1
            tag = "1d" if args.decomposition == "1d" else \
                f"grid{cfg['grid_rows']}x{cfg['grid_cols']}_bk{cfg['summa_bk']}"
            stem = logs / f"{index:03d}_{cfg['library']}_N{size}_{tag}"
            print(f"calibration {index}/{len(configs)}: {cfg} at N={size}", flush=True)
            job = RunnerJob(args, cfg, size, output, CALIBRATION_LABEL)
            stem.with_suffix(".command.txt").write_text(shlex.join(job.command()) + "\n")
            try:
                done, path, doc = job.run(cwd=root)
            except OSError as exc:
                tally["failed_launches"] += 1
                stem.with_suffix(".stderr").write_text(f"{exc}\n")
                continue
            stem.with_suffix(".stdout").write_text(done.stdout)
            stem.with_suffix(".stderr").write_text(done.stderr)
            if done.returncode != 0:
                tally["failed_launches"] += 1
                print(f"  runner exited {done.returncode}; see {stem}.*", file=sys.stderr)
                continue
            if doc is None:
                tally["missing_reports"] += 1
                continue
            if doc.get("failures"):
                tally["rejected_reports"] += 1
                continue
            median = _median_of(doc, cfg["library"])
            if median is None:
                tally["rejected_results"] += 1
                continue
            measured.append({"size": size, "median_s": median})
            tally["valid_measurements"] += 1
            print(f"  median {median:.6f} s ({path})", flush=True)
        if len({m["size"] for m in measured}) >= 2:
            model["candidates"].append({**cfg, "measurements": measured})
        else:
            print(f"  {cfg['library']}: fewer than two measured sizes, left out", file=sys.stderr)
        partial.parent.mkdir(parents=True, exist_ok=True)
        partial.write_text(json.dumps(model, indent=2) + "\n")
    print(f"calibration: {tally}; {len(model['candidates'])} of {len(configs)} configurations "
          "kept", flush=True)
    if not model["candidates"]:
        raise ValueError(f"no configuration has two measured sizes (partial file {partial}, "
                         f"logs {logs})")
    save_model(model, out)
    partial.unlink(missing_ok=True)
    print(f"calibration written to {out}")
    return 0


def _recovered_entry(doc: dict, args):
    """(candidate key, size, median) of one historical calibration report, or
    the reason it is rejected."""
    a = doc.get("command_args") or {}
    if not isinstance(a, dict) or a.get("label", doc.get("label")) != CALIBRATION_LABEL:
        return "not written by a calibration run"
    try:
        size, ranks = int(a["size"]), int(a["ranks"])
    except (KeyError, TypeError, ValueError):
        return "size or ranks missing"
    if (ranks, a.get("decomposition")) != (args.ranks, args.decomposition):
        return "other ranks or decomposition"
    if size <= 0 or (args.decomposition == "1d" and size % ranks):
        return "size does not fit the decomposition"
    libs = a.get("libraries")
    libs = [libs] if isinstance(libs, str) else libs
    if not isinstance(libs, list) or len(libs) != 1 or libs[0] not in LIBRARIES:
        return "not exactly one known library"
    shape = (None, None, None)
    if args.decomposition == "2d":
        try:
            shape = tuple(int(a[k]) for k in ("grid_rows", "grid_cols", "summa_bk"))
        except (KeyError, TypeError, ValueError):
            return "grid or panel width missing"
        rows, cols, bk = shape
        if min(shape) <= 0 or rows * cols != ranks or size % rows or size % cols:
            return "grid or panel width does not fit"
    if doc.get("failures"):
        return "the report lists failures"
    median = _median_of(doc, libs[0])
    if median is None:
        return "no validated positive median"
    return (libs[0], *shape), size, median


def recover_calibration(args) -> int:
    """A calibration file rebuilt from reports of earlier calibration runs.
    The reports cannot prove which GPU or executable produced them, so the
    caller must vouch with --assume-compatible."""
    root = Path(args.project_root).resolve()
    source = Path(args.input) if Path(args.input).is_absolute() else root / args.input
    reports = sorted(source.glob("*/benchmark.json"))
    if not reports:
        raise ValueError(f"{source} holds no */benchmark.json")
    if not args.assume_compatible:
        raise ValueError("old reports carry no proof of the GPU or executable they ran on; check "
                         "them and pass --assume-compatible")
    signature = fingerprint(root, args.decomposition, args.ranks, args.build_dir)
    by_key: dict[tuple, dict[int, list[float]]] = {}
    rejected: dict[str, int] = {}
    for path in reports:
        try:
            entry = _recovered_entry(json.loads(path.read_text()), args)
        except (OSError, ValueError) as exc:
            entry = f"unreadable ({type(exc).__name__})"
        if isinstance(entry, str):
            rejected[entry] = rejected.get(entry, 0) + 1
            continue
        k, size, median = entry
        by_key.setdefault(k, {}).setdefault(size, []).append(median)
    candidates = []
    for (lib, rows, cols, bk), per_size in sorted(by_key.items(), key=lambda kv: str(kv[0])):
        if len(per_size) < 2:
            rejected[f"{lib} {rows}x{cols} BK {bk}: one size only"] = 1
            continue
        candidates.append({"library": lib, "grid_rows": rows, "grid_cols": cols,
                           "summa_bk": bk,
                           "measurements": [{"size": n, "median_s": statistics.median(ts)}
                                            for n, ts in sorted(per_size.items())]})
    print(f"{len(reports)} report(s) read, {len(candidates)} configuration(s) recovered", flush=True)
    for why, count in sorted(rejected.items()):
        print(f"  left out ({count}): {why}", flush=True)
    if not candidates:
        raise ValueError("no configuration has validated medians at two sizes")
    model = {"schema": SCHEMA, "signature": signature,
             "sizes": sorted({m["size"] for c in candidates for m in c["measurements"]}),
             "candidates": candidates,
             "warning": ("rebuilt from earlier calibration reports vouched for with "
                         "--assume-compatible; predictions are estimates")}
    out = resolve_model(args.decomposition, args.ranks, args.model, root)
    save_model(model, out)
    print(f"recovered calibration written to {out}", flush=True)
    return 0


# ---- AUTO -----------------------------------------------------------------------------------

def validate_and_feedback(args, best: dict, report: dict, model: dict, model_path: Path) -> int:
    """Times the chosen configuration through the runner; with feedback its
    validated median joins that candidate's calibration (old file kept)."""
    job = RunnerJob(args, best, args.size, Path(args.output) / "validation", VALIDATION_LABEL)
    print("AUTO validation: " + shlex.join(job.command()), flush=True)
    done, path, doc = job.run(capture=False)
    if done.returncode != 0:
        report["validation"] = {"status": "FAILED", "exit_code": done.returncode}
        return 2
    measured = None if doc is None or doc.get("failures") else _median_of(doc, best["library"])
    if measured is None:
        report["validation"] = {"status": "FAILED",
                                "reason": "the runner left no clean report for this choice"}
        return 2
    predicted = best["predicted_s"]
    error = abs(measured - predicted)
    report["validation"] = {"status": "PASS", "median_s": measured, "predicted_s": predicted,
                            "absolute_error_s": error, "relative_error": error / measured,
                            "benchmark_report": str(path)}
    print(f"AUTO validation: predicted {predicted:.6f} s, measured {measured:.6f} s "
          f"({100 * error / measured:.2f}% off)", flush=True)
    if not args.auto_feedback:
        return 0
    target = next((c for c in model["candidates"] if key(c) == key(best)), None)
    if target is None:
        report["feedback"] = {"status": "SKIPPED", "reason": "choice not in the calibration"}
        return 0
    target["measurements"].append({"size": args.size, "median_s": measured, "source": str(path),
                                   "timestamp_utc": _now()})
    model["sizes"] = sorted(set(model.get("sizes", [])) | {args.size})
    save_model(model, model_path)
    report["feedback"] = {"status": "SAVED", "model": str(model_path),
                          "backup": str(model_path.with_suffix(model_path.suffix + ".bak"))}
    return 0


def _print_choice(args, report: dict, out: Path) -> None:
    best, speedup = report["best"], report["speedup_vs_MPI_predicted"]
    print(f"\nAUTO ({report['prediction_model']}): N={args.size}, {args.ranks} GPU(s), "
          f"{args.decomposition.upper()}")
    shape = "" if args.decomposition == "1d" else \
        f", grid {best['grid_rows']}x{best['grid_cols']}, BK {best['summa_bk']}"
    print(f"  choice: {best['library']} ({LIBRARY_NAMES[best['library']]}){shape}")
    print(f"  predicted {best['predicted_s']:.6f} s"
          + (f", {speedup:.2f}x the MMM prediction" if speedup else "")
          + (" (outside the calibrated sizes)" if best["extrapolated"] else ""))
    print(f"  status {report['status']}; report {out}")
    if "validation" in report:
        print(f"  validation {report['validation'].get('status')}")
    if "feedback" in report:
        print(f"  feedback {report['feedback']['status']}")


def auto(args) -> int:
    """--libraries AUTO: the calibrated candidates ranked at the requested size."""
    path = resolve_model(args.decomposition, args.ranks, getattr(args, "auto_model", None),
                         Path.cwd())
    if not path.is_file():
        print(f"AUTO: no calibration at {path}; create one with "
              "`python -m mgxbench_amd.predictive calibrate`", file=sys.stderr)
        return 2
    try:
        model = load_model(path)
        here = fingerprint(Path.cwd(), args.decomposition, args.ranks, args.build_dir)
        if not compatible_signature(model["signature"], here):
            raise ValueError(f"{path} was calibrated for other GPUs, another executable or "
                             "another allocation; calibrate again")
        grid = (args.grid_rows, args.grid_cols) if args.grid_rows and args.grid_cols else None
        ranking, method = select_predictions(model, args.size, args.ranks, args.decomposition,
                                             grid, getattr(args, "prediction_model",
                                                           "extrapolation"))
    except (ValueError, ImportError, OSError) as exc:
        print(f"AUTO: {exc}", file=sys.stderr)
        return 2
    best = ranking[0]
    same_shape = [r for r in ranking if r["library"] == "MMM"
                  and key(r)[1:] == key(best)[1:]]
    speedup = same_shape[0]["predicted_s"] / best["predicted_s"] if same_shape else None
    report = {"size": args.size, "ranks": args.ranks, "decomposition": args.decomposition,
              "architecture_id": here["id"], "best": best, "ranking": ranking,
              "speedup_vs_MPI_predicted": speedup, "prediction_model": method,
              "calibration_file": str(path), "status": "PREDICTED", "measured": False}
    status = 0
    if getattr(args, "auto_validate", False):
        status = validate_and_feedback(args, best, report, model, path)
        report["status"] = "VALIDATED" if status == 0 else "VALIDATION_FAILED"
    out = Path(args.output) / "autotuning_prediction.json"
    out.parent.mkdir(parents=True, exist_ok=True)
    out.write_text(json.dumps(report, indent=2) + "\n")
    _print_choice(args, report, out)
    return status


# ---- command line ---------------------------------------------------------------------------

def _grid_problem(args) -> str | None:
    if args.ranks <= 0:
        return "--ranks must be positive"
    rows, cols = getattr(args, "grid_rows", None), getattr(args, "grid_cols", None)
    if (rows is None) != (cols is None):
        return "give --grid-rows and --grid-cols together"
    if rows is not None:
        if args.decomposition != "2d":
            return "a process grid applies to --decomposition 2d only"
        if min(rows, cols) <= 0 or rows * cols != args.ranks:
            return f"a {rows}x{cols} grid does not hold {args.ranks} ranks"
        if args.command == "calibrate" and args.grids:
            return "give either --grids or --grid-rows/--grid-cols"
        if args.command == "auto" and (args.size % rows or args.size % cols):
            return f"N={args.size} does not split over a {rows}x{cols} grid"
    for g in (getattr(args, "grids", None) or []):
        try:
            r, c = (int(v) for v in g.lower().split("x"))
        except ValueError:
            return f"--grids entry {g!r} is not RxC"
        if min(r, c) <= 0 or r * c != args.ranks:
            return f"--grids entry {g!r} does not hold {args.ranks} ranks"
    return None


def build_parser() -> argparse.ArgumentParser:
    launcher = f"{shlex.quote(sys.executable)} -m mgxbench_amd.launch"
    top = argparse.ArgumentParser(description="calibrated choice of MGXBench run configurations")
    sub = top.add_subparsers(dest="command", required=True)

    def common(p, decomposition_required=True):
        p.add_argument("--decomposition", choices=("1d", "2d"), required=decomposition_required)
        p.add_argument("--ranks", type=int, required=True)

    cal = sub.add_parser("calibrate", help="time every configuration at a few sizes")
    cal.add_argument("--project-root", default=".")
    common(cal)
    cal.add_argument("--sizes", type=int, nargs="+", default=[2048, 4096, 8192])
    cal.add_argument("--libraries", nargs="+", choices=LIBRARIES, default=list(LIBRARIES))
    cal.add_argument("--bks", type=int, nargs="+", default=[128, 256], help="2D panel widths")
    cal.add_argument("--grids", nargs="+", default=None, help="2D grids, e.g. 2x4 4x2")
    cal.add_argument("--grid-rows", type=int, default=None)
    cal.add_argument("--grid-cols", type=int, default=None)
    cal.add_argument("--warmup", type=int, default=1)
    cal.add_argument("--repeats", type=int, default=3)
    cal.add_argument("--launcher", default=launcher)
    cal.add_argument("--launcher-args", default="")
    cal.add_argument("--build-dir", default=str(PKG / "bin"))
    cal.add_argument("--output", default="results/auto_calibration")
    cal.add_argument("--model", default=None, help="calibration file to write")

    rec = sub.add_parser("recover", help="rebuild a calibration file from earlier reports")
    rec.add_argument("--project-root", default=".")
    common(rec)
    rec.add_argument("--input", default="results/auto_calibration/runs")
    rec.add_argument("--model", default=None)
    rec.add_argument("--build-dir", default=str(PKG / "bin"))
    rec.add_argument("--assume-compatible", action="store_true",
                     help="vouch that the reports come from this GPU and executable")

    au = sub.add_parser("auto", help="the fastest calibrated configuration at one size")
    common(au)
    au.add_argument("--size", type=int, required=True)
    au.add_argument("--grid-rows", type=int, default=None)
    au.add_argument("--grid-cols", type=int, default=None)
    au.add_argument("--prediction-model", choices=METHODS, default="extrapolation")
    au.add_argument("--build-dir", default=str(PKG / "bin"))
    au.add_argument("--auto-model", default=None)
    au.add_argument("--auto-validate", action="store_true")
    au.add_argument("--auto-feedback", action="store_true")
    au.add_argument("--warmup", type=int, default=2)
    au.add_argument("--repeats", type=int, default=5)
    au.add_argument("--launcher", default=launcher)
    au.add_argument("--launcher-args", default="")
    au.add_argument("--timeout", type=int, default=600)
    au.add_argument("--output", default="results/auto_prediction")
    return top


def main(argv=None) -> int:
    parser = build_parser()
    args = parser.parse_args(argv)
    problem = _grid_problem(args)
    if args.command == "auto" and problem is None:
        if args.auto_feedback and not args.auto_validate:
            problem = "--auto-feedback only makes sense with --auto-validate"
        elif args.size <= 0:
            problem = "--size must be positive"
    if problem:
        parser.error(problem)
    handlers = {"calibrate": calibrate, "recover": recover_calibration, "auto": auto}
    try:
        return handlers[args.command](args)
    except ValueError as exc:
        print(f"{args.command}: {exc}", file=sys.stderr)
        return 2


if __name__ == "__main__":
    sys.exit(main())
