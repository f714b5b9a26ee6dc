"""Calibrated AUTO selection for run configurations (the role of the
reference's autotuning/predictive.py, built around its own object model).

`--libraries AUTO` chooses, for one matrix size, how a whole run is
configured -- the library code and, in 2D, the process grid and panel width
-- from short runs measured beforehand.  The pieces:

  Shape        one configuration of a run (library, grid rows / cols, panel).
  Curve        the calibrated (size, median seconds) points of one Shape.
  fit chains   each prediction method is an ordered chain of fits; a
               Shape's forecast is the first fit of the chain that applies
               to its Curve (PowerLaw, Cubic, Network).
  Calibration  the schema-1 calibration file and the Curves it holds.
  Calibrator   times every Shape at every calibration size through the runner.
  AutoSession  ranks the Shapes at the requested size, optionally times the
               winner (validation) and feeds the measurement back.

Kept from the reference because callers depend on it: the calibrate /
recover / auto command lines (predictive.py:604-659), the schema-1 file
{schema, signature, sizes, candidates: [{library, grid_rows, grid_cols,
summa_bk, measurements: [{size, median_s}]}], warning} (predictive.py:281-298,
367), the compatibility fields of the signature, the ranking and report keys
of autotuning_prediction.json, and the fit formulas (predictive.py:98-233):
a power law between the two calibrated sizes around the target (the two end
sizes outside the range), exponent clipped to [0, 4]; with four or more
sizes t = c0 + c1 x^2 + c2 x^3 (x = N / largest calibrated N) by least
squares, or by ridge (alpha 0.1) on the standardised (x^2, x^3) with a free
intercept; 'auto' takes the fit with the lower median error on each curve's
held-out largest size.  This build adds the 'mlp' method: one 3-layer
width-64 FP32 network over all curves, trained on the GPU by libmgx.

Hardware identity comes from rocm-smi, runs go through
`python -m mgxbench_amd.runner` with the one-rank launcher, and calibration
files live in mgxbench_amd/models/.

    python -m mgxbench_amd.predictive calibrate --decomposition 2d --ranks 1 --sizes 2048 4096 8192
    python -m mgxbench_amd.predictive auto --decomposition 2d --ranks 1 --size 16384
"""
from __future__ import annotations

import argparse
import bisect
import collections
import csv
import datetime
import hashlib
import io
import itertools
import json
import math
import os
import shlex
import shutil
import statistics
import subprocess
import sys
from dataclasses import dataclass, field
from pathlib import Path

from .runner import LIBRARY_NAMES

PKG = Path(__file__).resolve().parent
MODELS_DIR = PKG / "models"
SCHEMA = 1
METHODS = ("extrapolation", "ml", "auto", "mlp")
CALIBRATION_LABEL = "auto-calibration"
VALIDATION_LABEL = "auto-validation"
# signature fields that must agree for a calibration to be reused
BINDING_FIELDS = ("gpu_models", "ranks", "decomposition", "binary_sha256", "slurm_nodes")


class PredictionError(ValueError):
    """A calibration cannot answer the question asked of it."""


# ---- where the calibration runs ----------------------------------------------------------------

def _rocm_smi(args: list[str]) -> str:
    tool = shutil.which("rocm-smi")
    if tool is None:
        raise PredictionError("rocm-smi is not on PATH: the GPU model cannot be identified")
    try:
        done = subprocess.run([tool, *args], capture_output=True, text=True, timeout=15)
    except (OSError, subprocess.TimeoutExpired) as exc:
        raise PredictionError(f"rocm-smi {' '.join(args)} did not answer: {exc}") from None
    if done.returncode != 0:
        raise PredictionError(f"rocm-smi {' '.join(args)} exited {done.returncode}")
    return done.stdout


@dataclass(frozen=True)
class Platform:
    """GPU model names and driver versions as rocm-smi reports them.  A model
    name is the card series when one is reported, else the PCI model id,
    followed by the gfx target ("0x75a3 gfx950" on an MI355X)."""
    gpu_models: tuple[str, ...]
    driver_versions: tuple[str, ...]

    @classmethod
    def probe(cls) -> "Platform":
        table = _rocm_smi(["--showproductname", "--csv"])
        rows = csv.DictReader(io.StringIO("\n".join(l for l in table.splitlines() if "," in l)))
        names = set()
        for row in rows:
            if not (row.get("device") or "").lower().startswith("card"):
                continue
            said = [(row.get(c) or "").strip() for c in ("Card Series", "Card Model", "GFX Version")]
            said = [s for s in said if s and s.upper() != "N/A"]
            if said:
                names.add(said[0] if len(said) == 1 else f"{said[0]} {said[-1]}")
        if not names:
            raise PredictionError("rocm-smi lists no GPU")
        versions = set()
        for line in _rocm_smi(["--showdriverversion", "--csv"]).splitlines()[1:]:
            value = line.rpartition(",")[2].strip().strip('"')
            if "," in line and value:
                versions.add(value)
        return cls(tuple(sorted(names)), tuple(sorted(versions)))


def executable(root, build_dir, decomposition: str) -> Path:
    exe = Path(root) / build_dir / f"mm_{decomposition}"
    if not exe.is_file():
        raise PredictionError(f"{exe} does not exist; build it with `python -m mgxbench_amd.build`")
    return exe


def make_signature(exe: Path, decomposition: str, ranks: int, platform: Platform) -> dict:
    """What a calibration is valid for.  `id` hashes the binding fields;
    the other fields are informational."""
    binding = {"gpu_models": list(platform.gpu_models), "ranks": ranks,
               "decomposition": decomposition,
               "binary_sha256": hashlib.sha256(exe.read_bytes()).hexdigest(),
               "slurm_nodes": os.environ.get("SLURM_NNODES", "1")}
    tag = hashlib.sha256(json.dumps(binding, sort_keys=True).encode()).hexdigest()[:16]
    env = os.environ
    return dict(binding, driver_versions=list(platform.driver_versions),
                visible_devices=env.get("HIP_VISIBLE_DEVICES", env.get("ROCR_VISIBLE_DEVICES", "")),
                gpus_per_node=env.get("SLURM_GPUS_ON_NODE", ""), id=tag)


def signatures_agree(stored: dict, current: dict) -> bool:
    """True when every binding field matches; GPU names compare by the part
    before any comma, as a set."""
    def models(sig):
        return {str(v).partition(",")[0].strip() for v in sig.get("gpu_models", [])}
    if models(stored) != models(current):
        return False
    return all(str(stored.get(f)) == str(current.get(f)) for f in BINDING_FIELDS[1:])


def default_calibration_path(decomposition: str, ranks: int) -> Path:
    """mgxbench_amd/models/calibration-<d>-<p>gpu.json, or an older single
    calibration.json when only that one exists."""
    per_run = MODELS_DIR / f"calibration-{decomposition}-{ranks}gpu.json"
    older = MODELS_DIR / "calibration.json"
    return older if older.exists() and not per_run.exists() else per_run


def calibration_path(decomposition: str, ranks: int, given=None, base: Path | None = None) -> Path:
    if not given:
        return default_calibration_path(decomposition, ranks)
    p = Path(given)
    return p if p.is_absolute() or base is None else base / p


# ---- configurations ------------------------------------------------------------------------------

@dataclass(frozen=True)
class Shape:
    """One run configuration.  1D shapes carry only the library code."""
    library: str
    grid_rows: int | None = None
    grid_cols: int | None = None
    summa_bk: int | None = None

    @classmethod
    def of(cls, entry: dict) -> "Shape":
        return cls(entry["library"], entry.get("grid_rows"), entry.get("grid_cols"),
                   entry.get("summa_bk"))

    def fields(self) -> dict:
        return {"library": self.library, "grid_rows": self.grid_rows,
                "grid_cols": self.grid_cols, "summa_bk": self.summa_bk}

    def splits(self, size: int) -> bool:
        """The process grid divides an N x N matrix into whole blocks."""
        return self.grid_rows is None or not (size % self.grid_rows or size % self.grid_cols)

    def runner_flags(self) -> list[str]:
        if self.grid_rows is None:
            return []
        return ["--grid-rows", str(self.grid_rows), "--grid-cols", str(self.grid_cols),
                "--summa-bk", str(self.summa_bk)]

    def tag(self) -> str:
        return "1d" if self.grid_rows is None else \
            f"grid{self.grid_rows}x{self.grid_cols}_bk{self.summa_bk}"

    @staticmethod
    def grids_of(ranks: int) -> list[tuple[int, int]]:
        """rows x cols factorisations of `ranks`, rows ascending."""
        return [(r, ranks // r) for r in range(1, ranks + 1) if ranks % r == 0]

    @classmethod
    def space(cls, decomposition: str, ranks: int, sizes, libraries, panels,
              grids=None) -> list["Shape"]:
        """Every configuration worth calibrating: each library in 1D; in 2D
        each library on each positive panel width and each grid that holds
        `ranks` processes and splits every calibration size."""
        if decomposition == "1d":
            return [cls(lib) for lib in libraries]
        usable = [(r, c) for r, c in (cls.grids_of(ranks) if grids is None else grids)
                  if r * c == ranks and all(n % r == 0 and n % c == 0 for n in sizes)]
        widths = [w for w in panels if w > 0]
        return [cls(lib, r, c, w) for (r, c), w, lib in itertools.product(usable, widths, libraries)]


# ---- timing curves and the fits over them ------------------------------------------------------

class NotApplicable(Exception):
    """A fit cannot forecast this curve at this size."""


class Curve:
    """Calibrated points of one Shape: the median of the valid samples at
    each size (positive sizes, positive finite times), sizes ascending."""

    def __init__(self, samples):
        per_size = collections.defaultdict(list)
        for s in samples:
            n, t = int(s["size"]), float(s["median_s"])
            if n > 0 and t > 0 and math.isfinite(t):
                per_size[n].append(t)
        self.sizes = sorted(per_size)
        self.seconds = [statistics.median(per_size[n]) for n in self.sizes]

    def __len__(self) -> int:
        return len(self.sizes)

    def covers(self, size: int) -> bool:
        return bool(self.sizes) and self.sizes[0] <= size <= self.sizes[-1]

    def without_largest(self) -> tuple["Curve", int, float]:
        rest = Curve([])
        rest.sizes, rest.seconds = self.sizes[:-1], self.seconds[:-1]
        return rest, self.sizes[-1], self.seconds[-1]


class PowerLaw:
    """t = t0 (N / n0)^p through the two calibrated points bracketing N (the
    two first or last points outside the calibrated range), p in [0, 4]."""
    name = "piecewise-power-law-fallback"

    def forecast(self, curve: Curve, size: int, shape: Shape | None = None) -> float:
        if len(curve) < 2:
            raise NotApplicable("a power law needs two calibrated sizes")
        at = bisect.bisect_left(curve.sizes, size)
        if at < len(curve) and curve.sizes[at] == size:
            return curve.seconds[at]
        lo = min(max(at - 1, 0), len(curve) - 2)
        n0, n1 = curve.sizes[lo], curve.sizes[lo + 1]
        t0, t1 = curve.seconds[lo], curve.seconds[lo + 1]
        p = math.log(t1 / t0) / math.log(n1 / n0)
        return t0 * (size / n0) ** min(4.0, max(0.0, p))


class Cubic:
    """t = c0 + c1 x^2 + c2 x^3, x = N / (largest calibrated N), over four or
    more calibrated sizes.  Least squares, or with `ridge` the ridge solution
    on standardised features (population standard deviation, 1 where it is
    0) with an unpenalised intercept: w = (Z'Z + alpha I)^-1 Z'(t - mean t)."""

    def __init__(self, ridge: bool, alpha: float = 0.1):
        self.ridge, self.alpha = ridge, alpha
        self.name = "ridge-physical" if ridge else "physical-polynomial"

    def forecast(self, curve: Curve, size: int, shape: Shape | None = None) -> float:
        import numpy as np
        if len(curve) < 4:
            raise NotApplicable("the x^2 / x^3 fits need four calibrated sizes")
        top = float(curve.sizes[-1])
        x = np.asarray(curve.sizes, dtype=float) / top
        feats = np.column_stack([x * x, x * x * x])
        q = np.array([(size / top) ** 2, (size / top) ** 3])
        t = np.asarray(curve.seconds, dtype=float)
        if self.ridge:
            mu, sd = feats.mean(axis=0), feats.std(axis=0)
            sd[sd == 0] = 1.0
            z = (feats - mu) / sd
            w = np.linalg.solve(z.T @ z + self.alpha * np.eye(2), z.T @ (t - t.mean()))
            value = float(t.mean() + ((q - mu) / sd) @ w)
        else:
            coef, *_ = np.linalg.lstsq(np.column_stack([np.ones_like(x), feats]), t, rcond=None)
            value = float(coef[0] + q @ coef[1:])
        if not math.isfinite(value) or value <= 0:
            raise NotApplicable(f"{self.name} gives {value} s at N={size}")
        return value


class Network:
    """One MLP over every curve of a calibration: features (log2 N, library
    one-hot, log2 of grid rows / cols / panel), target log seconds; trained
    on the GPU by libmgx's forward / backward / Adam kernels, evaluated once
    for all shapes at the requested size."""
    name = "mlp-gpu"

    def __init__(self, epochs: int = 3000, seed: int = 3):
        self.epochs, self.seed = epochs, seed
        self._answers: dict[Shape, float] = {}

    @staticmethod
    def features(size: int, shape: Shape) -> list[float]:
        onehot = [1.0 if shape.library == code else 0.0 for code in LIBRARY_NAMES]
        dims = [math.log2(v or 1) for v in (shape.grid_rows, shape.grid_cols, shape.summa_bk)]
        return [math.log2(size)] + onehot + dims

    def prepare(self, curves: dict[Shape, Curve], size: int) -> None:
        import ctypes
        import numpy as np
        from . import _lib
        xs, ys = [], []
        for shape, curve in curves.items():
            for n, t in zip(curve.sizes, curve.seconds):
                xs.append(self.features(n, shape))
                ys.append(math.log(t))
        if len(ys) < 4:
            raise PredictionError("the network needs four calibrated points or more")
        X = np.array(xs)
        mu, sd = X.mean(axis=0), X.std(axis=0)
        sd[sd == 0] = 1.0
        train_x = np.ascontiguousarray((X - mu) / sd, dtype=np.float32)
        train_y = np.ascontiguousarray(ys, dtype=np.float32)
        dim = train_x.shape[1]
        lib = _lib.lib()
        theta = np.empty(lib.mgx_mlp_param_count(dim), dtype=np.float32)
        losses = np.empty(self.epochs, dtype=np.float32)
        _lib.call("mgx_mlp_init", theta.ctypes.data, dim, self.seed)
        _lib.call("mgx_mlp_train", theta.ctypes.data, dim, train_x.ctypes.data,
                  train_y.ctypes.data, len(ys), self.epochs, ctypes.c_float(3e-3),
                  ctypes.c_float(1e-5), losses.ctypes.data)
        order = list(curves)
        query = np.ascontiguousarray(
            (np.array([self.features(size, s) for s in order]) - mu) / sd, dtype=np.float32)
        logs = np.empty(len(order), dtype=np.float32)
        _lib.call("mgx_mlp_predict", theta.ctypes.data, dim, query.ctypes.data, len(order),
                  logs.ctypes.data)
        self._answers = {s: math.exp(float(v)) for s, v in zip(order, logs)}

    def forecast(self, curve: Curve, size: int, shape: Shape | None = None) -> float:
        if shape not in self._answers:
            raise NotApplicable("not part of the trained network's input")
        return self._answers[shape]


def fit_chain(method: str) -> list:
    """The fits a method tries, in order.  'extrapolation' falls back to the
    power law when the cubic does not apply; 'ml' has no fallback."""
    chains = {"extrapolation": lambda: [Cubic(ridge=False), PowerLaw()],
              "ml": lambda: [Cubic(ridge=True)],
              "mlp": lambda: [Network()]}
    if method not in chains:
        raise PredictionError(f"no prediction method {method!r} (choose from {', '.join(METHODS)})")
    return chains[method]()


# ---- the calibration file ----------------------------------------------------------------------

class Calibration:
    """A schema-1 calibration document and the Shapes / Curves it holds."""

    def __init__(self, doc: dict):
        if doc.get("schema") != SCHEMA:
            raise PredictionError(f"calibration schema {doc.get('schema')!r}; this reader "
                                  f"knows schema {SCHEMA}")
        self.doc = doc

    @classmethod
    def empty(cls, signature: dict, sizes, note: str) -> "Calibration":
        return cls({"schema": SCHEMA, "signature": signature, "sizes": sorted(sizes),
                    "candidates": [], "warning": note})

    @classmethod
    def read(cls, path) -> "Calibration":
        return cls(json.loads(Path(path).read_text()))

    def write(self, path: Path) -> Path | None:
        """Atomic replace; a file already there is kept as <name>.bak, whose
        path is returned."""
        path.parent.mkdir(parents=True, exist_ok=True)
        backup = None
        if path.exists():
            backup = path.with_suffix(path.suffix + ".bak")
            shutil.copy2(path, backup)
        staging = path.with_name(path.name + ".tmp")
        staging.write_text(json.dumps(self.doc, indent=2) + "\n")
        staging.replace(path)
        return backup

    @property
    def signature(self) -> dict:
        return self.doc["signature"]

    def add_shape(self, shape: Shape, samples: list[dict]) -> None:
        self.doc["candidates"].append(dict(shape.fields(), measurements=samples))

    def entries(self):
        for entry in self.doc["candidates"]:
            yield Shape.of(entry), entry

    def append_sample(self, shape: Shape, sample: dict) -> bool:
        for s, entry in self.entries():
            if s == shape:
                entry["measurements"].append(sample)
                self.doc["sizes"] = sorted(set(self.doc.get("sizes", [])) | {sample["size"]})
                return True
        return False

    def curves_for(self, size: int, ranks: int, decomposition: str, grid=None) -> dict[Shape, Curve]:
        """Curves of the shapes that can run an N = size problem on this
        rank count (optionally only on one grid)."""
        sig = self.signature
        if (sig.get("ranks"), sig.get("decomposition")) != (ranks, decomposition):
            raise PredictionError(f"calibrated for {sig.get('ranks')} rank(s) "
                                  f"{sig.get('decomposition')}, asked about {ranks} rank(s) "
                                  f"{decomposition}")
        if decomposition == "1d" and size % ranks:
            return {}
        out = {}
        for shape, entry in self.entries():
            if grid and (shape.grid_rows, shape.grid_cols) != tuple(grid):
                continue
            if decomposition == "2d" and not shape.splits(size):
                continue
            out[shape] = Curve(entry["measurements"])
        return out

    def holdout_score(self, fit) -> list[float]:
        """Relative errors of `fit` forecasting each curve's largest size from
        its smaller ones (curves of five sizes or more)."""
        errs = []
        for _, entry in self.entries():
            curve = Curve(entry["measurements"])
            if len(curve) < 5:
                continue
            rest, n, t = curve.without_largest()
            try:
                errs.append(abs(fit.forecast(rest, n) - t) / t)
            except (NotApplicable, ArithmeticError):
                continue
        return errs


def forecast(cal: Calibration, size: int, ranks: int, decomposition: str, grid=None,
             method: str = "extrapolation") -> tuple[list[dict], str]:
    """Ranking of the calibrated shapes at `size`, fastest first, and the
    method that made it ('auto' resolves to 'extrapolation' or 'ml' by the
    lower median hold-out error; 'extrapolation' when neither is scored)."""
    if size <= 0 or ranks <= 0:
        raise PredictionError("size and ranks must be positive")
    if method == "auto":
        scored = []
        for name in ("extrapolation", "ml"):
            errs = cal.holdout_score(fit_chain(name)[0])
            if errs:
                scored.append((statistics.median(errs), name))
        method = min(scored)[1] if scored else "extrapolation"
    chain = fit_chain(method)
    curves = cal.curves_for(size, ranks, decomposition, grid)
    usable = {s: c for s, c in curves.items() if len(c)}
    for fit in chain:
        if isinstance(fit, Network) and usable:
            fit.prepare(usable, size)
    ranking = []
    for shape, curve in usable.items():
        for fit in chain:
            try:
                seconds = fit.forecast(curve, size, shape)
            except (NotApplicable, ArithmeticError):
                continue
            ranking.append(dict(shape.fields(), predicted_s=seconds, algorithm=fit.name,
                                calibration_sizes=len(curve), extrapolated=not curve.covers(size)))
            break
    if not ranking:
        raise PredictionError(f"{method}: none of the calibrated configurations can be "
                              f"forecast at N={size}")
    ranking.sort(key=lambda r: r["predicted_s"])
    return ranking, method


# ---- runner reports ----------------------------------------------------------------------------

def validated_median(doc: dict, library: str) -> float | None:
    """The validated, positive, finite median_s of `library` in a runner
    report (None when there is none)."""
    for row in doc.get("results") or []:
        if isinstance(row, dict) and row.get("library") == library and row.get("validated") is True:
            m = row.get("median_s")
            if type(m) in (int, float) and math.isfinite(m) and m > 0:
                return float(m)
    return None


def report_request(doc: dict) -> dict:
    req = doc.get("command_args") or {}
    if not isinstance(req, dict):
        return {}
    libs = req.get("libraries")
    return dict(req, libraries=[libs] if isinstance(libs, str) else libs,
                label=req.get("label", doc.get("label")))


@dataclass
class RunnerCall:
    """One runner invocation (one shape, one size, one label) and the
    benchmark.json it writes under <output>/runs."""
    args: argparse.Namespace
    shape: Shape
    size: int
    output: Path
    label: str
    timeout: int | None = None

    def argv(self) -> list[str]:
        a = self.args
        cmd = [sys.executable, "-m", "mgxbench_amd.runner", "--decomposition", a.decomposition,
               "--size", str(self.size), "--ranks", str(a.ranks),
               "--libraries", self.shape.library, "--warmup", str(a.warmup),
               "--repeats", str(a.repeats), "--build-dir", str(a.build_dir),
               "--output", str(self.output), "--label", self.label]
        if a.launcher:
            cmd += ["--launcher", a.launcher]
        if a.launcher_args:
            cmd += ["--launcher-args", a.launcher_args]
        if self.timeout:
            cmd += ["--timeout", str(self.timeout)]
        return cmd + self.shape.runner_flags()

    def _stamps(self) -> dict:
        stamps = {}
        for p in (self.output / "runs").glob("*/benchmark.json"):
            st = p.stat()
            stamps[p.resolve()] = (st.st_mtime_ns, st.st_size)
        return stamps

    def answers_me(self, doc: dict) -> bool:
        req = report_request(doc)
        if (req.get("label"), req.get("size"), req.get("ranks"), req.get("decomposition"),
                req.get("libraries")) != (self.label, self.size, self.args.ranks,
                                          self.args.decomposition, [self.shape.library]):
            return False
        return self.shape.grid_rows is None or all(
            req.get(k) == v for k, v in self.shape.fields().items() if k != "library")

    def run(self, cwd=None, quiet=True):
        """(completed process, path of the newest report written for this
        call or None, that report or None)."""
        before = self._stamps()
        done = subprocess.run(self.argv(), cwd=cwd, capture_output=quiet, text=True)
        best = (None, None, None)
        for path, stamp in self._stamps().items():
            if before.get(path) == stamp:
                continue
            try:
                doc = json.loads(path.read_text())
            except (OSError, ValueError):
                continue
            if self.answers_me(doc) and (best[0] is None or stamp[0] > best[0]):
                best = (stamp[0], path, doc)
        return done, best[1], best[2]


# ---- calibrate -----------------------------------------------------------------------------------

class Calibrator:
    """Times every shape at every calibration size; keeps the shapes with
    validated medians at two sizes or more.  Each run leaves
    <output>/calibration_logs/<nnn>_<code>_N<size>_<shape>.{command.txt,
    stdout,stderr}; progress is saved to <calibration>.partial.json after
    each shape."""

    OUTCOMES = ("attempted", "valid_measurements", "failed_launches", "missing_reports",
                "rejected_reports", "rejected_results")

    def __init__(self, args, root: Path):
        self.args, self.root = args, root
        out = Path(args.output)
        self.output = out if out.is_absolute() else root / out
        self.logs = self.output / "calibration_logs"
        self.tally = dict.fromkeys(self.OUTCOMES, 0)

    def _measure(self, number: int, shape: Shape, size: int) -> float | None:
        self.tally["attempted"] += 1
        stem = self.logs / f"{number:03d}_{shape.library}_N{size}_{shape.tag()}"
        call = RunnerCall(self.args, shape, size, self.output, CALIBRATION_LABEL)
        stem.with_suffix(".command.txt").write_text(shlex.join(call.argv()) + "\n")
        try:
            done, path, doc = call.run(cwd=self.root)
        except OSError as exc:
            stem.with_suffix(".stderr").write_text(f"{exc}\n")
            self.tally["failed_launches"] += 1
            return None
        stem.with_suffix(".stdout").write_text(done.stdout)
        stem.with_suffix(".stderr").write_text(done.stderr)
        outcome, median = "valid_measurements", None
        if done.returncode != 0:
            outcome = "failed_launches"
        elif doc is None:
            outcome = "missing_reports"
        elif doc.get("failures"):
            outcome = "rejected_reports"
        else:
            median = validated_median(doc, shape.library)
            outcome = "rejected_results" if median is None else outcome
        self.tally[outcome] += 1
        if median is None:
            print(f"    no measurement ({outcome.replace('_', ' ')}); logs {stem}.*",
                  file=sys.stderr, flush=True)
        else:
            print(f"    {median:.6f} s  ({path})", flush=True)
        return median

    def run(self, shapes: list[Shape], sizes: list[int], cal: Calibration, target: Path) -> None:
        self.logs.mkdir(parents=True, exist_ok=True)
        partial = target.with_name(target.stem + ".partial.json")
        partial.parent.mkdir(parents=True, exist_ok=True)
        for number, shape in enumerate(shapes, 1):
            print(f"[{number}/{len(shapes)}] {shape.library} {shape.tag()}", flush=True)
            samples = []
            for size in sizes:
                print(f"  N={size}", flush=True)
                median = self._measure(number, shape, size)
                if median is not None:
                    samples.append({"size": size, "median_s": median})
            if len({s["size"] for s in samples}) >= 2:
                cal.add_shape(shape, samples)
            else:
                print(f"  {shape.library} {shape.tag()} dropped: measured at fewer than two "
                      "sizes", file=sys.stderr, flush=True)
            partial.write_text(json.dumps(cal.doc, indent=2) + "\n")
        kept = len(cal.doc["candidates"])
        print(f"calibration runs: " + ", ".join(f"{k} {v}" for k, v in self.tally.items())
              + f"; {kept} of {len(shapes)} configurations kept", flush=True)
        if not kept:
            raise PredictionError(f"every configuration was dropped (progress in {partial}, "
                                  f"logs in {self.logs})")
        cal.write(target)
        partial.unlink(missing_ok=True)
        print(f"calibration written to {target}", flush=True)


def parse_grids(items) -> list[tuple[int, int]] | None:
    if not items:
        return None
    return [tuple(int(v) for v in g.lower().split("x")) for g in items]


def calibrate_cmd(args) -> int:
    root = Path(args.project_root).resolve()
    sizes = sorted(set(args.sizes))
    if len(sizes) < 2 or sizes[0] <= 0:
        raise PredictionError("calibrate at two or more distinct positive sizes")
    if args.decomposition == "1d" and any(n % args.ranks for n in sizes):
        raise PredictionError("each 1D calibration size must be a multiple of --ranks")
    sig = make_signature(executable(root, args.build_dir, args.decomposition),
                         args.decomposition, args.ranks, Platform.probe())
    grids = [(args.grid_rows, args.grid_cols)] if args.grid_rows is not None \
        else parse_grids(args.grids)
    shapes = Shape.space(args.decomposition, args.ranks, sizes, args.libraries, args.bks, grids)
    if not shapes:
        raise PredictionError("no configuration fits these ranks, grids and sizes")
    cal = Calibration.empty(sig, sizes, "short calibration runs: forecasts are estimates, "
                                        "check a choice with --auto-validate")
    target = calibration_path(args.decomposition, args.ranks, args.model, root)
    Calibrator(args, root).run(shapes, sizes, cal, target)
    return 0


# ---- recover -----------------------------------------------------------------------------------

def classify_report(doc: dict, decomposition: str, ranks: int):
    """(shape, size, median) from one earlier calibration report, or the
    reason it cannot be used."""
    req = report_request(doc)
    if req.get("label") != CALIBRATION_LABEL:
        return "not a calibration run"
    try:
        size, their_ranks = int(req["size"]), int(req["ranks"])
    except (KeyError, TypeError, ValueError):
        return "no usable size / ranks"
    if (their_ranks, req.get("decomposition")) != (ranks, decomposition):
        return "another rank count or decomposition"
    libs = req.get("libraries")
    if not isinstance(libs, list) or len(libs) != 1 or libs[0] not in LIBRARY_NAMES:
        return "not exactly one known library code"
    shape = Shape(libs[0])
    if decomposition == "2d":
        try:
            rows, cols, bk = (int(req[k]) for k in ("grid_rows", "grid_cols", "summa_bk"))
        except (KeyError, TypeError, ValueError):
            return "no usable grid / panel width"
        shape = Shape(libs[0], rows, cols, bk)
        if min(rows, cols, bk) <= 0 or rows * cols != ranks:
            return "grid / panel width does not fit"
    if size <= 0 or not shape.splits(size) or (decomposition == "1d" and size % ranks):
        return "size does not fit the decomposition"
    if doc.get("failures"):
        return "the report records failures"
    median = validated_median(doc, shape.library)
    return "no validated positive median" if median is None else (shape, size, median)


def recover_cmd(args) -> int:
    """Rebuilds a calibration from reports of earlier calibration runs.  The
    reports cannot prove which GPU or executable produced them, so the
    caller vouches for them with --assume-compatible."""
    root = Path(args.project_root).resolve()
    source = Path(args.input) if Path(args.input).is_absolute() else root / args.input
    paths = sorted(source.glob("*/benchmark.json"))
    if not paths:
        raise PredictionError(f"no */benchmark.json under {source}")
    if not args.assume_compatible:
        raise PredictionError("earlier reports carry no proof of the GPU or executable they ran "
                              "on; check them and pass --assume-compatible")
    sig = make_signature(executable(root, args.build_dir, args.decomposition),
                         args.decomposition, args.ranks, Platform.probe())
    found = collections.defaultdict(lambda: collections.defaultdict(list))
    skipped = collections.Counter()
    for path in paths:
        try:
            verdict = classify_report(json.loads(path.read_text()), args.decomposition, args.ranks)
        except (OSError, ValueError) as exc:
            verdict = f"unreadable ({type(exc).__name__})"
        if isinstance(verdict, str):
            skipped[verdict] += 1
            continue
        shape, size, median = verdict
        found[shape][size].append(median)
    all_sizes = set()
    cal = Calibration.empty(sig, [], "rebuilt from earlier calibration reports vouched for with "
                                     "--assume-compatible; forecasts are estimates")
    for shape in sorted(found, key=lambda s: str(tuple(s.fields().values()))):
        per_size = found[shape]
        if len(per_size) < 2:
            skipped[f"{shape.library} {shape.tag()}: one size only"] += 1
            continue
        cal.add_shape(shape, [{"size": n, "median_s": statistics.median(per_size[n])}
                              for n in sorted(per_size)])
        all_sizes |= set(per_size)
    cal.doc["sizes"] = sorted(all_sizes)
    print(f"{len(paths)} report(s) read, {len(cal.doc['candidates'])} configuration(s) "
          "recovered", flush=True)
    for why, count in sorted(skipped.items()):
        print(f"  skipped {count}: {why}", flush=True)
    if not cal.doc["candidates"]:
        raise PredictionError("no configuration has validated medians at two sizes")
    target = calibration_path(args.decomposition, args.ranks, args.model, root)
    cal.write(target)
    print(f"recovered calibration written to {target}", flush=True)
    return 0


# ---- auto ------------------------------------------------------------------------------------------

@dataclass
class AutoSession:
    """--libraries AUTO at one size: rank, optionally validate and feed back,
    write <output>/autotuning_prediction.json and print the choice."""
    args: argparse.Namespace
    path: Path = field(init=False)
    cal: Calibration = field(init=False)
    report: dict = field(init=False, default_factory=dict)

    def rank(self) -> None:
        a = self.args
        self.path = calibration_path(a.decomposition, a.ranks, getattr(a, "auto_model", None),
                                     Path.cwd())
        if not self.path.is_file():
            raise FileNotFoundError(f"no calibration at {self.path}; make one with "
                                    "`python -m mgxbench_amd.predictive calibrate`")
        self.cal = Calibration.read(self.path)
        here = make_signature(executable(Path.cwd(), a.build_dir, a.decomposition),
                              a.decomposition, a.ranks, Platform.probe())
        if not signatures_agree(self.cal.signature, here):
            raise PredictionError(f"{self.path} belongs to other GPUs, another executable or "
                                  "another allocation; calibrate again")
        grid = (a.grid_rows, a.grid_cols) if a.grid_rows and a.grid_cols else None
        ranking, method = forecast(self.cal, a.size, a.ranks, a.decomposition, grid,
                                   getattr(a, "prediction_model", "extrapolation"))
        best = ranking[0]
        layout = (best["grid_rows"], best["grid_cols"], best["summa_bk"])
        twin = next((r for r in ranking if r["library"] == "MMM"
                     and (r["grid_rows"], r["grid_cols"], r["summa_bk"]) == layout), None)
        self.report = {"size": a.size, "ranks": a.ranks, "decomposition": a.decomposition,
                       "architecture_id": here["id"], "best": best, "ranking": ranking,
                       "speedup_vs_MPI_predicted":
                           twin["predicted_s"] / best["predicted_s"] if twin else None,
                       "prediction_model": method, "calibration_file": str(self.path),
                       "status": "PREDICTED", "measured": False}

    def validate(self) -> int:
        """Times the choice through the runner; with --auto-feedback its
        validated median joins that shape's curve (old file kept as .bak)."""
        a, best = self.args, self.report["best"]
        call = RunnerCall(a, Shape.of(best), a.size, Path(a.output) / "validation",
                          VALIDATION_LABEL, getattr(a, "timeout", None))
        print("AUTO validation: " + shlex.join(call.argv()), flush=True)
        done, where, doc = call.run(quiet=False)
        if done.returncode != 0:
            self.report["validation"] = {"status": "FAILED", "exit_code": done.returncode}
            return 2
        measured = None if doc is None or doc.get("failures") else \
            validated_median(doc, best["library"])
        if measured is None:
            self.report["validation"] = {"status": "FAILED",
                                         "reason": "the runner left no clean report for the choice"}
            return 2
        guess = best["predicted_s"]
        miss = abs(measured - guess)
        self.report["validation"] = {"status": "PASS", "median_s": measured, "predicted_s": guess,
                                     "absolute_error_s": miss, "relative_error": miss / measured,
                                     "benchmark_report": str(where)}
        print(f"AUTO validation: forecast {guess:.6f} s, measured {measured:.6f} s "
              f"({100 * miss / measured:.2f}% apart)", flush=True)
        if a.auto_feedback:
            sample = {"size": a.size, "median_s": measured, "source": str(where),
                      "timestamp_utc": datetime.datetime.now(datetime.timezone.utc).isoformat()}
            if self.cal.append_sample(Shape.of(best), sample):
                backup = self.cal.write(self.path)
                self.report["feedback"] = {"status": "SAVED", "model": str(self.path),
                                           "backup": str(backup)}
            else:
                self.report["feedback"] = {"status": "SKIPPED",
                                           "reason": "the choice is not in the calibration"}
        return 0

    def announce(self, out: Path) -> None:
        a, r = self.args, self.report
        best, speedup = r["best"], r["speedup_vs_MPI_predicted"]
        where = "" if a.decomposition == "1d" else \
            f" on a {best['grid_rows']}x{best['grid_cols']} grid, panels of {best['summa_bk']}"
        print(f"\nAUTO [{r['prediction_model']}] N={a.size}, {a.ranks} GPU(s), "
              f"{a.decomposition.upper()}: {best['library']} "
              f"({LIBRARY_NAMES[best['library']]}){where}")
        lines = [f"forecast {best['predicted_s']:.6f} s"
                 + (" (outside the calibrated sizes)" if best["extrapolated"] else "")]
        if speedup:
            lines.append(f"{speedup:.2f}x the MMM forecast on the same grid")
        lines.append(f"{r['status'].lower()}; report {out}")
        for part in ("validation", "feedback"):
            if part in r:
                lines.append(f"{part}: {r[part].get('status')}")
        print("  " + "\n  ".join(lines), flush=True)


def run_auto(args) -> int:
    session = AutoSession(args)
    try:
        session.rank()
    except (PredictionError, FileNotFoundError, ImportError, OSError) as exc:
        print(f"AUTO: {exc}", file=sys.stderr)
        return 2
    status = 0
    if getattr(args, "auto_validate", False):
        status = session.validate()
        session.report["status"] = "VALIDATED" if status == 0 else "VALIDATION_FAILED"
    out = Path(args.output) / "autotuning_prediction.json"
    out.parent.mkdir(parents=True, exist_ok=True)
    out.write_text(json.dumps(session.report, indent=2) + "\n")
    session.announce(out)
    return status


# ---- command line ------------------------------------------------------------------------------

def _argument_problem(args) -> str | None:
    if args.ranks <= 0:
        return "--ranks must be positive"
    rows, cols = getattr(args, "grid_rows", None), getattr(args, "grid_cols", None)
    listed = getattr(args, "grids", None) or []
    if (rows is None) != (cols is None):
        return "--grid-rows needs --grid-cols and the other way round"
    if rows is not None:
        if args.decomposition != "2d":
            return "a process grid applies to --decomposition 2d only"
        if min(rows, cols) <= 0 or rows * cols != args.ranks:
            return f"a {rows}x{cols} grid does not hold {args.ranks} ranks"
        if listed:
            return "pass --grids or --grid-rows/--grid-cols, not both"
        if args.command == "auto" and (args.size % rows or args.size % cols):
            return f"N={args.size} does not split over a {rows}x{cols} grid"
    for item in listed:
        parts = item.lower().split("x")
        if len(parts) != 2 or not all(p.isdigit() for p in parts):
            return f"--grids entry {item!r} is not RxC"
        r, c = int(parts[0]), int(parts[1])
        if min(r, c) <= 0 or r * c != args.ranks:
            return f"--grids entry {item!r} does not hold {args.ranks} ranks"
    if args.command == "auto":
        if args.size <= 0:
            return "--size must be positive"
        if args.auto_feedback and not args.auto_validate:
            return "--auto-feedback only makes sense with --auto-validate"
    return None


def build_parser() -> argparse.ArgumentParser:
    launcher = f"{shlex.quote(sys.executable)} -m mgxbench_amd.launch"
    bin_dir = str(PKG / "bin")
    top = argparse.ArgumentParser(description="calibrated choice of MGXBench run configurations")
    sub = top.add_subparsers(dest="command", required=True)
    spec = {
        "calibrate": ("time every configuration at a few sizes", [
            ("--project-root", dict(default=".")),
            ("--sizes", dict(type=int, nargs="+", default=[2048, 4096, 8192])),
            ("--libraries", dict(nargs="+", choices=list(LIBRARY_NAMES),
                                 default=list(LIBRARY_NAMES))),
            ("--bks", dict(type=int, nargs="+", default=[128, 256], help="2D panel widths")),
            ("--grids", dict(nargs="+", default=None, help="2D grids, e.g. 2x4 4x2")),
            ("--grid-rows", dict(type=int, default=None)),
            ("--grid-cols", dict(type=int, default=None)),
            ("--warmup", dict(type=int, default=1)),
            ("--repeats", dict(type=int, default=3)),
            ("--launcher", dict(default=launcher)),
            ("--launcher-args", dict(default="")),
            ("--build-dir", dict(default=bin_dir)),
            ("--output", dict(default="results/auto_calibration")),
            ("--model", dict(default=None, help="calibration file to write"))]),
        "recover": ("rebuild a calibration file from earlier reports", [
            ("--project-root", dict(default=".")),
            ("--input", dict(default="results/auto_calibration/runs")),
            ("--model", dict(default=None)),
            ("--build-dir", dict(default=bin_dir)),
            ("--assume-compatible", dict(action="store_true",
                                         help="vouch that the reports come from this GPU "
                                              "and executable"))]),
        "auto": ("the fastest calibrated configuration at one size", [
            ("--size", dict(type=int, required=True)),
            ("--grid-rows", dict(type=int, default=None)),
            ("--grid-cols", dict(type=int, default=None)),
            ("--prediction-model", dict(choices=METHODS, default="extrapolation")),
            ("--build-dir", dict(default=bin_dir)),
            ("--auto-model", dict(default=None)),
            ("--auto-validate", dict(action="store_true")),
            ("--auto-feedback", dict(action="store_true")),
            ("--warmup", dict(type=int, default=2)),
            ("--repeats", dict(type=int, default=5)),
            ("--launcher", dict(default=launcher)),
            ("--launcher-args", dict(default="")),
            ("--timeout", dict(type=int, default=600)),
            ("--output", dict(default="results/auto_prediction"))]),
    }
    for name, (help_text, options) in spec.items():
        p = sub.add_parser(name, help=help_text)
        p.add_argument("--decomposition", choices=("1d", "2d"), required=True)
        p.add_argument("--ranks", type=int, required=True)
        for flag, kw in options:
            p.add_argument(flag, **kw)
    return top


def main(argv=None) -> int:
    parser = build_parser()
    args = parser.parse_args(argv)
    problem = _argument_problem(args)
    if problem:
        parser.error(problem)
    if args.command == "auto":
        return run_auto(args)
    try:
        return {"calibrate": calibrate_cmd, "recover": recover_cmd}[args.command](args)
    except PredictionError as exc:
        print(f"{args.command}: {exc}", file=sys.stderr)
        return 2


if __name__ == "__main__":
    sys.exit(main())
