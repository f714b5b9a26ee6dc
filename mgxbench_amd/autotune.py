"""Autotuner with a learned cost model: the autotuning/predictive.py counterpart.

The reference autotuner (autotuning/predictive.py) times `benchmark.py` per
candidate (library, grid, BK) over several sizes and fits one runtime model
per candidate: a clipped log-log power law or an lstsq fit on (x^2, x^3)
("extrapolation", predictive.py:100-165), a StandardScaler + Ridge(alpha=0.1)
fit ("ml", :140-165) or a hold-out vote between them ("auto", :213-233).

Here the candidates are the knobs of the MI355X local product, which the
reference does not have: form (FP64 MFMA, INT8 digit-slice split with s
slices, or INT8 modular split with n moduli), the kernel block tile and the
SUMMA panel width.  Only FP64-accurate candidates are offered, calibrated on
the accuracy-stress inputs (random sign x U[1,2) x 2^U(-30,30), rows / columns
x 2^U(-10,10)) as well as on uniform ones: s = 8 slices (s = 7 is up to 1.9x
plain FP64's componentwise error on stress inputs at depth 4096), and moduli
counts whose scaling exponent at the split-product depth is at least MIN_BETA
(mgx_split_mod_beta): 15 moduli at group depths up to 4096, 16 at any depth.  One MLP (3 layers, width 64,
FP32, trained on the GPU by libmgx's own kernels) predicts log runtime from
log-shape features for every candidate at once; exhaustive timing of its
top-k candidates stays the fallback and the arbiter.

Subcommands (python -m mgxbench_amd.autotune ...):
  measure   time candidates on the GPU -> dataset JSON (checkpointed per shape)
  train     fit the MLP on the training shapes of a dataset -> model JSON
  evaluate  held-out MAPE and regret of the MLP vs the reference-style
            per-candidate power law and vs the fixed default configuration
  select    rank candidates for one shape (optionally time the top k)
  share     the same for one rank's share of a 1D / 2D run

The reference's run-level AUTO path (calibrated library / grid / BK choice)
is mgxbench_amd/predictive.py.
"""
from __future__ import annotations

import argparse
import ctypes
import hashlib
import json
import math
import os
import random
import sys
import time
from dataclasses import asdict, dataclass
from pathlib import Path

import numpy as np

from . import _lib

SIZES = list(range(4096, 32768 + 1, 4096))
PANELS = (128, 256, 512, 1024)
SLICES = (8,)
MODULI = (15, 16)
# quantisation to >= 52 bits per row / column keeps the modular form within
# plain FP64's normwise and componentwise error on stress inputs
MIN_BETA = 52
TILES = (_lib.TILE_LARGE, _lib.TILE_SMALL)
DEFAULT = None  # filled below: the reference's default configuration
SCHEMA = 1


@dataclass(frozen=True)
class Candidate:
    form: str        # "fp64" | "int8" | "mod"
    slices: int      # slices (int8), moduli (mod), 0 for fp64
    tile: int        # _lib.TILE_LARGE | _lib.TILE_SMALL
    panel: int       # SUMMA panel width (BK)

    def key(self) -> str:
        return f"{self.form}-s{self.slices}-t{self.tile}-b{self.panel}"


DEFAULT = Candidate("fp64", 0, _lib.TILE_LARGE, 256)   # MGXBENCH_SUMMA_BK default 256


FORM_CODE = {"fp64": _lib.FORM_FP64, "int8": _lib.FORM_INT8, "mod": _lib.FORM_INT8_MOD}


def mod_beta(moduli: int, k: int) -> int:
    return int(_lib.lib().mgx_split_mod_beta(moduli, k))


def step_k(c: Candidate, k: int) -> int:
    """Reduction depth of each launch of the panel loop: a group of whole
    panels (mgx_split_group_k with the form's default group)."""
    return int(_lib.lib().mgx_split_group_k(FORM_CODE[c.form], c.panel, 0, k))


def candidate_space(k: int) -> list[Candidate]:
    out = []
    for panel in PANELS:
        if panel > k:
            continue
        for tile in TILES:
            out.append(Candidate("fp64", 0, tile, panel))
            for s in SLICES:
                out.append(Candidate("int8", s, tile, panel))
            for nm in MODULI:
                c = Candidate("mod", nm, tile, panel)
                if mod_beta(nm, step_k(c, k)) >= MIN_BETA:
                    out.append(c)
    return out


def rank_shares(sizes=SIZES) -> list[tuple[int, int, int]]:
    """(m, n, k) of one rank of the 1x4, 2x2 and 4x1 grids at each global N."""
    shapes = []
    for N in sizes:
        for pr, pc in ((1, 4), (2, 2), (4, 1)):
            shapes.append((N // pr, N // pc, N))
    return shapes


def random_shapes(count: int, seed: int = 7, sizes=SIZES,
                  exclude=()) -> list[tuple[int, int, int]]:
    """`count` distinct (m, n, k) drawn from sizes^3, none of them in `exclude`."""
    seen = {tuple(s) for s in exclude}
    if count > len(sizes) ** 3 - len(seen & {(a, b, c) for a in sizes for b in sizes
                                              for c in sizes}):
        raise ValueError("not enough distinct shapes")
    rng = random.Random(seed)
    out = []
    while len(out) < count:
        s = (rng.choice(sizes), rng.choice(sizes), rng.choice(sizes))
        if s not in seen:
            seen.add(s)
            out.append(s)
    return out


FEATURE_NAMES = ["log2_m", "log2_n", "log2_k", "log2_panel", "log2_panels", "ragged_frac",
                 "int8", "slices", "products", "tile_small", "log2_mnk", "log2_mn", "mod",
                 "log2_step_k", "log2_steps", "wave_fill"]

# C block tile edge and concurrently resident workgroups (CU count x
# workgroups per CU) of each product kernel, for the wave-fill feature
_TILE_SLOTS = {("split", False): (256, 256), ("split", True): (128, 512),
               ("fp64", False): (128, 512), ("fp64", True): (64, 1024)}


def wave_fill(m: int, n: int, c) -> float:
    """Mean occupancy of the workgroup rounds of one product launch:
    (tiles / slots) / ceil(tiles / slots), 1 when the last round is full."""
    edge, slots = _TILE_SLOTS[("fp64" if c.form == "fp64" else "split",
                               c.tile == _lib.TILE_SMALL)]
    tiles = math.ceil(m / edge) * math.ceil(n / edge)
    rounds = math.ceil(tiles / slots)
    return tiles / (slots * rounds)


def features(m: int, n: int, k: int, c: Candidate) -> list[float]:
    panel = min(c.panel, k)
    panels = math.ceil(k / panel)
    ragged = (k - (panels - 1) * panel) / panel
    # products of the panel loop: one per panel (FP64) or per panel group
    sk = step_k(c, k)
    steps = math.ceil(k / sk)
    # INT8 products per output: slice pairs (int8) or moduli (mod)
    products = (c.slices * (c.slices + 1) / 2 if c.form == "int8"
                else float(c.slices) if c.form == "mod" else 0.0)
    return [math.log2(m), math.log2(n), math.log2(k), math.log2(panel), math.log2(panels), ragged,
            1.0 if c.form == "int8" else 0.0, c.slices / 8.0, products / 36.0,
            1.0 if c.tile == _lib.TILE_SMALL else 0.0, math.log2(m * n * k) / 3.0,
            math.log2(m * n) / 2.0, 1.0 if c.form == "mod" else 0.0,
            math.log2(sk), math.log2(steps), wave_fill(m, n, c)]


# ---- measurement (GPU) --------------------------------------------------------------------

class Timer:
    """Times candidates on device buffers sized for the largest shape."""

    def __init__(self, max_m: int, max_n: int, max_k: int, seed: int = 1):
        import torch
        import mgxbench_amd as mgx
        self.torch, self.mgx = torch, mgx
        dev = torch.device("cuda", 0)
        self.A = torch.empty(max_m * max_k, dtype=torch.float64, device=dev)
        self.B = torch.empty(max_k * max_n, dtype=torch.float64, device=dev)
        self.C = torch.empty(max_m * max_n, dtype=torch.float64, device=dev)
        mgx.fill(self.A, max_m, max_k, max_k, 1, seed=seed)
        mgx.fill(self.B, max_k, max_n, max_n, 1, seed=seed + 1)
        ws = max(max(mgx.split_workspace_bytes(mgx.FORM_INT8, max(SLICES), max_m, max_n, g),
                     mgx.split_workspace_bytes(mgx.FORM_INT8_MOD, max(MODULI), max_m, max_n, g))
                 for g in {step_k(Candidate("mod", 0, _lib.TILE_LARGE, p), max_k) for p in PANELS})
        self.ws = torch.empty(ws, dtype=torch.uint8, device=dev)
        self.ws_bytes = ws
        torch.cuda.synchronize()

    def time(self, m: int, n: int, k: int, c: Candidate, est_ms: float | None = None) -> float:
        cfg = _lib.TuneConfig(FORM_CODE[c.form], c.slices, c.tile, c.panel)
        est = est_ms if est_ms is not None else 2.0 * m * n * k / 50e12 * 1e3
        warmup = 1 if est < 50 else 0
        reps = int(min(5, max(1, 100 // max(est, 1))))
        ms = ctypes.c_double(0.0)
        # operands are views of the preallocated buffers with tight leading dims
        _lib.call("mgx_time_local_product", ctypes.byref(cfg), self.A.data_ptr(),
                  self.B.data_ptr(), self.C.data_ptr(), m, n, k, k, n, n, self.ws.data_ptr(),
                  self.ws_bytes, warmup, reps, ctypes.byref(ms))
        return ms.value


def measure(shapes, out_path: Path, log=print, time_budget_s: float | None = None) -> dict:
    """Times every candidate of every shape; resumes from out_path if present.
    With a time budget, stops before a shape that would likely overrun it
    (the dataset stays checkpointed; a later call resumes)."""
    doc = {"schema": SCHEMA, "kind": "mgx-autotune-dataset", "rows": [], "shapes_done": []}
    if out_path.exists():
        doc = json.loads(out_path.read_text())
    done = {tuple(s) for s in doc["shapes_done"]}
    todo = [s for s in shapes if tuple(s) not in done]
    if not todo:
        return doc
    mm = max(s[0] for s in todo)
    nn = max(s[1] for s in todo)
    kk = max(s[2] for s in todo)
    timer = Timer(mm, nn, kk)
    start = time.time()
    worst = 0.0
    for i, (m, n, k) in enumerate(todo):
        t0 = time.time()
        if time_budget_s is not None and (t0 - start) + 1.5 * worst > time_budget_s:
            log(f"[autotune] time budget reached after {i} of {len(todo)} shapes")
            break
        for c in candidate_space(k):
            ms = timer.time(m, n, k, c)
            doc["rows"].append({"m": m, "n": n, "k": k, **asdict(c), "ms": ms})
        doc["shapes_done"].append([m, n, k])
        tmp = out_path.with_suffix(".tmp")
        tmp.write_text(json.dumps(doc))
        os.replace(tmp, out_path)
        worst = max(worst, time.time() - t0)
        log(f"[autotune] shape {i + 1}/{len(todo)} m={m} n={n} k={k} "
            f"({len(candidate_space(k))} candidates, {time.time() - t0:.1f} s)")
    return doc


# ---- model ----------------------------------------------------------------------------------

def shape_split(shapes, holdout_frac: float = 0.25, seed: int = 11):
    """Deterministic split of shapes into (train, held-out)."""
    keyed = sorted(shapes, key=lambda s: hashlib.sha256(f"{seed}:{s}".encode()).hexdigest())
    n_hold = max(1, int(round(holdout_frac * len(keyed))))
    return sorted(keyed[n_hold:]), sorted(keyed[:n_hold])


def _rows_xy(rows):
    X = np.array([features(r["m"], r["n"], r["k"], Candidate(r["form"], r["slices"], r["tile"],
                                                             r["panel"])) for r in rows],
                 dtype=np.float64)
    y = np.log(np.array([r["ms"] for r in rows], dtype=np.float64))
    return X, y


class MLPModel:
    """Standardised features -> log milliseconds through libmgx's GPU MLP."""

    def __init__(self, mean, std, theta, meta=None):
        self.mean = np.asarray(mean, dtype=np.float64)
        self.std = np.asarray(std, dtype=np.float64)
        self.theta = np.ascontiguousarray(theta, dtype=np.float32)
        self.meta = meta or {}

    @property
    def in_dim(self) -> int:
        return len(self.mean)

    def _x(self, X) -> np.ndarray:
        return np.ascontiguousarray((np.asarray(X) - self.mean) / self.std, dtype=np.float32)

    def predict_log_ms(self, X) -> np.ndarray:
        Xs = self._x(X)
        out = np.empty(len(Xs), dtype=np.float32)
        _lib.call("mgx_mlp_predict", self.theta.ctypes.data, self.in_dim, Xs.ctypes.data,
                  len(Xs), out.ctypes.data)
        return out.astype(np.float64)

    def to_json(self) -> dict:
        return {"schema": SCHEMA, "kind": "mgx-mlp-cost-model", "features": FEATURE_NAMES,
                "hidden": 64, "mean": self.mean.tolist(), "std": self.std.tolist(),
                "theta": self.theta.astype(float).tolist(), **self.meta}

    @classmethod
    def from_json(cls, doc: dict) -> "MLPModel":
        if doc.get("schema") != SCHEMA or doc.get("kind") != "mgx-mlp-cost-model":
            raise ValueError("not an mgx MLP cost model (schema 1)")
        meta = {k: v for k, v in doc.items() if k not in ("mean", "std", "theta")}
        return cls(doc["mean"], doc["std"], doc["theta"], meta)


def train(rows, train_shapes, epochs: int = 4000, lr: float = 3e-3, seed: int = 5,
          weight_decay: float = 1e-5) -> tuple[MLPModel, np.ndarray]:
    tset = {tuple(s) for s in train_shapes}
    tr = [r for r in rows if (r["m"], r["n"], r["k"]) in tset]
    X, y = _rows_xy(tr)
    mean, std = X.mean(axis=0), X.std(axis=0)
    std[std == 0] = 1.0
    Xs = np.ascontiguousarray((X - mean) / std, dtype=np.float32)
    ys = np.ascontiguousarray(y, dtype=np.float32)
    D = Xs.shape[1]
    P = _lib.lib().mgx_mlp_param_count(D)
    theta = np.empty(P, dtype=np.float32)
    _lib.call("mgx_mlp_init", theta.ctypes.data, D, seed)
    hist = np.empty(epochs, dtype=np.float32)
    _lib.call("mgx_mlp_train", theta.ctypes.data, D, Xs.ctypes.data, ys.ctypes.data, len(ys),
              epochs, ctypes.c_float(lr), ctypes.c_float(weight_decay), hist.ctypes.data)
    model = MLPModel(mean, std, theta, {"train_shapes": [list(s) for s in train_shapes],
                                        "epochs": epochs, "lr": lr,
                                        "final_train_loss": float(hist[-1])})
    return model, hist


class PowerLawModel:
    """Reference-style baseline: per candidate, log t = a + b log(2mnk) with the
    exponent clipped to [0, 4] (the 'extrapolation' fit of predictive.py:100-127,
    applied to the problem's flop count)."""

    def __init__(self, rows, train_shapes):
        tset = {tuple(s) for s in train_shapes}
        by = {}
        for r in rows:
            if (r["m"], r["n"], r["k"]) not in tset:
                continue
            c = Candidate(r["form"], r["slices"], r["tile"], r["panel"])
            by.setdefault(c, []).append((math.log(2.0 * r["m"] * r["n"] * r["k"]), math.log(r["ms"])))
        self.fit = {}
        for c, pts in by.items():
            x = np.array([p[0] for p in pts])
            y = np.array([p[1] for p in pts])
            if len(pts) >= 2 and np.ptp(x) > 0:
                b = float(np.clip(np.polyfit(x, y, 1)[0], 0.0, 4.0))
            else:
                b = 1.0
            a = float(np.mean(y - b * x))
            self.fit[c] = (a, b)

    def predict_log_ms(self, m, n, k, cands) -> np.ndarray:
        x = math.log(2.0 * m * n * k)
        out = []
        for c in cands:
            a, b = self.fit.get(c, (0.0, 1.0))
            out.append(a + b * x)
        return np.array(out)


def evaluate(rows, model: MLPModel, train_shapes, heldout_shapes, topk: int = 3) -> dict:
    """Held-out prediction error and regret (measured t(pick) / t(best) - 1)
    over the current candidate space (rows of retired candidates are skipped)."""
    by_shape = {}
    space = {}
    for r in rows:
        if r["k"] not in space:
            space[r["k"]] = set(candidate_space(r["k"]))
        if Candidate(r["form"], r["slices"], r["tile"], r["panel"]) in space[r["k"]]:
            by_shape.setdefault((r["m"], r["n"], r["k"]), []).append(r)
    base = PowerLawModel(rows, train_shapes)
    ape, regret_mlp, regret_topk, regret_base, regret_default, per_shape = [], [], [], [], [], []
    for s in heldout_shapes:
        rs = by_shape.get(tuple(s), [])
        if not rs:
            continue
        cands = [Candidate(r["form"], r["slices"], r["tile"], r["panel"]) for r in rs]
        t = np.array([r["ms"] for r in rs])
        X = np.array([features(r["m"], r["n"], r["k"], c) for r, c in zip(rs, cands)])
        pred = np.exp(model.predict_log_ms(X))
        ape.extend(np.abs(pred - t) / t)
        best = t.min()
        order = np.argsort(pred)
        pick = t[order[0]]
        pick_k = t[order[:topk]].min()
        bpred = base.predict_log_ms(*s, cands)
        bpick = t[int(np.argmin(bpred))]
        dflt = next((ti for c, ti in zip(cands, t) if c == DEFAULT), None)
        regret_mlp.append(pick / best - 1)
        regret_topk.append(pick_k / best - 1)
        regret_base.append(bpick / best - 1)
        if dflt is not None:
            regret_default.append(dflt / best - 1)
        per_shape.append({"shape": list(s), "best": cands[int(np.argmin(t))].key(),
                          "mlp_pick": cands[int(order[0])].key(),
                          "regret_mlp": pick / best - 1, f"regret_top{topk}": pick_k / best - 1,
                          "regret_powerlaw": bpick / best - 1})
    def mean(x):
        return float(np.mean(x)) if len(x) else None
    return {
        "heldout_shapes": len(per_shape), "heldout_configs": len(ape),
        "mape": mean(ape), "median_ape": float(np.median(ape)) if ape else None,
        "regret_mlp_mean": mean(regret_mlp), "regret_mlp_max": float(np.max(regret_mlp)) if regret_mlp else None,
        f"regret_top{topk}_mean": mean(regret_topk),
        f"regret_top{topk}_max": float(np.max(regret_topk)) if regret_topk else None,
        "regret_powerlaw_mean": mean(regret_base),
        "regret_default_mean": mean(regret_default),
        "per_shape": per_shape,
    }


def select(model: MLPModel, m: int, n: int, k: int, topk: int = 3, timer: Timer | None = None):
    """Candidates ranked by predicted time; with a timer, the top k are timed
    and the fastest measured one is chosen."""
    cands = candidate_space(k)
    X = np.array([features(m, n, k, c) for c in cands])
    pred = np.exp(model.predict_log_ms(X))
    order = np.argsort(pred)
    ranked = [{"candidate": asdict(cands[i]), "key": cands[i].key(), "predicted_ms": float(pred[i])}
              for i in order]
    choice = ranked[0]
    if timer is not None:
        for r in ranked[:topk]:
            r["measured_ms"] = timer.time(m, n, k, Candidate(**r["candidate"]), r["predicted_ms"])
        choice = min(ranked[:topk], key=lambda r: r["measured_ms"])
    return choice, ranked


def share_shape(decomposition: str, size: int, ranks: int, grid=None) -> tuple[int, int, int]:
    if decomposition == "1d":
        return size // ranks, size, size
    pr, pc = grid
    return size // pr, size // pc, size


def default_model_path() -> Path:
    return Path(__file__).resolve().parent / "models" / "mlp_cost_model.json"


def share(args) -> int:
    """Pick (form, slices, tile, panel) for one rank's local product of a
    1D / 2D run (the `share` subcommand)."""
    from .grid import resolve_grid
    path = Path(args.auto_model) if args.auto_model else default_model_path()
    if not path.is_file():
        print(f"no cost model at {path}; train one with `python -m mgxbench_amd.autotune "
              "train`", file=sys.stderr)
        return 2
    model = MLPModel.from_json(json.loads(path.read_text()))
    grid = resolve_grid(args.ranks, args.grid_rows, args.grid_cols) if args.decomposition == "2d" else None
    m, n, k = share_shape(args.decomposition, args.size, args.ranks, grid)
    timer = Timer(m, n, k) if args.auto_validate else None
    choice, ranked = select(model, m, n, k, timer=timer)
    out = {"schema": SCHEMA, "decomposition": args.decomposition, "size": args.size,
           "ranks": args.ranks, "share": [m, n, k], "choice": choice, "ranking": ranked[:10],
           "model": str(path)}
    Path(args.output).mkdir(parents=True, exist_ok=True)
    dest = Path(args.output) / "autotuning_prediction.json"
    dest.write_text(json.dumps(out, indent=2))
    print(f"choice for share m={m} n={n} k={k}: {choice['key']} "
          f"(predicted {choice['predicted_ms']:.2f} ms)")
    print(f"prediction written to {dest}")
    if args.auto_feedback and timer is not None:
        ds = Path(args.auto_model).with_name("dataset.json") if args.auto_model else \
            default_model_path().with_name("dataset.json")
        doc = json.loads(ds.read_text()) if ds.is_file() else {"schema": SCHEMA, "rows": [],
                                                               "shapes_done": []}
        for r in ranked:
            if "measured_ms" in r:
                doc["rows"].append({"m": m, "n": n, "k": k, **r["candidate"], "ms": r["measured_ms"]})
        tmp = ds.with_suffix(".tmp")
        tmp.write_text(json.dumps(doc))
        os.replace(tmp, ds)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("measure")
    p.add_argument("--out", required=True)
    p.add_argument("--random-shapes", type=int, default=40)
    p.add_argument("--max-size", type=int, default=32768)
    p.add_argument("--time-budget", type=float, default=None,
                   help="seconds; stop before a shape that would overrun it")
    p.add_argument("--seed", type=int, default=7, help="random-shape seed")
    p.add_argument("--no-shares", action="store_true", help="skip the grid rank shares")
    p.add_argument("--exclude-dataset", default=None,
                   help="draw random shapes outside this dataset's shapes (a held-out set)")
    p = sub.add_parser("train")
    p.add_argument("--dataset", required=True)
    p.add_argument("--out", required=True)
    p.add_argument("--epochs", type=int, default=4000)
    p.add_argument("--lr", type=float, default=3e-3)
    p.add_argument("--holdout", type=float, default=0.25)
    p = sub.add_parser("evaluate")
    p.add_argument("--dataset", required=True)
    p.add_argument("--model", required=True)
    p.add_argument("--out", default=None)
    p.add_argument("--all-heldout", action="store_true",
                   help="every shape of --dataset is held out (a separately measured set)")
    p.add_argument("--train-dataset", default=None,
                   help="the training rows, for the power-law baseline of --all-heldout")
    p = sub.add_parser("select")
    p.add_argument("--model", required=True)
    p.add_argument("--m", type=int, required=True)
    p.add_argument("--n", type=int, required=True)
    p.add_argument("--k", type=int, required=True)
    p.add_argument("--validate", action="store_true")
    p = sub.add_parser("share", help="pick the local-product configuration of a rank share")
    p.add_argument("--decomposition", choices=("1d", "2d"), required=True)
    p.add_argument("--size", type=int, required=True)
    p.add_argument("--ranks", type=int, required=True)
    p.add_argument("--grid-rows", type=int, default=None)
    p.add_argument("--grid-cols", type=int, default=None)
    p.add_argument("--auto-model", default=None, help="cost model JSON (default: the committed one)")
    p.add_argument("--auto-validate", action="store_true", help="time the top candidates")
    p.add_argument("--auto-feedback", action="store_true",
                   help="append the timed candidates to the model's dataset")
    p.add_argument("--output", default="results/autotune")
    a = ap.parse_args(argv)
    if a.cmd == "share":
        return share(a)
    if a.cmd == "measure":
        sizes = [s for s in SIZES if s <= a.max_size]
        exclude = []
        if a.exclude_dataset:
            exclude = json.loads(Path(a.exclude_dataset).read_text())["shapes_done"]
        shapes = ([] if a.no_shares else rank_shares(sizes)) + random_shapes(
            a.random_shapes, seed=a.seed, sizes=sizes, exclude=exclude)
        doc = measure(shapes, Path(a.out), time_budget_s=a.time_budget)
        print(f"[autotune] {len(doc['rows'])} measured configurations in {a.out}")
    elif a.cmd == "train":
        doc = json.loads(Path(a.dataset).read_text())
        shapes = sorted({(r["m"], r["n"], r["k"]) for r in doc["rows"]})
        tr, ho = shape_split(shapes, a.holdout)
        model, hist = train(doc["rows"], tr, epochs=a.epochs, lr=a.lr)
        model.meta["heldout_shapes"] = [list(s) for s in ho]
        model.meta["metrics"] = evaluate(doc["rows"], model, tr, ho)
        Path(a.out).parent.mkdir(parents=True, exist_ok=True)
        Path(a.out).write_text(json.dumps(model.to_json()))
        m = model.meta["metrics"]
        print(f"[autotune] trained on {len(tr)} shapes; loss {hist[0]:.4f} -> {hist[-1]:.5f}; "
              f"held-out {m['heldout_shapes']} shapes: MAPE {m['mape']:.3f}, regret mlp "
              f"{m['regret_mlp_mean']:.4f} top3 {m['regret_top3_mean']:.4f} powerlaw "
              f"{m['regret_powerlaw_mean']:.4f} default {m['regret_default_mean']}")
    elif a.cmd == "evaluate":
        doc = json.loads(Path(a.dataset).read_text())
        model = MLPModel.from_json(json.loads(Path(a.model).read_text()))
        tr = [tuple(s) for s in model.meta["train_shapes"]]
        ho = [tuple(s) for s in model.meta["heldout_shapes"]]
        rows = doc["rows"]
        if a.all_heldout:
            ho = sorted({(r["m"], r["n"], r["k"]) for r in rows})
            if set(ho) & set(tr):
                ap.error("--all-heldout: the dataset shares shapes with the model's training set")
            if a.train_dataset:
                rows = rows + json.loads(Path(a.train_dataset).read_text())["rows"]
        res = evaluate(rows, model, tr, ho)
        text = json.dumps(res, indent=2)
        if a.out:
            Path(a.out).write_text(text)
        print(text)
    else:
        model = MLPModel.from_json(json.loads(Path(a.model).read_text()))
        timer = Timer(a.m, a.n, a.k) if a.validate else None
        choice, ranked = select(model, a.m, a.n, a.k, timer=timer)
        print(json.dumps({"choice": choice, "ranking": ranked[:8]}, indent=2))
    return 0


if __name__ == "__main__":
    sys.exit(main())
