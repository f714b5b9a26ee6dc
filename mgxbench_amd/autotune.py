"""Autotuner with a learned cost model: the autotuning/predictive.py counterpart.

The reference autotuner (autotuning/predictive.py) times `benchmark.py` per
candidate (library, grid, BK) over several sizes and fits one runtime model
per candidate: a clipped log-log power law or an lstsq fit on (x^2, x^3)
("extrapolation", predictive.py:100-165), a StandardScaler + Ridge(alpha=0.1)
fit ("ml", :140-165) or a hold-out vote between them ("auto", :213-233).

Here the candidates are the knobs of one rank's part of the SUMMA loop on an
MI355X, which the reference does not have: form (FP64 MFMA, INT8 digit-slice
split with s slices, or INT8 modular split with n moduli), the kernel block
tile, the SUMMA panel width and the columns per launch (a group of whole
panels).  Two targets are modelled:

  resident  every panel already in HBM (as after the panel broadcasts): the
            whole reduction is one launch (group 0) - FP64 still adds one
            partial per panel, a split product ignores the panel width;
  streamed  the loop of mgx_summa_run from pinned host strips: groups of
            1024 or 4096 columns, each group's copy overlapping the previous
            group's product.

Only FP64-accurate candidates are offered, calibrated on the accuracy-stress
inputs (random sign x U[1,2) x 2^U(-30,30), rows / columns x 2^U(-10,10)) as
well as on uniform ones: s = 8 slices (s = 7 is up to 1.9x plain FP64's
componentwise error on stress inputs), and moduli counts whose scaling
exponent at the launch depth is at least MIN_BETA (mgx_split_mod_beta): 15
moduli up to depth 16384, 16 at any depth.  One MLP (3 layers, width 64,
FP32, trained on the GPU by libmgx's own kernels) predicts log runtime from
log-shape features for every candidate and both targets at once; exhaustive
timing of its top-k candidates stays the fallback and the arbiter.

Subcommands (python -m mgxbench_amd.autotune ...):
  measure   time candidates on the GPU -> dataset JSON (checkpointed per shape)
  train     fit the MLP on the training shapes of a dataset -> model JSON
  evaluate  held-out MAPE and regret of the MLP against three methods without
            it: the reference-style per-candidate power law, the best single
            fixed candidate chosen on the training shapes, and the reference's
            default configuration
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

# shape sizes: 4096 ... 32768 in steps of 4096 (the rank shares of the 1x4 /
# 2x2 / 4x1 grids up to N = 32768) plus smaller ones, where the 128 tile and
# the plain FP64 form win
SIZES = (1024, 2048, 3072, 4096, 6144, 8192, 12288, 16384, 20480, 24576, 28672, 32768)
PANELS = (128, 256, 512, 1024)
SLICES = (8,)
MODULI = (15, 16)
# quantisation to >= 51 bits per row / column keeps the modular form within
# plain FP64's normwise and componentwise error on stress inputs (the
# guaranteed per-line exponent; spread-out lines get more)
MIN_BETA = 51
TILES = (_lib.TILE_LARGE, _lib.TILE_SMALL)
STREAM_GROUPS = (1024, 4096)
TARGETS = ("resident", "streamed")
SCHEMA = 2
PCIE_GBPS = 57.0     # host-to-device rate while a copy is active (pinned strips)


@dataclass(frozen=True)
class Candidate:
    form: str        # "fp64" | "int8" | "mod"
    slices: int      # slices (int8), moduli (mod), 0 for fp64
    tile: int        # _lib.TILE_LARGE | _lib.TILE_SMALL
    panel: int       # SUMMA panel width (BK)
    group: int = 0   # columns per launch (whole panels); 0: the whole k in one launch

    def key(self) -> str:
        return f"{self.form}-s{self.slices}-t{self.tile}-b{self.panel}-g{self.group}"


# the reference's default configuration (FP64, MGXBENCH_SUMMA_BK = 256), with
# the library's default launch grouping for each target
DEFAULT = {"resident": Candidate("fp64", 0, _lib.TILE_LARGE, 256, 0),
           "streamed": Candidate("fp64", 0, _lib.TILE_LARGE, 256, 1024)}

FORM_CODE = {"fp64": _lib.FORM_FP64, "int8": _lib.FORM_INT8, "mod": _lib.FORM_INT8_MOD}


def mod_beta(moduli: int, k: int) -> int:
    return int(_lib.lib().mgx_split_mod_beta(moduli, k))


def step_k(c: Candidate, k: int) -> int:
    """Reduction depth of each launch: the whole k (group 0) or a group of
    whole panels covering c.group columns (mgx_split_group_k)."""
    if c.group == 0:
        return k
    return int(_lib.lib().mgx_split_group_k(FORM_CODE[c.form], c.panel, c.group, k))


def candidate_space(k: int, target: str = "resident") -> list[Candidate]:
    if target not in TARGETS:
        raise ValueError(f"unknown target {target!r}")
    groups = (0,) if target == "resident" else STREAM_GROUPS
    out = []
    for panel in PANELS:
        if panel > k:
            continue
        for tile in TILES:
            for group in groups:
                out.append(Candidate("fp64", 0, tile, panel, group))
                for s in SLICES:
                    out.append(Candidate("int8", s, tile, panel, group))
                for nm in MODULI:
                    c = Candidate("mod", nm, tile, panel, group)
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
                 "log2_step_k", "log2_steps", "wave_fill", "streamed", "log2_h2d_ms",
                 "log2_copy_per_launch_ms"]

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


def features(m: int, n: int, k: int, c: Candidate, target: str = "resident") -> list[float]:
    panel = min(c.panel, k)
    panels = math.ceil(k / panel)
    ragged = (k - (panels - 1) * panel) / panel
    sk = step_k(c, k)
    steps = math.ceil(k / sk)
    # INT8 products per output: slice pairs (int8) or moduli (mod)
    products = (c.slices * (c.slices + 1) / 2 if c.form == "int8"
                else float(c.slices) if c.form == "mod" else 0.0)
    streamed = 1.0 if target == "streamed" else 0.0
    # host-to-device time of the strips, and of one launch's group, at the
    # copy rate (zero for resident operands)
    h2d_ms = 8.0 * (m * k + k * n) / (PCIE_GBPS * 1e6)
    per_launch_ms = 8.0 * (m + n) * sk / (PCIE_GBPS * 1e6)
    return [math.log2(m), math.log2(n), math.log2(k), math.log2(panel), math.log2(panels), ragged,
            1.0 if c.form == "int8" else 0.0, c.slices / 8.0, products / 36.0,
            1.0 if c.tile == _lib.TILE_SMALL else 0.0, math.log2(m * n * k) / 3.0,
            math.log2(m * n) / 2.0, 1.0 if c.form == "mod" else 0.0,
            math.log2(sk), math.log2(steps), wave_fill(m, n, c), streamed,
            streamed * math.log2(1.0 + h2d_ms), streamed * math.log2(1.0 + per_launch_ms)]


def _candidate(r: dict) -> Candidate:
    return Candidate(r["form"], r["slices"], r["tile"], r["panel"], r.get("group", 0))


# ---- datasets -------------------------------------------------------------------------------

COLUMNS = ("target", "m", "n", "k", "form", "slices", "tile", "panel", "group", "ms")


def read_dataset(path) -> dict:
    """A dataset file (rows stored as lists under `columns`) with rows as dicts."""
    doc = json.loads(Path(path).read_text())
    if doc.get("schema") != SCHEMA or doc.get("kind") != "mgx-autotune-dataset":
        raise ValueError(f"{path}: not an mgx autotune dataset (schema {SCHEMA})")
    cols = doc["columns"]
    doc["rows"] = [dict(zip(cols, r)) for r in doc["rows"]]
    return doc


def write_dataset(doc: dict, path) -> None:
    path = Path(path)
    packed = {**doc, "columns": list(COLUMNS),
              "rows": [[r[c] for c in COLUMNS] for r in doc["rows"]]}
    tmp = path.with_suffix(".tmp")
    tmp.write_text(json.dumps(packed, separators=(",", ":")))
    os.replace(tmp, path)


# ---- measurement (GPU) --------------------------------------------------------------------

class Timer:
    """Times candidates: resident ones on device buffers, streamed ones from
    pinned host strips, both sized for the largest shape."""

    def __init__(self, max_m: int, max_n: int, max_k: int, seed: int = 1,
                 streamed: bool = True):
        import torch
        import mgxbench_amd as mgx
        self.torch, self.mgx = torch, mgx
        dev = torch.device("cuda", 0)
        self.A = torch.empty(max_m * max_k, dtype=torch.float64, device=dev)
        self.B = torch.empty(max_k * max_n, dtype=torch.float64, device=dev)
        self.C = torch.empty(max_m * max_n, dtype=torch.float64, device=dev)
        mgx.fill(self.A, max_m, max_k, max_k, 1, seed=seed)
        mgx.fill(self.B, max_k, max_n, max_n, 1, seed=seed + 1)
        self.C.zero_()
        self.hA = self.hB = None
        if streamed:
            self.hA = torch.empty(max_m * max_k, dtype=torch.float64, pin_memory=True)
            self.hB = torch.empty(max_k * max_n, dtype=torch.float64, pin_memory=True)
            self.hA.copy_(self.A)
            self.hB.copy_(self.B)
        torch.cuda.synchronize()
        self._warm = set()

    def _resident(self, m, n, k, c, est_ms):
        cfg = _lib.TuneConfig(FORM_CODE[c.form], c.slices, c.tile, c.panel,
                              k if c.group == 0 else c.group)
        # one untimed run always (it may grow the library scratch); then at
        # least two timed ones
        warmup = 1
        reps = int(min(5, max(2, 200 // max(est_ms, 1))))
        ms = ctypes.c_double(0.0)
        # operands are views of the preallocated buffers with tight leading dims
        _lib.call("mgx_time_local_product", ctypes.byref(cfg), self.A.data_ptr(),
                  self.B.data_ptr(), self.C.data_ptr(), m, n, k, k, n, n, None, 0, warmup, reps,
                  ctypes.byref(ms))
        return ms.value

    def _streamed(self, m, n, k, c):
        if self.hA is None:
            raise ValueError("this timer holds no host strips")
        kw = dict(m=m, n=n, k=k, panel=c.panel, form=FORM_CODE[c.form], slices=c.slices or 8,
                  split_k=c.group, tile=c.tile)
        if (c.form, c.tile) not in self._warm:     # first use loads the kernels
            self.mgx.summa_run(self.hA.data_ptr(), k, self.hB.data_ptr(), n, self.C, n, **kw)
            self._warm.add((c.form, c.tile))
        # the faster of two loops (a short loop varies by a few percent); a
        # loop copying for more than 60 ms runs once
        loops = 2 if 8.0 * (m * k + k * n) / (PCIE_GBPS * 1e6) < 60.0 else 1
        return min(self.mgx.summa_run(self.hA.data_ptr(), k, self.hB.data_ptr(), n, self.C, n,
                                      **kw)["seconds"] for _ in range(loops)) * 1e3

    def time(self, m: int, n: int, k: int, c: Candidate, target: str = "resident",
             est_ms: float | None = None) -> float:
        if target == "streamed":
            return self._streamed(m, n, k, c)
        est = est_ms if est_ms is not None else 2.0 * m * n * k / 50e12 * 1e3
        return self._resident(m, n, k, c, est)


def measure(shapes, out_path: Path, targets=TARGETS, log=print,
            time_budget_s: float | None = None) -> dict:
    """Times every candidate of every shape and target; resumes from out_path
    if present.  Resident split products do not depend on the panel width, so
    each is timed once and its row repeated for every panel.  With a time
    budget, stops before a shape that would likely overrun it (the dataset
    stays checkpointed; a later call resumes)."""
    out_path = Path(out_path)
    doc = {"schema": SCHEMA, "kind": "mgx-autotune-dataset", "rows": [], "shapes_done": []}
    if out_path.exists():
        doc = read_dataset(out_path)
    done = {tuple(s) for s in doc["shapes_done"]}
    todo = [tuple(s) for s in shapes if tuple(s) not in done]
    if not todo:
        return doc
    timer = Timer(max(s[0] for s in todo), max(s[1] for s in todo), max(s[2] for s in todo),
                  streamed="streamed" in targets)
    start = time.time()
    worst = 0.0
    for i, (m, n, k) in enumerate(todo):
        t0 = time.time()
        if time_budget_s is not None and (t0 - start) + 1.5 * worst > time_budget_s:
            log(f"[autotune] time budget reached after {i} of {len(todo)} shapes")
            break
        count = 0
        for target in targets:
            shared = {}
            for c in candidate_space(k, target):
                if target == "resident" and c.form != "fp64":
                    key = (c.form, c.slices, c.tile)
                    if key not in shared:
                        shared[key] = timer.time(m, n, k, c, target)
                    ms = shared[key]
                else:
                    ms = timer.time(m, n, k, c, target)
                doc["rows"].append({"target": target, "m": m, "n": n, "k": k, **asdict(c),
                                    "ms": ms})
                count += 1
        doc["shapes_done"].append([m, n, k])
        write_dataset(doc, out_path)
        worst = max(worst, time.time() - t0)
        log(f"[autotune] shape {i + 1}/{len(todo)} m={m} n={n} k={k} "
            f"({count} rows, {time.time() - t0:.1f} s)")
    return doc


# ---- model ----------------------------------------------------------------------------------

def shape_split(shapes, holdout_frac: float = 0.25, seed: int = 11):
    """Deterministic split of shapes into (train, held-out)."""
    keyed = sorted(shapes, key=lambda s: hashlib.sha256(f"{seed}:{s}".encode()).hexdigest())
    n_hold = max(1, int(round(holdout_frac * len(keyed))))
    return sorted(keyed[n_hold:]), sorted(keyed[:n_hold])


def _rows_xy(rows):
    X = np.array([features(r["m"], r["n"], r["k"], _candidate(r), r.get("target", "resident"))
                  for r in rows], dtype=np.float64)
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
            raise ValueError(f"not an mgx MLP cost model (schema {SCHEMA})")
        if doc.get("features") != FEATURE_NAMES:
            raise ValueError("the cost model was trained on other features")
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


def _training_table(rows, train_shapes):
    """{target: {shape: {candidate: ms}}} over the training shapes."""
    tset = {tuple(s) for s in train_shapes}
    table: dict = {}
    for r in rows:
        s = (r["m"], r["n"], r["k"])
        if s in tset:
            table.setdefault(r.get("target", "resident"), {}).setdefault(s, {})[_candidate(r)] = \
                r["ms"]
    return table


class PowerLawModel:
    """Reference-style baseline: per candidate and target, log t = a + b
    log(2mnk) with the exponent clipped to [0, 4] (the 'extrapolation' fit of
    predictive.py:100-127, applied to the problem's flop count)."""

    def __init__(self, rows, train_shapes):
        self.fit = {}
        for target, shapes in _training_table(rows, train_shapes).items():
            by: dict = {}
            for (m, n, k), cands in shapes.items():
                for c, ms in cands.items():
                    by.setdefault(c, []).append((math.log(2.0 * m * n * k), math.log(ms)))
            for c, pts in by.items():
                x = np.array([p[0] for p in pts])
                y = np.array([p[1] for p in pts])
                b = float(np.clip(np.polyfit(x, y, 1)[0], 0.0, 4.0)) \
                    if len(pts) >= 2 and np.ptp(x) > 0 else 1.0
                self.fit[(target, c)] = (float(np.mean(y - b * x)), b)

    def predict_log_ms(self, m, n, k, cands, target: str = "resident") -> np.ndarray:
        x = math.log(2.0 * m * n * k)
        out = []
        for c in cands:
            a, b = self.fit.get((target, c), (0.0, 1.0))
            out.append(a + b * x)
        return np.array(out)


def best_fixed(rows, train_shapes) -> dict:
    """{target: candidate} - the single configuration a tuner without a model
    would ship: of the candidates timed on every training shape, the one with
    the lowest mean slowdown t / t_best there."""
    out = {}
    for target, shapes in _training_table(rows, train_shapes).items():
        everywhere = set.intersection(*(set(c) for c in shapes.values()))
        slow = {c: np.mean([cands[c] / min(cands.values()) for cands in shapes.values()])
                for c in everywhere}
        out[target] = min(slow, key=lambda c: (slow[c], c.key()))
    return out


def evaluate(rows, model, train_shapes, heldout_shapes, topk: int = 3) -> dict:
    """Held-out prediction error and regret (measured t(pick) / t(best) - 1)
    per target, for the MLP and the methods without it, over the current
    candidate space (rows of retired candidates are skipped)."""
    by: dict = {}
    space: dict = {}
    for r in rows:
        target = r.get("target", "resident")
        if (r["k"], target) not in space:
            space[(r["k"], target)] = set(candidate_space(r["k"], target))
        if _candidate(r) in space[(r["k"], target)]:
            by.setdefault((target, (r["m"], r["n"], r["k"])), []).append(r)
    power = PowerLawModel(rows, train_shapes)
    fixed = best_fixed(rows, train_shapes)
    out = {"targets": {}}
    for target in TARGETS:
        ape, reg, per_shape = [], {k: [] for k in ("mlp", "top", "powerlaw", "fixed",
                                                    "default")}, []
        for s in heldout_shapes:
            rs = by.get((target, tuple(s)), [])
            if not rs:
                continue
            cands = [_candidate(r) for r in rs]
            t = np.array([r["ms"] for r in rs])
            X = np.array([features(*s, c, target) for c in cands])
            pred = np.exp(model.predict_log_ms(X))
            ape.extend(np.abs(pred - t) / t)
            best = t.min()
            order = np.argsort(pred)
            picks = {"mlp": t[order[0]], "top": t[order[:topk]].min(),
                     "powerlaw": t[int(np.argmin(power.predict_log_ms(*s, cands, target)))]}
            for name, c0 in (("fixed", fixed.get(target)), ("default", DEFAULT[target])):
                hit = [ti for c, ti in zip(cands, t) if c == c0]
                if hit:
                    picks[name] = hit[0]
            for name, v in picks.items():
                reg[name].append(v / best - 1)
            per_shape.append({"shape": list(s), "best": cands[int(np.argmin(t))].key(),
                              "mlp_pick": cands[int(order[0])].key(),
                              **{f"regret_{k}": v / best - 1 for k, v in picks.items()}})

        def mean(x):
            return float(np.mean(x)) if len(x) else None

        def worst(x):
            return float(np.max(x)) if len(x) else None

        out["targets"][target] = {
            "heldout_shapes": len(per_shape), "heldout_configs": len(ape),
            "mape": mean(ape), "median_ape": float(np.median(ape)) if ape else None,
            "regret_mlp_mean": mean(reg["mlp"]), "regret_mlp_max": worst(reg["mlp"]),
            f"regret_top{topk}_mean": mean(reg["top"]), f"regret_top{topk}_max": worst(reg["top"]),
            "regret_powerlaw_mean": mean(reg["powerlaw"]),
            "regret_powerlaw_max": worst(reg["powerlaw"]),
            "fixed_candidate": fixed[target].key() if target in fixed else None,
            "regret_fixed_mean": mean(reg["fixed"]), "regret_fixed_max": worst(reg["fixed"]),
            "regret_default_mean": mean(reg["default"]),
            "per_shape": per_shape,
        }
    return out


def select(model: MLPModel, m: int, n: int, k: int, target: str = "streamed", topk: int = 3,
           timer: Timer | None = None):
    """Candidates ranked by predicted time; with a timer, the top k are timed
    and the fastest measured one is chosen."""
    cands = candidate_space(k, target)
    X = np.array([features(m, n, k, c, target) for c in cands])
    pred = np.exp(model.predict_log_ms(X))
    order = np.argsort(pred)
    ranked = [{"candidate": asdict(cands[i]), "key": cands[i].key(), "target": target,
               "predicted_ms": float(pred[i])} for i in order]
    choice = ranked[0]
    if timer is not None:
        for r in ranked[:topk]:
            r["measured_ms"] = timer.time(m, n, k, Candidate(**r["candidate"]), target,
                                          r["predicted_ms"])
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
    """Pick (form, slices, tile, panel, group) for one rank's part of a 1D /
    2D run (the `share` subcommand)."""
    from .grid import resolve_grid
    path = Path(args.auto_model) if args.auto_model else default_model_path()
    if not path.is_file():
        print(f"no cost model at {path}; train one with `python -m mgxbench_amd.autotune "
              "train`", file=sys.stderr)
        return 2
    model = MLPModel.from_json(json.loads(path.read_text()))
    grid = resolve_grid(args.ranks, args.grid_rows, args.grid_cols) if args.decomposition == "2d" else None
    m, n, k = share_shape(args.decomposition, args.size, args.ranks, grid)
    timer = Timer(m, n, k, streamed=args.target == "streamed") if args.auto_validate else None
    choice, ranked = select(model, m, n, k, args.target, timer=timer)
    out = {"schema": SCHEMA, "decomposition": args.decomposition, "size": args.size,
           "ranks": args.ranks, "share": [m, n, k], "target": args.target, "choice": choice,
           "ranking": ranked[:10], "model": str(path)}
    Path(args.output).mkdir(parents=True, exist_ok=True)
    dest = Path(args.output) / "autotuning_prediction.json"
    dest.write_text(json.dumps(out, indent=2))
    print(f"choice for share m={m} n={n} k={k} ({args.target}): {choice['key']} "
          f"(predicted {choice['predicted_ms']:.2f} ms)")
    print(f"prediction written to {dest}")
    if args.auto_feedback and timer is not None:
        ds = path.with_name("dataset.json")
        doc = read_dataset(ds) if ds.is_file() else {"schema": SCHEMA,
                                                     "kind": "mgx-autotune-dataset",
                                                     "rows": [], "shapes_done": []}
        for r in ranked:
            if "measured_ms" in r:
                doc["rows"].append({"target": args.target, "m": m, "n": n, "k": k,
                                    **r["candidate"], "ms": r["measured_ms"]})
        write_dataset(doc, ds)
    return 0


def _summary(res: dict) -> str:
    parts = []
    for target, m in res["targets"].items():
        if not m["heldout_shapes"]:
            continue
        parts.append(f"{target}: {m['heldout_shapes']} shapes, MAPE {m['mape']:.3f}, regret "
                     f"mlp {m['regret_mlp_mean']:.4f} top3 {m['regret_top3_mean']:.4f} "
                     f"powerlaw {m['regret_powerlaw_mean']:.4f} fixed "
                     f"{m['regret_fixed_mean']:.4f} ({m['fixed_candidate']}) default "
                     f"{m['regret_default_mean']:.4f}")
    return "; ".join(parts)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("measure")
    p.add_argument("--out", required=True)
    p.add_argument("--random-shapes", type=int, default=84)
    p.add_argument("--max-size", type=int, default=max(SIZES))
    p.add_argument("--targets", nargs="+", choices=TARGETS, default=list(TARGETS))
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
                   help="the training rows, for the baselines of --all-heldout")
    p = sub.add_parser("select")
    p.add_argument("--model", required=True)
    p.add_argument("--m", type=int, required=True)
    p.add_argument("--n", type=int, required=True)
    p.add_argument("--k", type=int, required=True)
    p.add_argument("--target", choices=TARGETS, default="streamed")
    p.add_argument("--validate", action="store_true")
    p = sub.add_parser("share", help="pick the configuration of a rank's share")
    p.add_argument("--decomposition", choices=("1d", "2d"), required=True)
    p.add_argument("--size", type=int, required=True)
    p.add_argument("--ranks", type=int, required=True)
    p.add_argument("--grid-rows", type=int, default=None)
    p.add_argument("--grid-cols", type=int, default=None)
    p.add_argument("--target", choices=TARGETS, default="streamed")
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
            exclude = read_dataset(a.exclude_dataset)["shapes_done"]
        shapes = ([] if a.no_shares else rank_shares(sizes)) + random_shapes(
            a.random_shapes, seed=a.seed, sizes=sizes, exclude=exclude)
        doc = measure(shapes, Path(a.out), targets=a.targets, time_budget_s=a.time_budget)
        print(f"[autotune] {len(doc['rows'])} measured configurations in {a.out}")
    elif a.cmd == "train":
        doc = read_dataset(a.dataset)
        shapes = sorted({(r["m"], r["n"], r["k"]) for r in doc["rows"]})
        tr, ho = shape_split(shapes, a.holdout)
        model, hist = train(doc["rows"], tr, epochs=a.epochs, lr=a.lr)
        model.meta["heldout_shapes"] = [list(s) for s in ho]
        model.meta["metrics"] = evaluate(doc["rows"], model, tr, ho)
        Path(a.out).parent.mkdir(parents=True, exist_ok=True)
        Path(a.out).write_text(json.dumps(model.to_json()))
        print(f"[autotune] trained on {len(tr)} shapes; loss {hist[0]:.4f} -> {hist[-1]:.5f}; "
              f"held-out {_summary(model.meta['metrics'])}")
    elif a.cmd == "evaluate":
        doc = read_dataset(a.dataset)
        model = MLPModel.from_json(json.loads(Path(a.model).read_text()))
        tr = [tuple(s) for s in model.meta["train_shapes"]]
        ho = [tuple(s) for s in model.meta["heldout_shapes"]]
        rows = doc["rows"]
        if a.all_heldout:
            ho = sorted({(r["m"], r["n"], r["k"]) for r in rows})
            if set(ho) & set(tr):
                ap.error("--all-heldout: the dataset shares shapes with the model's training set")
            if a.train_dataset:
                rows = rows + read_dataset(a.train_dataset)["rows"]
        res = evaluate(rows, model, tr, ho)
        text = json.dumps(res, indent=2)
        if a.out:
            Path(a.out).write_text(text)
        print(_summary(res))
    else:
        model = MLPModel.from_json(json.loads(Path(a.model).read_text()))
        timer = Timer(a.m, a.n, a.k, streamed=a.target == "streamed") if a.validate else None
        choice, ranked = select(model, a.m, a.n, a.k, a.target, timer=timer)
        print(json.dumps({"choice": choice, "ranking": ranked[:8]}, indent=2))
    return 0


if __name__ == "__main__":
    sys.exit(main())
