"""The committed learned cost model against its committed datasets.

mgxbench_amd/models/dataset.json holds the autotuner's measured times (every
FP64-accurate candidate of every shape, for both targets: resident operands
and the streamed SUMMA loop, timed on an MI355X);
mgxbench_amd/models/mlp_cost_model.json is the MLP trained on the GPU on the
training shapes of that dataset.  mgxbench_amd/models/heldout_dataset.json
holds 200 further shapes, drawn outside the training dataset and measured
separately.

On held-out shapes, for each target, the configuration the network ranks
first must beat every method without the network: the reference-style
per-candidate power law (autotuning/predictive.py:100-127), the best single
fixed candidate chosen on the training shapes, and the reference's default
configuration (FP64, BK = 256).

The CPU tests evaluate the network with the oracle's double-precision
forward pass; the GPU test with the library's own FP32 inference kernel.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from mgxbench_amd import autotune as at

MODELS = Path(__file__).resolve().parents[1] / "mgxbench_amd" / "models"


def _load():
    rows = at.read_dataset(MODELS / "dataset.json")["rows"]
    model = at.MLPModel.from_json(json.loads((MODELS / "mlp_cost_model.json").read_text()))
    tr = [tuple(s) for s in model.meta["train_shapes"]]
    ho = [tuple(s) for s in model.meta["heldout_shapes"]]
    return rows, model, tr, ho


class _OracleForward:
    """The committed network evaluated in double precision by the oracle."""

    def __init__(self, model, orc):
        self.model, self.orc = model, orc

    def predict_log_ms(self, X):
        Xs = (np.asarray(X, dtype=np.float64) - self.model.mean) / self.model.std
        return self.orc.mlp_forward(self.model.theta.astype(np.float64), Xs)


def _beats_the_methods_without_it(t):
    assert t["regret_top3_mean"] <= t["regret_mlp_mean"] + 1e-12
    assert t["regret_mlp_mean"] < t["regret_powerlaw_mean"]
    assert t["regret_mlp_mean"] < t["regret_fixed_mean"]
    assert t["regret_mlp_mean"] < t["regret_default_mean"]


def _check(res, stored):
    for target in at.TARGETS:
        t = res["targets"][target]
        assert t["heldout_shapes"] >= 30
        assert t["regret_mlp_mean"] <= 0.03
        assert t["mape"] <= 0.06
        _beats_the_methods_without_it(t)
        # agrees with the metrics stored at training time (FP32 on the GPU)
        s = stored["targets"][target]
        assert t["mape"] == pytest.approx(s["mape"], abs=0.01)
        assert t["regret_mlp_mean"] == pytest.approx(s["regret_mlp_mean"], abs=0.01)
        assert t["fixed_candidate"] == s["fixed_candidate"]


def test_dataset_covers_every_candidate_of_every_shape():
    rows, model, tr, ho = _load()
    shapes = {(r["m"], r["n"], r["k"]) for r in rows}
    assert set(tr) | set(ho) == shapes and not set(tr) & set(ho)
    for target in at.TARGETS:
        for (m, n, k) in shapes:
            keys = {at._candidate(r) for r in rows
                    if (r["m"], r["n"], r["k"]) == (m, n, k) and r["target"] == target}
            assert set(at.candidate_space(k, target)) <= keys, (target, m, n, k)
    assert all(r["ms"] > 0 for r in rows)
    assert model.in_dim == len(at.FEATURE_NAMES)
    # resident split products ignore the panel width: one timing per panel set
    for (m, n, k) in list(shapes)[:20]:
        t = {r["panel"]: r["ms"] for r in rows if (r["m"], r["n"], r["k"]) == (m, n, k)
             and r["target"] == "resident" and r["form"] == "mod" and r["slices"] == 15
             and r["tile"] == 1}
        assert len(set(t.values())) == 1


def test_heldout_regret_with_the_oracle_forward(orc):
    rows, model, tr, ho = _load()
    res = at.evaluate(rows, _OracleForward(model, orc), tr, ho)
    _check(res, model.meta["metrics"])


@pytest.mark.gpu
def test_heldout_regret_with_the_gpu_forward():
    rows, model, tr, ho = _load()
    res = at.evaluate(rows, model, tr, ho)
    _check(res, model.meta["metrics"])


def test_separate_heldout_set_with_the_oracle_forward(orc):
    rows, model, tr, _ = _load()
    held = at.read_dataset(MODELS / "heldout_dataset.json")["rows"]
    ho = sorted({(r["m"], r["n"], r["k"]) for r in held})
    assert len(ho) == 200 and not set(ho) & set(tr)
    res = at.evaluate(rows + held, _OracleForward(model, orc), tr, ho)
    for target in at.TARGETS:
        t = res["targets"][target]
        assert t["heldout_shapes"] == 200
        assert t["regret_mlp_mean"] <= 0.02
        assert t["mape"] <= 0.05
        _beats_the_methods_without_it(t)
        stored = model.meta["separate_heldout"][target]
        assert t["regret_mlp_mean"] == pytest.approx(stored["regret_mlp_mean"], abs=1e-6)
