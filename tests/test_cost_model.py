"""The committed learned cost model against its committed dataset.

mgxbench_amd/models/dataset.json holds the autotuner's measured local-product
times (every FP64-accurate candidate of every shape, timed on an MI355X);
mgxbench_amd/models/mlp_cost_model.json is the MLP trained on the GPU on the
training shapes of that dataset.  On the held-out shapes, the configuration
the network ranks first must be within 5% of the exhaustive best on average,
and it must beat the method without the network: the reference's fixed
default configuration (FP64, BK = 256).  The power-law per-candidate fit of
autotuning/predictive.py:100-127 is reported beside it.

The CPU test evaluates the network with the oracle's double-precision
forward pass; the GPU test with the library's own FP32 inference kernel.

mgxbench_amd/models/heldout_dataset.json holds 200 further shapes, drawn
outside the training dataset and measured separately (every candidate of
each): there the network's first pick must stay within 5% of the best on
every shape, and its mean regret must beat the power-law fit's.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from mgxbench_amd import autotune as at

MODELS = Path(__file__).resolve().parents[1] / "mgxbench_amd" / "models"


def _load():
    doc = json.loads((MODELS / "dataset.json").read_text())
    model = at.MLPModel.from_json(json.loads((MODELS / "mlp_cost_model.json").read_text()))
    tr = [tuple(s) for s in model.meta["train_shapes"]]
    ho = [tuple(s) for s in model.meta["heldout_shapes"]]
    return doc["rows"], model, tr, ho


class _OracleForward:
    """The committed network evaluated in double precision by the oracle."""

    def __init__(self, model, orc):
        self.model, self.orc = model, orc

    def predict_log_ms(self, X):
        Xs = (np.asarray(X, dtype=np.float64) - self.model.mean) / self.model.std
        return self.orc.mlp_forward(self.model.theta.astype(np.float64), Xs)


def _check(res, stored):
    assert res["heldout_shapes"] >= 8
    assert res["regret_mlp_mean"] <= 0.05
    assert res["regret_top3_mean"] <= res["regret_mlp_mean"] + 1e-12
    # what the network gains over the method without it
    assert res["regret_mlp_mean"] < res["regret_default_mean"]
    # agrees with the metrics stored at training time (FP32 on the GPU)
    assert res["mape"] == pytest.approx(stored["mape"], abs=0.01)
    assert res["regret_mlp_mean"] == pytest.approx(stored["regret_mlp_mean"], abs=0.01)


def test_dataset_covers_every_candidate_of_every_shape():
    rows, model, tr, ho = _load()
    shapes = {(r["m"], r["n"], r["k"]) for r in rows}
    assert set(tr) | set(ho) == shapes and not set(tr) & set(ho)
    for (m, n, k) in shapes:
        keys = {(r["form"], r["slices"], r["tile"], r["panel"]) for r in rows
                if (r["m"], r["n"], r["k"]) == (m, n, k)}
        want = {(c.form, c.slices, c.tile, c.panel) for c in at.candidate_space(k)}
        assert want <= keys, (m, n, k)
    assert all(r["ms"] > 0 for r in rows)
    assert model.in_dim == len(at.FEATURE_NAMES)


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
    doc = json.loads((MODELS / "heldout_dataset.json").read_text())
    ho = sorted({(r["m"], r["n"], r["k"]) for r in doc["rows"]})
    assert len(ho) == 200 and not set(ho) & set(tr)
    res = at.evaluate(rows + doc["rows"], _OracleForward(model, orc), tr, ho)
    assert res["heldout_shapes"] == 200
    assert res["regret_mlp_max"] <= 0.05
    assert res["regret_mlp_mean"] < res["regret_powerlaw_mean"]
    assert res["regret_top3_mean"] <= res["regret_mlp_mean"] + 1e-12
    assert res["mape"] <= 0.06
