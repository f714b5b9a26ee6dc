"""Autotuner logic on the CPU: candidate space, features, shape split, the
reference-style power-law baseline, regret bookkeeping (GPU-free; the MLP
itself is exercised by tests/test_gpu_mlp.py and test_gpu_autotune.py)."""
import math

import numpy as np
import pytest

from mgxbench_amd import autotune as at
from mgxbench_amd import _lib


def test_candidate_space():
    c = at.candidate_space(16384)
    mods = [x for x in c if x.form == "mod"]
    assert len(c) == len(at.PANELS) * len(at.TILES) * (1 + len(at.SLICES)) + len(mods)
    assert at.DEFAULT in c
    # only FP64-accurate candidates (calibrated on stress inputs): s = 8, and
    # moduli with beta >= MIN_BETA = 52 at the depth of each split product (a
    # group of whole panels, at most 4096 deep by default)
    assert {x.slices for x in c if x.form == "int8"} == {8}
    assert mods and all(at.mod_beta(x.slices, at.step_k(x, 16384)) >= at.MIN_BETA for x in mods)
    assert {x.slices for x in mods} == {15, 16}
    assert at.mod_beta(15, 4096) == 52 and at.mod_beta(15, 8192) == 51
    assert at.mod_beta(14, 1024) == 49
    assert {x.slices for x in at.candidate_space(1024) if x.form == "mod"} == {15, 16}
    small = at.candidate_space(200)     # panels wider than K are dropped
    assert {x.panel for x in small} == {128}
    assert len({x.key() for x in c}) == len(c)


def test_step_depth_is_one_panel_or_a_whole_panel_group():
    L = _lib.TILE_LARGE
    # FP64 groups of whole panels covering 2048 columns, one fused launch each
    assert at.step_k(at.Candidate("fp64", 0, L, 256), 16384) == 2048
    assert at.step_k(at.Candidate("fp64", 0, L, 1000), 16384) == 3000
    assert at.step_k(at.Candidate("fp64", 0, L, 256), 100) == 100
    assert at.step_k(at.Candidate("mod", 15, L, 256), 16384) == 4096
    assert at.step_k(at.Candidate("int8", 7, L, 1000), 16384) == 5000    # whole panels
    assert at.step_k(at.Candidate("mod", 15, L, 1024), 3000) == 3000     # capped at k
    lib = _lib.lib()
    assert lib.mgx_split_group_k(_lib.FORM_INT8_MOD, 128, 300, 16384) == 384
    assert lib.mgx_split_group_k(_lib.FORM_INT8_MOD, 0, 0, 16384) == -1
    assert lib.mgx_split_group_k(99, 128, 0, 16384) == -1


def test_rank_shares_are_the_grid_shares():
    shapes = at.rank_shares([16384])
    assert shapes == [(16384, 4096, 16384), (8192, 8192, 16384), (4096, 16384, 16384)]
    assert len(at.rank_shares()) == 24
    r = at.random_shapes(10)
    assert len(set(r)) == 10 and r == at.random_shapes(10)


def test_features_are_finite_and_distinguish_candidates():
    m, n, k = 8192, 8192, 16384
    feats = {c.key(): at.features(m, n, k, c) for c in at.candidate_space(k)}
    assert all(len(f) == len(at.FEATURE_NAMES) for f in feats.values())
    assert all(math.isfinite(x) for f in feats.values() for x in f)
    assert len({tuple(f) for f in feats.values()}) == len(feats)
    f = at.features(4096, 4096, 4096 + 100, at.Candidate("fp64", 0, _lib.TILE_LARGE, 1024))
    assert f[at.FEATURE_NAMES.index("log2_panels")] == pytest.approx(math.log2(5))
    assert f[at.FEATURE_NAMES.index("ragged_frac")] == pytest.approx(100 / 1024)


def test_shape_split_is_deterministic_and_disjoint():
    shapes = at.rank_shares() + at.random_shapes(40)
    tr, ho = at.shape_split(shapes, 0.25)
    assert tr == at.shape_split(shapes, 0.25)[0]
    assert not set(tr) & set(ho) and len(tr) + len(ho) == len(shapes)
    assert len(ho) == 16


def _synthetic_rows(shapes):
    # time = flops / rate(c) with candidate-dependent rates: an exact power law
    rows = []
    for (m, n, k) in shapes:
        for c in at.candidate_space(k):
            rate = 40.0 + (30.0 if c.form == "int8" else 0.0) - 2.0 * c.slices \
                + (5.0 if c.tile == _lib.TILE_LARGE else 0.0) + c.panel / 256.0
            rows.append({"m": m, "n": n, "k": k, "form": c.form, "slices": c.slices,
                         "tile": c.tile, "panel": c.panel, "ms": 2.0 * m * n * k / rate / 1e9})
    return rows


def test_power_law_baseline_recovers_exact_power_law():
    shapes = at.rank_shares()
    rows = _synthetic_rows(shapes)
    base = at.PowerLawModel(rows, shapes)
    m, n, k = 8192, 8192, 16384
    cands = at.candidate_space(k)
    pred = np.exp(base.predict_log_ms(m, n, k, cands))
    truth = {(r["form"], r["slices"], r["tile"], r["panel"]): r["ms"] for r in rows
             if (r["m"], r["n"], r["k"]) == (m, n, k)}
    for c, p in zip(cands, pred):
        assert p == pytest.approx(truth[(c.form, c.slices, c.tile, c.panel)], rel=1e-9)


class _OracleModel:
    """Predicts the true log time (perfect model) or a perturbed one."""

    def __init__(self, rows, bias=None):
        self.t = {(r["m"], r["n"], r["k"], r["form"], r["slices"], r["tile"], r["panel"]): r["ms"]
                  for r in rows}
        self.bias = bias

    def predict_log_ms(self, X):
        raise AssertionError("replaced per test")


def test_evaluate_regret_bookkeeping(monkeypatch):
    shapes = at.rank_shares()[:6]
    rows = _synthetic_rows(shapes)
    tr, ho = shapes[:4], shapes[4:]
    table = {}
    for r in rows:
        c = at.Candidate(r["form"], r["slices"], r["tile"], r["panel"])
        table[tuple(at.features(r["m"], r["n"], r["k"], c))] = math.log(r["ms"])

    class Perfect:
        def predict_log_ms(self, X):
            return np.array([table[tuple(x)] for x in np.asarray(X).tolist()])

    res = at.evaluate(rows, Perfect(), tr, ho, topk=3)
    assert res["heldout_shapes"] == 2
    assert res["mape"] == pytest.approx(0.0, abs=1e-12)
    assert res["regret_mlp_mean"] == 0.0 and res["regret_top3_mean"] == 0.0
    assert res["regret_powerlaw_mean"] == pytest.approx(0.0, abs=1e-9)
    assert res["regret_default_mean"] > 0.0     # fp64/large/256 is not the fastest here

    class Inverted:
        def predict_log_ms(self, X):
            return -np.array([table[tuple(x)] for x in np.asarray(X).tolist()])

    bad = at.evaluate(rows, Inverted(), tr, ho, topk=3)
    assert bad["regret_mlp_mean"] > 0.0


def test_share_shape():
    assert at.share_shape("1d", 32768, 4) == (8192, 32768, 32768)
    assert at.share_shape("2d", 16384, 4, (2, 2)) == (8192, 8192, 16384)


def test_random_shapes_exclusion_for_heldout_sets():
    from mgxbench_amd import autotune as at
    train = at.rank_shares() + at.random_shapes(56)
    held = at.random_shapes(200, seed=8, exclude=train)
    assert len(set(held)) == 200 and not set(held) & set(train)
    assert held == at.random_shapes(200, seed=8, exclude=train)       # deterministic
    assert at.random_shapes(56) == train[24:]                          # default stream unchanged
    space = len(at.SIZES) ** 3
    with pytest.raises(ValueError):                                    # one shape of the space excluded
        at.random_shapes(space, exclude=at.random_shapes(1))
    assert len(at.random_shapes(space, exclude=[(1, 2, 3)])) == space  # outside the space: no effect
