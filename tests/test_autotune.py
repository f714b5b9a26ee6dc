"""Autotuner logic on the CPU: candidate spaces of both targets, features,
shape split, the reference-style power-law baseline, the best fixed
candidate, regret bookkeeping and the dataset format (GPU-free; the MLP
itself is exercised by tests/test_gpu_mlp.py and tests/test_gpu_variants.py)."""
import math

import numpy as np
import pytest

from mgxbench_amd import autotune as at
from mgxbench_amd import _lib


def test_candidate_space():
    c = at.candidate_space(16384)
    mods = [x for x in c if x.form == "mod"]
    assert len(c) == len(at.PANELS) * len(at.TILES) * (1 + len(at.SLICES)) + len(mods)
    assert at.DEFAULT["resident"] in c and all(x.group == 0 for x in c)
    # only FP64-accurate candidates (calibrated on stress inputs): s = 8, and
    # moduli with beta >= MIN_BETA = 51 at the depth of each launch (the whole
    # k for resident operands, a group of whole panels when streamed)
    assert {x.slices for x in c if x.form == "int8"} == {8}
    assert mods and all(at.mod_beta(x.slices, at.step_k(x, 16384)) >= at.MIN_BETA for x in mods)
    assert {x.slices for x in mods} == {15, 16}
    assert at.mod_beta(15, 4096) == 52 and at.mod_beta(15, 8192) == 51
    assert at.mod_beta(15, 16384) == 51 and at.mod_beta(15, 32768) == 50
    assert at.mod_beta(14, 1024) == 49
    assert {x.slices for x in at.candidate_space(32768) if x.form == "mod"} == {16}
    assert {x.slices for x in at.candidate_space(1024) if x.form == "mod"} == {15, 16}
    small = at.candidate_space(200)     # panels wider than K are dropped
    assert {x.panel for x in small} == {128}
    assert len({x.key() for x in c}) == len(c)
    streamed = at.candidate_space(32768, "streamed")
    assert {x.group for x in streamed} == set(at.STREAM_GROUPS)
    assert at.DEFAULT["streamed"] in streamed
    # streamed launches are at most 4096 deep: 15 moduli stay FP64-accurate
    assert {x.slices for x in streamed if x.form == "mod"} == {15, 16}
    assert len(streamed) == 2 * len(at.PANELS) * len(at.TILES) * 4
    with pytest.raises(ValueError):
        at.candidate_space(1024, "elsewhere")


def test_step_depth_is_the_whole_k_or_a_whole_panel_group():
    L = _lib.TILE_LARGE
    # group 0: the whole reduction in one launch
    assert at.step_k(at.Candidate("fp64", 0, L, 256), 16384) == 16384
    assert at.step_k(at.Candidate("mod", 15, L, 1024, 0), 700) == 700
    # streamed groups: whole panels covering the group's columns, capped at k
    assert at.step_k(at.Candidate("fp64", 0, L, 256, 1024), 16384) == 1024
    assert at.step_k(at.Candidate("fp64", 0, L, 1000, 1024), 16384) == 2000
    assert at.step_k(at.Candidate("mod", 15, L, 256, 4096), 16384) == 4096
    assert at.step_k(at.Candidate("int8", 8, L, 1000, 4096), 16384) == 5000
    assert at.step_k(at.Candidate("mod", 15, L, 1024, 4096), 3000) == 3000
    lib = _lib.lib()
    assert lib.mgx_split_group_k(_lib.FORM_INT8_MOD, 128, 300, 16384) == 384
    assert lib.mgx_split_group_k(_lib.FORM_INT8_MOD, 0, 0, 16384) == -1
    assert lib.mgx_split_group_k(99, 128, 0, 16384) == -1


def test_rank_shares_are_the_grid_shares():
    shapes = at.rank_shares([16384])
    assert shapes == [(16384, 4096, 16384), (8192, 8192, 16384), (4096, 16384, 16384)]
    assert len(at.rank_shares()) == 36
    r = at.random_shapes(10)
    assert len(set(r)) == 10 and r == at.random_shapes(10)


def test_features_are_finite_and_distinguish_candidates():
    m, n, k = 8192, 8192, 16384
    for target in at.TARGETS:
        feats = {c.key(): at.features(m, n, k, c, target) for c in at.candidate_space(k, target)}
        assert all(len(f) == len(at.FEATURE_NAMES) for f in feats.values())
        assert all(math.isfinite(x) for f in feats.values() for x in f)
        assert len({tuple(f) for f in feats.values()}) == len(feats)
    assert len(at.FEATURE_NAMES) <= 24          # MLP_MAX_IN
    f = at.features(4096, 4096, 4096 + 100, at.Candidate("fp64", 0, _lib.TILE_LARGE, 1024))
    assert f[at.FEATURE_NAMES.index("log2_panels")] == pytest.approx(math.log2(5))
    assert f[at.FEATURE_NAMES.index("ragged_frac")] == pytest.approx(100 / 1024)
    assert f[at.FEATURE_NAMES.index("streamed")] == 0.0
    assert f[at.FEATURE_NAMES.index("log2_h2d_ms")] == 0.0
    c = at.Candidate("mod", 15, _lib.TILE_LARGE, 256, 1024)
    g = at.features(8192, 8192, 16384, c, "streamed")
    # 2 GiB of strips at 57 GB/s, 128 MiB per launch
    h2d = 8 * 2 * 8192 * 16384 / 57e6
    assert g[at.FEATURE_NAMES.index("log2_h2d_ms")] == pytest.approx(math.log2(1 + h2d))
    assert g[at.FEATURE_NAMES.index("log2_copy_per_launch_ms")] == \
        pytest.approx(math.log2(1 + 8 * 16384 * 1024 / 57e6))
    assert g[at.FEATURE_NAMES.index("log2_steps")] == pytest.approx(4.0)


def test_shape_split_is_deterministic_and_disjoint():
    shapes = at.rank_shares() + at.random_shapes(40)
    tr, ho = at.shape_split(shapes, 0.25)
    assert tr == at.shape_split(shapes, 0.25)[0]
    assert not set(tr) & set(ho) and len(tr) + len(ho) == len(shapes)
    assert len(ho) == round(0.25 * len(shapes))


def _synthetic_rows(shapes):
    # time = flops / rate(c) (+ a copy term when streamed) with
    # candidate-dependent rates: an exact power law per candidate and target
    rows = []
    for (m, n, k) in shapes:
        for target in at.TARGETS:
            for c in at.candidate_space(k, target):
                rate = 40.0 + (30.0 if c.form == "int8" else 0.0) - 2.0 * c.slices \
                    + (5.0 if c.tile == _lib.TILE_LARGE else 0.0) + c.panel / 256.0 \
                    + c.group / 1024.0
                ms = 2.0 * m * n * k / rate / 1e9 * (1.5 if target == "streamed" else 1.0)
                rows.append({"target": target, "m": m, "n": n, "k": k, "form": c.form,
                             "slices": c.slices, "tile": c.tile, "panel": c.panel,
                             "group": c.group, "ms": ms})
    return rows


def test_power_law_baseline_recovers_exact_power_law():
    shapes = at.rank_shares()
    rows = _synthetic_rows(shapes)
    base = at.PowerLawModel(rows, shapes)
    m, n, k = 8192, 8192, 16384
    for target in at.TARGETS:
        cands = at.candidate_space(k, target)
        pred = np.exp(base.predict_log_ms(m, n, k, cands, target))
        truth = {at._candidate(r): r["ms"] for r in rows
                 if (r["m"], r["n"], r["k"]) == (m, n, k) and r["target"] == target}
        for c, p in zip(cands, pred):
            assert p == pytest.approx(truth[c], rel=1e-9)


def test_best_fixed_candidate_is_chosen_on_training_shapes():
    shapes = at.rank_shares()
    rows = _synthetic_rows(shapes)
    fixed = at.best_fixed(rows, shapes)
    # the fastest rate in every shape: int8 s=8, large tile, widest panel / group
    assert fixed["resident"] == at.Candidate("int8", 8, _lib.TILE_LARGE, 1024, 0)
    assert fixed["streamed"] == at.Candidate("int8", 8, _lib.TILE_LARGE, 1024, 4096)
    # a candidate missing from one training shape cannot be the fixed choice
    pruned = [r for r in rows if not (r["form"] == "int8" and r["k"] == 16384)]
    assert at.best_fixed(pruned, shapes)["resident"].form != "int8"


def test_evaluate_regret_bookkeeping():
    shapes = at.rank_shares()[:6]
    rows = _synthetic_rows(shapes)
    tr, ho = shapes[:4], shapes[4:]
    table = {}
    for r in rows:
        table[tuple(at.features(r["m"], r["n"], r["k"], at._candidate(r), r["target"]))] = \
            math.log(r["ms"])

    class Perfect:
        def predict_log_ms(self, X):
            return np.array([table[tuple(x)] for x in np.asarray(X).tolist()])

    res = at.evaluate(rows, Perfect(), tr, ho, topk=3)
    for target in at.TARGETS:
        t = res["targets"][target]
        assert t["heldout_shapes"] == 2
        assert t["mape"] == pytest.approx(0.0, abs=1e-12)
        assert t["regret_mlp_mean"] == 0.0 and t["regret_top3_mean"] == 0.0
        assert t["regret_powerlaw_mean"] == pytest.approx(0.0, abs=1e-9)
        assert t["regret_fixed_mean"] == pytest.approx(0.0, abs=1e-12)
        assert t["regret_default_mean"] > 0.0     # fp64/large/256 is not the fastest here

    class Inverted:
        def predict_log_ms(self, X):
            return -np.array([table[tuple(x)] for x in np.asarray(X).tolist()])

    bad = at.evaluate(rows, Inverted(), tr, ho, topk=3)
    assert all(bad["targets"][t]["regret_mlp_mean"] > 0.0 for t in at.TARGETS)


def test_dataset_files_round_trip(tmp_path):
    rows = _synthetic_rows(at.rank_shares()[:2])
    doc = {"schema": at.SCHEMA, "kind": "mgx-autotune-dataset", "rows": rows,
           "shapes_done": [list(s) for s in at.rank_shares()[:2]]}
    at.write_dataset(doc, tmp_path / "d.json")
    back = at.read_dataset(tmp_path / "d.json")
    assert back["rows"] == rows and back["shapes_done"] == doc["shapes_done"]
    (tmp_path / "old.json").write_text('{"schema": 1, "rows": []}')
    with pytest.raises(ValueError):
        at.read_dataset(tmp_path / "old.json")


def test_share_shape():
    assert at.share_shape("1d", 32768, 4) == (8192, 32768, 32768)
    assert at.share_shape("2d", 16384, 4, (2, 2)) == (8192, 8192, 16384)


def test_random_shapes_exclusion_for_heldout_sets():
    from mgxbench_amd import autotune as at
    train = at.rank_shares() + at.random_shapes(84)
    held = at.random_shapes(200, seed=8, exclude=train)
    assert len(set(held)) == 200 and not set(held) & set(train)
    assert held == at.random_shapes(200, seed=8, exclude=train)       # deterministic
    assert at.random_shapes(84) == train[36:]                          # default stream unchanged
    space = len(at.SIZES) ** 3
    with pytest.raises(ValueError):                                    # one shape of the space excluded
        at.random_shapes(space, exclude=at.random_shapes(1))
    assert len(at.random_shapes(space, exclude=[(1, 2, 3)])) == space  # outside the space: no effect
