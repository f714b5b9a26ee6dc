"""The calibrated AUTO path (mgxbench_amd.predictive), on the CPU.

Pinned by closed forms and by the libraries the reference's fits use: the
piecewise log-log estimate against hand-computed power laws, the least
squares fit against an exact cubic, the closed-form ridge against scikit-learn's
StandardScaler + Ridge(alpha=0.1) (the reference's own 'ml' model,
autotuning/predictive.py:147-158), the 'auto' vote, candidate filtering,
recovery from reports, and a full calibrate -> auto cycle driving the runner
over a stand-in executable."""
import argparse
import json
import math
import sys
from pathlib import Path

import numpy as np
import pytest

from mgxbench_amd import predictive as pr

GPU = (["AMD Instinct MI355X"], ["6.10.5"])


@pytest.fixture
def fake_gpu(monkeypatch):
    monkeypatch.setattr(pr, "gpu_identity", lambda: GPU)


def _cand(lib, pts, gr=None, gc=None, bk=None):
    return {"library": lib, "grid_rows": gr, "grid_cols": gc, "summa_bk": bk,
            "measurements": [{"size": n, "median_s": t} for n, t in pts]}


def _model(cands, decomposition="1d", ranks=1):
    return {"schema": 1, "signature": {"ranks": ranks, "decomposition": decomposition},
            "sizes": sorted({m["size"] for c in cands for m in c["measurements"]}),
            "candidates": cands}


def test_estimate_power_law_interpolation_and_clipping():
    s = [{"size": 2048, "median_s": 1.0}, {"size": 4096, "median_s": 8.0},
         {"size": 8192, "median_s": 64.0}]
    assert pr.estimate(s, 4096) == (8.0, False)
    t, ext = pr.estimate(s, 6000)
    assert t == pytest.approx(8.0 * (6000 / 4096) ** 3) and not ext
    t, ext = pr.estimate(s, 16384)
    assert t == pytest.approx(512.0) and ext
    t, ext = pr.estimate(s, 1024)
    assert t == pytest.approx(1.0 / 8) and ext
    # the exponent is clipped to [0, 4] and applied from the lower size
    down = [{"size": 1000, "median_s": 2.0}, {"size": 2000, "median_s": 1.0}]
    assert pr.estimate(down, 4000)[0] == pytest.approx(2.0)
    steep = [{"size": 1000, "median_s": 1.0}, {"size": 2000, "median_s": 64.0}]
    assert pr.estimate(steep, 4000)[0] == pytest.approx(4.0 ** 4)
    # repeated sizes use their median; invalid samples are ignored
    rep = s + [{"size": 4096, "median_s": 10.0}, {"size": 4096, "median_s": 9.0},
               {"size": 512, "median_s": -1.0}]
    assert pr.estimate(rep, 4096)[0] == 9.0
    with pytest.raises(ValueError):
        pr.estimate(s[:1], 100)


def test_least_squares_recovers_an_exact_cubic():
    sizes = [1024, 2048, 3072, 4096, 6144]
    ref = max(sizes)
    f = lambda n: 0.5 + 2.0 * (n / ref) ** 2 + 3.0 * (n / ref) ** 3
    pts = [(n, f(n)) for n in sizes]
    assert pr.physical_fit(pts, 8192, "extrapolation") == pytest.approx(f(8192), rel=1e-9)
    with pytest.raises(ValueError):
        pr.physical_fit(pts[:3], 8192, "extrapolation")


def test_ridge_matches_scikit_learn():
    sk = pytest.importorskip("sklearn")
    from sklearn.linear_model import Ridge
    from sklearn.preprocessing import StandardScaler
    r = np.random.default_rng(3)
    for trial in range(5):
        sizes = sorted(r.choice(np.arange(1, 40) * 512, size=6, replace=False).tolist())
        times = [float(x) for x in r.uniform(0.1, 5.0, size=6)]
        pts = list(zip(sizes, times))
        target = int(r.integers(512, 30000))
        ref = max(sizes)
        X = np.column_stack([(np.array(sizes) / ref) ** 2, (np.array(sizes) / ref) ** 3])
        sc = StandardScaler()
        reg = Ridge(alpha=0.1, fit_intercept=True).fit(sc.fit_transform(X), times)
        xt = np.array([[(target / ref) ** 2, (target / ref) ** 3]])
        want = float(reg.predict(sc.transform(xt))[0])
        if want <= 0:
            with pytest.raises(ValueError):
                pr.physical_fit(pts, target, "ml")
            continue
        assert pr.physical_fit(pts, target, "ml") == pytest.approx(want, rel=1e-9)


def test_auto_vote_prefers_the_fit_with_lower_holdout_error():
    sizes = [1024, 2048, 3072, 4096, 5120, 6144]
    cubic = lambda n: 0.1 + 4.0 * (n / 6144) ** 3
    model = _model([_cand("MMM", [(n, cubic(n)) for n in sizes]),
                    _cand("NNN", [(n, 0.5 * cubic(n)) for n in sizes])])
    ranking, method = pr.select_predictions(model, 8192, 1, "1d", None, "auto")
    assert method == "extrapolation"        # least squares fits the cubic exactly
    assert [r["library"] for r in ranking] == ["NNN", "MMM"]
    assert ranking[0]["algorithm"] == "physical-polynomial" and ranking[0]["extrapolated"]
    # fewer than five sizes: no vote, the extrapolation method
    small = _model([_cand("MMM", [(1024, 1.0), (2048, 8.0)])])
    assert pr.select_predictions(small, 4096, 1, "1d", None, "auto")[1] == "extrapolation"
    with pytest.raises(ValueError):
        pr.select_predictions(small, 4096, 1, "1d", None, "nope")


def test_prediction_filters_and_fallbacks():
    cands = [_cand("MMM", [(1024, 1.0), (2048, 8.0)], 1, 4, 256),
             _cand("NNN", [(1024, 0.5), (2048, 4.0)], 2, 2, 256),
             _cand("SSS", [(1024, 0.4), (2048, 3.0)], 4, 1, 128)]
    model = _model(cands, "2d", 4)
    ranking = pr.predict_method(model, 4096, 4, "2d", None, "extrapolation")
    assert [r["library"] for r in ranking] == ["SSS", "NNN", "MMM"]
    assert all(r["algorithm"] == "piecewise-power-law-fallback" for r in ranking)
    only = pr.predict_method(model, 4096, 4, "2d", (2, 2), "extrapolation")
    assert [r["library"] for r in only] == ["NNN"]
    # N = 6 is not divisible by the 4-row grids
    assert [r["library"] for r in pr.predict_method(model, 6, 4, "2d", None, "extrapolation")] \
        == ["NNN"]
    with pytest.raises(ValueError):          # 'ml' needs >= 4 sizes per candidate
        pr.predict_method(model, 4096, 4, "2d", None, "ml")
    with pytest.raises(ValueError):
        pr.predict_method(model, 4096, 8, "2d", None, "extrapolation")


def test_candidates_and_grids():
    assert pr.factor_grids(4) == [(1, 4), (2, 2), (4, 1)]
    c1 = pr.candidate_configs("1d", 4, [4096], ["MMM", "NNN"], [256])
    assert [c["library"] for c in c1] == ["MMM", "NNN"] and c1[0]["grid_rows"] is None
    c2 = pr.candidate_configs("2d", 4, [4096, 6144], ["NNN"], [128, 256, 0])
    assert sorted((c["grid_rows"], c["grid_cols"], c["summa_bk"]) for c in c2) == \
        [(1, 4, 128), (1, 4, 256), (2, 2, 128), (2, 2, 256), (4, 1, 128), (4, 1, 256)]
    c3 = pr.candidate_configs("2d", 4, [6], ["NNN"], [128])   # 6 % 4 != 0
    assert [(c["grid_rows"], c["grid_cols"]) for c in c3] == [(2, 2)]
    assert pr.model_filename("2d", 4) == "calibration-2d-4gpu.json"


def _fake_exe(build: Path, decomposition="1d"):
    build.mkdir(parents=True, exist_ok=True)
    exe = build / ("mm_2d" if decomposition == "2d" else "mm_1d")
    # a stand-in executable: time = size^3 * per-library rate, always valid
    exe.write_text("#!" + sys.executable + "\n" + (
        "import sys\n"
        "dev, n, lib = sys.argv[1], int(sys.argv[2]), sys.argv[3]\n"
        "rate = {'MMM': 4e9, 'NNN': 8e9, 'SSS': 6e9}.get(lib, 1e9)\n"
        f"print('BENCHMARK_RESULT decomposition={decomposition.upper()} library=%s size=%d "
        "ranks=1 seconds=%.9f valid=1 loops=1' % (lib, n, n ** 3 / rate))\n"))
    exe.chmod(0o755)
    return exe


def _args(**kw):
    base = dict(project_root=".", decomposition="1d", ranks=1, sizes=[256, 512, 1024],
                libraries=["MMM", "NNN", "SSS"], bks=[128], grids=None, grid_rows=None,
                grid_cols=None, warmup=0, repeats=1,
                launcher=f"{sys.executable} -m mgxbench_amd.launch", launcher_args="",
                build_dir=None, output=None, model=None)
    base.update(kw)
    return argparse.Namespace(**base)


def test_calibrate_then_auto_through_the_runner(tmp_path, fake_gpu, monkeypatch):
    monkeypatch.chdir(Path(__file__).resolve().parents[1])
    build = tmp_path / "bin"
    _fake_exe(build)
    model_path = tmp_path / "models" / "cal.json"
    a = _args(build_dir=str(build), output=str(tmp_path / "cal"), model=str(model_path))
    assert pr.calibrate(a) == 0
    model = pr.load_model(model_path)
    assert model["signature"]["gpu_models"] == GPU[0]
    assert len(model["candidates"]) == 3
    for c in model["candidates"]:
        rate = {"MMM": 4e9, "NNN": 8e9, "SSS": 6e9}[c["library"]]
        assert [m["size"] for m in c["measurements"]] == [256, 512, 1024]
        for m in c["measurements"]:
            assert m["median_s"] == pytest.approx(m["size"] ** 3 / rate, rel=1e-6)
    # AUTO: rank at an uncalibrated size, report written
    auto = argparse.Namespace(decomposition="1d", ranks=1, size=2048, grid_rows=None,
                              grid_cols=None, prediction_model="extrapolation",
                              build_dir=str(build), auto_model=str(model_path),
                              auto_validate=True, auto_feedback=True, warmup=0, repeats=1,
                              launcher=a.launcher, launcher_args="", timeout=60,
                              output=str(tmp_path / "auto"))
    assert pr.auto(auto) == 0
    rep = json.loads((tmp_path / "auto" / "autotuning_prediction.json").read_text())
    assert rep["best"]["library"] == "NNN" and rep["status"] == "VALIDATED"
    assert rep["best"]["predicted_s"] == pytest.approx(2048 ** 3 / 8e9, rel=1e-6)
    assert rep["validation"]["relative_error"] < 1e-6
    assert rep["speedup_vs_MPI_predicted"] == pytest.approx(2.0, rel=1e-6)
    # feedback appended the validated size to the chosen candidate, backup kept
    again = pr.load_model(model_path)
    nnn = next(c for c in again["candidates"] if c["library"] == "NNN")
    assert [m["size"] for m in nnn["measurements"]][-1] == 2048
    assert model_path.with_suffix(".json.bak").exists()
    # a different executable no longer matches the calibration
    (build / "mm_1d").write_text((build / "mm_1d").read_text() + "\n# changed\n")
    assert pr.auto(auto) == 2


def test_recover_from_reports(tmp_path, fake_gpu):
    build = tmp_path / "bin"
    _fake_exe(build, "2d")
    runs = tmp_path / "runs"

    def report(i, lib, size, median, label="auto-calibration", grid=(1, 1, 256), failures=()):
        d = runs / f"r{i}"
        d.mkdir(parents=True)
        doc = {"command_args": {"label": label, "size": size, "ranks": 1, "decomposition": "2d",
                                "libraries": [lib], "grid_rows": grid[0], "grid_cols": grid[1],
                                "summa_bk": grid[2]},
               "failures": list(failures),
               "results": [{"library": lib, "validated": True, "median_s": median}]}
        (d / "benchmark.json").write_text(json.dumps(doc))

    report(0, "NNN", 1024, 0.1)
    report(1, "NNN", 2048, 0.8)
    report(2, "NNN", 2048, 1.0)          # same size twice: median of the two
    report(3, "MMM", 1024, 0.2)          # only one size -> rejected
    report(4, "MMM", 2048, 1.6, label="other")
    report(5, "SSS", 1024, 0.3, failures=[{"x": 1}])
    a = argparse.Namespace(project_root=str(tmp_path), decomposition="2d", ranks=1,
                           input=str(runs), model=str(tmp_path / "m.json"), build_dir=str(build),
                           assume_compatible=False)
    with pytest.raises(ValueError):
        pr.recover_calibration(a)
    a.assume_compatible = True
    assert pr.recover_calibration(a) == 0
    model = pr.load_model(tmp_path / "m.json")
    assert [c["library"] for c in model["candidates"]] == ["NNN"]
    assert model["candidates"][0]["measurements"] == [{"size": 1024, "median_s": 0.1},
                                                      {"size": 2048, "median_s": 0.9}]


def test_gpu_identity_from_rocm_smi(monkeypatch):
    # the MI355X reports no series; the PCI model id and the gfx target name it
    out = {"--showproductname": (
        "device,Card Series,Card Model,Card Vendor,Card SKU,Subsystem ID,Device Rev,Node ID,"
        "GUID,GFX Version\ncard0,N/A,0x75a3,Advanced Micro Devices Inc. [AMD/ATI],N/A,0x0000,"
        "0x00,1,28206,gfx950\n"),
        "--showdriverversion": 'name, value\n"Driver version", "6.18.44"\n'}
    monkeypatch.setattr(pr, "_rocm_smi", lambda args: out[args[0]])
    assert pr.gpu_identity() == (["0x75a3 gfx950"], ["6.18.44"])


def test_cli_argument_rules():
    with pytest.raises(SystemExit):
        pr.main(["auto", "--decomposition", "1d", "--ranks", "1", "--size", "64",
                 "--auto-feedback"])
    with pytest.raises(SystemExit):
        pr.main(["auto", "--decomposition", "1d", "--ranks", "2", "--size", "64",
                 "--grid-rows", "1", "--grid-cols", "2"])
    with pytest.raises(SystemExit):
        pr.main(["calibrate", "--decomposition", "2d", "--ranks", "4", "--grids", "3x2"])
