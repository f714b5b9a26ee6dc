"""The calibrated AUTO path (mgxbench_amd.predictive), on the CPU.

Pinned by closed forms and by the libraries the reference's fits use: the
power law against hand-computed values, the least-squares cubic against an
exact cubic, the closed-form ridge against scikit-learn's StandardScaler +
Ridge(alpha=0.1) (the reference's 'ml' model, autotuning/predictive.py:147-158),
the 'auto' vote, shape filtering, recovery from reports, and a full
calibrate -> auto cycle driving the runner over a stand-in executable."""
import argparse
import json
import sys
from pathlib import Path

import numpy as np
import pytest

from mgxbench_amd import predictive as pr

GPU = pr.Platform(("AMD Instinct MI355X",), ("6.10.5",))


@pytest.fixture
def fake_gpu(monkeypatch):
    monkeypatch.setattr(pr.Platform, "probe", classmethod(lambda cls: GPU))


def _entry(lib, pts, gr=None, gc=None, bk=None):
    return {"library": lib, "grid_rows": gr, "grid_cols": gc, "summa_bk": bk,
            "measurements": [{"size": n, "median_s": t} for n, t in pts]}


def _cal(entries, decomposition="1d", ranks=1):
    return pr.Calibration({"schema": 1, "signature": {"ranks": ranks,
                                                      "decomposition": decomposition},
                           "sizes": sorted({m["size"] for e in entries
                                            for m in e["measurements"]}),
                           "candidates": entries})


def _curve(pts):
    return pr.Curve([{"size": n, "median_s": t} for n, t in pts])


def test_power_law_interpolation_extrapolation_and_clipping():
    law = pr.PowerLaw()
    c = _curve([(2048, 1.0), (4096, 8.0), (8192, 64.0)])
    assert law.forecast(c, 4096) == 8.0
    assert law.forecast(c, 6000) == pytest.approx(8.0 * (6000 / 4096) ** 3)
    assert c.covers(6000) and not c.covers(16384) and not c.covers(1024)
    assert law.forecast(c, 16384) == pytest.approx(512.0)
    assert law.forecast(c, 1024) == pytest.approx(1.0 / 8)
    # the exponent is clipped to [0, 4] and applied from the lower size
    assert law.forecast(_curve([(1000, 2.0), (2000, 1.0)]), 4000) == pytest.approx(2.0)
    assert law.forecast(_curve([(1000, 1.0), (2000, 64.0)]), 4000) == pytest.approx(4.0 ** 4)
    # repeated sizes use their median; invalid samples are ignored
    rep = _curve([(2048, 1.0), (4096, 8.0), (8192, 64.0), (4096, 10.0), (4096, 9.0),
                  (512, -1.0), (0, 1.0), (1024, float("inf"))])
    assert rep.sizes == [2048, 4096, 8192] and law.forecast(rep, 4096) == 9.0
    with pytest.raises(pr.NotApplicable):
        law.forecast(_curve([(2048, 1.0)]), 100)


def test_least_squares_recovers_an_exact_cubic():
    sizes = [1024, 2048, 3072, 4096, 6144]
    f = lambda n: 0.5 + 2.0 * (n / 6144) ** 2 + 3.0 * (n / 6144) ** 3
    c = _curve([(n, f(n)) for n in sizes])
    assert pr.Cubic(ridge=False).forecast(c, 8192) == pytest.approx(f(8192), rel=1e-9)
    with pytest.raises(pr.NotApplicable):
        pr.Cubic(ridge=False).forecast(_curve([(n, f(n)) for n in sizes[:3]]), 8192)


def test_ridge_matches_scikit_learn():
    pytest.importorskip("sklearn")
    from sklearn.linear_model import Ridge
    from sklearn.preprocessing import StandardScaler
    r = np.random.default_rng(3)
    for _ in range(5):
        sizes = sorted(r.choice(np.arange(1, 40) * 512, size=6, replace=False).tolist())
        times = [float(x) for x in r.uniform(0.1, 5.0, size=6)]
        target = int(r.integers(512, 30000))
        top = max(sizes)
        X = np.column_stack([(np.array(sizes) / top) ** 2, (np.array(sizes) / top) ** 3])
        sc = StandardScaler()
        reg = Ridge(alpha=0.1, fit_intercept=True).fit(sc.fit_transform(X), times)
        want = float(reg.predict(sc.transform([[(target / top) ** 2, (target / top) ** 3]]))[0])
        c = _curve(list(zip(sizes, times)))
        if want <= 0:
            with pytest.raises(pr.NotApplicable):
                pr.Cubic(ridge=True).forecast(c, target)
            continue
        assert pr.Cubic(ridge=True).forecast(c, target) == pytest.approx(want, rel=1e-9)


def test_auto_vote_prefers_the_fit_with_lower_holdout_error():
    sizes = [1024, 2048, 3072, 4096, 5120, 6144]
    cubic = lambda n: 0.1 + 4.0 * (n / 6144) ** 3
    cal = _cal([_entry("MMM", [(n, cubic(n)) for n in sizes]),
                _entry("NNN", [(n, 0.5 * cubic(n)) for n in sizes])])
    ranking, method = pr.forecast(cal, 8192, 1, "1d", None, "auto")
    assert method == "extrapolation"        # least squares fits the cubic exactly
    assert [r["library"] for r in ranking] == ["NNN", "MMM"]
    assert ranking[0]["algorithm"] == "physical-polynomial" and ranking[0]["extrapolated"]
    assert ranking[0]["calibration_sizes"] == 6
    assert ranking[0]["predicted_s"] == pytest.approx(0.5 * cubic(8192), rel=1e-9)
    # fewer than five sizes: nothing to vote on, the extrapolation method
    small = _cal([_entry("MMM", [(1024, 1.0), (2048, 8.0)])])
    assert pr.forecast(small, 4096, 1, "1d", None, "auto")[1] == "extrapolation"
    with pytest.raises(pr.PredictionError):
        pr.forecast(small, 4096, 1, "1d", None, "nope")


def test_ridge_vote_wins_on_noisy_curves():
    # curves whose largest point a cubic through the others overshoots;
    # the ridge shrinks toward the mean and misses by less
    r = np.random.default_rng(5)
    sizes = [1000, 2000, 3000, 4000, 5000, 6000]
    entries = []
    for lib in ("MMM", "NNN", "SSS"):
        t = [1.0 + 0.3 * r.standard_normal() for _ in sizes[:-1]] + [1.0]
        entries.append(_entry(lib, list(zip(sizes, [abs(v) + 0.1 for v in t]))))
    cal = _cal(entries)
    errs = {m: np.median(cal.holdout_score(pr.fit_chain(m)[0])) for m in ("extrapolation", "ml")}
    want = min(errs, key=lambda m: (errs[m], m))
    assert pr.forecast(cal, 7000, 1, "1d", None, "auto")[1] == want


def test_forecast_filters_and_fallbacks():
    cal = _cal([_entry("MMM", [(1024, 1.0), (2048, 8.0)], 1, 4, 256),
                _entry("NNN", [(1024, 0.5), (2048, 4.0)], 2, 2, 256),
                _entry("SSS", [(1024, 0.4), (2048, 3.0)], 4, 1, 128)], "2d", 4)
    ranking, _ = pr.forecast(cal, 4096, 4, "2d", None, "extrapolation")
    assert [r["library"] for r in ranking] == ["SSS", "NNN", "MMM"]
    assert all(r["algorithm"] == "piecewise-power-law-fallback" for r in ranking)
    assert set(ranking[0]) == {"library", "grid_rows", "grid_cols", "summa_bk", "predicted_s",
                               "algorithm", "calibration_sizes", "extrapolated"}
    only, _ = pr.forecast(cal, 4096, 4, "2d", (2, 2), "extrapolation")
    assert [r["library"] for r in only] == ["NNN"]
    # N = 6 is not divisible by the 4-row and 4-column grids
    assert [r["library"] for r in pr.forecast(cal, 6, 4, "2d", None, "extrapolation")[0]] \
        == ["NNN"]
    with pytest.raises(pr.PredictionError):          # 'ml' needs >= 4 sizes per curve
        pr.forecast(cal, 4096, 4, "2d", None, "ml")
    with pytest.raises(pr.PredictionError):          # calibrated for 4 ranks
        pr.forecast(cal, 4096, 8, "2d", None, "extrapolation")
    with pytest.raises(pr.PredictionError):          # 1D size not a multiple of the ranks
        pr.forecast(_cal([_entry("MMM", [(1024, 1.0), (2048, 8.0)])], "1d", 4), 4098, 4, "1d")


def test_shape_space_and_grids():
    assert pr.Shape.grids_of(4) == [(1, 4), (2, 2), (4, 1)]
    one = pr.Shape.space("1d", 4, [4096], ["MMM", "NNN"], [256])
    assert one == [pr.Shape("MMM"), pr.Shape("NNN")]
    two = pr.Shape.space("2d", 4, [4096, 6144], ["NNN"], [128, 256, 0])
    assert sorted((s.grid_rows, s.grid_cols, s.summa_bk) for s in two) == \
        [(1, 4, 128), (1, 4, 256), (2, 2, 128), (2, 2, 256), (4, 1, 128), (4, 1, 256)]
    three = pr.Shape.space("2d", 4, [6], ["NNN"], [128])   # 6 % 4 != 0
    assert [(s.grid_rows, s.grid_cols) for s in three] == [(2, 2)]
    assert pr.Shape.space("2d", 4, [4096], ["NNN"], [128], grids=[(2, 2), (3, 1)]) == \
        [pr.Shape("NNN", 2, 2, 128)]
    assert pr.default_calibration_path("2d", 4).name == "calibration-2d-4gpu.json"
    assert pr.Shape("NNN", 2, 2, 256).runner_flags() == ["--grid-rows", "2", "--grid-cols", "2",
                                                          "--summa-bk", "256"]


def test_signature_agreement():
    sig = {"gpu_models": ["0x75a3 gfx950"], "ranks": 1, "decomposition": "2d",
           "binary_sha256": "ab", "slurm_nodes": "1", "driver_versions": ["x"], "id": "1"}
    assert pr.signatures_agree(sig, dict(sig, driver_versions=["y"], id="2"))
    assert pr.signatures_agree(sig, dict(sig, gpu_models=["0x75a3 gfx950, extra"]))
    assert not pr.signatures_agree(sig, dict(sig, binary_sha256="cd"))
    assert not pr.signatures_agree(sig, dict(sig, ranks=2))
    assert not pr.signatures_agree(sig, dict(sig, gpu_models=["other"]))


def _fake_exe(build: Path, decomposition="1d"):
    build.mkdir(parents=True, exist_ok=True)
    exe = build / ("mm_2d" if decomposition == "2d" else "mm_1d")
    # a stand-in executable: time = size^3 / per-library rate, always valid
    exe.write_text("#!" + sys.executable + "\n" + (
        "import sys\n"
        "dev, n, lib = sys.argv[1], int(sys.argv[2]), sys.argv[3]\n"
        "rate = {'MMM': 4e9, 'NNN': 8e9, 'SSS': 6e9}.get(lib, 1e9)\n"
        f"print('BENCHMARK_RESULT decomposition={decomposition.upper()} library=%s size=%d "
        "ranks=1 seconds=%.9f valid=1 loops=1' % (lib, n, n ** 3 / rate))\n"))
    exe.chmod(0o755)
    return exe


def _args(**kw):
    base = dict(command="calibrate", project_root=".", decomposition="1d", ranks=1,
                sizes=[256, 512, 1024], libraries=["MMM", "NNN", "SSS"], bks=[128], grids=None,
                grid_rows=None, grid_cols=None, warmup=0, repeats=1,
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
    assert pr.calibrate_cmd(a) == 0
    cal = pr.Calibration.read(model_path)
    assert cal.signature["gpu_models"] == list(GPU.gpu_models)
    assert len(cal.doc["candidates"]) == 3
    assert not model_path.with_name("cal.partial.json").exists()
    logs = sorted(p.name for p in (tmp_path / "cal" / "calibration_logs").iterdir())
    assert "001_MMM_N256_1d.command.txt" in logs and "003_SSS_N1024_1d.stdout" in logs
    for shape, entry in cal.entries():
        rate = {"MMM": 4e9, "NNN": 8e9, "SSS": 6e9}[shape.library]
        assert [m["size"] for m in entry["measurements"]] == [256, 512, 1024]
        for m in entry["measurements"]:
            assert m["median_s"] == pytest.approx(m["size"] ** 3 / rate, rel=1e-6)
    # AUTO: rank at an uncalibrated size, validate, feed back, write the report
    auto = argparse.Namespace(command="auto", decomposition="1d", ranks=1, size=2048,
                              grid_rows=None, grid_cols=None, prediction_model="extrapolation",
                              build_dir=str(build), auto_model=str(model_path),
                              auto_validate=True, auto_feedback=True, warmup=0, repeats=1,
                              launcher=a.launcher, launcher_args="", timeout=60,
                              output=str(tmp_path / "auto"))
    assert pr.run_auto(auto) == 0
    rep = json.loads((tmp_path / "auto" / "autotuning_prediction.json").read_text())
    assert rep["best"]["library"] == "NNN" and rep["status"] == "VALIDATED"
    assert rep["best"]["predicted_s"] == pytest.approx(2048 ** 3 / 8e9, rel=1e-6)
    assert rep["validation"]["relative_error"] < 1e-6
    assert rep["speedup_vs_MPI_predicted"] == pytest.approx(2.0, rel=1e-6)
    assert rep["feedback"]["status"] == "SAVED"
    # feedback appended the validated size to the chosen curve, backup kept
    again = pr.Calibration.read(model_path)
    nnn = next(e for s, e in again.entries() if s.library == "NNN")
    assert [m["size"] for m in nnn["measurements"]][-1] == 2048
    assert 2048 in again.doc["sizes"]
    assert model_path.with_suffix(".json.bak").exists()
    # a different executable no longer matches the calibration
    (build / "mm_1d").write_text((build / "mm_1d").read_text() + "\n# changed\n")
    assert pr.run_auto(auto) == 2


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
    report(3, "MMM", 1024, 0.2)          # only one size -> left out
    report(4, "MMM", 2048, 1.6, label="other")
    report(5, "SSS", 1024, 0.3, failures=[{"x": 1}])
    report(6, "SSS", 2048, 0.3, grid=(1, 1, 0))
    a = argparse.Namespace(project_root=str(tmp_path), decomposition="2d", ranks=1,
                           input=str(runs), model=str(tmp_path / "m.json"), build_dir=str(build),
                           assume_compatible=False)
    with pytest.raises(pr.PredictionError):
        pr.recover_cmd(a)
    a.assume_compatible = True
    assert pr.recover_cmd(a) == 0
    cal = pr.Calibration.read(tmp_path / "m.json")
    assert [s.library for s, _ in cal.entries()] == ["NNN"]
    assert cal.doc["candidates"][0]["measurements"] == [{"size": 1024, "median_s": 0.1},
                                                        {"size": 2048, "median_s": 0.9}]
    assert cal.doc["sizes"] == [1024, 2048]


def test_classify_report_reasons():
    def doc(**kw):
        args = {"label": "auto-calibration", "size": 1024, "ranks": 4, "decomposition": "2d",
                "libraries": "NNN", "grid_rows": 2, "grid_cols": 2, "summa_bk": 128}
        args.update(kw)
        return {"command_args": args,
                "results": [{"library": "NNN", "validated": True, "median_s": 0.5}]}
    assert pr.classify_report(doc(), "2d", 4) == (pr.Shape("NNN", 2, 2, 128), 1024, 0.5)
    for bad in (dict(label="x"), dict(size="?"), dict(ranks=2), dict(libraries=["NNN", "MMM"]),
                dict(grid_rows=1), dict(summa_bk=None), dict(size=1023)):
        assert isinstance(pr.classify_report(doc(**bad), "2d", 4), str), bad


def test_platform_from_rocm_smi(monkeypatch):
    # the MI355X reports no series; the PCI model id and the gfx target name it
    out = {"--showproductname": (
        "device,Card Series,Card Model,Card Vendor,Card SKU,Subsystem ID,Device Rev,Node ID,"
        "GUID,GFX Version\ncard0,N/A,0x75a3,Advanced Micro Devices Inc. [AMD/ATI],N/A,0x0000,"
        "0x00,1,28206,gfx950\n"),
        "--showdriverversion": 'name, value\n"Driver version", "6.18.44"\n'}
    monkeypatch.setattr(pr, "_rocm_smi", lambda args: out[args[0]])
    assert pr.Platform.probe() == pr.Platform(("0x75a3 gfx950",), ("6.18.44",))


def test_cli_argument_rules():
    for argv in (["auto", "--decomposition", "1d", "--ranks", "1", "--size", "64",
                  "--auto-feedback"],
                 ["auto", "--decomposition", "1d", "--ranks", "2", "--size", "64",
                  "--grid-rows", "1", "--grid-cols", "2"],
                 ["auto", "--decomposition", "2d", "--ranks", "4", "--size", "66",
                  "--grid-rows", "4", "--grid-cols", "1"],
                 ["calibrate", "--decomposition", "2d", "--ranks", "4", "--grids", "3x2"],
                 ["calibrate", "--decomposition", "2d", "--ranks", "4", "--grids", "2by2"],
                 ["calibrate", "--decomposition", "2d", "--ranks", "4", "--grids", "2x2",
                  "--grid-rows", "2", "--grid-cols", "2"]):
        with pytest.raises(SystemExit):
            pr.main(argv)
