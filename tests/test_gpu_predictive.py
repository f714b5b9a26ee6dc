"""The 'mlp' prediction method of the calibrated AUTO path: one network over
all calibrated curves, trained on the GPU by libmgx's kernels on the
calibration medians, ranks the shapes at an uncalibrated size."""
import pytest

from mgxbench_amd import predictive as pr

pytestmark = pytest.mark.gpu


def test_mlp_ranks_calibrated_shapes():
    rates = {"MMM": 2e9, "YYY": 3e9, "NNN": 8e9, "SSS": 6e9}
    entries = []
    for lib, rate in rates.items():
        for bk in (128, 256):
            r = rate * (1.0 + 0.1 * (bk == 256))
            entries.append({"library": lib, "grid_rows": 1, "grid_cols": 1, "summa_bk": bk,
                            "measurements": [{"size": n, "median_s": n ** 3 / r}
                                             for n in (1024, 2048, 3072, 4096, 6144)]})
    cal = pr.Calibration({"schema": 1, "signature": {"ranks": 1, "decomposition": "2d"},
                          "sizes": [1024, 2048, 3072, 4096, 6144], "candidates": entries})
    ranking, method = pr.forecast(cal, 5120, 1, "2d", None, "mlp")
    assert method == "mlp"
    assert len(ranking) == 8 and all(r["algorithm"] == "mlp-gpu" for r in ranking)
    assert (ranking[0]["library"], ranking[0]["summa_bk"]) == ("NNN", 256)
    want = 5120 ** 3 / (8e9 * 1.1)
    assert ranking[0]["predicted_s"] == pytest.approx(want, rel=0.1)
    assert [r["library"] for r in ranking[-2:]] == ["MMM", "MMM"]


def test_calibrate_then_auto_on_the_real_executable(tmp_path, monkeypatch):
    """calibrate -> auto --auto-validate --auto-feedback through the runner
    and the one-rank launcher on the real mm_1d: a schema-1 calibration
    with a rocm-smi signature, a ranked AUTO report and a validated,
    fed-back measurement."""
    import json
    import shutil
    from pathlib import Path
    if shutil.which("rocm-smi") is None:
        pytest.skip("rocm-smi is not on PATH")
    monkeypatch.chdir(Path(__file__).resolve().parents[1])
    model = tmp_path / "cal.json"
    rc = pr.main(["calibrate", "--decomposition", "1d", "--ranks", "1", "--sizes", "512", "1024",
                  "--libraries", "MMM", "NNN", "--warmup", "0", "--repeats", "1",
                  "--output", str(tmp_path / "cal"), "--model", str(model)])
    assert rc == 0
    cal = pr.Calibration.read(model)
    assert cal.signature["gpu_models"] and len(cal.signature["binary_sha256"]) == 64
    assert sorted(s.library for s, _ in cal.entries()) == ["MMM", "NNN"]
    rc = pr.main(["auto", "--decomposition", "1d", "--ranks", "1", "--size", "2048",
                  "--auto-model", str(model), "--auto-validate", "--auto-feedback",
                  "--warmup", "0", "--repeats", "1", "--output", str(tmp_path / "auto")])
    assert rc == 0
    rep = json.loads((tmp_path / "auto" / "autotuning_prediction.json").read_text())
    assert rep["status"] == "VALIDATED" and rep["validation"]["status"] == "PASS"
    assert rep["best"]["library"] in ("MMM", "NNN") and rep["best"]["predicted_s"] > 0
    assert rep["feedback"]["status"] == "SAVED"
    sizes = {m["size"] for s, e in pr.Calibration.read(model).entries()
             if s.library == rep["best"]["library"] for m in e["measurements"]}
    assert 2048 in sizes
