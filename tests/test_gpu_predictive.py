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
