"""Tile variants, the autotuner's timing entry point and a small end-to-end
autotune cycle (measure -> GPU-trained MLP -> select) through the C ABI."""
import ctypes

import numpy as np
import pytest

from tests.gpu_util import announce, empty, to_dev, to_host

pytestmark = pytest.mark.gpu

U = 2.0 ** -53


@pytest.mark.parametrize("tile", [1, 2])
@pytest.mark.parametrize("m,n,k", [(65, 130, 77), (300, 200, 513)])
def test_dgemm_tiles(mgx, orc, tile, m, n, k):
    announce(f"dgemm_ex tile={tile} m={m} n={n} k={k}")
    A = orc.fill(m, k, 1, seed=5)
    B = orc.fill(k, n, 1, seed=6)
    # keep the device tensors referenced for the whole call (a temporary's
    # memory could be handed to the next allocation)
    dA, dB = to_dev(A), to_dev(B)
    dC = empty(m, n)
    mgx._lib.call("mgx_dgemm_ex", tile, 0, dA.data_ptr(), dB.data_ptr(),
                  dC.data_ptr(), m, n, k, k, n, n, 0)
    C = to_host(dC)
    ref = orc.gemm_kernel_order(A, B)
    gamma = k * U / (1 - k * U)
    assert np.all(np.abs(C - ref) <= 2 * gamma * orc.gemm_abs_rows(A, B))
    # accumulate form
    C0 = orc.fill(m, n, 1, seed=7)
    dC = to_dev(C0)
    mgx._lib.call("mgx_dgemm_ex", tile, 1, dA.data_ptr(), dB.data_ptr(),
                  dC.data_ptr(), m, n, k, k, n, n, 0)
    ref2 = orc.gemm_accumulate_order(A, B, C0, k, 0, k)
    absp = orc.gemm_abs_rows(A, B)
    assert np.all(np.abs(to_host(dC) - ref2) <= 2 * gamma * absp + 2 * U * (np.abs(C0) + 2 * absp))


@pytest.mark.parametrize("tile", [1, 2])
@pytest.mark.parametrize("m,n,k,s", [(130, 127, 129, 8), (257, 300, 700, 6), (64, 64, 64, 9)])
def test_split_tiles_bit_exact(mgx, orc, tile, m, n, k, s):
    announce(f"split_ex tile={tile} m={m} n={n} k={k} s={s}")
    A = orc.fill(m, k, 1, seed=m + k)
    B = orc.fill(k, n, 1, seed=n + k)
    dA, dB = to_dev(A), to_dev(B)
    dC = empty(m, n)
    mgx._lib.call("mgx_split_gemm_ex", mgx.FORM_INT8, s, tile, dA.data_ptr(),
                  dB.data_ptr(), dC.data_ptr(), m, n, k, k, n, n, 0, None, 0, 0)
    assert np.array_equal(to_host(dC), orc.split_gemm_int8(A, B, s), equal_nan=True)


def test_time_local_product_result_and_timing(mgx, orc):
    # The timed panel loop leaves C = A.B (first panel or panel group
    # overwrites, later ones add), for every form and split group.
    r = np.random.default_rng(5)
    m, n, k = 256, 192, 1000
    announce("time_local_product")
    A = r.integers(-30, 30, size=(m, k)).astype(np.float64)
    B = r.integers(-30, 30, size=(k, n)).astype(np.float64)
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    dA, dB = to_dev(A), to_dev(B)
    for form, slices, split_k in ((0, 0, 0), (1, 8, 0), (1, 8, 300), (3, 15, 0), (3, 15, 300)):
        for tile in (1, 2):
            dC = empty(m, n)
            cfg = mgx._lib.TuneConfig(form, slices, tile, 256, split_k)
            ms = ctypes.c_double()
            mgx._lib.call("mgx_time_local_product", ctypes.byref(cfg), dA.data_ptr(),
                          dB.data_ptr(), dC.data_ptr(), m, n, k, k, n, n, None, 0, 1, 2,
                          ctypes.byref(ms))
            assert ms.value > 0
            assert np.array_equal(to_host(dC), exact)


def test_autotune_cycle_small(mgx, tmp_path):
    from mgxbench_amd import autotune as at
    announce("autotune small cycle")
    shapes = [(512, 512, 1024), (1024, 512, 1024), (512, 1024, 2048), (1024, 1024, 1024),
              (768, 512, 1536), (512, 768, 1024)]
    doc = at.measure(shapes, tmp_path / "ds.json", log=lambda *_: None)
    rows = doc["rows"]
    assert len(rows) == sum(len(at.candidate_space(k, t)) for _, _, k in shapes
                            for t in at.TARGETS)
    assert all(r["ms"] > 0 for r in rows)
    # resident split products do not depend on the panel width: timed once
    res_mod = {r["ms"] for r in rows if r["target"] == "resident" and r["form"] == "mod"
               and r["slices"] == 15 and r["tile"] == 1 and (r["m"], r["n"], r["k"]) == shapes[0]}
    assert len(res_mod) == 1
    # resume: nothing left to measure
    again = at.measure(shapes, tmp_path / "ds.json", log=lambda *_: None)
    assert len(again["rows"]) == len(rows)
    tr, ho = shapes[:4], shapes[4:]
    model, hist = at.train(rows, tr, epochs=600)
    assert hist[-1] < hist[0]
    res = at.evaluate(rows, model, tr, ho)
    for t in at.TARGETS:
        assert res["targets"][t]["heldout_shapes"] == 2
        assert res["targets"][t]["mape"] is not None
        assert res["targets"][t]["fixed_candidate"] is not None
    timer = at.Timer(1024, 1024, 2048)
    for target in at.TARGETS:
        choice, ranked = at.select(model, 512, 512, 1024, target, topk=3, timer=timer)
        assert "measured_ms" in choice and len(ranked) == len(at.candidate_space(1024, target))
    # the model round-trips through JSON
    m2 = at.MLPModel.from_json(model.to_json())
    X = np.array([at.features(512, 512, 1024, c, "streamed")
                  for c in at.candidate_space(1024, "streamed")])
    assert np.allclose(m2.predict_log_ms(X), model.predict_log_ms(X))
