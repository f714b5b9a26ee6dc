"""Every split-form candidate the autotuner offers is FP64-accurate on the
accuracy-stress inputs, not only on uniform ones.

Inputs: random sign x U[1,2) x 2^U(-30,30), rows of A and columns of B further
x 2^U(-10,10) (the accuracy-stress configuration), m = n = 512 and
K = 4096, 16384, 32768.  Each distinct (form, count, launch depth) of the
candidates of both targets runs the autotuner's local product
(mgx_time_local_product: split products over the candidate's launches - the
whole k for resident operands, groups of whole panels when streamed), and
its normwise and componentwise errors against the GPU double-double product
over all rows must not exceed those of the plain FP64 MFMA product of the
same inputs.  The monolithic modular product (one split product over the
whole depth, as bench.py times it) is held to the same bar at K = 4096 and
16384 (beta 52 and 51).
"""
import ctypes

import numpy as np
import pytest

from tests.gpu_util import announce

pytestmark = pytest.mark.gpu


def _stress(mgx, torch, m, n, k, seed):
    dev = torch.device("cuda", 0)
    A = torch.empty((m, k), dtype=torch.float64, device=dev)
    B = torch.empty((k, n), dtype=torch.float64, device=dev)
    mgx.fill(A, m, k, k, 2, seed=seed)
    mgx.fill(B, k, n, n, 2, seed=seed + 1)
    re = torch.empty(m, dtype=torch.int32, device=dev)
    ce = torch.empty(n, dtype=torch.int32, device=dev)
    mgx.fill_exponents(re, m, seed + 2, -10, 10)
    mgx.fill_exponents(ce, n, seed + 3, -10, 10)
    mgx.scale_rows_cols(A, m, k, k, row_exp=re)
    mgx.scale_rows_cols(B, k, n, n, col_exp=ce)
    return A, B


@pytest.mark.parametrize("k", [4096, 16384, 32768])
def test_autotuner_split_candidates_no_worse_than_plain(mgx, k):
    import torch
    from mgxbench_amd import autotune as at
    m = n = 512
    announce(f"candidate accuracy on stress inputs m=n={m} k={k}")
    A, B = _stress(mgx, torch, m, n, k, seed=31 + k)
    C = torch.empty((m, n), dtype=torch.float64, device=A.device)
    rows = np.arange(m)
    mgx.ABMultiply(A, B, C, m, n, k, k, n, n)
    plain = mgx.error_report(A, B, C, m, n, k, k, n, n, rows)
    assert plain["normwise"] < 1e-13
    seen = set()
    for c in at.candidate_space(k, "resident") + at.candidate_space(k, "streamed"):
        if c.form == "fp64":
            continue
        depth = at.step_k(c, k)
        if (c.form, c.slices, depth) in seen:
            continue
        seen.add((c.form, c.slices, depth))
        cfg = mgx._lib.TuneConfig(at.FORM_CODE[c.form], c.slices, c.tile, c.panel,
                                  c.group or k)
        ms = ctypes.c_double()
        C.fill_(float("nan"))
        mgx._lib.call("mgx_time_local_product", ctypes.byref(cfg), A.data_ptr(), B.data_ptr(),
                      C.data_ptr(), m, n, k, k, n, n, None, 0, 0, 1, ctypes.byref(ms))
        rep = mgx.error_report(A, B, C, m, n, k, k, n, n, rows)
        print(f"k={k} {c.key()} depth {depth}: normwise {rep['normwise']:.3e} "
              f"(plain {plain['normwise']:.3e}) componentwise {rep['componentwise']:.3e} "
              f"(plain {plain['componentwise']:.3e})", flush=True)
        assert rep["normwise"] <= plain["normwise"], c.key()
        assert rep["componentwise"] <= plain["componentwise"], c.key()
    assert {f for f, _, _ in seen} == {"int8", "mod"}
    if k <= 16384:
        mgx.split_gemm(A, B, C, m, n, k, k, n, n, form=mgx.FORM_INT8_MOD, slices=15)
        rep = mgx.error_report(A, B, C, m, n, k, k, n, n, rows)
        assert rep["normwise"] <= plain["normwise"]
        assert rep["componentwise"] <= plain["componentwise"]
    del A, B, C
    torch.cuda.empty_cache()
