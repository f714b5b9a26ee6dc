"""The device input generators of the accuracy-stress configuration
(mgx_fill_exponents, mgx_scale_rows_cols, mgx_fill kind 2) against the
oracle's twins, which tests/test_oracle_generators.py pins to the published
splitmix64 outputs and to exact power-of-two scaling.  Device and oracle must
agree bit for bit, and the device results must have the closed-form
properties themselves (range, exact scaling)."""
import numpy as np
import pytest

from tests.gpu_util import announce, to_host

pytestmark = pytest.mark.gpu

SEED0_OUT1 = 0xE220A8397B1DCDAF


@pytest.mark.parametrize("count,seed,lo,hi", [
    (1, 0, -10, 10), (4096, 72, -10, 10), (8192, 3, -30, 30), (1000, 9, 4, 4),
    (100003, 11, -10, 10),
])
def test_fill_exponents_device(mgx, orc, count, seed, lo, hi):
    import torch
    announce(f"fill_exponents count={count} seed={seed} [{lo},{hi}]")
    e = torch.empty(count, dtype=torch.int32, device="cuda")
    mgx.fill_exponents(e, count, seed, lo, hi)
    got = to_host(e)
    assert np.array_equal(got, orc.fill_exponents(count, seed, lo, hi))
    assert got.min() >= lo and got.max() <= hi
    if seed == 0:
        assert got[0] == lo + (SEED0_OUT1 >> 11) % (hi - lo + 1)


@pytest.mark.parametrize("rows,cols,ld", [(37, 53, 53), (300, 129, 200), (1, 4096, 4096)])
def test_scale_rows_cols_device(mgx, orc, rows, cols, ld):
    import torch
    announce(f"scale_rows_cols {rows}x{cols} ld={ld}")
    X = orc.fill(rows, cols, 2, seed=71)
    re = orc.fill_exponents(rows, 72, -10, 10)
    ce = orc.fill_exponents(cols, 74, -10, 10)
    dX = torch.full((rows, ld), 7.0, dtype=torch.float64, device="cuda")
    dX[:, :cols] = torch.from_numpy(X).cuda()
    dre = torch.from_numpy(re).cuda()
    dce = torch.from_numpy(ce).cuda()
    for row_exp, col_exp in ((dre, dce), (dre, None), (None, dce), (None, None)):
        dY = dX.clone()
        mgx.scale_rows_cols(dY, rows, cols, ld, row_exp=row_exp, col_exp=col_exp)
        Y = to_host(dY)
        want = orc.scale_rows_cols(X, row_exp=None if row_exp is None else re,
                                   col_exp=None if col_exp is None else ce)
        assert np.array_equal(Y[:, :cols], want)
        assert np.all(Y[:, cols:] == 7.0)         # outside the window untouched
        e = np.zeros((rows, cols), dtype=np.int64)
        if row_exp is not None:
            e += re[:, None]
        if col_exp is not None:
            e += ce[None, :]
        # exact: dividing by the power of two gives the input back
        assert np.array_equal(np.ldexp(Y[:, :cols], -e), X)


def test_stress_inputs_device_equal_oracle(mgx, orc):
    """The C4 stress operands as the accuracy sweep builds them on the device."""
    import torch
    from mgxbench_amd.accuracy_sweep import stress_inputs
    n, seed = 512, 1
    announce("stress inputs n=512")
    A, B = stress_inputs(mgx, torch, n, seed)
    wantA = orc.scale_rows_cols(orc.fill(n, n, 2, seed=seed),
                                row_exp=orc.fill_exponents(n, seed + 2, -10, 10))
    wantB = orc.scale_rows_cols(orc.fill(n, n, 2, seed=seed + 1),
                                col_exp=orc.fill_exponents(n, seed + 3, -10, 10))
    assert np.array_equal(to_host(A), wantA)
    assert np.array_equal(to_host(B), wantB)
