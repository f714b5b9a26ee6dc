"""One rank's share of the 1D row-block product at P = 4, N = 32768 (A_i
8192x32768, B 32768x32768 resident in HBM, C_i 8192x32768), uniform[-1,1)
inputs, through the C ABI.

Size-independent checks at the full size: the FP64-accurate split forms
(15 moduli at beta = 50, INT8 s = 8) are no worse than the plain FP64 MFMA
product against the double-double product on sampled rows, and INT8 s = 6, 7,
8 gain about 8 bits per slice; the plain product is within the FP64 gate of it; and the
modular form's result does not depend on the tile (both tiles compute the
same integers and round once)."""
import numpy as np
import pytest

from tests.gpu_util import announce

pytestmark = pytest.mark.gpu


def test_share_n32768_p4_split_forms_no_worse_than_plain(mgx):
    import torch
    m, n, k = 8192, 32768, 32768
    announce("1D share N=32768 P=4: plain vs modular vs digit slices")
    dev = torch.device("cuda", 0)
    A = torch.empty((m, k), dtype=torch.float64, device=dev)
    B = torch.empty((k, n), dtype=torch.float64, device=dev)
    C = torch.empty((m, n), dtype=torch.float64, device=dev)
    mgx.fill(A, m, k, k, 1, seed=11)
    mgx.fill(B, k, n, n, 1, seed=12)
    rows = np.linspace(0, m - 1, 8).astype(np.int64)

    mgx.ABMultiply(A, B, C, m, n, k, k, n, n, 32)
    plain = mgx.error_report(A, B, C, m, n, k, k, n, n, rows)
    assert plain["normwise"] < 1e-13
    assert mgx.lib().mgx_split_mod_beta(15, k) == 50

    mgx.split_gemm(A, B, C, m, n, k, k, n, n, form=mgx.FORM_INT8_MOD, slices=15)
    mod = mgx.error_report(A, B, C, m, n, k, k, n, n, rows)
    sample = C[:: 1024, :: 1024].cpu().numpy()
    mgx._lib.call("mgx_split_gemm_ex", mgx.FORM_INT8_MOD, 15, mgx._lib.TILE_SMALL, A.data_ptr(),
                  B.data_ptr(), C.data_ptr(), m, n, k, k, n, n, 0, None, 0, 0)
    assert np.array_equal(C[:: 1024, :: 1024].cpu().numpy(), sample)

    # INT8 digit slices s = 6, 7, 8 (the configuration's range): each slice
    # adds 8 bits until FP64 rounding; s = 8 is no worse than plain
    sl = {}
    for s in (6, 7, 8):
        mgx.split_gemm(A, B, C, m, n, k, k, n, n, form=mgx.FORM_INT8, slices=s)
        sl[s] = mgx.error_report(A, B, C, m, n, k, k, n, n, rows)
    print(f"normwise plain {plain['normwise']:.3e} mod15 {mod['normwise']:.3e} " +
          " ".join(f"int8 s={s} {sl[s]['normwise']:.3e}" for s in sl), flush=True)
    assert sl[7]["normwise"] < sl[6]["normwise"] / 64
    assert sl[8]["normwise"] < sl[7]["normwise"]
    assert mod["normwise"] <= plain["normwise"]
    assert mod["componentwise"] <= plain["componentwise"]
    assert sl[8]["normwise"] <= plain["normwise"]
    assert sl[8]["componentwise"] <= plain["componentwise"]
    del A, B, C
    torch.cuda.empty_cache()
