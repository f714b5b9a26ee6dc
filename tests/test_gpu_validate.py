"""The constant-result validation (mgx_validate_constant, the counterpart of
validate_C_kernel / validate_matrix_C, 1D-row-block/mulmat_kernel.cu:165-326)
on failing data, against the oracle's restatement of the rule
(orc_validate_constant).

An element is invalid iff |C - e| > abs_tol AND |C - e| / |e| > rel_tol (the
relative error is 0 when e == 0), so an element over only one tolerance is
valid.  Planted cases: over the absolute tolerance only, over the relative
tolerance only, over both, NaN (compares false: valid; the maxima skip it),
+-Inf (invalid, maxima Inf) and e == 0 (never invalid).  The invalid count and
the maximum absolute / relative errors must equal the oracle's exactly: both
take fabs(C - e) and its quotient by |e| in FP64, and the maxima do not depend
on the order of the reduction.  Shapes: the 1D rank share at P = 4, N = 32768
(C_i 8192 x 32768, 2^28 elements) and ragged ones.
"""
import numpy as np
import pytest

from tests.gpu_util import announce

pytestmark = pytest.mark.gpu

ABS_TOL, REL_TOL = 1.0e-9, 1.0e-12


def _plant(C, expected, rng, count):
    """Writes `count` planted values of every class at random distinct
    positions of the device matrix C (torch, used for memory only); returns the
    positions per class."""
    import torch
    total = C.numel()
    e = abs(expected)
    classes = {
        # |err| just above abs_tol and at most rel_tol * |e| (needs |e| > 1e-3)
        "abs_only": lambda: expected + (1.5 * ABS_TOL if e * REL_TOL > 1.5 * ABS_TOL else 0.0),
        # |err| below abs_tol but above rel_tol * |e| (needs |e| < 1e3)
        "rel_only": lambda: expected + (0.5 * ABS_TOL if e * REL_TOL < 0.5 * ABS_TOL else 0.0),
        "both": lambda: expected + float(rng.uniform(1e-6, 10.0)) * rng.choice([-1.0, 1.0]),
        "nan": lambda: float("nan"),
        "inf": lambda: float(rng.choice([np.inf, -np.inf])),
    }
    pos = rng.choice(total, size=min(total, count * len(classes)), replace=False)
    flat = C.view(-1)
    out = {}
    for i, (name, make) in enumerate(classes.items()):
        idx = pos[i::len(classes)]
        vals = torch.tensor([make() for _ in idx], dtype=torch.float64)
        flat[torch.from_numpy(idx).to(C.device)] = vals.to(C.device)
        out[name] = idx
    return out


def _compare(mgx, orc, C, rows, cols, expected):
    import torch
    torch.cuda.synchronize()
    got = mgx.validate_matrix_C(C, rows, cols, expected, ABS_TOL, REL_TOL, rank=1, details=True)
    bad, max_abs, max_rel = orc.validate_constant(C.cpu().numpy(), expected, ABS_TOL, REL_TOL)
    assert got["invalid"] == bad
    assert got["pass"] == (1 if bad == 0 else 0)
    # exact equality, Inf included
    assert got["max_abs"] == max_abs or (np.isnan(got["max_abs"]) and np.isnan(max_abs))
    assert got["max_rel"] == max_rel or (np.isnan(got["max_rel"]) and np.isnan(max_rel))
    return got


@pytest.mark.parametrize("rows,cols,expected", [
    (1, 1, 2048.0), (3, 5, 6.0), (257, 1023, 65536.0), (1000, 1000, 1e-3), (511, 129, 0.0),
    (64, 4099, -7.5),
])
def test_planted_ragged(mgx, orc, rows, cols, expected):
    import torch
    announce(f"validate planted {rows}x{cols} e={expected}")
    rng = np.random.default_rng(rows * 131 + cols)
    C = torch.full((rows, cols), expected, dtype=torch.float64, device="cuda")
    got = _compare(mgx, orc, C, rows, cols, expected)
    assert got["invalid"] == 0 and got["max_abs"] == 0.0
    planted = _plant(C, expected, rng, count=max(1, rows * cols // 50))
    got = _compare(mgx, orc, C, rows, cols, expected)
    n_both = len(planted["both"]) + len(planted["inf"])
    if expected == 0.0:
        assert got["invalid"] == 0 and got["max_rel"] == 0.0   # relative error is 0
    elif rows * cols >= 5:
        assert got["invalid"] == n_both
        assert got["max_abs"] == np.inf


def test_planted_one_class_at_a_time(mgx, orc):
    import torch
    rows, cols, e = 300, 700, 65536.0
    announce("validate one class at a time")
    for value, invalid in ((e + 1.5e-9, 0), (e + 1e-6, 1), (e - 1e-6, 1), (float("nan"), 0),
                           (float("inf"), 1), (-float("inf"), 1), (np.nextafter(e, 0), 0)):
        C = torch.full((rows, cols), e, dtype=torch.float64, device="cuda")
        C[123, 456] = value
        got = _compare(mgx, orc, C, rows, cols, e)
        assert got["invalid"] == invalid, value
    # relative-only over tolerance at a small expected value is valid
    e = 1e-3
    C = torch.full((rows, cols), e, dtype=torch.float64, device="cuda")
    C[0, 0] = e + 5e-10
    got = _compare(mgx, orc, C, rows, cols, e)
    assert got["invalid"] == 0 and got["max_rel"] > REL_TOL and got["max_abs"] < ABS_TOL


def test_planted_1d_share_n32768(mgx, orc):
    """C_i of rank 0 at P = 4, N = 32768 (8192 x 32768), expected 2N."""
    import torch
    rows, cols = 8192, 32768
    expected = 2.0 * 32768
    announce("validate planted 1D share 8192x32768")
    C = torch.empty((rows, cols), dtype=torch.float64, device="cuda")
    mgx.fill(C, rows, cols, cols, 0, value=expected)
    rng = np.random.default_rng(2024)
    planted = _plant(C, expected, rng, count=400)
    got = _compare(mgx, orc, C, rows, cols, expected)
    assert got["invalid"] == len(planted["both"]) + len(planted["inf"])
    # the first and the last element of the share
    flat = C.view(-1)
    flat[:] = expected
    flat[0] = expected + 3.0
    flat[-1] = expected - 4.0
    got = _compare(mgx, orc, C, rows, cols, expected)
    assert got["invalid"] == 2 and got["max_abs"] == 4.0 and got["max_rel"] == 4.0 / expected
    del C
    torch.cuda.empty_cache()
