"""The accuracy sweep on stress inputs (mgxbench_amd.accuracy_sweep) at a small
size: errors against the GPU double-double product fall with every added
slice or modulus until they reach FP64 rounding, the BF16 form gives the INT8
form's numbers, and the FP64-accurate settings are no worse than plain FP64.
(The double-double reference itself is pinned against the host oracle in
test_gpu_summa.py.)"""
import pytest

pytestmark = pytest.mark.gpu


def test_sweep_n512_stress():
    from mgxbench_amd.accuracy_sweep import sweep
    res = sweep(512, rows=0, int8=range(2, 10), bf16=range(2, 6), moduli=range(6, 17),
                log=lambda s: print(s, flush=True))
    assert res["rows_checked"] == 512
    by = {(r["form"], r["count"]): r for r in res["results"]}
    plain = by[("fp64", 0)]["normwise"]
    int8 = [by[("int8", s)]["normwise"] for s in range(2, 10)]
    # each slice adds 8 bits until the error reaches FP64 rounding
    for s in range(2, 6):
        assert int8[s - 1] < int8[s - 2] / 16, (s, int8)
    # s = 7 leaves about 54 bits per row / column maximum: on these inputs
    # at k = 512 that is ~3x plain FP64's error; s >= 8 is below it
    assert all(by[("int8", s)]["normwise"] <= plain for s in (8, 9))
    for s in range(2, 6):
        for key in ("normwise", "componentwise", "max_abs"):
            assert by[("bf16", s)][key] == by[("int8", s)][key]
    mods = [by[("mod", m)]["normwise"] for m in range(6, 17) if ("mod", m) in by]
    assert len(mods) >= 8
    # each modulus adds about 4 bits to both operands until FP64 rounding
    assert all(b < a for a, b in zip(mods, mods[1:]) if a > 1e-14), mods
    assert mods[-1] < mods[0] / 1000
    assert by[("mod", 16)]["normwise"] <= plain


@pytest.mark.parametrize("n,rows", [(4096, 0), (8192, 256)])
def test_sweep_stress_full_size(n, rows):
    """The accuracy-stress configuration at its own sizes: N = 4096 over all
    rows, N = 8192 over 256 evenly spaced rows (all columns), INT8 s = 3..9,
    BF16 s = 2..5, moduli 8..16.  Bars: every slice adds about 8 bits and
    every modulus about 4 until FP64 rounding; BF16 equals INT8 exactly; the
    FP64-accurate settings (s >= 8, 15 and 16 moduli) are no worse than the
    plain FP64 MFMA product, normwise and componentwise."""
    from mgxbench_amd.accuracy_sweep import sweep
    res = sweep(n, rows=rows, int8=range(3, 10), bf16=range(2, 6), moduli=range(8, 17),
                log=lambda s: print(s, flush=True))
    assert res["rows_checked"] == (n if rows == 0 else rows)
    by = {(r["form"], r["count"]): r for r in res["results"]}
    plain = by[("fp64", 0)]
    assert plain["normwise"] < 1e-14
    for s in range(3, 7):
        assert by[("int8", s + 1)]["normwise"] < by[("int8", s)]["normwise"] / 64, s
    for s in range(3, 6):
        for key in ("normwise", "componentwise", "max_abs"):
            assert by[("bf16", s)][key] == by[("int8", s)][key]
    for m in range(8, 15):
        assert by[("mod", m + 1)]["normwise"] < by[("mod", m)]["normwise"] / 4, m
    for key in ("normwise", "componentwise"):
        for form, count in (("int8", 8), ("int8", 9), ("mod", 15), ("mod", 16)):
            assert by[(form, count)][key] <= plain[key], (form, count, key)
