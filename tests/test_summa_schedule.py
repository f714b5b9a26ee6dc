"""The SUMMA loop's group schedule (mgx_summa_groups, a host-only query of
libmgx, no GPU call): hand-computed schedules, the rule restated
independently, and the invariants every schedule must keep (widths sum to
k, whole panels except the ragged last one, uniform groups unless the
tapered schedule is asked for, whose tail appears only when the loop has
more than one full group)."""
import itertools

import pytest

import mgxbench_amd as mgx

FORMS = (mgx.FORM_FP64, mgx.FORM_INT8, mgx.FORM_INT8_MOD, mgx.FORM_BF16)


def _default_width(form):
    return 1024


def _expected(form, panel, k, split_k):
    """The schedule from its definition, written as a walk over panels."""
    p = min(panel, k)
    panels = [min(p, k - k0) for k0 in range(0, k, p)]
    want = abs(split_k) if split_k != 0 else _default_width(form)
    per = min(-(-want // p), len(panels))           # whole panels per group
    sizes = []
    taper = split_k < 0
    if taper and len(panels) > per:
        t = max(1, per // 4)
        tail = [t] * min(4, per // t)
        body = len(panels) - sum(tail)
        sizes = [min(per, body - i) for i in range(0, body, per)] + tail
    else:
        sizes = [min(per, len(panels) - i) for i in range(0, len(panels), per)]
    out, i = [], 0
    for s in sizes:
        out.append(sum(panels[i:i + s]))
        i += s
    return out


def test_hand_computed_schedules():
    # default: uniform groups of whole panels covering 1024 columns
    assert mgx.summa_groups(mgx.FORM_INT8_MOD, 256, 16384) == [1024] * 16
    assert mgx.summa_groups(mgx.FORM_FP64, 256, 16384) == [1024] * 16
    assert mgx.summa_groups(mgx.FORM_FP64, 1000, 16384) == [2000] * 8 + [384]
    assert mgx.summa_groups(mgx.FORM_INT8_MOD, 128, 600) == [600]
    assert mgx.summa_groups(mgx.FORM_INT8_MOD, 256, 6000) == [1024] * 5 + [880]
    # tapered schedule on wider groups: the last group's panels split in four
    assert mgx.summa_groups(mgx.FORM_INT8_MOD, 256, 16384, -4096) == [4096] * 3 + [1024] * 4
    assert mgx.summa_groups(mgx.FORM_INT8, 1024, 16384, -4096) == [4096] * 3 + [1024] * 4
    assert mgx.summa_groups(mgx.FORM_FP64, 256, 16384, -2048) == [2048] * 7 + [512] * 4
    assert mgx.summa_groups(mgx.FORM_INT8_MOD, 256, 6000, -4096) == [2048, 1024, 1024, 1024, 880]
    # k within one group: no taper
    assert mgx.summa_groups(mgx.FORM_INT8_MOD, 256, 4096, -4096) == [4096]
    # explicit group width: uniform
    assert mgx.summa_groups(mgx.FORM_INT8_MOD, 128, 1000, 256) == [256, 256, 256, 232]
    # panel wider than k
    assert mgx.summa_groups(mgx.FORM_FP64, 4096, 1000) == [1000]


@pytest.mark.parametrize("form", FORMS)
def test_schedule_matches_rule_and_invariants(form):
    for panel, k, split_k in itertools.product((1, 7, 64, 128, 256, 300, 1024, 5000),
                                               (1, 63, 256, 1000, 4096, 6000, 16384, 33000),
                                               (0, -3, -1000, -4096, 1, 200, 256, 4096, 10 ** 6)):
        got = mgx.summa_groups(form, panel, k, split_k)
        assert got == _expected(form, panel, k, split_k), (form, panel, k, split_k)
        p = min(panel, k)
        assert sum(got) == k and all(w > 0 for w in got)
        assert all(w % p == 0 for w in got[:-1])
        if split_k > 0:
            g = mgx.split_group_k(form, panel, k, split_k)
            assert got[:-1] == [g] * (len(got) - 1) and got[-1] <= g
        else:
            g = mgx.split_group_k(form, panel, k, -split_k)
            assert max(got) <= g
            if k > g and split_k < 0:   # quarter-group tail
                assert got[-1] <= max(p, g // 4)


def test_invalid_schedules():
    lib = mgx._lib.lib()
    assert lib.mgx_summa_groups(mgx.FORM_FP64, 0, 0, 100, None, 0) == -1
    assert lib.mgx_summa_groups(mgx.FORM_FP64, 128, 0, 0, None, 0) == -1
    assert lib.mgx_summa_groups(99, 128, 0, 100, None, 0) == -1
    with pytest.raises(mgx.MgxError):
        mgx.summa_groups(mgx.FORM_FP64, -1, 100)
