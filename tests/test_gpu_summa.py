"""SUMMA panel streaming (mgx_summa_run) and the GPU accuracy report, through
the C ABI.

The panel loop must equal the reference's: C += A[:, k0:k0+kc] . B[k0:k0+kc, :]
per panel in order (2D-block-matrix/mmb.cu:239-361), ragged last panel
included.  Integer inputs make every order exact, so those cases compare bit
for bit; real inputs are compared with the per-panel reference-order oracle
within 2 gamma_kc |A||B| per panel.  Split forms multiply groups of whole
panels, each group bit-exact against the oracle's model at the group's depth.
Overlapped and serialized runs, graph and stream runs, pinned and pageable
strips must agree bit for bit.
"""
import ctypes

import numpy as np
import pytest

from tests.gpu_util import announce, empty, to_dev, to_host

pytestmark = pytest.mark.gpu


class PinnedArray:
    """A numpy view of library-allocated pinned host memory."""

    def __init__(self, mgx, rows, cols):
        self.mgx = mgx
        ptr = ctypes.c_void_p()
        mgx._lib.call("mgx_host_alloc", ctypes.byref(ptr), rows * cols * 8)
        self.ptr = ptr.value
        buf = (ctypes.c_double * (rows * cols)).from_address(self.ptr)
        self.array = np.ctypeslib.as_array(buf).reshape(rows, cols)

    def free(self):
        self.mgx._lib.call("mgx_host_free", self.ptr)


def _run(mgx, A, B, panel, *, pinned, form=0, slices=8, graph=False, serialize=False,
         copy_mode=0, C0=None, split_k=0):
    m, k = A.shape
    n = B.shape[1]
    dC = to_dev(C0) if C0 is not None else empty(m, n, fill=0.0)
    if pinned:
        pa, pb = PinnedArray(mgx, m, k), PinnedArray(mgx, k, n)
        pa.array[:] = A
        pb.array[:] = B
        try:
            st = mgx.summa_run(pa.ptr, k, pb.ptr, n, dC, n, m=m, n=n, k=k, panel=panel, form=form,
                               slices=slices, use_graph=graph, serialize=serialize,
                               copy_mode=copy_mode, split_k=split_k)
        finally:
            pa.free()
            pb.free()
    else:
        A = np.ascontiguousarray(A)
        B = np.ascontiguousarray(B)
        st = mgx.summa_run(A.ctypes.data, k, B.ctypes.data, n, dC, n, m=m, n=n, k=k, panel=panel,
                           form=form, slices=slices, serialize=serialize, copy_mode=copy_mode,
                           split_k=split_k)
    assert st["panels"] == (k + min(panel, k) - 1) // min(panel, k)
    assert st["host_pinned"] == (1 if pinned else 0)
    return to_host(dC), st


@pytest.mark.parametrize("pinned,graph", [(True, False), (True, True), (False, False)])
def test_integer_panels_exact(mgx, pinned, graph):
    r = np.random.default_rng(1)
    m, n, k, panel = 300, 260, 1000, 256      # ragged last panel of 232
    announce(f"summa integer pinned={pinned} graph={graph}")
    A = r.integers(-60, 60, size=(m, k)).astype(np.float64)
    B = r.integers(-60, 60, size=(k, n)).astype(np.float64)
    C, st = _run(mgx, A, B, panel, pinned=pinned, graph=graph)
    assert np.array_equal(C, (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64))


def test_real_panels_match_reference_order(mgx, orc):
    m, n, k, panel = 257, 199, 700, 128
    announce("summa uniform reference order")
    A = orc.fill(m, k, 1, seed=11)
    B = orc.fill(k, n, 1, seed=12)
    C, _ = _run(mgx, A, B, panel, pinned=True)
    ref = np.zeros((m, n))
    bound = np.zeros((m, n))
    u = 2.0 ** -53
    for k0 in range(0, k, panel):
        kc = min(panel, k - k0)
        Ap = np.ascontiguousarray(A[:, k0:k0 + kc])
        Bp = np.ascontiguousarray(B[k0:k0 + kc])
        # packed panel, lda = kc (2D-block-matrix/mmb.cu:356-358)
        ref = orc.gemm_accumulate_order(Ap, Bp, ref, kc, 0, kc)
        absp = orc.gemm_abs_rows(Ap, Bp)
        bound += 2 * (kc * u / (1 - kc * u)) * absp + 4 * u * (np.abs(ref) + absp)
    assert np.all(np.abs(C - ref) <= bound)


@pytest.mark.parametrize("form", ["fp64", "mod"])
def test_panel_width_1024(mgx, orc, form):
    """The widest panel of the 2x2 campaign (b = 1024) with a ragged last
    panel (k = 5000: four panels of 1024 and one of 904): per-panel reference
    order bound (FP64) or the grouped oracle model (15 moduli), and graph,
    stream and pageable runs equal bit for bit."""
    m, n, k, panel = 384, 320, 5000, 1024
    announce(f"summa b=1024 form={form}")
    A = orc.fill(m, k, 1, seed=41)
    B = orc.fill(k, n, 1, seed=42)
    code, sl = (0, 8) if form == "fp64" else (mgx.FORM_INT8_MOD, 15)
    C, st = _run(mgx, A, B, panel, pinned=True, form=code, slices=sl)
    assert st["panels"] == 5
    C_graph, _ = _run(mgx, A, B, panel, pinned=True, graph=True, form=code, slices=sl)
    C_page, _ = _run(mgx, A, B, panel, pinned=False, form=code, slices=sl)
    assert np.array_equal(C, C_graph) and np.array_equal(C, C_page)
    if form == "fp64":
        ref = np.zeros((m, n))
        bound = np.zeros((m, n))
        u = 2.0 ** -53
        for k0 in range(0, k, panel):
            kc = min(panel, k - k0)
            Ap = np.ascontiguousarray(A[:, k0:k0 + kc])
            Bp = np.ascontiguousarray(B[k0:k0 + kc])
            ref = orc.gemm_accumulate_order(Ap, Bp, ref, kc, 0, kc)
            absp = orc.gemm_abs_rows(Ap, Bp)
            bound += 2 * (kc * u / (1 - kc * u)) * absp + 4 * u * (np.abs(ref) + absp)
        assert np.all(np.abs(C - ref) <= bound)
    else:
        # default groups of 1024 columns: every panel is its own split product
        groups = mgx.summa_groups(code, panel, k)
        assert groups == [1024, 1024, 1024, 1024, 904]
        assert st["groups"] == 5
        ref = _grouped(A, B, groups, lambda a, b, C: orc.crt_gemm(a, b, 15, C=C))
        assert np.array_equal(C, ref)


def test_overlapped_equals_serialized_and_graph(mgx, orc):
    m, n, k, panel = 320, 288, 1500, 512
    announce("summa overlap vs serialize vs graph")
    A = orc.fill(m, k, 1, seed=21)
    B = orc.fill(k, n, 1, seed=22)
    C_over, _ = _run(mgx, A, B, panel, pinned=True)
    C_ser, _ = _run(mgx, A, B, panel, pinned=True, serialize=True)
    C_graph, _ = _run(mgx, A, B, panel, pinned=True, graph=True)
    C_page, _ = _run(mgx, A, B, panel, pinned=False)
    C_sync, _ = _run(mgx, A, B, panel, pinned=True, copy_mode=1)
    for other in (C_ser, C_graph, C_page, C_sync):
        assert np.array_equal(C_over, other)


def test_graph_replays_are_identical(mgx, orc):
    m, n, k, panel = 200, 200, 600, 256
    announce("summa graph repeat")
    A = orc.fill(m, k, 1, seed=31)
    B = orc.fill(k, n, 1, seed=32)
    first, _ = _run(mgx, A, B, panel, pinned=True, graph=True)
    for _ in range(2):
        again, _ = _run(mgx, A, B, panel, pinned=True, graph=True)
        assert np.array_equal(first, again)


def _grouped(A, B, groups, product):
    """The oracle's split product per group (uniform groups of `groups`
    columns, or the listed widths), accumulated into C in group order."""
    k = A.shape[1]
    if isinstance(groups, int):
        groups = [min(groups, k - g0) for g0 in range(0, k, groups)]
    assert sum(groups) == k
    ref = np.zeros((A.shape[0], B.shape[1]))
    g0 = 0
    for w in groups:
        ref = product(np.ascontiguousarray(A[:, g0:g0 + w]), np.ascontiguousarray(B[g0:g0 + w]),
                      C=ref)
        g0 += w
    return ref


@pytest.mark.parametrize("split_k,group", [(0, 600), (256, 256), (300, 384), (100, 128)])
@pytest.mark.parametrize("form_name", ["int8", "mod", "bf16"])
def test_split_forms_multiply_whole_panel_groups(mgx, orc, form_name, split_k, group):
    # panels of 128 (ragged last of 88); each group of whole panels is one
    # split product, bit-exact against the oracle's model at that depth (the
    # bf16 form gives the INT8 form's result)
    m, n, k, panel = 190, 170, 600, 128
    announce(f"summa {form_name} split_k={split_k}")
    form = {"int8": mgx.FORM_INT8, "mod": mgx.FORM_INT8_MOD, "bf16": mgx.FORM_BF16}[form_name]
    assert mgx.split_group_k(form, panel, k, split_k) == group
    # FP64 groups whole panels by the same rule (default 1024 columns)
    assert mgx.split_group_k(mgx.FORM_FP64, panel, k, split_k) == group
    uniform = [min(group, k - g0) for g0 in range(0, k, group)]
    assert mgx.summa_groups(form, panel, k, split_k) == uniform
    A = orc.fill(m, k, 1, seed=41)
    B = orc.fill(k, n, 1, seed=42)
    if form_name in ("int8", "bf16"):
        C, st = _run(mgx, A, B, panel, pinned=True, form=form, slices=8, split_k=split_k)
        ref = _grouped(A, B, group, lambda a, b, C: orc.split_gemm_int8(a, b, 8, C=C))
    else:
        C, st = _run(mgx, A, B, panel, pinned=True, form=form, slices=15, split_k=split_k)
        ref = _grouped(A, B, group, lambda a, b, C: orc.crt_gemm(a, b, 15, C=C))
    assert st["panels"] == 5
    assert np.array_equal(C, ref)


def test_split_groups_agree_across_streaming_modes(mgx, orc):
    # groups of 2 panels, ragged last panel inside the last group; pinned DMA,
    # pageable staging, host-synchronised copies, serialized and graph runs
    m, n, k, panel = 160, 150, 1000, 128
    announce("summa modular groups: streaming modes")
    A = orc.fill(m, k, 1, seed=45)
    B = orc.fill(k, n, 1, seed=46)
    kw = dict(form=mgx.FORM_INT8_MOD, slices=15, split_k=256)
    C0, _ = _run(mgx, A, B, panel, pinned=True, **kw)
    assert np.array_equal(C0, _grouped(A, B, 256, lambda a, b, C: orc.crt_gemm(a, b, 15, C=C)))
    for extra in (dict(pinned=False), dict(pinned=True, graph=True),
                  dict(pinned=True, serialize=True), dict(pinned=True, copy_mode=1),
                  dict(pinned=False, copy_mode=1)):
        C1, _ = _run(mgx, A, B, panel, **extra, **kw)
        assert np.array_equal(C0, C1), extra


def test_padded_host_strips_grouped_copies(mgx, orc):
    # pinned strips with lda > k and ldb > n: each split group arrives as one
    # strided A copy and one B copy; the pageable run copies panel by panel;
    # both equal the oracle's grouped model bit for bit
    m, n, k, panel, lda, ldb = 150, 140, 700, 128, 700 + 24, 140 + 8
    announce("summa modular groups: padded host strips")
    A = orc.fill(m, k, 1, seed=47)
    B = orc.fill(k, n, 1, seed=48)
    ref = _grouped(A, B, 256, lambda a, b, C: orc.crt_gemm(a, b, 15, C=C))
    pa, pb = PinnedArray(mgx, m, lda), PinnedArray(mgx, k, ldb)
    try:
        pa.array[:] = 7.0
        pb.array[:] = 7.0
        pa.array[:, :k] = A
        pb.array[:, :n] = B
        dC = empty(m, n, fill=0.0)
        st = mgx.summa_run(pa.ptr, lda, pb.ptr, ldb, dC, n, m=m, n=n, k=k, panel=panel,
                           form=mgx.FORM_INT8_MOD, slices=15, split_k=256)
        C_pinned = to_host(dC)
    finally:
        pa.free()
        pb.free()
    assert st["panels"] == 6 and st["host_pinned"] == 1
    Ap = np.full((m, lda), 7.0)
    Bp = np.full((k, ldb), 7.0)
    Ap[:, :k] = A
    Bp[:, :n] = B
    dC = empty(m, n, fill=0.0)
    mgx.summa_run(Ap.ctypes.data, lda, Bp.ctypes.data, ldb, dC, n, m=m, n=n, k=k, panel=panel,
                  form=mgx.FORM_INT8_MOD, slices=15, split_k=256)
    assert np.array_equal(C_pinned, ref)
    assert np.array_equal(to_host(dC), ref)


def test_invalid_summa_arguments(mgx):
    A = np.zeros((4, 4))
    dC = empty(4, 4)
    with pytest.raises(mgx.MgxError):      # graph capture needs pinned strips
        mgx.summa_run(A.ctypes.data, 4, A.ctypes.data, 4, dC, 4, m=4, n=4, k=4, panel=2,
                      use_graph=True)
    with pytest.raises(mgx.MgxError):
        mgx.summa_run(A.ctypes.data, 4, A.ctypes.data, 4, dC, 4, m=4, n=4, k=4, panel=0)
    for threads in (-1, 65):
        with pytest.raises(mgx.MgxError):
            mgx.summa_run(A.ctypes.data, 4, A.ctypes.data, 4, dC, 4, m=4, n=4, k=4, panel=2,
                          host_threads=threads)


# ---- accuracy report -------------------------------------------------------------------

def test_error_report_matches_oracle(mgx, orc):
    m, n, k = 150, 130, 400
    announce("error report vs oracle")
    A = orc.fill(m, k, 1, seed=51)
    B = orc.fill(k, n, 1, seed=52)
    dA, dB = to_dev(A), to_dev(B)
    dC = empty(m, n)
    mgx.ABMultiply(dA, dB, dC, m, n, k, k, n, n, 32)
    C = to_host(dC)
    rows = np.array([0, 7, 149, 64])
    rep = mgx.error_report(dA, dB, dC, m, n, k, k, n, n, rows)
    hi, lo = orc.gemm_dd_rows(A, B, rows)
    d = (C[rows] - hi) - lo
    assert rep["rows_checked"] == 4
    assert rep["normwise"] == pytest.approx(np.linalg.norm(d) / np.linalg.norm(hi), rel=1e-9)
    assert rep["max_abs"] == pytest.approx(np.abs(d).max(), rel=1e-12)
    absp = orc.gemm_abs_rows(A, B, rows)
    assert rep["componentwise"] == pytest.approx((np.abs(d) / absp).max(), rel=1e-9)


def test_error_report_planted_error(mgx, orc):
    m, n, k = 64, 64, 64
    announce("error report planted")
    r = np.random.default_rng(2)
    A = r.integers(-9, 9, size=(m, k)).astype(np.float64)
    B = r.integers(-9, 9, size=(k, n)).astype(np.float64)
    C = A @ B                                    # exact
    C[5, 6] += 0.5
    rep = mgx.error_report(to_dev(A), to_dev(B), to_dev(C), m, n, k, k, n, n, np.arange(m))
    assert rep["max_abs"] == 0.5
    C[5, 6] = np.nan
    rep = mgx.error_report(to_dev(A), to_dev(B), to_dev(C), m, n, k, k, n, n, np.arange(m))
    assert np.isinf(rep["normwise"])


@pytest.mark.parametrize("form_name", ["mod", "int8"])
def test_tapered_schedule(mgx, orc, form_name):
    # 24 panels of 256 (ragged last of 112), tapered schedule on groups of 16
    # panels: one group of 8 panels, then four of 4; each launch is one split
    # product, bit-exact against the oracle's model at its own depth; pinned,
    # graph and pageable runs agree
    m, n, k, panel = 96, 80, 6000, 256
    announce(f"summa {form_name} tapered schedule")
    form, sl = ((mgx.FORM_INT8_MOD, 15) if form_name == "mod" else (mgx.FORM_INT8, 8))
    groups = mgx.summa_groups(form, panel, k, -4096)
    assert groups == [2048, 1024, 1024, 1024, 880]
    A = orc.fill(m, k, 1, seed=61)
    B = orc.fill(k, n, 1, seed=62)
    C, st = _run(mgx, A, B, panel, pinned=True, form=form, slices=sl, split_k=-4096)
    assert st["groups"] == 5
    if form_name == "mod":
        ref = _grouped(A, B, groups, lambda a, b, C: orc.crt_gemm(a, b, 15, C=C))
    else:
        ref = _grouped(A, B, groups, lambda a, b, C: orc.split_gemm_int8(a, b, 8, C=C))
    assert np.array_equal(C, ref)
    C_graph, _ = _run(mgx, A, B, panel, pinned=True, graph=True, form=form, slices=sl,
                      split_k=-4096)
    C_page, _ = _run(mgx, A, B, panel, pinned=False, form=form, slices=sl, split_k=-4096)
    assert np.array_equal(C, C_graph) and np.array_equal(C, C_page)


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_pageable_packing_threads(mgx, orc, threads):
    # pageable strips packed by 1, 3 or 8 host threads into staging units of
    # four panels (ragged last panel, units straddling no group boundary):
    # FP64 integer inputs exact, modular groups equal to the pinned run
    r = np.random.default_rng(5 + threads)
    m, n, k, panel = 300, 260, 3000, 256
    announce(f"summa pageable packing threads={threads}")
    A = r.integers(-60, 60, size=(m, k)).astype(np.float64)
    B = r.integers(-60, 60, size=(k, n)).astype(np.float64)
    dC = empty(m, n, fill=0.0)
    st = mgx.summa_run(A.ctypes.data, k, B.ctypes.data, n, dC, n, m=m, n=n, k=k, panel=panel,
                       host_threads=threads)
    assert st["host_pinned"] == 0 and st["host_threads"] == threads
    assert st["groups"] == len(mgx.summa_groups(mgx.FORM_FP64, panel, k))
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    assert np.array_equal(to_host(dC), exact)
    Ar = orc.fill(m, k, 1, seed=70 + threads)
    Br = orc.fill(k, n, 1, seed=80 + threads)
    C_pin, _ = _run(mgx, Ar, Br, panel, pinned=True, form=mgx.FORM_INT8_MOD, slices=15)
    dC = empty(m, n, fill=0.0)
    mgx.summa_run(Ar.ctypes.data, k, Br.ctypes.data, n, dC, n, m=m, n=n, k=k, panel=panel,
                  form=mgx.FORM_INT8_MOD, slices=15, host_threads=threads)
    assert np.array_equal(to_host(dC), C_pin)


@pytest.mark.parametrize("form_name", ["fp64", "mod"])
def test_copies_overlap_products(mgx, form_name):
    # 4096 x 4096 share, k = 8192 from pinned strips: 512 MiB of panels
    # (about 9 ms at PCIe rates) against 2-4 ms of products.  The overlapped
    # loop must beat the serialized one (host waits after every copy and
    # every product) by a clear margin: it hides most of the products under
    # the copies.  Both loops run twice; the faster run of each counts.
    m = n = 4096
    k, panel = 8192, 256
    announce(f"summa overlap timing {form_name}")
    form, sl = (mgx.FORM_FP64, 8) if form_name == "fp64" else (mgx.FORM_INT8_MOD, 15)
    pa, pb = PinnedArray(mgx, m, k), PinnedArray(mgx, k, n)
    try:
        r = np.random.default_rng(3)
        pa.array[:] = r.uniform(-1, 1, size=(m, k))
        pb.array[:] = r.uniform(-1, 1, size=(k, n))
        dC = empty(m, n, fill=0.0)
        times = {}
        for serialize in (False, True):
            times[serialize] = min(
                mgx.summa_run(pa.ptr, k, pb.ptr, n, dC, n, m=m, n=n, k=k, panel=panel,
                              form=form, slices=sl, serialize=serialize)["seconds"]
                for _ in range(2))
    finally:
        pa.free()
        pb.free()
    print(f"{form_name}: overlapped {times[False] * 1e3:.2f} ms, serialized "
          f"{times[True] * 1e3:.2f} ms", flush=True)
    assert times[False] < 0.9 * times[True]


@pytest.mark.parametrize("panel", [256, 512, 1024])
def test_share_2x2_n16384_streamed(mgx, panel):
    """The 2x2 rank share at N = 16384 (C 8192 x 8192, K = 16384) as the
    streamed loop, in panels of b = 256, 512 and 1024 from pinned host strips
    (1 GiB each).  FP64: bit-identical to the resident panel-fused launch over
    the same panels (one partial per panel, added in order).  Modular form,
    15 moduli: bit-identical to split products of the loop's panel groups
    accumulated on the device in order, and no worse than plain FP64 against
    the double-double product on sampled rows."""
    import torch
    m = n = 8192
    k = 16384
    announce(f"summa 2x2 share N=16384 streamed b={panel}")
    dev = torch.device("cuda", 0)
    dA = torch.empty((m, k), dtype=torch.float64, device=dev)
    dB = torch.empty((k, n), dtype=torch.float64, device=dev)
    mgx.fill(dA, m, k, k, 1, seed=21)
    mgx.fill(dB, k, n, n, 1, seed=22)
    pa, pb = PinnedArray(mgx, m, k), PinnedArray(mgx, k, n)
    try:
        torch.from_numpy(pa.array).copy_(dA)
        torch.from_numpy(pb.array).copy_(dB)
        rows = np.arange(0, m, 1024)
        out = {}
        for form, sl in ((mgx.FORM_FP64, 8), (mgx.FORM_INT8_MOD, 15)):
            dC = torch.zeros((m, n), dtype=torch.float64, device=dev)
            st = mgx.summa_run(pa.ptr, k, pb.ptr, n, dC, n, m=m, n=n, k=k, panel=panel,
                               form=form, slices=sl)
            assert st["panels"] == k // panel and st["host_pinned"] == 1
            out[form] = dC
    finally:
        pa.free()
        pb.free()
    ref = torch.empty((m, n), dtype=torch.float64, device=dev)
    mgx._lib.call("mgx_dgemm_panels", 0, 0, dA.data_ptr(), dB.data_ptr(), ref.data_ptr(),
                  m, n, k, panel, k, n, n, 0)
    assert torch.equal(out[mgx.FORM_FP64], ref)
    ref.zero_()
    k0 = 0
    for g in mgx.summa_groups(mgx.FORM_INT8_MOD, panel, k):
        mgx._lib.call("mgx_split_gemm_ex", mgx.FORM_INT8_MOD, 15, 0, dA.data_ptr() + 8 * k0,
                      dB.data_ptr() + 8 * k0 * n, ref.data_ptr(), m, n, g, k, n, n, 1, None,
                      0, 0)
        k0 += g
    assert k0 == k
    assert torch.equal(out[mgx.FORM_INT8_MOD], ref)
    plain = mgx.error_report(dA, dB, out[mgx.FORM_FP64], m, n, k, k, n, n, rows)
    mod = mgx.error_report(dA, dB, out[mgx.FORM_INT8_MOD], m, n, k, k, n, n, rows)
    print(f"b={panel}: normwise plain {plain['normwise']:.3e} modular {mod['normwise']:.3e}",
          flush=True)
    assert plain["normwise"] < 1e-14
    assert mod["normwise"] <= plain["normwise"]
    assert mod["componentwise"] <= plain["componentwise"]
    del dA, dB, out, ref
    torch.cuda.empty_cache()
