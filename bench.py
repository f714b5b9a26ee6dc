#!/usr/bin/env python3
"""Headline benchmark: FP64-accurate TFLOP/s at N=16384, split-precision MFMA
vs native FP64 MFMA, on one MI355X.

Workload (one step): one rank's local product of the 2D SUMMA 2x2 grid at
N=16384, C_ij (8192 x 8192) = A_ik (8192 x 16384) . B_kj (16384 x 8192), FP64,
inputs uniform[-1,1) from the splitmix64 stream (A seed 1, B seed 2), all
operands resident in HBM.  Every form is credited F = 2 M N K flop.

  * split (the `value`): libmgx INT8 modular product (Ozaki scheme II) with
    `--moduli` moduli, timed end to end (exponent scan, residue extraction,
    the per-modulus INT8 products, exact reconstruction into C);
  * split_slices: libmgx INT8 digit-slice product (Ozaki scheme I) with
    `--slices` slices, timed the same way;
  * plain: libmgx FP64 MFMA product on the same inputs.

Each split form's error against a double-double A.B on sampled rows is
reported next to the plain kernel's.  The dominant kernel of the headline
form is the INT8 residue-product kernel; its duration is read from HIP
events recorded on the launch stream around it inside the timed region.  The
CPU baseline is the oracle's restatement of the reference kernel's
arithmetic (sequential fma chains) run with OpenMP on the host cores over a
row slab of the same product.

Prints ONE JSON line on stdout.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

METRIC = "FP64-accurate TFLOP/s at N=16384: split-precision MFMA vs native FP64 MFMA"


def workload_name(m: int, n: int, k: int) -> str:
    """The campaign share these dimensions are, if any."""
    if (m, n, k) == (8192, 8192, 16384):
        return "2D SUMMA 2x2 rank share at N=16384"
    if (m, n, k) == (8192, 32768, 32768):
        return "1D row-block rank share at P=4, N=32768"
    return "local product"


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--moduli", type=int, default=15,
                    help="moduli of the headline INT8 modular form")
    ap.add_argument("--slices", type=int, default=7,
                    help="slices of the INT8 digit-slice form (secondary line)")
    ap.add_argument("--m", type=int, default=8192)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--k", type=int, default=16384)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="target CPU time of the CPU-baseline row slab")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--bf16", action="store_true",
                    help="also time the BF16 digit-slice form (same slices as --slices)")
    ap.add_argument("--no-summa", action="store_true",
                    help="skip the host-streamed SUMMA panel report")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one device copy of this many MiB first (known-size reference "
                         "for hardware byte counters)")
    return ap.parse_args()


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(m: int, n: int, k: int, target_s: float) -> dict:
    """Oracle port of the reference product, timed on a row slab (host cores)."""
    import oracle  # test infrastructure: used only as the CPU baseline
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    B = oracle.fill(k, n, 1, seed=2)

    def run(rows: int) -> float:
        A = oracle.fill(rows, k, 1, seed=1)  # the first `rows` rows of A
        t0 = time.perf_counter()
        oracle.gemm_kernel_order(A, B, nthreads=threads)
        return time.perf_counter() - t0

    r0 = 8
    t0 = run(r0)
    rows = int(max(8, min(m, r0 * target_s / max(t0, 1e-3))))
    rows = max(8, rows // 8 * 8)
    t = run(rows)
    tflops = 2.0 * rows * n * k / t / 1e12
    full_s = 2.0 * m * n * k / (tflops * 1e12)
    return {
        "value": tflops,
        "unit": "TFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{rows} rows x {n} cols x K={k} of the {m}x{n}x{k} share "
                   f"({t:.2f} s); reference kernel arithmetic (sequential fma chains), "
                   f"OpenMP {threads} threads; full share extrapolated {full_s:.1f} s"),
    }


def profiled_traffic(m: int, n: int, k: int, nmod: int):
    """HBM bytes per launch of the residue-product kernel from the committed
    counter summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), when it covers
    this exact workload; None otherwise."""
    import re
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"pmc_split_forms_{m}x{n}x{k}.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    hit = re.search(rf"traffic_bytes_per_launch int8_product_kernel modular {nmod} moduli "
                    rf"{m}x{n}x{k}: (\d+)", text)
    return int(hit.group(1)) if hit else None


def summa_report(mgx, torch, A, B, m, n, k, nmod) -> dict:
    """The same rank share as a SUMMA panel loop (mgx_summa_run), per panel
    width and form:
      streamed_ms            panels from pinned host strips (strided DMA,
                             hipGraph-captured loop, the default group schedule);
      streamed_pageable_ms   pageable strips packed on the host into pinned
                             staging inside the timed loop (the reference's
                             timed region), by the library's host threads;
      hbm_resident_ms        every panel already in HBM: the whole loop is one
                             group (one launch; FP64 still one partial per panel);
      hbm_resident_grouped_ms  resident operands in the streamed loop's
                             uniform group width (its compute side)."""
    import ctypes
    log("summa: pinned and pageable host strips")
    hA = torch.empty((m, k), dtype=torch.float64, pin_memory=True)
    hB = torch.empty((k, n), dtype=torch.float64, pin_memory=True)
    hA.copy_(A)
    hB.copy_(B)
    pA = torch.empty((m, k), dtype=torch.float64)     # pageable copies
    pB = torch.empty((k, n), dtype=torch.float64)
    pA.copy_(hA)
    pB.copy_(hB)
    torch.cuda.synchronize()
    dC = torch.empty((m, n), dtype=torch.float64, device=A.device)
    rep = {"form": f"int8 modular, {nmod} moduli", "h2d_bytes": 8 * (m * k + k * n), "panels": {}}

    def resident(form, sl, panel, split_k):
        cfg = mgx._lib.TuneConfig(form, sl, mgx._lib.TILE_DEFAULT, panel, split_k)
        ms = ctypes.c_double()
        mgx._lib.call("mgx_time_local_product", ctypes.byref(cfg), A.data_ptr(), B.data_ptr(),
                      dC.data_ptr(), m, n, k, k, n, n, None, 0, 1, 2, ctypes.byref(ms))
        return ms.value

    for panel in (256, 512, 1024):
        row = {}
        for form_name, form, sl in (("split", mgx.FORM_INT8_MOD, nmod), ("plain", mgx.FORM_FP64, 0)):
            dC.zero_()
            torch.cuda.synchronize()
            st = mgx.summa_run(hA.data_ptr(), k, hB.data_ptr(), n, dC, n, m=m, n=n, k=k,
                               panel=panel, form=form, slices=sl or 8, use_graph=True)
            dC.zero_()
            torch.cuda.synchronize()
            sp = mgx.summa_run(pA.data_ptr(), k, pB.data_ptr(), n, dC, n, m=m, n=n, k=k,
                               panel=panel, form=form, slices=sl or 8)
            landed = resident(form, sl, panel, k)
            grouped = resident(form, sl, panel, 0)
            streamed_ms = st["seconds"] * 1e3
            row[form_name] = {
                "groups": st["groups"],
                "streamed_ms": round(streamed_ms, 3),
                "streamed_pageable_ms": round(sp["seconds"] * 1e3, 3),
                "pageable_host_threads": sp["host_threads"],
                "hbm_resident_ms": round(landed, 3),
                "hbm_resident_grouped_ms": round(grouped, 3),
                "h2d_GBps": round(rep["h2d_bytes"] / (streamed_ms * 1e-3) / 1e9, 2),
                "tflops_streamed": round(2.0 * m * n * k / (streamed_ms * 1e-3) / 1e12, 3),
            }
            log(f"summa b={panel} {form_name}: streamed {streamed_ms:.2f} ms "
                f"({st['groups']} groups), pageable {sp['seconds'] * 1e3:.2f} ms "
                f"({sp['host_threads']} host threads), HBM-resident {landed:.2f} ms "
                f"(grouped {grouped:.2f} ms)")
        rep["panels"][str(panel)] = row
    del hA, hB, pA, pB
    return rep


def vendor_fp64(torch, A, B, stream, timed, warmup, steps) -> dict:
    """Sanity roof for the plain denominator: the vendor FP64 GEMM (torch.matmul
    in float64, which dispatches to hipBLASLt / rocBLAS) on the same operands."""
    with torch.cuda.stream(stream):
        out = torch.empty((A.shape[0], B.shape[1]), dtype=torch.float64, device=A.device)
        for _ in range(warmup):
            torch.matmul(A, B, out=out)
        torch.cuda.synchronize()
        ms = timed(lambda i: torch.matmul(A, B, out=out), steps)
    F = 2.0 * A.shape[0] * B.shape[1] * A.shape[1]
    del out
    return {"library": "torch.matmul float64 (hipBLASLt / rocBLAS)", "ms_per_step": round(ms, 4),
            "tflops": round(F / (ms * 1e-3) / 1e12, 3)}


def mfma_ceiling(mgx, torch, stream, iters=12000, reps=3):
    """INT8 matrix-core rate of an MFMA-only loop shaped like the product
    kernel (mgx_mfma_i8_ceiling: one 8-wave workgroup per CU), with random
    operand bytes re-read from LDS every round and with constant (all-ones)
    register operands, in TOP/s (hipEvent timing on `stream`).  The random
    rate is what the part sustains on data like the residue planes; the
    dense peak assumes the full clock."""
    blocks = torch.cuda.get_device_properties(0).multi_processor_count
    out = torch.empty(blocks * 512, dtype=torch.int32, device="cuda")
    lib = mgx._lib
    ops = lib.lib().mgx_mfma_i8_ceiling_ops(blocks, iters)
    rates = {}
    for name, rnd in (("random_lds_operands_tops", 1), ("constant_operands_tops", 0)):
        lib.call("mgx_mfma_i8_ceiling", out.data_ptr(), blocks, 2, rnd, stream.cuda_stream)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            lib.call("mgx_mfma_i8_ceiling", out.data_ptr(), blocks, iters, rnd, stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        rates[name] = round(ops / (ms * 1e-3) / 1e12, 2)
        if not rnd and not bool(torch.all(out == 8192 * iters)):
            raise RuntimeError("mfma ceiling probe: constant-operand result is wrong")
    rates.update({"workgroups": blocks, "mfma_per_wave": iters * 32,
                  "kind": ("v_mfma_i32_16x16x64_i8-only loop, 8 waves per CU, fragments "
                           "re-read from LDS every round (no global loads, no barriers)")})
    return rates

def main() -> int:
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU's local product; --gpus {args.gpus} is not supported "
              "(the multi-GPU exchange is not part of this build)", file=sys.stderr)
        return 2
    import torch
    import mgxbench_amd as mgx

    m, n, k = args.m, args.n, args.k
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    log(f"allocating m={m} n={n} k={k} moduli={args.moduli} slices={args.slices}")
    A = torch.empty((m, k), dtype=torch.float64, device=dev)
    B = torch.empty((k, n), dtype=torch.float64, device=dev)
    C = torch.empty((m, n), dtype=torch.float64, device=dev)
    Cp = torch.empty((m, n), dtype=torch.float64, device=dev)
    mgx.fill(A, m, k, k, 1, seed=1)
    mgx.fill(B, k, n, n, 1, seed=2)
    forms = {"mod": (mgx.FORM_INT8_MOD, args.moduli), "slices": (mgx.FORM_INT8, args.slices)}
    if args.bf16:
        forms["bf16"] = (mgx.FORM_BF16, args.slices)
    ws_bytes = max(mgx.split_workspace_bytes(f, p, m, n, k) for f, p in forms.values())
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    F = 2.0 * m * n * k
    if args.copy_probe_mib > 0:
        src = torch.ones(args.copy_probe_mib * 131072, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe_mib} MiB read + {args.copy_probe_mib} MiB written")
        del src, dst

    def timed(step, steps):
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0.record(stream)
        for i in range(steps):
            step(i)
        t1.record(stream)
        torch.cuda.synchronize()
        return t0.elapsed_time(t1) / steps

    # plain FP64 MFMA
    log("plain warmup")
    for _ in range(args.warmup):
        mgx.ABMultiply(A, B, Cp, m, n, k, k, n, n, 32, stream=stream)
    torch.cuda.synchronize()
    log("plain timed")
    ms_plain = timed(lambda i: mgx.ABMultiply(A, B, Cp, m, n, k, k, n, n, 32, stream=stream),
                     args.steps)
    log(f"plain {ms_plain:.3f} ms/step")
    rows = list(range(0, m, max(1, m // 16)))[:16]
    acc_plain = mgx.error_report(A, B, Cp, m, n, k, k, n, n, rows)
    acc_plain.pop("rows_checked", None)
    vendor = vendor_fp64(torch, A, B, stream, timed, args.warmup, max(3, args.steps // 4))
    log(f"vendor fp64 {vendor['ms_per_step']:.3f} ms/step ({vendor['tflops']:.1f} TFLOP/s)")

    def measure(name):
        form, planes = forms[name]
        # phase events [scan, operands, product, (reconstruction) end] inside the timed region
        events = [[torch.cuda.Event(enable_timing=True) for _ in range(5)]
                  for _ in range(args.steps)]
        for evs in events:  # materialise the HIP events before the timed region
            for e in evs:
                e.record(stream)

        def step(i=None):
            if i is None:
                mgx.split_gemm(A, B, C, m, n, k, k, n, n, form=form, slices=planes, workspace=ws,
                               workspace_bytes=ws_bytes, stream=stream)
            else:
                mgx.split_gemm_events(A, B, C, m, n, k, k, n, n, events[i][:4], form=form,
                                      slices=planes, workspace=ws, workspace_bytes=ws_bytes,
                                      stream=stream)
                events[i][4].record(stream)

        log(f"{name} warmup")
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        log(f"{name} timed")
        ms = timed(step, args.steps)

        def avg(a, b):
            return sum(ev[a].elapsed_time(ev[b]) for ev in events) / args.steps

        phases = {"exponent_scan": avg(0, 1), "operands": avg(1, 2), "int8_products": avg(2, 3),
                  "reconstruction": avg(3, 4)}
        log(f"{name} {ms:.3f} ms/step " + " ".join(f"{p} {v:.3f}" for p, v in phases.items()))
        diff = float(torch.linalg.vector_norm(C - Cp) / torch.linalg.vector_norm(Cp))
        acc = mgx.error_report(A, B, C, m, n, k, k, n, n, rows)
        acc.pop("rows_checked", None)
        products = planes if name == "mod" else planes * (planes + 1) // 2
        keep = C.clone() if args.bf16 and name in ("slices", "bf16") else None
        return {"ms": ms, "phases": phases, "diff": diff, "acc": acc, "products": products,
                "result": keep,
                "planes": planes}

    res = {name: measure(name) for name in forms}
    head = res["mod"]
    log(f"normwise error modular {head['acc']['normwise']:.3e} slices "
        f"{res['slices']['acc']['normwise']:.3e} plain {acc_plain['normwise']:.3e}")

    tf_plain = F / (ms_plain * 1e-3) / 1e12
    tf_split = F / (head["ms"] * 1e-3) / 1e12
    int8_ops = head["products"] * F
    achieved = int8_ops / (head["phases"]["int8_products"] * 1e-3) / 1e12
    peak_i8 = mgx.peak(mgx.FORM_INT8)
    beta = mgx._lib.lib().mgx_split_mod_beta(args.moduli, k)
    sl = res["slices"]
    out = {
        "metric": METRIC,
        "value": round(tf_split, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{workload_name(m, n, k)}: C {m}x{n} += A {m}x{k} . "
                         f"B {k}x{n}, FP64 uniform[-1,1), HBM-resident"),
            "m": m, "n": n, "k": k,
            "form": (f"int8 modular split (Ozaki scheme II), {args.moduli} moduli, "
                     f"beta={beta} ({args.moduli} INT8 products)"),
            "parallelism": "single GPU",
        },
        "plain_fp64": {
            "tflops": round(tf_plain, 3),
            "ms_per_step": round(ms_plain, 4),
            "frac_of_fp64_mfma_peak": round(tf_plain / mgx.peak(mgx.FORM_FP64), 4),
            "vendor_fp64_roof": vendor,
            "vs_vendor_fp64": round(vendor["ms_per_step"] / ms_plain, 4),
        },
        "speedup_vs_plain": round(tf_split / tf_plain, 4),
        "split_phase_ms": {p: round(v, 4) for p, v in head["phases"].items()},
        "split_vs_plain_rel_fro": head["diff"],
        "accuracy_vs_double_double": {"rows_sampled": len(rows), "split": head["acc"],
                                      "plain": acc_plain, "split_slices": sl["acc"]},
        "split_no_worse_than_plain": head["acc"]["normwise"] <= acc_plain["normwise"],
        "split_slices": {
            "form": (f"int8 digit-slice split (Ozaki scheme I), s={sl['planes']} "
                     f"({sl['products']} slice-pair products)"),
            "tflops": round(F / (sl["ms"] * 1e-3) / 1e12, 3),
            "ms_per_step": round(sl["ms"], 4),
            "speedup_vs_plain": round(ms_plain / sl["ms"], 4),
            "phase_ms": {p: round(v, 4) for p, v in sl["phases"].items() if p != "reconstruction"},
            "no_worse_than_plain": sl["acc"]["normwise"] <= acc_plain["normwise"],
        },
        **({"split_bf16": {
            "form": (f"bf16 digit-slice split, s={res['bf16']['planes']} (the INT8 form's "
                     "digits on the bf16 matrix cores; same result)"),
            "tflops": round(F / (res["bf16"]["ms"] * 1e-3) / 1e12, 3),
            "ms_per_step": round(res["bf16"]["ms"], 4),
            "speedup_vs_plain": round(ms_plain / res["bf16"]["ms"], 4),
            "bf16_products_tflops": round(res["bf16"]["products"] * F /
                                          (res["bf16"]["phases"]["int8_products"] * 1e-3) / 1e12,
                                          2),
            "bitwise_equal_to_int8_form": bool(torch.equal(res["bf16"]["result"],
                                                           res["slices"]["result"])),
        }} if "bf16" in res else {}),
        "roofline": {
            "kernel": "int8_product_kernel (per-modulus INT8 residue products)",
            "bound": "mfma",
            "achieved": round(achieved, 2),
            "peak": peak_i8,
            "unit": "TFLOP/s",
            "frac": round(achieved / peak_i8, 4),
            "traffic": profiled_traffic(m, n, k, args.moduli),
            "traffic_source": "profiles/pmc_split_forms_8192x8192x16384.md (rocprofv3 FETCH_SIZE "
                              "x2.0 + WRITE_SIZE x1.0, calibrated on a 1 GiB copy)",
        },
    }
    log("mfma ceiling probe")
    ceiling = mfma_ceiling(mgx, torch, stream)
    out["roofline"]["mfma_ceiling"] = ceiling
    out["roofline"]["frac_of_random_operand_ceiling"] = round(
        achieved / ceiling["random_lds_operands_tops"], 4)
    if not args.no_summa:
        out["summa_streamed"] = summa_report(mgx, torch, A, B, m, n, k, args.moduli)
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(m, n, k, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
