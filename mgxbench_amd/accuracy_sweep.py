"""Accuracy sweep of the split-precision forms on stress inputs.

Inputs (the accuracy-stress definition of mgx_fill kind 2): every entry is a
random sign x a mantissa uniform in [1, 2) x 2^U(-30, 30); rows of A and
columns of B are further scaled by 2^U(-10, 10) (mgx_fill_exponents,
mgx_scale_rows_cols).  For each form and slice / modulus count the product
C = A.B is compared with a double-double evaluation of A.B on the GPU
(mgx_error_report: Dot2 with fma-based TwoProduct) over the checked rows:

  normwise      ||C - AB||_F / ||AB||_F
  componentwise max |C - AB|_ij / (|A||B|)_ij

Forms: plain FP64 MFMA, INT8 digit slices s, BF16 digit slices s (the same
digits on the bf16 matrix cores, so the same numbers as INT8 with the same
s) and INT8 modular with n moduli.

    python -m mgxbench_amd.accuracy_sweep [--n 4096] [--rows 0 (all)]
        [--int8 3..9] [--bf16 2..5] [--moduli 8..16] [--seed 1] [--json out.json]
"""
from __future__ import annotations

import argparse
import json
import sys


def _span(text: str) -> list[int]:
    if ".." in text:
        lo, hi = text.split("..")
        return list(range(int(lo), int(hi) + 1))
    return [int(x) for x in text.split(",") if x]


def stress_inputs(mgx, torch, n: int, seed: int):
    """Device A, B (n x n) of the accuracy-stress class."""
    dev = torch.device("cuda", 0)
    A = torch.empty((n, n), dtype=torch.float64, device=dev)
    B = torch.empty((n, n), dtype=torch.float64, device=dev)
    mgx.fill(A, n, n, n, 2, seed=seed)
    mgx.fill(B, n, n, n, 2, seed=seed + 1)
    re = torch.empty(n, dtype=torch.int32, device=dev)
    ce = torch.empty(n, dtype=torch.int32, device=dev)
    mgx.fill_exponents(re, n, seed + 2, -10, 10)
    mgx.fill_exponents(ce, n, seed + 3, -10, 10)
    mgx.scale_rows_cols(A, n, n, n, row_exp=re)
    mgx.scale_rows_cols(B, n, n, n, col_exp=ce)
    return A, B


def sweep(n: int, rows: int = 0, int8=range(3, 10), bf16=range(2, 6), moduli=range(8, 17),
          seed: int = 1, log=print) -> dict:
    import numpy as np
    import torch
    import mgxbench_amd as mgx

    A, B = stress_inputs(mgx, torch, n, seed)
    C = torch.empty((n, n), dtype=torch.float64, device=A.device)
    if rows <= 0 or rows >= n:
        check = np.arange(n)
    else:
        check = np.unique(np.linspace(0, n - 1, rows).astype(np.int64))

    def measure(label, count, products, run):
        run()
        torch.cuda.synchronize()
        rep = mgx.error_report(A, B, C, n, n, n, n, n, n, check)
        row = {"form": label, "count": count, "mfma_products": products,
               "normwise": rep["normwise"], "componentwise": rep["componentwise"],
               "max_abs": rep["max_abs"]}
        log(f"[accuracy] N={n} {label:6s} {count:3d}  normwise {row['normwise']:.3e}  "
            f"componentwise {row['componentwise']:.3e}")
        return row

    out = []
    out.append(measure("fp64", 0, 1, lambda: mgx.ABMultiply(A, B, C, n, n, n, n, n, n)))
    for s in int8:
        out.append(measure("int8", s, s * (s + 1) // 2, lambda s=s: mgx.split_gemm(
            A, B, C, n, n, n, n, n, n, form=mgx.FORM_INT8, slices=s)))
    for s in bf16:
        out.append(measure("bf16", s, s * (s + 1) // 2, lambda s=s: mgx.split_gemm(
            A, B, C, n, n, n, n, n, n, form=mgx.FORM_BF16, slices=s)))
    for nm in moduli:
        if mgx.lib().mgx_split_mod_beta(nm, n) < 0:
            continue
        out.append(measure("mod", nm, nm, lambda nm=nm: mgx.split_gemm(
            A, B, C, n, n, n, n, n, n, form=mgx.FORM_INT8_MOD, slices=nm)))
    return {"n": n, "rows_checked": int(len(check)), "seed": seed,
            "inputs": "sign x U[1,2) x 2^U(-30,30), rows of A / columns of B x 2^U(-10,10)",
            "reference": "GPU double-double (Dot2) product on the checked rows",
            "results": out}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--rows", type=int, default=0, help="rows checked (0: all)")
    ap.add_argument("--int8", default="3..9")
    ap.add_argument("--bf16", default="2..5")
    ap.add_argument("--moduli", default="8..16")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    if a.n <= 0:
        ap.error("--n must be positive")
    res = sweep(a.n, a.rows, _span(a.int8), _span(a.bf16), _span(a.moduli), a.seed,
                log=lambda s: print(s, file=sys.stderr, flush=True))
    text = json.dumps(res, indent=2)
    if a.json:
        with open(a.json, "w") as f:
            f.write(text)
    print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
