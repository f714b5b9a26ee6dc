"""Process-grid selection for 2D SUMMA (counterpart of configuration/grid.py:4-19).

Explicit rows/cols must factor `ranks`; otherwise the most square factor
pair with rows <= cols is chosen.
"""
from __future__ import annotations

import math


def resolve_grid(ranks: int, rows: int | None = None, cols: int | None = None) -> tuple[int, int]:
    if ranks <= 0:
        raise ValueError("ranks must be positive")
    for name, v in (("rows", rows), ("cols", cols)):
        if v is not None and v <= 0:
            raise ValueError(f"grid {name} must be positive")
    if rows is not None and cols is not None:
        if rows * cols != ranks:
            raise ValueError("grid rows * grid cols must equal ranks")
        return rows, cols
    if rows is not None:
        if ranks % rows:
            raise ValueError("grid rows must divide ranks")
        return rows, ranks // rows
    if cols is not None:
        if ranks % cols:
            raise ValueError("grid cols must divide ranks")
        return ranks // cols, cols
    r = math.isqrt(ranks)
    while ranks % r:
        r -= 1
    return r, ranks // r
