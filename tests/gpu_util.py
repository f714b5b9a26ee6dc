"""Device-memory helpers for the GPU tests (torch is used only for allocation,
copies and streams; every computation goes through libmgx's C ABI)."""
import sys

import numpy as np


def to_dev(x: np.ndarray):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def to_host(t) -> np.ndarray:
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def empty(rows, cols, fill=np.nan):
    import torch
    return torch.full((rows, cols), fill, dtype=torch.float64, device="cuda")


def announce(msg: str) -> None:
    """Names the case about to run on the GPU (flushed, so a fault still shows it)."""
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def rel_fro(x: np.ndarray, ref: np.ndarray) -> float:
    d = np.linalg.norm((x - ref).ravel())
    r = np.linalg.norm(ref.ravel())
    return float(d / r) if r > 0 else float(d)
