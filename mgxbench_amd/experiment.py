"""Run identities, environment probes and sample statistics for the runner.

Counterpart of the reference's configuration/experiment.py (probe 19-26,
collect_environment 75-99, new_run 102-106, sample_statistics 109-118) with
ROCm probes in place of the NVIDIA ones.  The statistics use the same
stdlib procedure, so confidence bounds match the reference's for the same
samples: a percentile bootstrap of the mean with random.Random(0),
2000 resamples drawn by Random.choices, bounds at sorted indices
int(0.025 * 2000) and min(1999, int(0.975 * 2000)).
"""
from __future__ import annotations

import csv
import datetime
import hashlib
import io
import json
import os
import platform
import random
import re
import shutil
import statistics
import subprocess
import uuid
from pathlib import Path

BOOTSTRAP = 2000


def probe(command: list[str]) -> str | None:
    """stdout+stderr of a short command, or None when absent / failing / slow."""
    if not shutil.which(command[0]):
        return None
    try:
        res = subprocess.run(command, capture_output=True, text=True, timeout=8)
    except (OSError, subprocess.TimeoutExpired):
        return None
    if res.returncode != 0:
        return None
    return (res.stdout + res.stderr).strip()[:16000]


def _gpu_inventory() -> list[dict] | None:
    """GPUs from `rocm-smi --showproductname --showbus --csv` (header-keyed)."""
    raw = probe(["rocm-smi", "--showproductname", "--showbus", "--csv"])
    if raw is None:
        return None
    lines = [l for l in raw.splitlines() if l.strip() and "," in l]
    if not lines:
        return None
    rows = list(csv.DictReader(io.StringIO("\n".join(lines))))
    out = []
    for row in rows:
        device = row.get("device") or row.get("Device") or ""
        if not device.lower().startswith("card"):
            continue
        out.append({
            "index": device[4:],
            "name": row.get("Card Series") or row.get("Card series") or row.get("Card Model"),
            "sku": row.get("Card SKU"),
            "gfx_version": row.get("GFX Version"),
            "pci_bus_id": row.get("PCI Bus"),
        })
    return out or None


def _header_version(path: Path, prefix: str) -> str | None:
    if not path.is_file():
        return None
    text = path.read_text(errors="replace")
    parts = []
    for part in ("MAJOR", "MINOR", "PATCH"):
        m = re.search(rf"^\s*#\s*define\s+{prefix}{part}\s+([0-9]+)", text, re.M)
        if not m:
            return None
        parts.append(m.group(1))
    return ".".join(parts)


def _library_versions() -> dict:
    rocm = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
    rccl = rocm / "include" / "rccl" / "rccl.h"
    hip = rocm / "include" / "hip" / "hip_version.h"
    hip_ver = None
    if hip.is_file():
        t = hip.read_text(errors="replace")
        v = [re.search(rf"#define\s+HIP_VERSION_{p}\s+([0-9]+)", t) for p in ("MAJOR", "MINOR", "PATCH")]
        if all(v):
            hip_ver = ".".join(x.group(1) for x in v)
    return {
        "rccl_header_version": _header_version(rccl, "NCCL_"),
        "rccl_header_path": str(rccl) if rccl.is_file() else None,
        "hip_header_version": hip_ver,
        "note": ("read from the installed headers: the versions this build compiles against, "
                 "which need not be the libraries loaded at run time"),
    }


def collect_environment(metadata_path: str | None = None, build_dir: str | None = None) -> dict:
    data = {
        "hostname": platform.node(),
        "platform": platform.platform(),
        "python": platform.python_version(),
        "gpu": probe(["rocm-smi", "--showproductname"]),
        "gpu_topology": probe(["rocm-smi", "--showtopo"]),
        "gpu_inventory": _gpu_inventory(),
        "library_versions": _library_versions(),
        "hip": probe(["hipcc", "--version"]),
        "mpi": probe(["mpirun", "--version"]),
        "interconnect": probe(["ibstat"]),
        "modules": os.environ.get("LOADEDMODULES"),
        "runtime": {k: v for k, v in os.environ.items()
                    if k.startswith(("MM_PIPELINE_", "NCCL_", "RCCL_", "MGX_", "HSA_", "HIP_"))
                    and not any(x in k.upper() for x in ("TOKEN", "KEY", "PASSWORD", "SECRET"))},
        "allocation": {k: os.environ[k] for k in ("SLURM_JOB_NODELIST", "SLURM_NNODES",
                                                   "SLURM_GPUS_ON_NODE") if k in os.environ},
        "software_note": ("library_versions come from headers; record the loaded runtime "
                          "versions separately when they matter"),
    }
    if metadata_path:
        supplied = json.loads(Path(metadata_path).read_text())
        if not isinstance(supplied, dict):
            raise ValueError(f"{metadata_path}: expected a JSON object of environment fields")
        data["user"] = supplied
    if build_dir:
        data["binaries"] = {}
        for name in ("mm_1d", "mm_2d"):
            p = Path(build_dir) / name
            data["binaries"][name] = hashlib.sha256(p.read_bytes()).hexdigest() if p.is_file() else None
    # strict identity: distinct hardware / runtime environments never mix
    data["id"] = hashlib.sha256(json.dumps(data, sort_keys=True).encode()).hexdigest()[:16]
    return data


def new_run(output: str | Path) -> tuple[Path, str]:
    """A fresh, never reused run directory output/runs/<utc time>-<random>."""
    stamp = datetime.datetime.now(datetime.timezone.utc).strftime("%Y%m%dT%H%M%S.%fZ")
    run_id = f"{stamp}-{uuid.uuid4().hex[:8]}"
    dest = Path(output) / "runs" / run_id
    dest.mkdir(parents=True, exist_ok=False)
    return dest, run_id


def sample_statistics(samples: list[float], bootstrap: int = BOOTSTRAP) -> dict:
    mean = statistics.mean(samples)
    low = high = None
    if len(samples) > 1:
        rng = random.Random(0)
        means = sorted(statistics.mean(rng.choices(samples, k=len(samples)))
                       for _ in range(bootstrap))
        low = means[int(0.025 * bootstrap)]
        high = means[min(bootstrap - 1, int(0.975 * bootstrap))]
    return {"cv_pct": 100.0 * statistics.pstdev(samples) / mean,
            "ci95_low_s": low, "ci95_high_s": high}
