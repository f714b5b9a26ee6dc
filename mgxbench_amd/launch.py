"""Single-process stand-in for the MPI launcher the runner builds commands for.

The runner launches `<launcher> <launcher args> -np 1 exe args... [: -np 1 exe
args...]` (Open MPI MPMD syntax, one segment per rank).  This build computes
one GPU's local product, so exactly one segment with -np 1 is accepted; it is
run as a child process and its exit code returned.  More ranks are refused
with exit code 2 and a message: the multi-rank exchange is not part of this
build.

    python -m mgxbench_amd.launch -np 1 mgxbench_amd/bin/mm_1d auto 4096 NNN
"""
from __future__ import annotations

import subprocess
import sys


def split_segments(argv: list[str]) -> list[list[str]]:
    segments, cur = [], []
    for a in argv:
        if a == ":":
            segments.append(cur)
            cur = []
        else:
            cur.append(a)
    segments.append(cur)
    return [s for s in segments if s]


def parse_segment(seg: list[str]) -> tuple[int, list[str]]:
    """(-np count, command) of one MPMD segment; other launcher options are dropped."""
    np_count = 1
    i = 0
    while i < len(seg) and seg[i].startswith("-"):
        if seg[i] in ("-np", "-n", "--np") and i + 1 < len(seg):
            np_count = int(seg[i + 1])
            i += 2
        else:
            i += 1
    return np_count, seg[i:]


def main(argv: list[str] | None = None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if argv in (["-h"], ["--help"]):
        print(__doc__.strip())
        return 0
    segments = [parse_segment(s) for s in split_segments(argv)]
    total = sum(n for n, _ in segments)
    if not segments or not segments[0][1]:
        print("mgxbench_amd.launch: no command given", file=sys.stderr)
        return 2
    if total != 1:
        print(f"mgxbench_amd.launch: {total} ranks requested; this build runs one rank "
              "(one GPU's local product) per launch", file=sys.stderr)
        return 2
    return subprocess.run(segments[0][1]).returncode


if __name__ == "__main__":
    sys.exit(main())
