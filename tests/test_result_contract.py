"""The BENCHMARK_RESULT format of mm_1d / mm_2d, checked in their sources (no
GPU needed): the unconditional part of the line is the reference's format
string, and the build's extra keys are printed only behind the
extras_requested() knob check.  The GPU driver tests run the executables and
match the printed line against the same format.

Reference format strings (the contract benchmark.py parses):
  1D-row-block/mmb.cu:964
    "BENCHMARK_RESULT decomposition=1D library=%s size=%d ranks=%d seconds=%.9f valid=%d loops=%d\\n"
  2D-block-matrix/mmb.cu:398
    "BENCHMARK_RESULT decomposition=2D library=%s size=%d ranks=%d seconds=%.9f valid=%d loops=1 grid=%dx%d\\n"
"""
import re
from pathlib import Path

import pytest

CSRC = Path(__file__).resolve().parents[1] / "mgxbench_amd" / "csrc"

REFERENCE = {
    "mm_1d.cpp": "BENCHMARK_RESULT decomposition=1D library=%s size=%d ranks=%d seconds=%.9f "
                 "valid=%d loops=%d",
    "mm_2d.cpp": "BENCHMARK_RESULT decomposition=2D library=%s size=%d ranks=%d seconds=%.9f "
                 "valid=%d loops=1 grid=%dx%d",
}


def _printf_calls(text):
    """(start offset, format string) of every printf call: the adjacent string
    literals of its first argument, concatenated."""
    out = []
    for m in re.finditer(r"printf\(\s*((?:\"(?:[^\"\\]|\\.)*\"\s*)+)", text):
        lits = re.findall(r"\"((?:[^\"\\]|\\.)*)\"", m.group(1))
        out.append((m.start(), "".join(lits)))
    return out


@pytest.mark.parametrize("src", sorted(REFERENCE))
def test_result_line_is_reference_format(src):
    text = (CSRC / src).read_text()
    calls = _printf_calls(text)
    heads = [(pos, f) for pos, f in calls if f.startswith("BENCHMARK_RESULT")]
    assert len(heads) == 1, heads
    pos, fmt = heads[0]
    assert fmt == REFERENCE[src]
    # the calls after it up to the newline: extra keys behind the knob check,
    # then the terminating newline
    tail = [(p, f) for p, f in calls if p > pos]
    extra_pos, extra_fmt = tail[0]
    newline_pos, newline_fmt = tail[1]
    assert extra_fmt.startswith(" form=%s slices=%d")
    assert newline_fmt == "\\n"
    assert extra_pos < newline_pos
    # the extra-keys printf is the body of `if (extras_requested())` / `if (extras)`
    guard = text[pos:extra_pos].rstrip()
    guard = guard[: guard.rfind("std::")].rstrip() if guard.endswith("std::") else guard
    assert re.search(r"if \(extras(_requested\(\))?\)$", guard), guard[-120:]
