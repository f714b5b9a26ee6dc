"""The modular form's arithmetic (mgxbench_amd/csrc/modular.hpp), built for
the host with g++ and checked against exact 128-bit integers
(tests/modular_host_check.cpp): Chinese-remainder reconstruction at every
scale, next to +-P/2 where the quotient estimate can miss, every
representative for two moduli, exact rounding ties up to 2^117, the
general and narrow residue extraction, and the per-line exponent search
against its definition.  The same header is compiled into
the GPU kernels; the GPU parity tests compare those against the oracle."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_modular_arithmetic_against_exact_integers(tmp_path):
    exe = tmp_path / "modular_host_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(ROOT / "mgxbench_amd" / "csrc"),
                    str(ROOT / "tests" / "modular_host_check.cpp"), "-o", str(exe)],
                   check=True)
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert run.returncode == 0, run.stdout + run.stderr
    m = re.search(r"reconstructions (\d+) residues (\d+) line exponents (\d+) bad 0", run.stdout)
    assert m, run.stdout
    assert int(m.group(1)) > 1_000_000 and int(m.group(2)) > 5_000_000
    assert int(m.group(3)) > 100_000
