"""The triangular solve's division by a precomputed reciprocal (Markstein's
correction, trsv_flow_kernel's div_known) gives the correctly rounded
quotient: restated in C with the host's IEEE operations and compared with
a / b bit for bit over random and edge-mantissa operands in the range the
device code uses it for."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs a C compiler")
def test_division_by_a_correctly_rounded_reciprocal_is_correctly_rounded(tmp_path):
    exe = str(tmp_path / "div_known")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "native", "div_known.c"), "-lm"],
                   check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    tried, bad = int(out[0]), int(out[1])
    assert tried > 10_000_000 and bad == 0
