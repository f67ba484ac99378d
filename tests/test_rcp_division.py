"""The one-division quotient helper (bes::RcpDiv, used for the gradient's
divisions by the cell volume) gives the bitwise result of the division, so the
gradient stays the reference's.  Compiled for the host with hipcc from the
kernel header itself; Markstein's theorem says it cannot differ inside the
ranges it uses the reciprocal for, the samples check that and the edge cases
outside them."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not available")
def test_rcp_division_is_the_division(tmp_path):
    exe = tmp_path / "rcp_division"
    src = os.path.join(HERE, "cpp", "rcp_division.hip")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off", "-o", str(exe), src],
                   check=True, capture_output=True)
    res = subprocess.run([str(exe), "100000000"], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "mismatches 0" in res.stdout
