"""The kernels' 32-bit division by a fixed lattice extent (sgld::FastDiv,
sgld::NodeDiv) equals integer division: compiled for the host from the same
header and checked exhaustively over divisors 1..5000."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="needs hipcc for the HIP host headers")
def test_fastdiv_equals_integer_division(tmp_path):
    exe = str(tmp_path / "fastdiv_check")
    subprocess.run(["hipcc", "-O2", "-std=c++17", os.path.join(HERE, "fastdiv_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok"), out.stdout
