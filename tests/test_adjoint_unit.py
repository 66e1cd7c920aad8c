"""Per-point adjoints of the hydro arithmetic against finite differences
of the forward functions (host build of camr_amd/csrc/adjoint_math.h)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_pointwise_adjoints_match_finite_differences(tmp_path):
    exe = tmp_path / "adjoint_unit"
    cmd = [HIPCC, "-x", "hip", "-O2", "-std=c++17", "-ffp-contract=off",
           "--offload-arch=gfx950", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "camr_amd", "csrc"), "-o", str(exe),
           os.path.join(ROOT, "tests", "adjoint_unit.cpp")]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    print(out.stdout)
    assert out.returncode == 0 and "ADJOINT CHECK OK" in out.stdout, out.stdout[-3000:]
