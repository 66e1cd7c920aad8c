"""The embedded-boundary hydro adjoint against central finite differences,
on the host: tests/eb_adjoint_host.cpp runs the cut-cell stages and their
adjoints point by point (the same host/device functions the batched
kernels run) on one box, for regular and plane-cut geometries with
interior, wall and outflow domain faces and PLM orders 1, 2 and 4."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_eb_adjoint_matches_finite_differences_on_host(tmp_path):
    exe = tmp_path / "eb_adjoint_host"
    cmd = [HIPCC, "-x", "hip", "-O2", "-std=c++17", "-ffp-contract=off",
           "--offload-arch=gfx950", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "camr_amd", "csrc"), "-o", str(exe),
           os.path.join(ROOT, "tests", "eb_adjoint_host.cpp")]
    subprocess.run(cmd, check=True, capture_output=True, timeout=900)
    out = subprocess.run([str(exe), "150"], capture_output=True, text=True, timeout=900)
    print(out.stdout)
    assert out.returncode == 0 and "EB ADJOINT CHECK OK" in out.stdout, out.stdout[-3000:]
