"""CPU check of the adjoint arithmetic (no GPU): the device headers of
hydro-bench_amd/csrc are compiled for the host and the hand-written reverse
mode of every slot-flux branch and of the cell update is compared with the
dual-number tangent of the same code (must agree to rounding everywhere) and
with central finite differences of the fp64 flux (may disagree only where
finite differences are unreliable: near-dry depths and divergent boundary
iterations, a tiny fraction)."""
import os
import re
import subprocess

import pytest

from conftest import ROOT


def test_adjoint_equals_tangent_and_fd(tmp_path):
    exe = tmp_path / "ad_check"
    src = os.path.join(ROOT, "tests", "native", "ad_check.cpp")
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "hydro-bench_amd", "csrc"),
           "-I" + os.path.join(ROOT, "include"), src, "-o", str(exe)]
    if not os.path.exists(cmd[0]):
        pytest.skip("hipcc not available")
    subprocess.run(cmd, check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    print(out.stdout)
    m = re.search(r"(\d+) adjoint/tangent comparisons, (\d+) failures; (\d+) finite-difference checks, (\d+) disagree",
                  out.stdout)
    assert m, out.stdout
    n, fails, nfd, fdbad = map(int, m.groups())
    assert n > 100000 and fails == 0 and out.returncode == 0
    assert fdbad <= 1e-4 * nfd
