"""CPU check of the fixed-exponent pow evaluations of the device code
(hydro-bench_amd/csrc/pow_fixed.hpp), compiled for the host with the
device's error budget (reciprocal and seed errors put in): within 0.51 ulp of
the true x^1.5 and x^0.33333 for every normal result, within 1 ulp for
subnormal results, pow's values at 0, infinities, negatives and NaN, and
correctly rounded for all but a small fraction of inputs."""
import os
import re
import subprocess

import pytest

from conftest import ROOT


def test_pow_fixed_accuracy(tmp_path):
    exe = tmp_path / "pow_check"
    src = os.path.join(ROOT, "tests", "native", "pow_check.cpp")
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "hydro-bench_amd", "csrc"), src, "-o", str(exe)]
    if not os.path.exists(cmd[0]):
        pytest.skip("hipcc not available")
    subprocess.run(cmd, check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    rows = re.findall(r"(\w+) max_ulp ([\d.]+) at \S+ subnormal_max_ulp ([\d.]+) n (\d+) not_rn (\d+) specials_ok (\d)",
                      out.stdout)
    assert len(rows) == 2 and out.returncode == 0, out.stdout
    for name, mx, sub, n, not_rn, ok in rows:
        assert float(mx) <= 0.51, name
        assert float(sub) <= 1.0, name
        assert int(ok) == 1, name
        assert int(not_rn) <= 2e-3 * int(n), name
