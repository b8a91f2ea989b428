"""CPU: the device source's two-kernel step, compiled for the host, against
the oracle over whole meshes.

tests/native/host_step.cpp runs ext_flux_slot / ext_update_cell (the bodies
of the GPU kernels k_calculate_flux / k_update_cell, hydro-bench_amd/csrc/
ext_step.hpp) over every slot and cell on the host.  On the channel turned by
0.5 rad with every boundary kind (KLAS 10, 1, 4, 3, 5, 6, 7 and a code with no
handler; breaches open under variant B), each of 30 steps must give the
oracle's fluxes, KLAS and state: fp64 within 1e-12 normwise relative (pow
results may differ in the last bit), fp32 within 1e-6, over the finite
entries, with infinite ones (an opening breach's momentum flux) identical.  The GPU build of the
same functions is compared with this host build by the GPU tests, so the two
checks together tie the GPU kernels to the oracle everywhere.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, cached_mesh, rare_kinds_mesh, rel_err


@pytest.fixture(scope="module")
def host_step(tmp_path_factory):
    from hydro_amd import _lib

    _lib.load()
    out = tmp_path_factory.mktemp("host_step") / "libhost_step.so"
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "hydro-bench_amd", "csrc"),
           "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "native", "host_step.cpp"),
           "-Wl,--as-needed", "-o", str(out)]
    if not os.path.exists(cmd[0]):
        pytest.skip("hipcc not available")
    subprocess.run(cmd, check=True, capture_output=True)
    return ctypes.CDLL(str(out))


def _diff(a, b):
    """Normwise relative difference over the finite entries; non-finite
    entries (a breach opening gives an infinite FLUX1) must match exactly."""
    fin = np.isfinite(b)
    if not np.array_equal(a[~fin], b[~fin], equal_nan=True):
        return np.inf
    return rel_err(a[fin], b[fin]) if fin.any() else 0.0


def _desc(host, arrays, variant):
    from hydro_amd import _lib

    d = _lib.MeshDesc()
    d.abi_version = _lib.ABI_VERSION
    d.dtype = _lib.F64 if host.dtype == np.float64 else _lib.F32
    d.variant = variant
    d.CELL, d.NDAYS, d.NHQ, d.MDT = host.CELL, host.NDAYS, host.NHQ, host.MDT
    d.DT, d.HM1, d.HM2, d.JL = host.DT, host.HM1, host.HM2, host.JL
    for name, _ in d._fields_:
        if name in arrays:
            setattr(d, name, arrays[name].ctypes.data)
    return d


@pytest.mark.parametrize("variant", [0, 1], ids=["A", "B"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_device_source_on_host_equals_oracle(host_step, oracle, dtype, variant):
    host = rare_kinds_mesh(cached_mesh("channel", dtype=dtype, rotate=0.5), breach_fires=(variant == 1))
    arrays = {k: v.copy() for k, v in host.arrays.items()}
    d = _desc(host, arrays, variant)
    orc = oracle.OracleRun(host, variant)
    tol = 1e-12 if dtype == np.float64 else 1e-6
    spd = host.steps_per_day
    day, kt = 0, 1
    worst = 0.0
    for step in range(1, 31):
        host_step.host_calculate_flux(ctypes.byref(d), kt, day)
        orc.calculate_flux(kt, day)
        for k in ("FLUX0", "FLUX1", "FLUX2", "FLUX3"):
            e = _diff(arrays[k], orc[k])
            worst = max(worst, e)
            assert e <= tol, f"step {step} {k}: {e:.3e}"
        assert np.array_equal(arrays["KLAS"], orc["KLAS"]), f"step {step} KLAS"
        host_step.host_update_cell(ctypes.byref(d))
        orc.update_cell()
        for k in ("H", "U", "V", "Z", "W"):
            e = _diff(arrays[k], orc[k])
            worst = max(worst, e)
            assert e <= tol, f"step {step} {k}: {e:.3e}"
        kt += 1
        if kt >= spd:
            kt, day = 1, day + 1
    if variant == 1:
        assert np.any(arrays["KLAS"] != host.arrays["KLAS"])   # breaches opened
    print(f"{np.dtype(dtype).name} variant {variant}: worst normwise difference over 30 steps {worst:.2e}")
