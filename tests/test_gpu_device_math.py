"""GPU: the device build of the timestep's functions equals their host build.

The device functions (Osher flux over all 16 cases, the slot flux of every
kind with both variants, the paired edge solve, the cell update, the
fixed-exponent pows), in fp64 and fp32, and the reverse-mode functions of the
adjoint (the Osher flux's reverse over all 16 cases, the slot flux's reverse
for every kind that has one, the cell update's reverse), are evaluated on
random inputs by a
gfx950 kernel and by the same source compiled for the host
(tests/native/device_check.hip).  The host build is checked against the
oracle by the CPU tests; this test catches the device compiler lowering a
construct differently (one such case, a two-case switch losing a negation in
the Osher flux, is why it exists).  Bar: the Osher flux bit for bit; every
other result within 4 ulps (the fixed-exponent pows use the device's
approximate reciprocal; for the cell update, whose momentum update can lose
bits to cancellation, within 1e-12 (1 + |H| + |U| + |V|) absolute), with at
most 0.1% of results not bit-identical.
"""
import ctypes
import os

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

NAMES = ["osher f64", "slot f64 A", "slot f64 B", "pair f64", "update f64", "slot f32 A", "slot f32 B", "pair f32",
         "update f32", "pow", "osher_bar f64", "slot_flux_bar f64", "update_bar f64"]
EXACT = {"osher f64", "osher_bar f64"}   # no pow inside: the device build must give the host's bits


def test_device_functions_equal_host_build():
    from hydro_amd import _lib

    _lib.load()   # torch's HIP runtime first: the check library then shares it
    path = os.path.join(ROOT, "tests", "native", "_build", "libdevice_check.so")
    assert os.path.exists(path), "device check library missing: run __graft_entry__.build()"
    lib = ctypes.CDLL(path)
    lib.dc_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double),
                           ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    lib.dc_report.restype = ctypes.c_char_p
    assert lib.dc_groups() == len(NAMES)
    n = 1 << 20
    for g, name in enumerate(NAMES):
        mx, inexact, bad = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
        rc = lib.dc_run(g, n, 1000 + g, ctypes.byref(mx), ctypes.byref(inexact), ctypes.byref(bad))
        assert rc == 0, f"{name}: HIP error {rc}"
        print(f"{name}: {n} cases, max {mx.value:.0f} ulp, {inexact.value} results not bit-identical", flush=True)
        if bad.value:
            print(lib.dc_report().decode(), flush=True)
        assert bad.value == 0, f"{name}: {bad.value} results differ by more than 4 ulps"
        assert inexact.value <= 1e-3 * n * 5, name
        if name in EXACT:
            assert inexact.value == 0, f"{name} differs from the host build"
