"""The C ABI library without a GPU: it loads, exports every symbol that
include/hydro.h declares, validates its arguments, and its host-side
routines (geometry preprocessing, RCM order) work.  No kernel is launched."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, cached_mesh


def _declared():
    src = open(os.path.join(ROOT, "include", "hydro.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(hydro_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    import hydro_amd._lib as L

    lib = L.load()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/hydro.h but not exported"
    assert set(names) == set(L.SIGNATURES), "ctypes signatures out of sync with include/hydro.h"
    assert lib.hydro_abi_version() == L.ABI_VERSION


def test_create_rejects_bad_descriptors():
    import hydro_amd._lib as L

    lib = L.load()
    h = ctypes.c_void_p()
    assert lib.hydro_mesh_create(None, ctypes.byref(h)) == L.ERR_INVALID
    d = L.MeshDesc()
    d.abi_version = L.ABI_VERSION + 7
    assert lib.hydro_mesh_create(ctypes.byref(d), ctypes.byref(h)) == L.ERR_INVALID
    assert b"abi_version" in lib.hydro_last_error()
    d.abi_version = L.ABI_VERSION
    d.dtype = 5
    assert lib.hydro_mesh_create(ctypes.byref(d), ctypes.byref(h)) == L.ERR_INVALID
    d.dtype = L.F64
    d.CELL, d.NDAYS, d.MDT, d.DT, d.HM1, d.HM2 = 10, 1, 60, 1.0, 0.001, 0.01
    # host memory is not accepted where device pointers are required
    buf = np.zeros(64)
    for name, _ in d._fields_:
        if name not in ("abi_version", "dtype", "variant", "CELL", "NDAYS", "NHQ", "MDT", "index_real",
                        "DT", "HM1", "HM2", "JL"):
            setattr(d, name, buf.ctypes.data)
    assert lib.hydro_mesh_create(ctypes.byref(d), ctypes.byref(h)) == L.ERR_INVALID
    assert b"device pointer" in lib.hydro_last_error()
    assert not h.value
    d.index_real = 4                      # only the NAC and KLAS bits exist
    assert lib.hydro_mesh_create(ctypes.byref(d), ctypes.byref(h)) == L.ERR_INVALID
    assert b"index_real" in lib.hydro_last_error()


def test_null_handle_calls_fail_cleanly():
    import hydro_amd._lib as L

    lib = L.load()
    assert lib.hydro_step_n(None, 1, 0, None) == L.ERR_INVALID
    assert lib.hydro_calculate_flux(None, 1, 0, None) == L.ERR_INVALID
    assert lib.hydro_sync(None, None) == L.ERR_INVALID
    assert lib.hydro_regroup(None, None) == L.ERR_INVALID
    pc, mc = ctypes.c_int(), ctypes.c_int()
    assert lib.hydro_plan_grouping(None, ctypes.byref(pc), ctypes.byref(mc), None) == L.ERR_INVALID
    assert lib.hydro_mesh_destroy(None) == L.OK


def test_precalculate_validates_node_ids():
    import hydro_amd._lib as L

    lib = L.load()
    z = np.zeros(16)
    nap = np.array([1, 2, 9, 0], np.int32)  # node 9 of 3
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    st = lib.hydro_precalculate(L.F64, 1, 3, p(nap), p(z), p(z), p(z), p(z), p(z), 0.001, *[p(z)] * 9)
    assert st == L.ERR_INVALID


def test_rcm_order_is_permutation_and_reduces_bandwidth():
    import hydro_amd._lib as L

    lib = L.load()
    host = cached_mesh("channel_6675")
    perm = np.zeros(host.CELL, np.int32)
    L.check(lib.hydro_rcm_order(host.CELL, host.NAC.ctypes.data_as(ctypes.c_void_p),
                                perm.ctypes.data_as(ctypes.c_void_p)))
    assert np.array_equal(np.sort(perm), np.arange(host.CELL))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(host.CELL)
    a = np.repeat(np.arange(host.CELL), 4)
    b = host.NAC - 1
    m = b >= 0
    bw_orig = np.abs(a[m] - b[m]).max()
    bw_rcm = np.abs(inv[a[m]] - inv[b[m]]).max()
    assert bw_rcm < bw_orig / 4, (bw_orig, bw_rcm)


def test_fp32_precalculate_runs_in_float():
    """The fp32 mesh's geometry is computed in float (like the reference's
    fp32 build), not rounded from the fp64 geometry."""
    h64 = cached_mesh("channel_6675")
    h32 = cached_mesh("channel_6675", dtype=np.float32)
    assert h32.SIDE.dtype == np.float32
    # float coordinates of a few km carry ~1e-4 m ulps, so edge lengths agree to ~1e-5
    assert np.allclose(h32.SIDE, h64.SIDE, rtol=1e-4)
    assert np.any(h32.SIDE != h64.SIDE.astype(np.float32))


def test_device_mesh_fails_loudly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from hydro_amd.device import DeviceMesh

    with pytest.raises(RuntimeError, match="HIP device"):
        DeviceMesh(cached_mesh("channel_6675"))
