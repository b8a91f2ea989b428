"""GPU: the allocation-free diagnostics entry point (hydro_diagnostics_noalloc)
gives exactly what hydro_diagnostics gives, through the C ABI, on two calls
in a row, the second on a smaller mesh (its fixed buffer keeps the first
call's records: the second must not read them)."""
import ctypes

import numpy as np
import pytest

from conftest import cached_mesh

pytestmark = pytest.mark.gpu


def _both(dm, lib, L, torch):
    n = len(L.DIAG_FIELDS)
    a = torch.full((n,), -1.0, dtype=torch.float64, device=dm.device)
    b = torch.full((n,), -2.0, dtype=torch.float64, device=dm.device)
    L.check(lib.hydro_diagnostics_noalloc(dm.h, ctypes.c_void_p(a.data_ptr()), dm.stream()))
    L.check(lib.hydro_diagnostics(dm.h, ctypes.c_void_p(b.data_ptr()), dm.stream()))
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


def test_noalloc_equals_allocating_entry_point():
    import torch

    from hydro_amd import _lib as L
    from hydro_amd.device import DeviceMesh

    lib = L.load()
    big = cached_mesh("floodplain_207234")
    small = cached_mesh("channel_6675")
    results = []
    for host, steps in ((big, 6), (small, 9)):          # second call: a smaller input
        dm = DeviceMesh(host, device=0)
        dm.step(steps, "two_kernel")
        a, b = _both(dm, lib, L, torch)
        dm.close()
        print(host.name, host.CELL, dict(zip(L.DIAG_FIELDS, a)), flush=True)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (host.name, a, b)
        results.append(a)
    assert results[0][0] != results[1][0]            # different meshes, different masses
