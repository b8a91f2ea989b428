"""GPU: the drop-in boundary as the reference's callers hold it.

* NAC and KLAS as Real arrays (the CUDA SideView, side.hpp:29-43) or KLAS
  only as Real (the Taichi port, F2_hydro_taichi_fp64.py:46-47) give the
  int32 results bit for bit in every mode, on a mesh with every boundary kind
  and opening breaches; the library's KLAS 7 -> 0 rewrites reach the caller's
  Real KLAS.  Non-whole values are rejected.
* run(days, backend, mesh, steps) with T2's signature: step_fn(on_step)
  calls back after every synchronised step and matches the oracle step for
  step; step_fn without a callback runs the whole loop; sync_fn makes the
  tensors current; the NDAYS cap stops the loop as the reference's does.
"""
import ctypes

import numpy as np
import pytest

from conftest import cached_mesh, rare_kinds_mesh, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("index_real", [3, 2], ids=["nac_klas_real", "klas_real"])
@pytest.mark.parametrize("mode", ["two_kernel", "fused", "persistent"])
def test_real_typed_nac_klas(index_real, mode):
    import hydro_amd
    from hydro_amd.device import DeviceMesh

    host = rare_kinds_mesh(cached_mesh("channel_6675"), breach_fires=True)
    tune = hydro_amd.default_tune(materialize_flux=1)
    out = {}
    for ir in (0, index_real):
        dm = DeviceMesh(host, device=0, tune=tune, index_real=ir)
        dm.step(40, mode)
        out[ir] = dm.fetch("H", "U", "V", "Z", "W", "KLAS", "FLUX0", "FLUX1", "FLUX2", "FLUX3")
        dm.close()
    a, b = out[0], out[index_real]
    for k in ("H", "U", "V", "Z", "W", "FLUX0", "FLUX1", "FLUX2", "FLUX3"):
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k
    assert b["KLAS"].dtype == host.dtype
    assert np.array_equal(a["KLAS"], b["KLAS"].astype(np.int32))
    opened = (host.KLAS == 7) & (a["KLAS"] == 0)
    assert opened.sum() > 0                                     # breaches opened and the caller sees it


def test_real_typed_index_rejects_fractions():
    import torch

    from hydro_amd import _lib
    from hydro_amd.device import DeviceMesh

    host = cached_mesh("channel_6675")
    dm = DeviceMesh(host, device=0, index_real=_lib.NAC_REAL)
    bad = dm.t["NAC"].clone()
    bad[17] = bad[17] + 0.5
    d = _lib.MeshDesc.from_buffer_copy(dm.desc)
    d.NAC = bad.data_ptr()
    torch.cuda.synchronize()
    h = ctypes.c_void_p()
    assert dm.lib.hydro_mesh_create(ctypes.byref(d), ctypes.byref(h)) == _lib.ERR_INVALID
    assert b"NAC[17]" in dm.lib.hydro_last_error()
    dm.close()


@pytest.mark.parametrize("mode", ["two_kernel", "fused"])
def test_run_with_the_taichi_signature(mode, oracle):
    import torch

    from hydro_amd.device import run

    host = cached_mesh("channel_6675")
    step_fn, sync_fn, H, U, V, Z, F0, F1, F2, F3 = run(days=1, backend="cuda", mesh=host, mode=mode)
    per_day = host.steps_per_day - 1
    assert step_fn.total_steps == per_day
    seen = {}
    r = oracle.OracleRun(host)
    checks = {1, 7, per_day}

    def on_step(s):
        if s in checks:
            seen[s] = (H.cpu().numpy().copy(), Z.cpu().numpy().copy())

    step_fn(on_step=on_step)
    assert sorted(seen) == sorted(checks)
    done = 0
    for s in sorted(checks):
        r.step(s - done)
        done = s
        assert rel_err(seen[s][0], r["H"]) <= 1e-12, s
        assert rel_err(seen[s][1], r["Z"]) <= 1e-12, s
    dm = step_fn.mesh
    assert dm.get_time()[:2] == (1, 1)
    # a second call starts again at day 0 from the state it finds
    step_fn()
    sync_fn()
    assert dm.get_time()[:2] == (1, 1)
    r2 = oracle.OracleRun(host)
    r2.step(per_day)
    r2.day.value, r2.kt.value = 0, 1      # the second call: day 0 again, from the state after the first
    r2.step(per_day)
    assert rel_err(H.cpu().numpy(), r2["H"]) <= 1e-12
    if mode == "two_kernel":
        assert torch.count_nonzero(F0).item() > 0      # live flux tensors of the last flux pass
    dm.close()


def test_run_stops_at_the_ndays_cap_and_refuses_a_cpu_backend():
    from hydro_amd.device import run

    host = cached_mesh("channel_6675")
    step_fn, sync_fn, *_ = run(backend="cuda", mesh=host, steps=10 ** 6)
    assert step_fn.total_steps == host.max_steps
    step_fn.mesh.close()
    step_fn, *_ = run(days=10 ** 6, mesh=host)
    assert step_fn.total_steps == host.max_steps
    step_fn.mesh.close()
    with pytest.raises(ValueError):
        run(backend="cpu", mesh=host)
