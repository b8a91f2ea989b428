"""CPU checks of the TILES scheme's host-side pieces.

* pair_flux (one Riemann solve for both slots of an edge) equals slot_flux
  on each side bit for bit, over every wet/dry ladder branch of both sides,
  fp64 and fp32, variants A and B (tests/native/pair_check.cpp, compiled for
  the host from the device headers);
* hydro_tile_order returns a permutation whose tiles are full and compact:
  most cell-graph edges join two cells of one tile.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, cached_mesh


def test_pair_flux_equals_slot_flux_both_sides(tmp_path):
    exe = tmp_path / "pair_check"
    src = os.path.join(ROOT, "tests", "native", "pair_check.cpp")
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "hydro-bench_amd", "csrc"), src, "-o", str(exe)]
    if not os.path.exists(cmd[0]):
        pytest.skip("hipcc not available")
    subprocess.run(cmd, check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    m = re.search(r"(\d+) pair checks, (\d+) mismatches, (\d+) branches under-covered", out.stdout)
    assert m and out.returncode == 0, out.stdout
    n, bad, missing = map(int, m.groups())
    assert n >= 800000 and bad == 0 and missing == 0


def _edges(NAC):
    C = NAC.shape[0]
    a = np.repeat(np.arange(C), 4)
    b = NAC.reshape(-1) - 1
    keep = b >= 0
    return a[keep], b[keep]


@pytest.mark.parametrize("name,block", [("channel_6675", 256), ("floodplain_24020", 128),
                                        ("floodplain_24020", 256)])
def test_tile_order_full_compact_tiles(name, block):
    from hydro_amd import _lib

    host = cached_mesh(name)
    lib = _lib.load()
    C = host.CELL
    nac = np.ascontiguousarray(host.arrays["NAC"], dtype=np.int32)
    perm = np.empty(C, np.int32)
    _lib.check(lib.hydro_tile_order(C, nac.ctypes.data_as(ctypes.c_void_p), block,
                                    perm.ctypes.data_as(ctypes.c_void_p)))
    assert np.array_equal(np.sort(perm), np.arange(C))          # a permutation
    pos = np.empty(C, np.int64)
    pos[perm] = np.arange(C)
    tile = pos // block
    a, b = _edges(nac.reshape(C, 4))
    inner = float((tile[a] == tile[b]).mean())
    rcm = np.empty(C, np.int32)
    _lib.check(lib.hydro_rcm_order(C, nac.ctypes.data_as(ctypes.c_void_p), rcm.ctypes.data_as(ctypes.c_void_p)))
    rpos = np.empty(C, np.int64)
    rpos[rcm] = np.arange(C)
    inner_rcm = float((rpos[a] // block == rpos[b] // block).mean())
    print(f"{name} block {block}: edges inside a tile {inner:.3f} (RCM blocks {inner_rcm:.3f})")
    assert inner > 0.75 and inner > inner_rcm
