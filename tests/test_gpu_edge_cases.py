"""GPU: small and degenerate meshes, every execution mode, against the CPU
oracle.

Structured strips and grids of quads (some split into triangles), built
here: one cell (its four slots are an inflow, an outflow and two walls),
two and three cells, a 257-cell single-row strip (one cell more than a
tile), 63/64/65-cell grids around a wavefront, and a grid that starts dry
except for the water its inflow edge brings in.  Each runs 30 steps (across
a day boundary, so the boundary series move) in the two-kernel, fused,
graph, persistent and graph-captured two-kernel modes: all modes give the
same bits, within 1e-12 normwise of the oracle in fp64 (1e-6 in fp32).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

FIELDS = ("H", "U", "V", "Z", "W")
MODES = ("two_kernel", "fused", "graph", "persistent", "graph_two_kernel")


def grid_mesh(nx, ny, tri_every=0, dry=False, dtype=np.float64, seed=5):
    """nx x ny unit-square cells of 10 m (every tri_every-th split into two
    triangles), inflow (KLAS 10) on the left edge, a fixed level (KLAS 1) on
    the right, walls elsewhere."""
    import hydro_amd.mesh as M

    rng = np.random.default_rng(seed)
    h = 10.0
    nid = lambda i, j: j * (nx + 1) + i
    cells = []
    for j in range(ny):
        for i in range(nx):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            if tri_every and (i + j * nx) % tri_every == tri_every - 1:
                cells.append([a, b, c, -1])
                cells.append([a, c, d, -1])
            else:
                cells.append([a, b, c, d])
    cells = np.array(cells, dtype=np.int64)
    C = len(cells)
    xs, ys = np.meshgrid(np.arange(nx + 1) * h, np.arange(ny + 1) * h)
    XP, YP = xs.ravel().astype(np.float64), ys.ravel().astype(np.float64)
    NAP, NAC, valid, n1, n2 = M._topology(cells, len(XP))
    nv = np.where(cells[:, 3] >= 0, 4, 3)
    idx = np.where(cells >= 0, cells, 0)
    w = (np.arange(4)[None, :] < nv[:, None]).astype(np.float64)
    xc = (XP[idx] * w).sum(1) / nv
    yc = (YP[idx] * w).sum(1) / nv
    mx = 0.5 * (XP[np.maximum(n1, 0)] + XP[np.maximum(n2, 0)])
    bnd = valid & (NAC == 0)
    Lx = nx * h
    bed = 5.0 + 1e-3 * (Lx - xc) + 0.05 * rng.standard_normal(C)
    depth = np.zeros(C) if dry else 0.5 + 0.1 * rng.random(C)
    MDT, DT, NDAYS = 20, 1.0, 3
    raw = dict(CELL=C, NOD=len(XP), XP=XP, YP=YP, NAP=NAP, NAC=NAC, bnd=bnd, mx=mx, xc=xc, yc=yc)
    KLAS, qs, zs = M._klas_series(raw, Lx, mx < 1e-9, mx > Lx - 1e-9, lambda t: 2.0 + 0.01 * t,
                                  float(bed.mean() + 0.4), 0, rng, NDAYS, MDT)
    raw.update(ZBC=bed, Zinit=bed + depth, n=np.full(C, 0.03), KLAS=KLAS, q_series=qs, z_series=zs,
               NDAYS=NDAYS, MDT=MDT, DT=DT, HM1=0.001, HM2=0.01, JL=0.0)
    return M._finish(f"grid_{nx}x{ny}", raw, dtype)


CASES = [
    ("one cell", dict(nx=1, ny=1)),
    ("two cells", dict(nx=2, ny=1)),
    ("three cells, triangles", dict(nx=2, ny=1, tri_every=2)),
    ("63 cells", dict(nx=9, ny=7)),
    ("64 cells", dict(nx=8, ny=8)),
    ("65 cells", dict(nx=13, ny=5)),
    ("257-cell strip", dict(nx=257, ny=1)),
    ("triangles and quads", dict(nx=11, ny=9, tri_every=3)),
    ("dry start", dict(nx=12, ny=6, dry=True)),
]


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_small_meshes_all_modes(oracle, name, kw, dtype):
    import hydro_amd
    from hydro_amd.device import DeviceMesh

    host = grid_mesh(dtype=dtype, **kw)
    steps = 30
    assert host.max_steps >= steps
    orc = oracle.OracleRun(host)
    orc.step(steps)
    res = {}
    for mode in MODES:
        tune = hydro_amd.default_tune(block=256)
        tune.graph_chunk = 4
        dm = DeviceMesh(host, device=0, tune=tune)
        dm.step(steps, mode)
        res[mode] = dm.fetch(*FIELDS)
        dm.close()
    tol = 1e-12 if dtype == np.float64 else 1e-6
    for k in FIELDS:
        err = rel_err(res["two_kernel"][k], orc[k])
        assert err <= tol, f"{name} {k} vs oracle {err:.2e}"
        for mode in MODES[1:]:
            assert np.array_equal(res[mode][k].view(np.uint8), res["two_kernel"][k].view(np.uint8)), \
                f"{name} {mode} {k}"
    if kw.get("dry"):
        # water entered through the inflow edge only
        assert np.count_nonzero(res["two_kernel"]["H"] > host.HM1) > 0
