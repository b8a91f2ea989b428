"""Host-side meshes in the reference layout, and the seeded synthetic meshes.

A :class:`HostMesh` holds numpy arrays laid out exactly like the reference's
``MeshData``/``MeshView`` (cuda_native_impl/hydro-cal-src/src/mesh.cpp,
cell.cpp:11-147, side.cpp:11-52; taichi_impl/mesh_loader_f2.py:61-81):
cell arrays of length CELL, slot arrays of length 4*CELL with slot
``4*cell + j``, boundary series ``[jt*CELL + cell]``, rating curves
``[cell*NHQ + k]``.  NAC and KLAS are int32.

Geometry is derived from node coordinates by the library's native
``hydro_precalculate`` (MeshData::preCalculate, mesh.cpp:280-372) in the
mesh's own precision, like the reference's fp32 build.  Boundary series follow
``take_boundary_for_two_d`` (mesh.cpp:374-466): each series value passes
through a float temporary and a discharge is split over the cells of its
group by a float division, then DQT/DZT are per-step increments over K0 =
MDT/DT steps.

The synthetic generators build jittered-point Delaunay triangulations
(scipy.spatial.Delaunay) and merge triangle pairs into quads to reach an exact
cell count.  Sizes and boundary layouts follow the workloads named in
:data:`CONFIGS`.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import os
from typing import Dict, Optional

import numpy as np

from . import _lib

CELL_FIELDS_R = ("H", "U", "V", "Z", "W", "ZBC", "ZB1", "AREA", "FNC", "BoundaryFeature")
SLOT_FIELDS_R = ("SIDE", "COSF", "SINF", "SLCOS", "SLSIN", "FLUX0", "FLUX1", "FLUX2", "FLUX3")
SERIES_FIELDS = ("QT", "DQT", "ZT", "DZT")


@dataclasses.dataclass
class HostMesh:
    """One mesh in the reference layout (host numpy arrays)."""

    name: str
    dtype: np.dtype
    CELL: int
    NDAYS: int
    NHQ: int
    MDT: int
    DT: float
    HM1: float
    HM2: float
    JL: float
    arrays: Dict[str, np.ndarray]
    # generator-side extras (not part of the kernel contract)
    xc: Optional[np.ndarray] = None
    yc: Optional[np.ndarray] = None

    def __getattr__(self, k):
        arrays = self.__dict__.get("arrays")
        if arrays is not None and k in arrays:
            return arrays[k]
        raise AttributeError(k)

    @property
    def steps_per_day(self) -> int:
        return int(self.MDT / self.DT)

    @property
    def max_steps(self) -> int:
        """Steps the reference loop can run before the NDAYS cap (benchmark.cu:89-90)."""
        return self.NDAYS * max(self.steps_per_day - 1, 0)

    def copy(self) -> "HostMesh":
        c = dataclasses.replace(self)
        c.arrays = {k: v.copy() for k, v in self.arrays.items()}
        return c

    def state(self):
        return {k: self.arrays[k].copy() for k in ("H", "U", "V", "Z", "W")}

    @property
    def n_edges(self) -> int:
        """Unique mesh edges (interior edges once, boundary edges once)."""
        nac = self.arrays["NAC"]
        side = self.arrays["SIDE"]
        real = side > 0
        interior = int(np.count_nonzero(real & (nac > 0)))
        boundary = int(np.count_nonzero(real & (nac == 0)))
        return interior // 2 + boundary


# ---------------------------------------------------------------------------
# triangulation helpers


def _jittered_points(Lx, Ly, nx, ny, rng, jitter=0.3):
    xs = np.linspace(0.0, Lx, nx)
    ys = np.linspace(0.0, Ly, ny)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    dx, dy = Lx / (nx - 1), Ly / (ny - 1)
    jx = rng.uniform(-jitter, jitter, X.shape) * dx
    jy = rng.uniform(-jitter, jitter, X.shape) * dy
    col = np.arange(nx)[None, :].repeat(ny, 0)
    row = np.arange(ny)[:, None].repeat(nx, 1)
    on_x = (col == 0) | (col == nx - 1)
    on_y = (row == 0) | (row == ny - 1)
    jx[on_x] = 0.0          # left/right sides stay on x = 0 / Lx
    jy[on_y] = 0.0          # bottom/top sides stay on y = 0 / Ly
    return np.stack([(X + jx).ravel(), (Y + jy).ravel()], axis=1)


def _delaunay(pts):
    from scipy.spatial import Delaunay

    tri = Delaunay(pts).simplices.astype(np.int64)
    a, b, c = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
    cross = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    flip = cross < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    cross = np.abs(cross)
    keep = cross > 1e-9 * np.median(cross)
    return tri[keep]


def _merge_pairs(pts, tri, n_merge, rng):
    """Merge n_merge adjacent triangle pairs into convex quads, best shaped first.

    Returns an int64 array [C, 4] of CCW node ids (-1 in the 4th column of a
    triangle), in the order of the triangles (a quad takes the place of its
    first triangle).
    """
    T = len(tri)
    if n_merge <= 0:
        return np.concatenate([tri, np.full((T, 1), -1, dtype=np.int64)], axis=1)
    N = len(pts)
    # half-edges (t, k): edge from tri[t,k] to tri[t,(k+1)%3]
    va = tri.ravel()
    vb = tri[:, [1, 2, 0]].ravel()
    key = np.minimum(va, vb) * N + np.maximum(va, vb)
    order = np.argsort(key, kind="stable")
    ks = key[order]
    same = np.nonzero(ks[1:] == ks[:-1])[0]
    h1, h2 = order[same], order[same + 1]
    t1, k1 = h1 // 3, h1 % 3
    t2, k2 = h2 // 3, h2 % 3
    # t1's shared edge runs b -> c, its third vertex is a; t2 adds d
    b = tri[t1, k1]
    c = tri[t1, (k1 + 1) % 3]
    a = tri[t1, (k1 + 2) % 3]
    d = tri[t2, (k2 + 2) % 3]
    quad = np.stack([a, b, d, c], axis=1)
    P = pts[quad]                                  # (E, 4, 2)
    e_in = P - np.roll(P, 1, axis=1)               # edge into each corner
    e_out = np.roll(P, -1, axis=1) - P             # edge out of each corner
    cross = e_in[..., 0] * e_out[..., 1] - e_in[..., 1] * e_out[..., 0]
    convex = np.all(cross > 0, axis=1)
    cosang = -(e_in * e_out).sum(-1) / (np.linalg.norm(e_in, axis=-1) * np.linalg.norm(e_out, axis=-1))
    ang = np.arccos(np.clip(cosang, -1, 1))
    score = np.abs(ang - math.pi / 2).max(axis=1)  # distance from a rectangle
    score = score + 1e-9 * rng.random(len(score))  # seeded tie break
    cand = np.nonzero(convex)[0]
    cand = cand[np.argsort(score[cand], kind="stable")]
    used = np.zeros(T, dtype=bool)
    cells = np.concatenate([tri, np.full((T, 1), -1, dtype=np.int64)], axis=1)
    drop = np.zeros(T, dtype=bool)
    count = 0
    t1l, t2l = t1.tolist(), t2.tolist()
    for e in cand.tolist():
        if count >= n_merge:
            break
        x, y = t1l[e], t2l[e]
        if used[x] or used[y]:
            continue
        used[x] = used[y] = True
        lo, hi = (x, y) if x < y else (y, x)
        cells[lo] = quad[e]
        drop[hi] = True
        count += 1
    if count < n_merge:
        raise ValueError(f"only {count} of {n_merge} quad merges possible")
    return cells[~drop]


def _topology(cells, n_nodes):
    """NAP [4*C] (1-based, triangle 4th = 0), NAC [4*C] (1-based, 0 = none),
    and per-slot edge endpoints for boundary classification."""
    C = len(cells)
    NAP = (cells + 1).astype(np.int32)                # -1 padding becomes 0
    nv = np.where(cells[:, 3] >= 0, 4, 3).astype(np.int32)
    cid = np.repeat(np.arange(C), 4)
    j = np.tile(np.arange(4), C)
    valid = j < nv[cid]
    n1 = NAP.ravel() - 1
    nxt = np.where(j + 1 < nv[cid], j + 1, 0)
    n2 = NAP[cid, nxt] - 1
    key = np.where(valid, np.minimum(n1, n2).astype(np.int64) * n_nodes + np.maximum(n1, n2), -1)
    idx = np.nonzero(valid)[0]
    order = idx[np.argsort(key[idx], kind="stable")]
    ks = key[order]
    NAC = np.zeros(4 * C, dtype=np.int32)
    pair = np.nonzero(ks[1:] == ks[:-1])[0]
    s1, s2 = order[pair], order[pair + 1]
    NAC[s1] = cid[s2] + 1
    NAC[s2] = cid[s1] + 1
    return NAP.ravel(), NAC, valid, n1, n2


# ---------------------------------------------------------------------------
# boundary series (take_boundary_for_two_d, mesh.cpp:374-466)


def _increments(lib, V, rows, C, K0, dtype):
    """Per-step increments of a column-major series (hydro_boundary_increments)."""
    D = np.zeros_like(V)
    _lib.check(lib.hydro_boundary_increments(_lib.F64 if dtype == np.float64 else _lib.F32, C, rows, rows, K0,
                                             V.ctypes.data_as(ctypes.c_void_p), D.ctypes.data_as(ctypes.c_void_p)))
    return D


def _take_boundary(lib, bd, C, rows, MDT, DT, HM2, dtype, ZBC, SIDE, KLAS):
    """Series and rating curves of a .DAT directory (hydro_take_boundary)."""
    keep = []

    def arr(a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        keep.append(a)
        return a.ctypes.data_as(ctypes.c_void_p) if a.size else None

    NHQ = int(bd["NHQ"])
    d = _lib.BoundaryDesc(dtype=_lib.F64 if dtype == np.float64 else _lib.F32, CELL=C, NDAYS=int(bd["NDAYS"]),
                          rows=rows, MDT=int(MDT), NHQ=NHQ, DT=float(DT), HM2=float(HM2))
    for key in ("NZ", "NQ", "NZQ"):
        nfiles, cell, group = bd[key]
        setattr(d, key, len(cell))
        setattr(d, "N" + key + "0", int(nfiles))
        setattr(d, "MB" + key[1:], arr(cell))
        setattr(d, "N" + key, arr(group))
    d.ZBC = ZBC.ctypes.data_as(ctypes.c_void_p)
    d.SIDE = SIDE.ctypes.data_as(ctypes.c_void_p)
    KLAS = np.ascontiguousarray(KLAS, dtype=np.int32)
    d.KLAS = KLAS.ctypes.data_as(ctypes.c_void_p)
    out = {k: np.zeros(rows * C, dtype=dtype) for k in ("QT", "DQT", "ZT", "DZT")}
    out["NHQ1"] = np.zeros(C, dtype=np.int32)
    out["ZW"] = np.zeros(max(C * NHQ, 1), dtype=dtype)
    out["QW"] = np.zeros(max(C * NHQ, 1), dtype=dtype)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    _lib.check(lib.hydro_take_boundary(ctypes.byref(d), os.fsencode(bd["dir"]), p(out["QT"]), p(out["DQT"]),
                                       p(out["ZT"]), p(out["DZT"]), p(out["NHQ1"]), p(out["ZW"]), p(out["QW"])))
    return out


def _finish(name, raw, dtype, lib=None) -> HostMesh:
    """Raw generator output -> HostMesh in `dtype`, via the native preCalculate."""
    dtype = np.dtype(dtype)
    R = dtype.type
    lib = lib or _lib.load()
    C, NOD = raw["CELL"], raw["NOD"]
    XP = raw["XP"].astype(dtype)
    YP = raw["YP"].astype(dtype)
    XYMIN = np.array([XP.min(), YP.min()], dtype=np.float64)   # preCalculate shifts XP, YP by these
    ZBC = raw["ZBC"].astype(dtype)
    Z = raw["Zinit"].astype(dtype)
    FNC0 = raw["n"].astype(dtype)
    NAP = np.ascontiguousarray(raw["NAP"], dtype=np.int32)
    out = {k: np.zeros(4 * C, dtype=dtype) for k in ("SIDE", "SINF", "COSF", "SLCOS", "SLSIN")}
    outc = {k: np.zeros(C, dtype=dtype) for k in ("AREA", "ZB1", "H", "FNC")}
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    st = lib.hydro_precalculate(
        _lib.F64 if dtype == np.float64 else _lib.F32, C, NOD, p(NAP), p(XP), p(YP), p(ZBC), p(Z), p(FNC0),
        float(raw["HM1"]), p(out["SIDE"]), p(out["SINF"]), p(out["COSF"]), p(out["SLCOS"]), p(out["SLSIN"]),
        p(outc["AREA"]), p(outc["ZB1"]), p(outc["H"]), p(outc["FNC"]))
    _lib.check(st)
    NDAYS = raw["NDAYS"]
    K0 = int(raw["MDT"] / raw["DT"])
    QT = np.zeros(NDAYS * C, dtype=dtype)
    ZT = np.zeros_like(QT)
    for cell, series in raw.get("q_series", {}).items():   # one value per row and cell
        QT[np.arange(NDAYS) * C + cell] = series.astype(dtype)
    for cell, series in raw.get("z_series", {}).items():
        ZT[np.arange(NDAYS) * C + cell] = series.astype(dtype)
    DQT = _increments(lib, QT, NDAYS, C, K0, dtype)
    DZT = _increments(lib, ZT, NDAYS, C, K0, dtype)
    for k in ("QT", "DQT", "ZT", "DZT"):   # series already laid out by the caller
        if k in raw:
            {"QT": QT, "DQT": DQT, "ZT": ZT, "DZT": DZT}[k][:] = raw[k].astype(dtype)
    NHQ = int(raw.get("NHQ", 0))
    if "bounde" in raw:   # a .DAT directory: series and rating curves read by the library
        tb = _take_boundary(lib, raw["bounde"], C, NDAYS, raw["MDT"], raw["DT"], raw["HM2"], dtype, ZBC,
                            out["SIDE"], raw["KLAS"])
        QT, DQT, ZT, DZT = tb["QT"], tb["DQT"], tb["ZT"], tb["DZT"]
        raw = dict(raw, NHQ1=tb["NHQ1"], ZW=tb["ZW"], QW=tb["QW"])
    arrays = dict(
        H=outc["H"], U=raw.get("U0", np.zeros(C)).astype(dtype), V=raw.get("V0", np.zeros(C)).astype(dtype),
        Z=Z, W=np.zeros(C, dtype=dtype), ZBC=ZBC, ZB1=outc["ZB1"], AREA=outc["AREA"], FNC=outc["FNC"],
        BoundaryFeature=raw.get("BF", np.zeros(C)).astype(dtype),
        NHQ1=np.ascontiguousarray(raw.get("NHQ1", np.zeros(C, np.int32)), dtype=np.int32),
        ZW=raw.get("ZW", np.zeros(max(C * NHQ, 1))).astype(dtype),
        QW=raw.get("QW", np.zeros(max(C * NHQ, 1))).astype(dtype),
        QT=QT, DQT=DQT, ZT=ZT, DZT=DZT,
        NAC=np.ascontiguousarray(raw["NAC"], dtype=np.int32), KLAS=np.ascontiguousarray(raw["KLAS"], dtype=np.int32),
        SIDE=out["SIDE"], COSF=out["COSF"], SINF=out["SINF"], SLCOS=out["SLCOS"], SLSIN=out["SLSIN"],
        FLUX0=np.zeros(4 * C, dtype=dtype), FLUX1=np.zeros(4 * C, dtype=dtype),
        FLUX2=np.zeros(4 * C, dtype=dtype), FLUX3=np.zeros(4 * C, dtype=dtype),
        NAP=NAP, XP=XP, YP=YP, XYMIN=XYMIN,
    )
    return HostMesh(name=name, dtype=dtype, CELL=C, NDAYS=NDAYS, NHQ=NHQ, MDT=int(raw["MDT"]),
                    DT=float(raw["DT"]), HM1=float(raw["HM1"]), HM2=float(raw["HM2"]), JL=float(raw.get("JL", 0.0)),
                    arrays=arrays, xc=raw.get("xc"), yc=raw.get("yc"))


def _grid_for(target_cells, aspect, quad_frac):
    """Grid nx x ny whose Delaunay has >= target/(1-quad_frac) triangles."""
    T_need = target_cells / (1.0 - quad_frac)
    ny = max(3, int(math.sqrt(T_need / (2.0 * aspect))))
    while True:
        nx = max(3, int(round(ny * aspect)))
        T = 2 * nx * ny - 2 - (2 * (nx - 1) + 2 * (ny - 1))
        if T >= T_need:
            return nx, ny
        ny += 1


def _gaussians(xc, yc, centers, amps, sigmas):
    z = np.zeros_like(xc)
    for (cx, cy), a, s in zip(centers, amps, sigmas):
        z += a * np.exp(-((xc - cx) ** 2 + (yc - cy) ** 2) / (2.0 * s * s))
    return z


def _raw_mesh(Lx, Ly, target_cells, quad_frac, rng, aspect=None):
    aspect = aspect or (Lx / Ly)
    nx, ny = _grid_for(target_cells, aspect, quad_frac)
    pts = _jittered_points(Lx, Ly, nx, ny, rng)
    tri = _delaunay(pts)
    n_merge = len(tri) - target_cells
    cells = _merge_pairs(pts, tri, n_merge, rng)
    assert len(cells) == target_cells, (len(cells), target_cells)
    NAP, NAC, valid, n1, n2 = _topology(cells, len(pts))
    C = len(cells)
    nv = np.where(cells[:, 3] >= 0, 4.0, 3.0)
    idx = np.where(cells >= 0, cells, 0)
    w = (cells >= 0).astype(np.float64)
    cx = (pts[idx, 0] * w).sum(1) / nv
    cy = (pts[idx, 1] * w).sum(1) / nv
    mx = 0.5 * (pts[n1, 0] + pts[n2, 0])
    my = 0.5 * (pts[n1, 1] + pts[n2, 1])
    bnd = valid & (NAC == 0)
    return dict(pts=pts, NAP=NAP, NAC=NAC, valid=valid, bnd=bnd, mx=mx, my=my, xc=cx, yc=cy, CELL=C,
                NOD=len(pts), XP=pts[:, 0].copy(), YP=pts[:, 1].copy())


def _klas_series(raw, Lx, inflow_mask, outflow_mask, q_of_t, z_out, n_extra_q0, rng, NDAYS, MDT):
    """KLAS per slot and the boundary series of the cells that need them."""
    C = raw["CELL"]
    KLAS = np.zeros(4 * C, dtype=np.int32)
    bnd = raw["bnd"]
    KLAS[bnd] = 4
    KLAS[bnd & inflow_mask] = 10
    KLAS[bnd & outflow_mask] = 1
    # extra discharge edges with QT = 0, on wall slots of otherwise plain cells
    cell_of = np.arange(4 * C) // 4
    special = np.zeros(C, dtype=bool)
    special[cell_of[KLAS == 10]] = True
    special[cell_of[KLAS == 1]] = True
    walls = np.nonzero((KLAS == 4))[0]
    walls = walls[~special[cell_of[walls]]]
    _, first = np.unique(cell_of[walls], return_index=True)
    walls = walls[first]
    extra = rng.choice(walls, size=min(n_extra_q0, len(walls)), replace=False) if n_extra_q0 else []
    KLAS[extra] = 10
    in_slots = np.nonzero(bnd & inflow_mask)[0]
    n_in = len(in_slots)
    t = np.arange(NDAYS) * MDT
    q_series = {}
    if n_in:
        # float temporary, then a float division over the group (mesh.cpp:442-445)
        qt = np.array([q_of_t(tt) for tt in t], dtype=np.float32) / np.float32(n_in)
        for cell in np.unique(cell_of[in_slots]):
            q_series[int(cell)] = qt.astype(np.float64)
    for s in extra:
        q_series[int(cell_of[s])] = np.zeros(NDAYS)
    z_series = {}
    out_slots = np.nonzero(bnd & outflow_mask)[0]
    zt = np.full(NDAYS, np.float32(z_out), dtype=np.float32).astype(np.float64)
    for cell in np.unique(cell_of[out_slots]):
        z_series[int(cell)] = zt
    return KLAS, q_series, z_series


def make_channel_mesh(seed=1001, cells=6675, steps=100, DT=1.0, MDT=60, dtype=np.float64, lib=None, rotate=0.0):
    """Plain forward, F1-style: a 3 km x 1 km channel of mixed triangles and quads.

    Bed: 1e-3 planar slope plus 5 Gaussian bumps of 0.5 m; initial depth
    1.0 +- 0.05 m at rest; Manning n uniform in [0.02, 0.035]; walls (KLAS 4)
    on the sides, an upstream discharge set (KLAS 10) carrying
    50 +- 20 m^3/s with a 3600 s period, a fixed downstream level (KLAS 1)
    and 8 more KLAS 10 edges with QT = 0.  `rotate` (radians) turns the
    finished geometry about the origin, so that no boundary normal has a
    zero component.
    """
    rng = np.random.default_rng(seed)
    Lx, Ly = 3000.0, 1000.0
    raw = _raw_mesh(Lx, Ly, cells, 0.43, rng)
    xc, yc = raw["xc"], raw["yc"]
    centers = np.stack([rng.uniform(300, Lx - 300, 5), rng.uniform(150, Ly - 150, 5)], 1)
    bed = 10.0 + 1e-3 * (Lx - xc) + _gaussians(xc, yc, centers, np.full(5, 0.5), rng.uniform(60, 150, 5))
    h0 = 1.0 + rng.uniform(-0.05, 0.05, cells)
    n = rng.uniform(0.02, 0.035, cells)
    spd = int(MDT / DT)
    NDAYS = math.ceil(steps / max(spd - 1, 1)) + 1
    eps = 1e-6 * Lx
    inflow = raw["mx"] < eps
    outflow = raw["mx"] > Lx - eps
    out_cells = np.unique((np.nonzero(raw["bnd"] & outflow)[0]) // 4)
    z_out = float(np.mean(bed[out_cells]) + 1.0)
    q = lambda t: 50.0 + 20.0 * math.sin(2.0 * math.pi * t / 3600.0)
    KLAS, qs, zs = _klas_series(raw, Lx, inflow, outflow, q, z_out, 8, rng, NDAYS, MDT)
    raw.update(ZBC=bed, Zinit=bed + h0, n=n, KLAS=KLAS, q_series=qs, z_series=zs, NDAYS=NDAYS, MDT=MDT, DT=DT,
               HM1=0.001, HM2=0.01, JL=0.0)
    if rotate:
        c, s = math.cos(rotate), math.sin(rotate)
        for kx, ky in (("XP", "YP"), ("xc", "yc")):
            x, y = raw[kx].copy(), raw[ky].copy()
            raw[kx], raw[ky] = c * x - s * y, s * x + c * y
    return _finish(f"channel_{cells}", raw, dtype, lib)


def make_floodplain_mesh(seed=2002, cells=207234, steps=900, DT=0.5, MDT=60, dtype=np.float64, lib=None,
                         Lx=20000.0, Ly=8000.0, q_peak=400.0, n_bumps=12, n_channel=0.03, n_plain=0.05,
                         slope=2e-4, depth_offset=0.0):
    """Plain forward, F2-style: a 20 km x 8 km floodplain with a meandering
    main channel carved 3 m into the bed; depths 0-4 m with wet/dry fronts;
    Manning n 0.03 in the channel and 0.05 on the floodplain.

    Mostly triangles (a few hundred quads make the cell count exact).
    Upstream discharge (KLAS 10) enters through the channel mouth, the
    downstream side holds a fixed level (KLAS 1), walls elsewhere.
    ``depth_offset`` raises the initial level and the downstream level by
    that many metres (a deeper flood on the same bed).
    """
    rng = np.random.default_rng(seed)
    raw = _raw_mesh(Lx, Ly, cells, 0.0015, rng, aspect=Lx / Ly)
    xc, yc = raw["xc"], raw["yc"]
    nb = int(n_bumps)
    centers = np.stack([rng.uniform(0.05 * Lx, 0.95 * Lx, nb), rng.uniform(0.05 * Ly, 0.95 * Ly, nb)], 1)
    plain = 10.0 + slope * (Lx - xc)
    bumps = _gaussians(xc, yc, centers, rng.uniform(0.3, 1.5, nb), rng.uniform(0.03 * Lx, 0.06 * Lx, nb) * 0.5)
    phase = rng.uniform(0, 2 * math.pi)
    wav = 0.3 * Lx
    ycl = 0.5 * Ly + 0.19 * Ly * np.sin(2 * math.pi * xc / wav + phase)
    hw = max(0.015 * Ly, 60.0)
    dist = np.abs(yc - ycl)
    carve = 3.0 * np.clip(1.0 - (dist / hw) ** 2, 0.0, None)
    in_channel = dist < hw
    bed = plain + bumps * (~in_channel) - carve
    level = plain + 0.8 + depth_offset
    n = np.where(in_channel, n_channel, n_plain)
    spd = int(MDT / DT)
    NDAYS = math.ceil(steps / max(spd - 1, 1)) + 1
    eps = 1e-6 * Lx
    my = raw["my"]
    y_mouth = 0.5 * Ly + 0.19 * Ly * math.sin(phase)
    inflow = (raw["mx"] < eps) & (np.abs(my - y_mouth) < hw)
    outflow = raw["mx"] > Lx - eps
    z_out = float(10.0 + 0.8 + depth_offset)
    q = lambda t: q_peak * (0.75 + 0.25 * math.sin(2.0 * math.pi * t / 3600.0))
    KLAS, qs, zs = _klas_series(raw, Lx, inflow, outflow, q, z_out, 0, rng, NDAYS, MDT)
    raw.update(ZBC=bed, Zinit=level, n=n, KLAS=KLAS, q_series=qs, z_series=zs, NDAYS=NDAYS, MDT=MDT, DT=DT,
               HM1=0.001, HM2=0.01, JL=0.0)
    return _finish(f"floodplain_{cells}", raw, dtype, lib)


def ensemble_specs(seed=3003, n=128, lo=2000, hi=207234):
    """Per-member parameters of the batched ensemble: cell counts log-uniform
    in [lo, hi]; domains scaled to keep the floodplain's cell size; each member
    its own bed (slope plus 1-8 bumps), Manning pair and inflow peak
    (10-500 m^3/s)."""
    rng = np.random.default_rng(seed)
    cells = np.exp(rng.uniform(math.log(lo), math.log(hi), n)).astype(int)
    specs = []
    for i, c in enumerate(cells.tolist()):
        area = c * (20000.0 * 8000.0 / 207234)
        Ly = math.sqrt(area / 2.5)
        specs.append(dict(seed=int(seed * 1000 + i), cells=int(c), Lx=2.5 * Ly, Ly=Ly,
                          q_peak=float(rng.uniform(10, 500)), n_bumps=int(rng.integers(1, 9)),
                          n_channel=float(rng.uniform(0.025, 0.035)), n_plain=float(rng.uniform(0.04, 0.06)),
                          slope=float(rng.uniform(1e-4, 4e-4))))
    return specs


def _ensemble_member(args):
    spec, steps, dtype = args
    return make_floodplain_mesh(steps=steps, DT=0.5, dtype=dtype, **spec)


def make_ensemble(seed=3003, n=128, lo=2000, hi=207234, steps=500, dtype=np.float32, workers=None, threads=None):
    """The members of the batched ensemble, built in parallel processes
    (`workers`, forked: only before the process uses the GPU) or threads of
    this process (`threads`: safe at any time; the builders release the
    interpreter lock in their array work).  Every member is built from its
    own seed, so the result does not depend on either."""
    specs = ensemble_specs(seed, n, lo, hi)
    jobs = [(sp, steps, dtype) for sp in specs]
    if threads:
        from multiprocessing.pool import ThreadPool

        with ThreadPool(int(threads)) as pool:
            return pool.map(_ensemble_member, jobs, chunksize=1)
    workers = workers or min(16, os.cpu_count() or 1)
    if workers <= 1 or n == 1:
        return [_ensemble_member(j) for j in jobs]
    import multiprocessing as mp

    with mp.get_context("fork").Pool(workers) as pool:
        return pool.map(_ensemble_member, jobs, chunksize=1)


# Workloads (sizes and time steps).  "channel_6675" is the CPU-runnable
# case; "floodplain_207234" is the throughput workload.
CONFIGS = {
    "channel_6675": dict(maker=make_channel_mesh, kw=dict(seed=1001, cells=6675, steps=100, DT=1.0), steps=100),
    "floodplain_207234": dict(maker=make_floodplain_mesh, kw=dict(seed=2002, cells=207234, steps=900, DT=0.5),
                              steps=900),
    "floodplain_24020": dict(maker=make_floodplain_mesh,
                             kw=dict(seed=4004, cells=24020, steps=500, DT=1.0, Lx=6800.0, Ly=2720.0, q_peak=60.0),
                             steps=500),
}


def make_config(name: str, dtype=np.float64, lib=None, **over) -> HostMesh:
    cfg = CONFIGS[name]
    kw = dict(cfg["kw"])
    kw.update(over)
    return cfg["maker"](dtype=dtype, lib=lib, **kw)
