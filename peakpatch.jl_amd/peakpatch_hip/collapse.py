"""Collapse tables: the file format (src/EllipsoidalCollapse/CollapseTable.jl:
81-122), GPU generation (make_table, :33-55) and analytic test tables.

Layout: 3 x Int32 (Nx, Ny, Nz), 6 x Float32 (X1, X2, Y1, Y2, Z1, Z2), then
Nx*Ny*Nz Float32 values, x fastest (Fortran stream order).  Values are
1 + z_collapse, or -1 where the ellipsoid does not collapse.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class CollapseTable:
    table: np.ndarray          # (Nx, Ny, Nz), indexed [ix, iy, iz]
    X1: float
    X2: float
    Y1: float
    Y2: float
    Z1: float
    Z2: float

    @property
    def bounds(self):
        return (self.X1, self.X2, self.Y1, self.Y2, self.Z1, self.Z2)


def read_homeltab(path: str) -> CollapseTable:
    """read_homeltab (CollapseTable.jl:102-122): bounds widened Float32 -> Float64."""
    raw = open(path, "rb").read()
    nx, ny, nz = (int(v) for v in np.frombuffer(raw[:12], "<i4"))
    b = np.frombuffer(raw[12:36], "<f4").astype(np.float64)
    need = 36 + 4 * nx * ny * nz
    if len(raw) < need:
        raise ValueError(f"{path}: {len(raw)} bytes, header needs {need}")
    data = np.frombuffer(raw[36:need], "<f4").reshape((nx, ny, nz), order="F").copy()
    return CollapseTable(data, *(float(v) for v in b))


def write_homeltab(path: str, ct: CollapseTable):
    """write_homeltab (CollapseTable.jl:85-100)."""
    nx, ny, nz = ct.table.shape
    with open(path, "wb") as f:
        f.write(np.asarray([nx, ny, nz], "<i4").tobytes())
        f.write(np.asarray(ct.bounds, "<f4").tobytes())
        f.write(np.asarray(ct.table, "<f4").ravel(order="F").tobytes())


DEFAULT_GRID = (50, 20, 20)
DEFAULT_BOUNDS = (math.log10(1.5), math.log10(8.0), 0.0, 0.5, -1.0 + 1e-4, 1.0 - 1e-4)


def generate(ctx, cosmology=None, shape=DEFAULT_GRID, bounds=DEFAULT_BOUNDS, install: bool = False) -> CollapseTable:
    """make_table with the Fortran-compatible RK4 solver
    (CollapseTable.jl:9-55, EllipsoidalCollapse.jl:527-624) on the GPU of
    ctx, for the cosmology's linear growth tables (Cosmology.jl:148-270).
    The grid and bounds default to CollapseTableParams().  With install the
    table also becomes ctx's collapse table."""
    import ctypes as C

    from . import _lib as L
    from .cosmology import CosmologyParams, GrowthTables
    c = cosmology or CosmologyParams()
    g = GrowthTables(c)
    cols = [np.ascontiguousarray(a, np.float64) for a in (g.atab, g.Datab, g.dlnD)]
    gr = L.Growth(len(cols[0]), *(a.ctypes.data for a in cols))
    nx, ny, nz = shape
    out = np.zeros(nx * ny * nz, np.float32)
    L.check(ctx.lib.pp_make_collapse_table(ctx.h, c.Om, c.OL, C.byref(gr), nx, ny, nz, *bounds, out.ctypes.data,
                                           1 if install else 0))
    return CollapseTable(out.reshape((nx, ny, nz), order="F"), *(float(b) for b in bounds))


def constant_table(value: float, nx=10, ny=10, nz=10, bounds=(-2.0, 2.0, 0.0, 1.0, 0.0, 2.0)) -> CollapseTable:
    """Constant analytic table (the reference GPU tests' synthetic table)."""
    return CollapseTable(np.full((nx, ny, nz), value, np.float32), *bounds)


def sinusoidal_table(nx=11, ny=11, nz=11, bounds=(-2.0, 2.0, 0.0, 1.0, 0.0, 2.0)) -> CollapseTable:
    """Structured analytic table 0.5 + 0.3 sin(0.3 i) + 0.2 cos(0.4 j) + 0.01 k
    (0-based i, j, k), the reference benchmark's synthetic table."""
    i = np.arange(nx, dtype=np.float32)[:, None, None]
    j = np.arange(ny, dtype=np.float32)[None, :, None]
    k = np.arange(nz, dtype=np.float32)[None, None, :]
    t = (np.float32(0.5) + np.float32(0.3) * np.sin(i * np.float32(0.3)) +
         np.float32(0.2) * np.cos(j * np.float32(0.4)) + np.float32(0.1) * k * np.float32(0.1))
    return CollapseTable(t.astype(np.float32), *bounds)
