"""Collapse-table files (src/EllipsoidalCollapse/CollapseTable.jl:268-310).

Layout: 3 x Int32 (Nx, Ny, Nz), 6 x Float32 (X1, X2, Y1, Y2, Z1, Z2), then
Nx*Ny*Nz Float32 values, x fastest (Fortran stream order).  Values are
1 + z_collapse, or -1 where the ellipsoid does not collapse.
"""
from __future__ import annotations

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
    """read_homeltab (CollapseTable.jl:290-310): bounds widened Float32 -> Float64."""
    raw = open(path, "rb").read()
    nx, ny, nz = (int(v) for v in np.frombuffer(raw[:12], "<i4"))
    b = np.frombuffer(raw[12:36], "<f4").astype(np.float64)
    need = 36 + 4 * nx * ny * nz
    if len(raw) < need:
        raise ValueError(f"{path}: {len(raw)} bytes, header needs {need}")
    data = np.frombuffer(raw[36:need], "<f4").reshape((nx, ny, nz), order="F").copy()
    return CollapseTable(data, *(float(v) for v in b))


def write_homeltab(path: str, ct: CollapseTable):
    """write_homeltab (CollapseTable.jl:273-288)."""
    nx, ny, nz = ct.table.shape
    with open(path, "wb") as f:
        f.write(np.asarray([nx, ny, nz], "<i4").tobytes())
        f.write(np.asarray(ct.bounds, "<f4").tobytes())
        f.write(np.asarray(ct.table, "<f4").ravel(order="F").tobytes())


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
