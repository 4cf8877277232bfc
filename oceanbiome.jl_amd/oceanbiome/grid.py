"""Grids and fields in the Oceananigans layout used by the C ABI.

A field's parent array has shape (Nz+2Hz, Ny+2Hy, Nx+2Hx) in C order, i.e.
the Oceananigans (i, j, k) column-major layout with i fastest and k = 1 at
the bottom.  Interior cell (i, j, k) (1-based) is parent[k-1+Hz, j-1+Hy, i-1+Hx].
"""
from fractions import Fraction

import numpy as np

from . import units


class RectilinearGrid:
    """Regular RectilinearGrid (Oceananigans): size, extent (or z interval),
    halo and topology.  Flat dimensions have size 1 and halo 0.

    Vertical coordinates follow Oceananigans' regular spacing: faces
    z_F[k] = z_bottom + (k-1) Δz and centres z_F[k] + Δz/2, each the
    correctly rounded value of the exact rational position (Julia's range
    arithmetic).  A z given as Nz+1 face positions makes a stretched grid
    with centres (z_F[k] + z_F[k+1]) / 2 and Δz = z_F[k+1] - z_F[k]."""

    def __init__(self, size, extent=None, z=None, halo=(3, 3, 3),
                 topology=("Periodic", "Periodic", "Bounded"), x=None, y=None):
        if len(size) == 2:  # (Nx, Nz) for a (Bounded, Flat, Bounded) grid
            size = (size[0], 1, size[1])
        self.Nx, self.Ny, self.Nz = (int(s) for s in size)
        self.topology = tuple(topology)
        hx, hy, hz = halo
        self.Hx = 0 if self.topology[0] == "Flat" else int(hx)
        self.Hy = 0 if self.topology[1] == "Flat" else int(hy)
        self.Hz = int(hz)
        if extent is not None:
            if len(extent) == 2:
                extent = (extent[0], 1.0, extent[1])
            Lx, Ly, Lz = extent
            self.x = (0.0, float(Lx))
            self.y = (0.0, float(Ly))
            self.z = (-float(Lz), 0.0)
        else:
            self.x = tuple(float(v) for v in (x or (0.0, 1.0)))
            self.y = tuple(float(v) for v in (y or (0.0, 1.0)))
            self.z = tuple(float(v) for v in z)
        if len(self.z) == 2:
            z0, z1 = (Fraction(v) for v in self.z)
            dz = (z1 - z0) / self.Nz
            self.zf = np.array([float(z0 + k * dz) for k in range(self.Nz + 1)])
            self.zc = np.array([float(z0 + k * dz + dz / 2) for k in range(self.Nz)])
            self.dz = np.full(self.Nz, float(dz))
        else:
            if len(self.z) != self.Nz + 1:
                raise ValueError("z must be an interval or Nz+1 face positions")
            zf = np.array(self.z, dtype=np.float64)
            if np.any(np.diff(zf) <= 0):
                raise ValueError("face positions must increase")
            self.zf = zf
            self.zc = (zf[:-1] + zf[1:]) / 2
            self.dz = np.diff(zf)
        x0, x1 = (Fraction(v) for v in self.x)
        dx = (x1 - x0) / self.Nx
        self.xc = np.array([float(x0 + i * dx + dx / 2) for i in range(self.Nx)])
        y0, y1 = (Fraction(v) for v in self.y)
        dy = (y1 - y0) / self.Ny
        self.yc = np.array([float(y0 + j * dy + dy / 2) for j in range(self.Ny)])

    @property
    def parent_shape(self):
        return (self.Nz + 2 * self.Hz, self.Ny + 2 * self.Hy, self.Nx + 2 * self.Hx)

    @property
    def n_interior(self):
        return self.Nx * self.Ny * self.Nz

    def interior(self, parent):
        """View of the interior (k, j, i) of a parent array."""
        return parent[self.Hz:self.Hz + self.Nz, self.Hy:self.Hy + self.Ny, self.Hx:self.Hx + self.Nx]

    def new_field(self, dtype=np.float64, fill=0.0):
        return np.full(self.parent_shape, fill, dtype=dtype)

    def znodes(self):
        return self.zc.copy()

    def __repr__(self):
        return (f"RectilinearGrid({self.Nx}×{self.Ny}×{self.Nz}, halo=({self.Hx},{self.Hy},{self.Hz}), "
                f"topology={self.topology})")


def ColumnGrid(Nz, dz, halo=3):
    """A single water column of Nz cells of thickness dz (1×1×Nz)."""
    return RectilinearGrid(size=(1, 1, Nz), z=(-Nz * dz, 0.0), halo=(0, 0, halo),
                           topology=("Flat", "Flat", "Bounded"))


__all__ = ["RectilinearGrid", "ColumnGrid", "units"]
