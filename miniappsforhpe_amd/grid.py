"""Padded-grid geometry shared by every operator (the `axis` / `cart_volume_regular` rules of the reference).

A volume holds an nx x ny x nz interior with `ghost` points on each side of every axis; the x extent is padded up to
a multiple of `align` floats (128 B for align=32). Index (ix, iy, iz) in array coordinates maps to
(iz*ldy + iy)*ldx + ix. See infrastructure/axis.cpp:19-27 (trailing padding), :78-100 (ghosts) and
utils/volume_index.h:22 (x-fastest indexing) in the reference.
"""
from __future__ import annotations

from dataclasses import dataclass


def roundup(n: int, m: int) -> int:
    return ((n + m - 1) // m) * m if m > 0 else n


@dataclass(frozen=True)
class Grid:
    nx: int
    ny: int
    nz: int
    ghost: int = 4
    align: int = 32

    def __post_init__(self):
        if min(self.nx, self.ny, self.nz) < 1 or self.ghost < 0:
            raise ValueError(f"invalid grid {self}")

    @property
    def ldx(self) -> int:
        return roundup(self.nx + 2 * self.ghost, self.align)

    @property
    def ldy(self) -> int:
        return self.ny + 2 * self.ghost

    @property
    def ldz(self) -> int:
        return self.nz + 2 * self.ghost

    @property
    def shape(self) -> tuple[int, int, int]:
        """Allocated (z, y, x) shape."""
        return (self.ldz, self.ldy, self.ldx)

    @property
    def ntot(self) -> int:
        return self.ldx * self.ldy * self.ldz

    @property
    def npoints(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def ranges(self) -> tuple[int, int, int, int, int, int]:
        """Inclusive interior ranges (ixbeg, ixend, iybeg, iyend, izbeg, izend) in array coordinates."""
        g = self.ghost
        return (g, g + self.nx - 1, g, g + self.ny - 1, g, g + self.nz - 1)

    def interior(self, vol):
        """Interior view of a (ldz, ldy, ldx) array or tensor."""
        g = self.ghost
        return vol[g : g + self.nz, g : g + self.ny, g : g + self.nx]

    def nbytes(self, itemsize: int = 4) -> int:
        return self.ntot * itemsize
