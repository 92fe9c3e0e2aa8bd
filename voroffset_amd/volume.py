"""Dexel volume container: the CompressedVolume layout as a flat CSR.

Mirrors ``vor3d::CompressedVolume`` (src/vor3d/CompressedVolume.h:10-41,
CompressedVolumeBase.h:14-49):

* a 2-D grid ``nx x ny`` of z-columns, column index ``x + nx*y``
  (CompressedVolume.h:28-29);
* each column a sorted flat list of endpoints ``[b0, e0, b1, e1, ...]`` in
  dexel units of absolute z (Dexelize.cpp:204); pairs are read as (k, k+1) and
  a trailing odd value is ignored (CompressedVolume.cpp:25-41);
* geometry: ``origin`` (already shifted by ``-padding*spacing`` on all three
  axes), ``extent``, ``spacing`` and ``padding`` (CompressedVolume.cpp:11-23).

Instead of ``vector<vector<double>>`` the columns live in one CSR pair:
``off`` (int64, ``nx*ny + 1`` entries) and ``z`` (float64 or float32).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class CompressedVolume:
    nx: int
    ny: int
    off: np.ndarray  # int64 [nx*ny + 1]
    z: np.ndarray  # float64 (or float32) endpoints, absolute dexel units
    origin: tuple = (0.0, 0.0, 0.0)
    extent: tuple = (0.0, 0.0, 0.0)
    spacing: float = 1.0
    padding: int = 0
    meta: dict = field(default_factory=dict)

    # -- construction -----------------------------------------------------
    @classmethod
    def from_bbox(cls, bbox_min, extent, spacing, padding):
        """Empty volume over a bounding box (CompressedVolume.cpp:11-23)."""
        origin = tuple(float(b) - padding * spacing for b in bbox_min)
        nx = int(math.ceil(extent[0] / spacing) + 2 * padding)
        ny = int(math.ceil(extent[1] / spacing) + 2 * padding)
        return cls(nx, ny, np.zeros(nx * ny + 1, np.int64), np.zeros(0, np.float64),
                   origin, tuple(float(e) for e in extent), float(spacing), int(padding))

    @classmethod
    def from_columns(cls, nx, ny, columns, **geom):
        """Build from a list of per-column endpoint sequences (index x + nx*y)."""
        if len(columns) != nx * ny:
            raise ValueError("need nx*ny columns")
        lens = np.fromiter((len(c) for c in columns), np.int64, count=nx * ny)
        off = np.zeros(nx * ny + 1, np.int64)
        np.cumsum(lens, out=off[1:])
        z = np.concatenate([np.asarray(c, np.float64) for c in columns]) if off[-1] else np.zeros(0)
        return cls(nx, ny, off, z.astype(np.float64), **geom)

    def like(self, off, z):
        """A volume with this volume's geometry and new data."""
        return CompressedVolume(self.nx, self.ny, off, z, self.origin, self.extent,
                                self.spacing, self.padding, dict(self.meta))

    # -- geometry ---------------------------------------------------------
    @property
    def ncols(self) -> int:
        return self.nx * self.ny

    def z_range(self):
        """[z_lo, z_hi] used by erosion (Voronoi.cpp:10-11, VoronoiVorPower.cpp:28-29)."""
        zlo = self.origin[2] / self.spacing
        zhi = self.origin[2] / self.spacing + 2 * self.padding + self.extent[2] / self.spacing
        return zlo, zhi

    def dexel_center(self, x, y):
        """(CompressedVolumeBase.cpp:5-11)."""
        return ((x + 0.5) * self.spacing + self.origin[0], (y + 0.5) * self.spacing + self.origin[1])

    # -- access -----------------------------------------------------------
    def at(self, x, y):
        c = x + self.nx * y
        return self.z[self.off[c]:self.off[c + 1]]

    def column(self, c):
        return self.z[self.off[c]:self.off[c + 1]]

    def columns(self):
        return [self.z[self.off[c]:self.off[c + 1]] for c in range(self.ncols)]

    def num_segments(self) -> int:
        """(CompressedVolume.cpp:75-83): pairs per column, odd tail ignored."""
        return int(np.sum((self.off[1:] - self.off[:-1]) // 2))

    def get_volume(self) -> float:
        """(CompressedVolume.cpp:61-74)."""
        s = self.spacing ** 3
        total = 0.0
        lens = self.off[1:] - self.off[:-1]
        if np.all(lens % 2 == 0):
            return float(s * np.sum(self.z[1::2] - self.z[0::2]))
        for c in range(self.ncols):
            col = self.column(c)
            for k in range(0, len(col) - 1, 2):
                total += s * (col[k + 1] - col[k])
        return total

    def crop(self, x0, y0, w, h):
        """The w x h sub-grid starting at column (x0, y0), with this volume's
        geometry (its z-range, which erosion depends on, is unchanged)."""
        if not (0 <= x0 and 0 <= y0 and w >= 0 and h >= 0 and x0 + w <= self.nx and y0 + h <= self.ny):
            raise ValueError("crop outside the grid")
        xs = np.arange(x0, x0 + w)
        cols = (xs[None, :] + self.nx * np.arange(y0, y0 + h)[:, None]).ravel()
        lens = self.off[cols + 1] - self.off[cols]
        off = np.zeros(cols.size + 1, np.int64)
        np.cumsum(lens, out=off[1:])
        idx = np.repeat(self.off[cols] - off[:-1], lens) + np.arange(off[-1])
        return CompressedVolume(w, h, off, self.z[idx].copy(), self.origin, self.extent, self.spacing,
                                self.padding, dict(self.meta))

    # -- text format (CompressedVolume.cpp:116-152) -------------------------
    def save(self, path):
        with open(path, "w") as f:
            f.write("%r %r %r\n" % tuple(map(float, self.origin)))
            f.write("%r %r %r\n" % tuple(map(float, self.extent)))
            f.write(f"{self.nx} {self.ny}\n{self.padding}\n{float(self.spacing)!r}\n")
            for c in range(self.ncols):
                col = self.column(c)
                f.write(str(len(col)) + "".join(" " + repr(float(v)) for v in col) + "\n")

    @classmethod
    def load(cls, path):
        with open(path) as f:
            tok = f.read().split()
        it = iter(tok)
        origin = tuple(float(next(it)) for _ in range(3))
        extent = tuple(float(next(it)) for _ in range(3))
        nx, ny = int(next(it)), int(next(it))
        padding = int(next(it))
        spacing = float(next(it))
        cols = []
        for _ in range(nx * ny):
            n = int(next(it))
            cols.append([float(next(it)) for _ in range(n)])
        return cls.from_columns(nx, ny, cols, origin=origin, extent=extent, spacing=spacing,
                                padding=padding)

    # -- checks -----------------------------------------------------------
    def validate(self):
        """Columns even-length and sorted; raises ValueError otherwise."""
        lens = self.off[1:] - self.off[:-1]
        if np.any(lens % 2):
            raise ValueError("odd column length")
        if self.z.size > 1:
            d = np.diff(self.z)
            inner = np.ones(self.z.size - 1, bool)
            starts = self.off[1:-1]
            inner[starts[(starts > 0) & (starts < self.z.size)] - 1] = False
            if np.any(d[inner] < 0):
                raise ValueError("unsorted column")
