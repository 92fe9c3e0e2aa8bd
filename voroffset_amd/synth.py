"""Seeded synthetic dexel volumes for parity tests and the throughput bench.

* ``poisson_columns``: a direct dexel volume; each column holds Poisson(lam)
  segments whose gaps and lengths are exponential (mean ``mean_len`` dexels),
  endpoints sorted and clipped to ``[0, zmax]``.  With 1024x1024, lam=6,
  mean 40, zmax 1024 this is the direct-volume throughput workload.
* ``ball_union``: an analytic union of random balls, rasterised exactly per
  dexel centre (the closed form the rasterizer is checked against).
* ``batch_specs``: the uneven-volume batch (side log-uniform, lam, radius and
  op drawn per volume).
"""

from __future__ import annotations

import math

import numpy as np

from .volume import CompressedVolume

OPS = ("dilation", "erosion", "opening", "closing")


def poisson_columns(nx, ny, lam=6.0, mean_len=40.0, zmax=1024.0, seed=1234,
                    z_margin=None, lam_per_column=None):
    """Columns of Poisson(lam) segments with Exp(mean_len) gaps and lengths.

    The volume's z-range is [-z_margin, zmax + z_margin] (origin_z =
    -z_margin, extent_z = zmax + 2*z_margin, padding 0), so that erosion and
    closing by radii below ``z_margin - 1`` keep every endpoint inside the
    complement range.  Returns a CompressedVolume with float64 endpoints.
    """
    rng = np.random.default_rng(seed)
    n = nx * ny
    lam_arr = lam if lam_per_column is None else lam_per_column
    k = rng.poisson(lam_arr, size=n).astype(np.int64)
    total = int(k.sum())
    gaps = rng.exponential(mean_len, size=total)
    lens = rng.exponential(mean_len, size=total)
    # z runs from 0 in every column: cumulative sum of gap + length, restarted per column
    col_of = np.repeat(np.arange(n), k)
    csum = np.cumsum(gaps + lens)
    first = np.zeros(n + 1, np.int64)
    np.cumsum(k, out=first[1:])
    base = np.concatenate([[0.0], csum])[first[:-1]][col_of]
    end = csum - base
    start = end - lens
    keep = start < zmax
    start, end, col_of = start[keep], np.minimum(end[keep], zmax), col_of[keep]
    cnt = np.bincount(col_of, minlength=n).astype(np.int64)
    off = np.zeros(n + 1, np.int64)
    np.cumsum(2 * cnt, out=off[1:])
    z = np.empty(2 * start.size)
    z[0::2] = start
    z[1::2] = end
    margin = 80.0 if z_margin is None else float(z_margin)
    vol = CompressedVolume(nx, ny, off, z, origin=(0.0, 0.0, -margin),
                           extent=(float(nx), float(ny), zmax + 2 * margin), spacing=1.0,
                           padding=0)
    vol.meta.update(kind="poisson", lam=lam, mean_len=mean_len, zmax=zmax, seed=seed)
    return vol


def ball_union(n_balls, grid, seed=7, rmin=3.0, rmax=12.0, pad=None):
    """Union of random balls, rasterised exactly at dexel centres.

    Grid ``grid x grid``, z-extent ``grid``; balls lie inside the box with a
    margin so that erosion/closing up to ``pad - 1`` stays in range.
    """
    rng = np.random.default_rng(seed)
    pad = grid // 8 if pad is None else pad
    inner = grid - 2 * pad
    cx = rng.uniform(pad + rmax, pad + inner - rmax, n_balls)
    cy = rng.uniform(pad + rmax, pad + inner - rmax, n_balls)
    cz = rng.uniform(pad + rmax, pad + inner - rmax, n_balls)
    rr = rng.uniform(rmin, rmax, n_balls)
    cols = []
    for y in range(grid):
        for x in range(grid):
            px, py = x + 0.5, y + 0.5
            d2 = (cx - px) ** 2 + (cy - py) ** 2
            hit = d2 < rr ** 2
            segs = sorted((cz[i] - math.sqrt(rr[i] ** 2 - d2[i]), cz[i] + math.sqrt(rr[i] ** 2 - d2[i]))
                          for i in np.nonzero(hit)[0])
            merged = []
            for a, b in segs:
                if merged and a <= merged[-1]:
                    merged[-1] = max(merged[-1], b)
                else:
                    merged.extend([a, b])
            cols.append(merged)
    vol = CompressedVolume.from_columns(grid, grid, cols, origin=(0.0, 0.0, 0.0),
                                        extent=(float(grid), float(grid), float(grid)),
                                        spacing=1.0, padding=0)
    vol.meta.update(kind="balls", n_balls=n_balls, seed=seed)
    return vol


def batch_specs(n_volumes=256, seed=4321, side_range=(64, 512), lam_range=(2.0, 10.0),
                r_range=(2.0, 32.0)):
    """Per-volume (side, lam, radius, op, seed) of the uneven batch."""
    rng = np.random.default_rng(seed)
    sides = np.exp(rng.uniform(math.log(side_range[0]), math.log(side_range[1]), n_volumes))
    sides = np.round(sides).astype(int)
    lams = rng.uniform(*lam_range, n_volumes)
    radii = rng.uniform(*r_range, n_volumes)
    ops = rng.integers(0, 4, n_volumes)
    seeds = rng.integers(0, 2 ** 31 - 1, n_volumes)
    return [dict(side=int(s), lam=float(l), radius=float(r), op=OPS[int(o)], seed=int(sd))
            for s, l, r, o, sd in zip(sides, lams, radii, ops, seeds)]


def to_float32_centered(vol):
    """FP32 copy of the endpoints (values as stored on the device, relative
    to nothing: absolute dexel units rounded to float32) and the FP64 shadow
    made of exactly those values."""
    z32 = vol.z.astype(np.float32)
    return vol.like(vol.off.copy(), z32.astype(np.float64))
