"""Comparison helpers shared by the parity tests."""

import numpy as np

from voroffset_amd._native import storage_shift


def compare_csr(off_a, z_a, off_b, z_b):
    """(columns whose segment counts differ, max endpoint difference over the others)."""
    la, lb = np.diff(off_a), np.diff(off_b)
    bad = np.nonzero(la != lb)[0]
    same = la == lb
    if not np.any(same & (la > 0)):
        return bad, 0.0
    ia = np.concatenate([np.arange(off_a[c], off_a[c + 1]) for c in np.nonzero(same & (la > 0))[0]])
    ib = np.concatenate([np.arange(off_b[c], off_b[c + 1]) for c in np.nonzero(same & (lb > 0))[0]])
    return bad, float(np.max(np.abs(z_a[ia] - z_b[ib])))


def narrow_gap_mismatches(off_a, z_a, off_b, z_b, cols, gap=1e-5):
    """True when every count mismatch in `cols` is explained by features
    narrower than `gap`: closing such gaps and then removing such segments
    in both columns makes them equal in count.  FP32 storage cannot resolve
    either: a gap below its rounding step is stored as width 0 (and kept),
    and the complement of such a gap, in an erosion or closing, is a segment
    of width 0 or one step (the exact result has a gap or a point there, or
    nothing)."""
    def fuse(z):
        out = []
        for k in range(0, len(z), 2):
            if out and z[k] - out[-1] < gap:
                out[-1] = max(out[-1], z[k + 1])
            else:
                out.extend([z[k], z[k + 1]])
        return [v for k in range(0, len(out), 2) if out[k + 1] - out[k] >= gap for v in out[k:k + 2]]
    for c in cols:
        if len(fuse(z_a[off_a[c]:off_a[c + 1]])) != len(fuse(z_b[off_b[c]:off_b[c + 1]])):
            return False
    return True


def quantize_fp32(vol):
    """The float64 values the FP32 device path sees: z rounded to float32
    relative to the integer centre of the z-range."""
    zlo, zhi = vol.z_range()
    shift = storage_shift(zlo, zhi)
    zq = (vol.z - shift).astype(np.float32).astype(np.float64) + shift
    return vol.like(vol.off.copy(), zq)


def crop(vol, x0, y0, w, h):
    """The w x h sub-grid of `vol` starting at (x0, y0), with vol's geometry
    (the z-range, which erosion depends on, is unchanged)."""
    from voroffset_amd.volume import CompressedVolume
    xs = np.arange(x0, x0 + w)
    cols = (xs[None, :] + vol.nx * np.arange(y0, y0 + h)[:, None]).ravel()
    lens = vol.off[cols + 1] - vol.off[cols]
    off = np.zeros(cols.size + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    idx = np.repeat(vol.off[cols] - off[:-1], lens) + np.arange(off[-1])
    return CompressedVolume(w, h, off, vol.z[idx].copy(), vol.origin, vol.extent, vol.spacing, vol.padding,
                            dict(vol.meta))


def csr_valid(off, z, zero_gaps=False):
    """Every column even-length, each segment non-inverted, consecutive
    segments of a column separated by a positive gap (touching ones fused).
    zero_gaps=True also accepts gaps of width exactly 0: float32 storage
    rounds a gap narrower than the rounding step to that, and keeps the two
    segments (as the FP64 result has them)."""
    lens = np.diff(off)
    if np.any(lens % 2) or off[-1] != z.size:
        return False
    if z.size == 0:
        return True
    a, b = z[0::2], z[1::2]
    if np.any(b < a) or not np.all(np.isfinite(z)):
        return False
    seg_col = np.repeat(np.arange(lens.size), lens // 2)
    same = seg_col[1:] == seg_col[:-1]
    gap = a[1:][same] - b[:-1][same]
    return bool(np.all(gap >= 0) if zero_gaps else np.all(gap > 0))


def zero_gaps(off, z):
    """Number of gaps of width exactly 0 between consecutive segments of a column."""
    lens = np.diff(off)
    if z.size == 0:
        return 0
    seg_col = np.repeat(np.arange(lens.size), lens // 2)
    same = seg_col[1:] == seg_col[:-1]
    return int(np.count_nonzero(z[2::2][same] == z[1:-1:2][same]))


def inside(inner_off, inner_z, outer_off, outer_z, tol=0.0):
    """(number of segments of `inner` not inside one segment of `outer` in the
    same column, after shrinking each inner segment by `tol` at both ends).
    Exact comparisons: a merge of the two sorted CSR lists, no key arithmetic."""
    n = inner_off.size - 1
    ci = np.repeat(np.arange(n), np.diff(inner_off) // 2)
    co = np.repeat(np.arange(n), np.diff(outer_off) // 2)
    a, b = inner_z[0::2] + tol, inner_z[1::2] - tol
    keep = a <= b
    a, b, ci = a[keep], b[keep], ci[keep]
    if a.size == 0:
        return 0
    s, e = outer_z[0::2], outer_z[1::2]
    col = np.concatenate([co, ci])
    zz = np.concatenate([s, a])
    kind = np.concatenate([np.zeros(s.size, np.int8), np.ones(a.size, np.int8)])
    order = np.lexsort((kind, zz, col))
    # outer segments are in (column, z) order, so the index of the last outer
    # start seen is a running maximum
    tag = np.where(order < s.size, order, -1)
    last = np.maximum.accumulate(tag)
    at_inner = order >= s.size
    j = last[at_inner]
    ii = order[at_inner] - s.size
    jj = np.maximum(j, 0)
    ok = (j >= 0) & (co[jj] == ci[ii]) & (e[jj] >= b[ii]) if s.size else np.zeros(ii.size, bool)
    return int(np.count_nonzero(~ok))


def window_columns(off, z, nx, x0, y0, w, h):
    """(off, z) of the w x h window at (x0, y0) of a CSR volume with nx columns per row."""
    xs = np.arange(x0, x0 + w)
    cols = (xs[None, :] + nx * np.arange(y0, y0 + h)[:, None]).ravel()
    lens = off[cols + 1] - off[cols]
    o = np.zeros(cols.size + 1, np.int64)
    np.cumsum(lens, out=o[1:])
    idx = np.repeat(off[cols] - o[:-1], lens) + np.arange(o[-1])
    return o, z[idx]


def reach(op, r):
    """Columns from the crop edge beyond which a cropped input gives the full
    volume's result: the operation reads input within r of a column (twice
    for closing and opening); the ring of an erosion sits one column outside."""
    k = int(np.floor(r))
    return 2 * k + 1 if op in ("closing", "opening") else k + 1


OPS = ["dilation", "erosion", "closing", "opening"]


def random_case(seed, degenerate):
    """(op, volume, radius) of a seeded random case: `degenerate` columns may
    hold zero-length and touching segments, the others neither."""
    from voroffset_amd.volume import CompressedVolume

    rng = np.random.default_rng(seed)
    if degenerate:
        nx, ny = int(rng.integers(1, 33)), int(rng.integers(1, 33))
        r = float(np.exp(rng.uniform(np.log(0.3), np.log(12))))
    else:
        nx, ny = int(rng.integers(1, 97)), int(rng.integers(1, 97))
        kind = rng.random()
        r = (float(rng.integers(1, 41)) if kind < 0.3 else float(rng.integers(0, 40)) + 0.5 if kind < 0.5
             else float(np.exp(rng.uniform(np.log(0.3), np.log(40)))))
    op = OPS[int(rng.integers(0, 4))]
    zext = float(rng.integers(150, 500))
    margin = int(np.ceil(r)) + 3 if op in ("dilation", "closing") else 2
    lam = float(rng.uniform(0, 10))
    grid = degenerate or rng.random() < 0.5
    cols = []
    for _ in range(nx * ny):
        v = rng.uniform(margin, zext - margin, 2 * int(rng.poisson(lam)))
        if grid:
            v = np.round(v * 4) / 4
        v.sort()
        if not degenerate:  # drop zero-length segments, fuse touching ones
            segs = []
            for a, b in zip(v[0::2], v[1::2]):
                if a >= b:
                    continue
                if segs and a <= segs[-1][1]:
                    segs[-1][1] = max(segs[-1][1], b)
                else:
                    segs.append([a, b])
            v = [x for s in segs for x in s]
        cols.append(list(v))
    vol = CompressedVolume.from_columns(nx, ny, cols, origin=(0.0, 0.0, 0.0), extent=(float(nx), float(ny), zext),
                                        spacing=1.0, padding=0)
    return op, vol, r
