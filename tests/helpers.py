"""Comparison helpers shared by the parity tests."""

import numpy as np


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
    """True when every count mismatch in `cols` is explained by gaps narrower than `gap`
    (closing such gaps in both columns makes them equal in count)."""
    def fuse(z):
        out = []
        for k in range(0, len(z), 2):
            if out and z[k] - out[-1] < gap:
                out[-1] = max(out[-1], z[k + 1])
            else:
                out.extend([z[k], z[k + 1]])
        return out
    for c in cols:
        if len(fuse(z_a[off_a[c]:off_a[c + 1]])) != len(fuse(z_b[off_b[c]:off_b[c + 1]])):
            return False
    return True


def quantize_fp32(vol):
    """The float64 values the FP32 device path sees: z rounded to float32
    relative to the integer centre of the z-range."""
    zlo, zhi = vol.z_range()
    shift = float(np.round(0.5 * (zlo + zhi)))
    zq = (vol.z - shift).astype(np.float32).astype(np.float64) + shift
    return vol.like(vol.off.copy(), zq)
