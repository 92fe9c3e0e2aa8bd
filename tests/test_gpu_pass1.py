"""The pass-1 intermediate on its own (vo_pass1), against the input and the
oracle's pass 1.

The reference's pass 1 builds, per column, z-disjoint triples (y1, y2, r)
from the half-space Voronoi sweeps of the x-line (src/vor3d/Voronoi2D.cpp:
699-739 emits them, MorphologyOperators.cpp:37-95 merges forward and
backward).  What pass 2 consumes is their envelope
    env(z) = max over triples (r^2 - dist_z(z, [y1, y2])^2),
which must equal the envelope of the input within reach,
    max over segments [a, b] of lines y' with |y - y'| <= R of
        R^2 - (y - y')^2 - dist_z(z, [a, b])^2
(both are the squared dilation height at z).  The GPU
keeps every piece of "nearest covering line" and the reference drops pieces
its sweep finds dominated, so the triples and their counts differ while the
envelopes agree.  Checked here per column at every endpoint, between
endpoints and beyond the ends, for R in {1, 4.5, 16, 33, 47.5, 70} in both
storage precisions; plus the triples' own invariants (sorted, disjoint,
r = sqrt(R^2 - dy^2) for an integer dy <= R) and the counts m_gpu beside
m_oracle.  FP32: the triples' endpoints are copies of the float32-rounded
input endpoints, so the identity holds for that input with no extra error.
"""

import numpy as np
import pytest

import oracle
from voroffset_amd import synth

from .helpers import crop, quantize_fp32

pytestmark = pytest.mark.gpu


def _env(z, lo, hi, r2):
    """max_k (r2_k - dist(z, [lo_k, hi_k])^2) for every z (-inf without pieces)."""
    if lo.size == 0:
        return np.full(z.shape, -np.inf)
    d = np.maximum(np.maximum(lo[None, :] - z[:, None], z[:, None] - hi[None, :]), 0.0)
    return np.max(r2[None, :] - d * d, axis=1)


def _samples(*arrays, pad):
    pts = np.unique(np.concatenate([a for a in arrays if a.size] or [np.zeros(1)]))
    mids = 0.5 * (pts[1:] + pts[:-1])
    return np.concatenate([pts, mids, pts - 1e-3, pts + 1e-3, [pts[0] - pad, pts[-1] + pad]])


def check_column_envelopes(vol, R, t_off, y1, y2, r, cols):
    """Envelope identity for the columns `cols` (indices x + nx*y); returns
    the largest difference seen."""
    Ri = int(np.floor(R))
    worst = 0.0
    for c in cols:
        x, y = c % vol.nx, c // vol.nx
        s, e = t_off[c], t_off[c + 1]
        tl, th, tr = y1[s:e], y2[s:e], r[s:e]
        # triples: sorted, disjoint (pieces may touch), radii from integer offsets
        assert np.all(tl <= th) and np.all(th[:-1] <= tl[1:]), (x, y)
        dy2 = R * R - tr * tr
        assert np.all(np.abs(dy2 - np.round(dy2)) <= 1e-6 * max(1.0, R * R)), (x, y)
        assert np.all((np.round(dy2) >= 0) & (np.round(dy2) <= Ri * Ri)), (x, y)
        lo, hi, w = [], [], []
        for yy in range(max(0, y - Ri), min(vol.ny, y + Ri + 1)):
            col = vol.at(x, yy)
            lo.append(col[0::2])
            hi.append(col[1::2])
            w.append(np.full(col.size // 2, R * R - (yy - y) ** 2))
        lo, hi, w = np.concatenate(lo), np.concatenate(hi), np.concatenate(w)
        z = _samples(tl, th, lo, hi, pad=R + 2)
        a, b = _env(z, tl, th, tr * tr), _env(z, lo, hi, w)
        assert np.array_equal(np.isinf(a), np.isinf(b)), (x, y)
        fin = np.isfinite(a)
        if np.any(fin):
            diff = float(np.max(np.abs(a[fin] - b[fin])))
            assert diff <= 1e-9 * (1.0 + R * R), (x, y, diff)
            worst = max(worst, diff)
    return worst


RADII = [1.0, 4.5, 16.0, 33.0, 47.5, 70.0]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("R", RADII)
def test_pass1_envelope_identity(hip_fp64, hip_fp32, prec, R):
    hip = hip_fp64 if prec == "fp64" else hip_fp32
    vol = synth.poisson_columns(22, 150, lam=6, zmax=512, seed=int(R * 10) + 3)
    seen = vol if prec == "fp64" else quantize_fp32(vol)
    zlo, zhi = vol.z_range()
    t_off, y1, y2, r = hip.ctx.pass1_host(prec, R, vol.nx, vol.ny, vol.off, vol.z, zlo, zhi)
    assert t_off.size == vol.ncols + 1 and t_off[-1] == y1.size
    rng = np.random.default_rng(int(R))
    cols = rng.choice(vol.ncols, 300, replace=False)
    check_column_envelopes(seen, R, t_off, y1, y2, r, cols)
    # counts beside the reference's culled pass 1 (oracle), and the two
    # intermediates' envelopes against each other on the same columns
    o_off, o1, o2, orr = oracle.pass1(seen, R, threads=4)
    m_gpu, m_ref = y1.size / vol.ncols, o1.size / vol.ncols
    print(f"R={R} {prec}: m_gpu={m_gpu:.2f} m_oracle={m_ref:.2f}")
    assert 0.8 * m_ref <= m_gpu <= 1.5 * m_ref
    for c in cols[:60]:
        zs = _samples(y1[t_off[c]:t_off[c + 1]], o1[o_off[c]:o_off[c + 1]], pad=R + 2)
        a = _env(zs, y1[t_off[c]:t_off[c + 1]], y2[t_off[c]:t_off[c + 1]], r[t_off[c]:t_off[c + 1]] ** 2)
        b = _env(zs, o1[o_off[c]:o_off[c + 1]], o2[o_off[c]:o_off[c + 1]], orr[o_off[c]:o_off[c + 1]] ** 2)
        fin = np.isfinite(a)
        assert np.array_equal(fin, np.isfinite(b))
        assert np.all(np.abs(a[fin] - b[fin]) <= 1e-9 * (1.0 + R * R))


def test_pass1_full_size_counts(hip_fp32):
    """1024x1024, Poisson(6), R=16 (the direct-volume workload): the envelope
    identity on a 16x16 window, and the mean counts m_gpu and m_oracle."""
    R = 16.0
    vol = synth.poisson_columns(1024, 1024, lam=6, zmax=1024, seed=1234)
    seen = quantize_fp32(vol)
    zlo, zhi = vol.z_range()
    t_off, y1, y2, r = hip_fp32.ctx.pass1_host("fp32", R, vol.nx, vol.ny, vol.off, vol.z, zlo, zhi)
    cols = [(300 + j) * vol.nx + 500 + i for j in range(16) for i in range(16)]
    check_column_envelopes(seen, R, t_off, y1, y2, r, cols)
    o_off = oracle.pass1(seen, R, threads=16)[0]
    m_gpu, m_ref = y1.size / vol.ncols, (o_off[-1]) / vol.ncols
    print(f"1024^2 R=16: m_gpu={m_gpu:.3f} m_oracle={m_ref:.3f}")
    assert 0.8 * m_ref <= m_gpu <= 1.5 * m_ref
    # a crop's intermediate equals the full volume's away from the crop edge
    # in y (pass 1 reads lines within R along y only)
    c = crop(vol, 480, 280, 64, 64)
    c_off, c1, c2, cr = hip_fp32.ctx.pass1_host("fp32", R, c.nx, c.ny, c.off, c.z, zlo, zhi)
    for j in range(16, 48):
        for i in range(64):
            a = (280 + j) * vol.nx + 480 + i
            b = j * 64 + i
            assert np.array_equal(y1[t_off[a]:t_off[a + 1]], c1[c_off[b]:c_off[b + 1]])
            assert np.array_equal(r[t_off[a]:t_off[a + 1]], cr[c_off[b]:c_off[b + 1]])
