"""Parity metrics (calculateXor, get_volume, numSegments): the oracle pinned
by set identities and pointwise membership, the device kernels against it
(the xor composition moves endpoints only, so it is bit-exact)."""

import numpy as np
import pytest

import oracle
from voroffset_amd import synth
from voroffset_amd.volume import CompressedVolume


def _vol(nx, ny, cols, zext=64.0):
    return CompressedVolume.from_columns(nx, ny, cols, origin=(0.0, 0.0, 0.0),
                                         extent=(float(nx), float(ny), zext), spacing=1.0, padding=0)


def _cols(off, z):
    return [z[off[c]:off[c + 1]] for c in range(len(off) - 1)]


def _inside(col, t):
    return any(col[i] < t < col[i + 1] for i in range(0, len(col) - 1, 2))


def _intersection(a, b):
    tot, i, j = 0.0, 0, 0
    while i + 1 < len(a) and j + 1 < len(b):
        lo, hi = max(a[i], b[j]), min(a[i + 1], b[j + 1])
        tot += max(0.0, hi - lo)
        if a[i + 1] < b[j + 1]:
            i += 2
        else:
            j += 2
    return tot


def test_xor_identities():
    a = synth.poisson_columns(12, 9, lam=4, zmax=60, seed=3, z_margin=2)
    b = synth.poisson_columns(12, 9, lam=4, zmax=60, seed=4, z_margin=2)
    empty = a.like(np.zeros_like(a.off), np.zeros(0))
    off, z = oracle.xor(a, a)
    assert z.size == 0
    off, z = oracle.xor(a, empty)
    assert np.array_equal(z, a.z) and np.array_equal(off, a.off)
    off, z = oracle.xor(a, b)
    m_ab, _ = oracle.measure(off, z, 1.0)
    m_a, _ = oracle.measure(a.off, a.z, 1.0)
    m_b, _ = oracle.measure(b.off, b.z, 1.0)
    inter = sum(_intersection(a.column(c), b.column(c)) for c in range(a.ncols))
    assert m_ab == pytest.approx(m_a + m_b - 2 * inter, abs=1e-9)
    # pointwise: inside the xor iff inside exactly one of a, b
    rng = np.random.default_rng(0)
    cx = _cols(off, z)
    for c in range(a.ncols):
        for t in rng.uniform(-2, 62, 40):
            assert _inside(cx[c], t) == (_inside(a.column(c), t) != _inside(b.column(c), t))


def test_xor_touching_and_points():
    a = _vol(3, 1, [[0.0, 5.0], [2.0, 2.0, 4.0, 9.0], [1.0, 3.0]])
    b = _vol(3, 1, [[5.0, 10.0], [], [1.0, 3.0 + 5e-11]])
    off, z = oracle.xor(a, b)
    cols = _cols(off, z)
    assert list(cols[0]) == [0.0, 10.0]          # A and B touch at 5: the closure fuses
    assert list(cols[1]) == [4.0, 9.0]           # the isolated point vanishes
    assert list(cols[2]) == []                   # the 5e-11 sliver is dropped


def test_measure_counts_pairs():
    v = _vol(2, 1, [[0.0, 2.0, 3.0, 7.5], [1.0, 4.0, 6.0]], zext=10.0)  # odd tail ignored
    m, s = oracle.measure(v.off, v.z, 2.0)
    assert s == 3 and m == pytest.approx(8 * (2.0 + 4.5 + 3.0))


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_hip_xor_and_measure(prec, hip_fp64, hip_fp32):
    hip = hip_fp64 if prec == "fp64" else hip_fp32
    for seed in range(3):
        a = synth.poisson_columns(40, 33, lam=5, zmax=300, seed=seed, z_margin=3)
        b = synth.poisson_columns(40, 33, lam=5, zmax=300, seed=seed + 50, z_margin=3)
        if prec == "fp32":
            from .helpers import quantize_fp32
            a, b = quantize_fp32(a), quantize_fp32(b)
        vol, res = hip.calculate_xor(a, b)
        off, z = oracle.xor(a, b)
        assert np.array_equal(res.off, off)
        assert np.array_equal(res.z, z) if prec == "fp64" else np.max(np.abs(res.z - z), initial=0) <= 1e-4
        m, s = oracle.measure(off, z, a.spacing)
        assert vol == pytest.approx(m, rel=1e-9, abs=1e-9)
        gm, gs = hip.get_volume(a)
        om, os_ = oracle.measure(a.off, a.z, a.spacing)
        assert gs == os_ and gm == pytest.approx(om, rel=1e-9)


@pytest.mark.gpu
def test_hip_xor_of_morphology_against_oracle_is_empty(hip_fp64):
    vol = synth.ball_union(8, 40, seed=5, rmin=4, rmax=9, pad=8)
    res, _, _ = hip_fp64.closing(vol, 3.0)
    ro, rz = oracle.morph(vol, "closing", 3.0)
    m, x = hip_fp64.calculate_xor(res, res.like(ro, rz))
    assert x.z.size == 0 and m == 0.0


def _hex_reference(vol, ox, oy, s):
    """dexel_dump's corners (Dexelize.cpp:318-340) in numpy: min + b * (max - min)."""
    verts = []
    for y in range(vol.ny):
        for x in range(vol.nx):
            col = vol.at(x, y)
            for i in range(0, len(col) - 1, 2):
                lo = np.array([ox + x * s, oy + y * s, col[i] * s])
                hi = np.array([ox + (x + 1) * s, oy + (y + 1) * s, col[i + 1] * s])
                for b in ((0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1), (0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0)):
                    verts.append([lo[d] + b[d] * (hi[d] - lo[d]) for d in range(3)])
    return np.array(verts).reshape(-1, 3)


@pytest.mark.gpu
def test_hip_dexel_mesh_matches_reference_formula():
    from voroffset_amd import _native
    ctx = _native.Context(0)
    try:
        vol = synth.poisson_columns(9, 7, lam=3, zmax=50, seed=8)
        ox, oy, s = -1.3, 0.7, 0.37
        vert, quad = ctx.dexel_mesh(vol.nx, vol.ny, vol.off, vol.z, ox, oy, s)
        assert np.array_equal(vert, _hex_reference(vol, ox, oy, s))
        nseg = vol.num_segments()
        assert quad.shape == (6 * nseg, 4) and quad.max() == 8 * nseg - 1
        # every quad is planar and its normal points away from its box's centre
        v = vert[quad]
        centre = vert.reshape(-1, 8, 3).mean(axis=1).repeat(6, axis=0)
        normal = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
        assert np.all(np.einsum("ij,ij->i", normal, v[:, 0] - centre) > 0)
        empty_v, empty_q = ctx.dexel_mesh(0, 0, np.zeros(1, np.int64), np.zeros(0), 0.0, 0.0, 1.0)
        assert empty_v.shape == (0, 3) and empty_q.shape == (0, 4)
    finally:
        ctx.close()
