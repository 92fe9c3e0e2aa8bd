"""GPU parity: libvoroffset_hip.so (through its C ABI) against the CPU oracle.

Bar: FP64 storage -- segment counts identical and endpoints within 1e-9
dexel; FP32 storage -- against the oracle run on the same float32-rounded
input, counts identical except where only gaps narrower than 1e-5 dexel
differ, endpoints within 1e-4 dexel (the stored outputs are float32 relative
to the z-range centre, so the rounding alone is up to 3.1e-5).
"""

import numpy as np
import pytest

import oracle
from voroffset_amd import synth
from voroffset_amd.volume import CompressedVolume

from .helpers import compare_csr, narrow_gap_mismatches, quantize_fp32

pytestmark = pytest.mark.gpu

RANDOM_CASES = [
    # nx, ny, lam, zmax, r
    (24, 24, 3, 256, 1.0),
    (20, 31, 6, 512, 2.5),
    (33, 17, 4, 256, 5.0),
    (40, 40, 6, 1024, 16.0),
    (16, 64, 10, 512, 7.5),
    (64, 9, 2, 512, 3.0),
    (50, 50, 6, 1024, 12.25),
    # larger radii: two coverage words per lane with lanes walking their own
    # events (R <= 40), two words with every lane stepping through every
    # event, four words (R >= 64); windows taller than one 64-column tile
    (12, 150, 5, 512, 33.0),
    (10, 180, 6, 512, 47.5),
    (8, 220, 4, 512, 70.0),
]


def check_fp64(got, ref, tol=1e-9):
    bad, err = compare_csr(got[0], got[1], ref[0], ref[1])
    assert len(bad) == 0, f"count mismatch in columns {bad[:8]}"
    assert err <= tol, err


def check_fp32(got, ref, tol=1e-4):
    bad, err = compare_csr(got[0], got[1], ref[0], ref[1])
    assert narrow_gap_mismatches(got[0], got[1], ref[0], ref[1], bad), f"count mismatch in {bad[:8]}"
    assert err <= tol, err
    return len(bad), err


@pytest.mark.parametrize("case", RANDOM_CASES, ids=lambda c: "x".join(map(str, c)))
def test_dilation_fp64_random(hip_fp64, case):
    nx, ny, lam, zmax, r = case
    vol = synth.poisson_columns(nx, ny, lam=lam, zmax=zmax, seed=nx * 1000 + ny)
    res, t1, t2 = hip_fp64.dilation(vol, r)
    check_fp64((res.off, res.z), oracle.morph(vol, "dilation", r))


@pytest.mark.parametrize("case", RANDOM_CASES, ids=lambda c: "x".join(map(str, c)))
def test_dilation_fp32_random(hip_fp32, case):
    nx, ny, lam, zmax, r = case
    vol = synth.poisson_columns(nx, ny, lam=lam, zmax=zmax, seed=nx * 1000 + ny + 1)
    res, _, _ = hip_fp32.dilation(vol, r)
    check_fp32((res.off, res.z), oracle.morph(quantize_fp32(vol), "dilation", r))


@pytest.mark.parametrize("op", ["dilation", "erosion", "closing", "opening"])
@pytest.mark.parametrize("r", [2.0, 3.5, 6.0])
def test_ops_balls_fp64(hip_fp64, op, r):
    vol = synth.ball_union(10, 48, seed=int(r * 10), rmin=4, rmax=10, pad=9)
    res, _, _ = hip_fp64.apply(op, vol, r)
    check_fp64((res.off, res.z), oracle.morph(vol, op, r))


@pytest.mark.parametrize("op", ["dilation", "erosion", "closing", "opening"])
def test_ops_balls_fp32(hip_fp32, op):
    vol = synth.ball_union(12, 56, seed=3, rmin=4, rmax=11, pad=10)
    res, _, _ = hip_fp32.apply(op, vol, 4.0)
    check_fp32((res.off, res.z), oracle.morph(quantize_fp32(vol), op, 4.0))


@pytest.mark.parametrize("op", ["erosion", "closing", "opening"])
def test_ops_random_fp64(hip_fp64, op):
    vol = synth.poisson_columns(30, 26, lam=6, zmax=300, seed=9, z_margin=12)
    for r in (1.0, 1.5, 4.0):
        res, _, _ = hip_fp64.apply(op, vol, r)
        check_fp64((res.off, res.z), oracle.morph(vol, op, r))


def _vol(nx, ny, cols, zext=64.0):
    return CompressedVolume.from_columns(nx, ny, cols, origin=(0.0, 0.0, 0.0),
                                         extent=(float(nx), float(ny), zext), spacing=1.0, padding=0)


EDGE_CASES = {
    "all_empty": lambda: _vol(9, 7, [[] for _ in range(63)]),
    "single_column": lambda: _vol(1, 1, [[3.0, 9.0, 12.0, 20.0]]),
    "one_row": lambda: _vol(37, 1, [[float(i), float(i) + 2.5] if i % 3 else [] for i in range(37)]),
    "one_col": lambda: _vol(1, 37, [[float(i % 5), float(i % 5) + 1.0] for i in range(37)]),
    "points": lambda: _vol(11, 11, [[5.0 + (i % 4), 5.0 + (i % 4)] if i % 2 else [] for i in range(121)]),
    "touching": lambda: _vol(8, 8, [[1.0, 4.0, 4.0, 6.0, 6.0, 9.0] if i % 3 == 0 else [2.0, 3.0]
                                    for i in range(64)]),
    "many_segments": lambda: _vol(5, 5, [[float(k) for k in range(0, 400, 2)] if i == 12 else []
                                         for i in range(25)], zext=400.0),
}


@pytest.mark.parametrize("name", sorted(EDGE_CASES))
@pytest.mark.parametrize("r", [0.0, 0.5, 1.0, 2.0, 3.7, 33.0, 70.0])
def test_edge_cases_fp64(hip_fp64, name, r):
    vol = EDGE_CASES[name]()
    res, _, _ = hip_fp64.dilation(vol, r)
    check_fp64((res.off, res.z), oracle.morph(vol, "dilation", r))


def _dense_column_volume():
    # one column of 600 segments (length 0.5, gaps 1.0): more than 64 blocks of
    # pass-1 triples in that column, the case the block scans take in windows
    # or without skipping
    cols = [[] for _ in range(25)]
    cols[12] = [v for k in range(600) for v in (1.5 * k, 1.5 * k + 0.5)]
    return _vol(5, 5, cols, zext=900.0)


@pytest.mark.parametrize("r", [0.3, 40.0])
def test_dense_column_fp64(hip_fp64, r):
    vol = _dense_column_volume()
    res, _, _ = hip_fp64.dilation(vol, r)
    check_fp64((res.off, res.z), oracle.morph(vol, "dilation", r))


@pytest.mark.parametrize("r", [0.3, 40.0])
def test_dense_column_fp32(hip_fp32, r):
    vol = quantize_fp32(_dense_column_volume())
    res, _, _ = hip_fp32.dilation(vol, r)
    check_fp32((res.off, res.z), oracle.morph(vol, "dilation", r))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_dense_column_painter_large_lists(prec, monkeypatch):
    """The painter form of pass 1 (the form above R = 255; forced here) keeps
    its piece lists in LDS: a column of 600 segments overflows the default
    lists, and the dilation reruns pass 1 with the large-list form (one
    wavefront per CU) instead of failing.  (At R = 40 every column of this
    5 x 5 grid reaches the 600 segments, and the painter's per-chunk output
    arena in LDS is the limit: the z-sweep form, the default up to R = 255,
    takes that case, test_dense_column_fp64.)"""
    from voroffset_amd.morpho import VoronoiMorphoHip
    monkeypatch.setenv("VO_PASS1", "paint")
    hip = VoronoiMorphoHip(prec)
    monkeypatch.delenv("VO_PASS1")
    try:
        vol = _dense_column_volume()
        src = vol if prec == "fp64" else quantize_fp32(vol)
        for r in (0.3,):
            res, _, _ = hip.dilation(src, r)
            ref = oracle.morph(src, "dilation", r)
            (check_fp64 if prec == "fp64" else check_fp32)((res.off, res.z), ref)
    finally:
        hip.close()


def test_empty_grid(hip_fp64):
    vol = _vol(0, 0, [])
    res, _, _ = hip_fp64.dilation(vol, 2.0)
    assert res.nx == 0 and res.ny == 0 and res.z.size == 0


def test_rejects_unsorted_column(hip_fp64):
    from voroffset_amd._native import VoError
    vol = _vol(2, 1, [[5.0, 1.0], []])
    with pytest.raises(VoError):
        hip_fp64.dilation(vol, 1.0)


def test_rejects_negative_radius(hip_fp64):
    from voroffset_amd._native import VoError
    vol = _vol(2, 1, [[1.0, 5.0], []])
    with pytest.raises(VoError):
        hip_fp64.dilation(vol, -1.0)


def test_full_size_window_fp32(hip_fp32):
    """1024x1024, Poisson(6), r=16 (the direct-volume workload): the GPU's
    full-size result equals the oracle's on a window, since a column depends
    only on input within the radius."""
    r, W, x0, y0 = 16.0, 24, 500, 300
    vol = synth.poisson_columns(1024, 1024, lam=6, zmax=1024, seed=1234)
    res, _, _ = hip_fp32.dilation(vol, r)
    m = int(r)
    xs, ys = range(x0 - m, x0 + W + m), range(y0 - m, y0 + W + m)
    cols = [vol.at(x, y) for y in ys for x in xs]
    crop = quantize_fp32(CompressedVolume.from_columns(len(xs), len(ys), cols, origin=vol.origin,
                                                       extent=vol.extent, spacing=1.0, padding=0))
    ro, rz = oracle.morph(crop, "dilation", r, threads=4)
    got_cols, ref_cols = [], []
    for y in range(W):
        for x in range(W):
            got_cols.append(res.at(x0 + x, y0 + y))
            c = (x + m) + len(xs) * (y + m)
            ref_cols.append(rz[ro[c]:ro[c + 1]])
    g = CompressedVolume.from_columns(W, W, got_cols)
    f = CompressedVolume.from_columns(W, W, ref_cols)
    check_fp32((g.off, g.z), (f.off, f.z))
    # size-independent properties over the whole volume
    lens = np.diff(res.off)
    assert np.all(lens % 2 == 0)
    d = np.diff(res.z)
    starts = res.off[1:-1]
    inner = np.ones(res.z.size - 1, bool)
    inner[starts[(starts > 0) & (starts < res.z.size)] - 1] = False
    assert np.all(d[inner] > 0)  # sorted, disjoint, touching fused


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("op", ["dilation", "erosion", "closing", "opening"])
def test_sweep_axis_does_not_change_results(prec, op, monkeypatch):
    """Pass 1 along x (contiguous windows) or along y (the reference's
    axis): the separable decomposition is exact either way and every
    endpoint is the same a - h / b + h from the same table, so the results
    are identical bit for bit."""
    from voroffset_amd import _native
    vol = synth.poisson_columns(70, 45, lam=5, zmax=400, seed=17, z_margin=20)
    zlo, zhi = vol.z_range()
    out = []
    for axis in ("x", "y"):
        monkeypatch.setenv("VO_AXIS", axis)
        ctx = _native.Context(0)
        try:
            out.append(ctx.morph_host(op, prec, 6.5, vol.nx, vol.ny, vol.off, vol.z, zlo, zhi)[:2])
        finally:
            ctx.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    ref = oracle.morph(vol if prec == "fp64" else quantize_fp32(vol), op, 6.5)
    (check_fp64 if prec == "fp64" else check_fp32)(out[0], ref)
