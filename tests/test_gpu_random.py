"""Seeded random volumes through the HIP path, both storage precisions.

Two families:

* regular columns (no zero-length segments, no touching segments): every
  operation against the oracle's `ours`, at grid sizes from 1x1 to 96x96
  (several 64-column tiles), radii 0.3 .. 40 (integer, half-integer and
  arbitrary), endpoints on a quarter-dexel grid or continuous;
* degenerate columns (zero-length segments, touching segments) against the
  oracle's brute force.  On these the reference's `ours` (as restated by the
  oracle) loses a zero-length segment that lies inside a farther line's
  segment in pass 1 (its merge keeps the larger-radius pieces by position
  and drops the point); the HIP path computes the exact morphology, which
  is what the reference's brute-force method computes
  (`test_point_inside_farther_segment` below is the smallest such case).
"""

import numpy as np
import pytest

import oracle
from voroffset_amd.volume import CompressedVolume

from .helpers import compare_csr, narrow_gap_mismatches, quantize_fp32, random_case

pytestmark = pytest.mark.gpu

def _check(hip, prec, op, vol, r, method):
    src = vol if prec == "fp64" else quantize_fp32(vol)
    got, _, _ = hip.apply(op, vol, r)
    ref = oracle.morph(src, op, r, method=method)
    bad, err = compare_csr(got.off, got.z, *ref)
    if prec == "fp64":
        assert len(bad) == 0 and err <= 1e-9, (op, r, vol.nx, vol.ny, bad[:5], err)
    else:
        assert narrow_gap_mismatches(got.off, got.z, ref[0], ref[1], bad) and err <= 1e-4, (op, r, bad[:5], err)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_random_regular_columns(prec, hip_fp64, hip_fp32):
    hip = hip_fp64 if prec == "fp64" else hip_fp32
    for seed in range(40):
        op, vol, r = random_case(1000 + seed, degenerate=False)
        _check(hip, prec, op, vol, r, "ours")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_random_degenerate_columns(prec, hip_fp64, hip_fp32):
    hip = hip_fp64 if prec == "fp64" else hip_fp32
    for seed in range(20):
        op, vol, r = random_case(2000 + seed, degenerate=True)
        _check(hip, prec, op, vol, r, "brute_force")


def test_point_inside_farther_segment(hip_fp64):
    """Line x = 1: [203.5, 223.5] at y = 0, a point at 209.25 at y = 1,
    [197.75, 227.75] at y = 2; R = 1.025.  The point reaches the columns
    (0, 1) and (2, 1) with half-length sqrt(R^2 - 1); the segments do not
    (their lines are at distance sqrt(2) > R).  Exact morphology (the HIP
    path, the oracle's brute force): [209.25 -+ 0.2216] in both columns; the
    oracle's `ours` drops the point in pass 1 and leaves them empty."""
    cols = [[] for _ in range(9)]
    cols[1], cols[4], cols[7] = [203.5, 223.5], [209.25, 209.25], [197.75, 227.75]
    vol = CompressedVolume.from_columns(3, 3, cols, origin=(0.0, 0.0, 0.0), extent=(3.0, 3.0, 285.0),
                                        spacing=1.0, padding=0)
    r = 1.025
    got, _, _ = hip_fp64.apply("dilation", vol, r)
    h = np.sqrt(r * r - 1.0)
    for c in (3, 5):
        assert np.allclose(got.z[got.off[c]:got.off[c + 1]], [209.25 - h, 209.25 + h], atol=1e-12), c
    brute = oracle.morph(vol, "dilation", r, method="brute_force")
    bad, err = compare_csr(got.off, got.z, *brute)
    assert len(bad) == 0 and err <= 1e-12
    ours = oracle.morph(vol, "dilation", r)
    assert ours[0][4] == ours[0][3] and ours[0][6] == ours[0][5]  # columns 3 and 5 empty in `ours`


def _large_radius_volume(op, r):
    from voroffset_amd import synth
    if op == "dilation":
        return synth.poisson_columns(37, 45, lam=5, zmax=900, seed=int(r), z_margin=int(r) + 4)
    if op == "closing":  # the erosion step empties everything within R of the grid's border
        return synth.poisson_columns(280, 270, lam=1, zmax=900, seed=7, z_margin=int(r) + 4)
    n = 330  # erosion of a box: columns farther than R from the border keep [100 + R, 900 - R]
    return CompressedVolume.from_columns(n, n, [[100.0, 900.0]] * (n * n), origin=(0.0, 0.0, 0.0),
                                         extent=(float(n), float(n), 1000.0), spacing=1.0, padding=0)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("op,r", [("dilation", 140.5), ("dilation", 300.0), ("closing", 131.0), ("erosion", 150.0)])
def test_radii_beyond_the_z_sweep(prec, op, r, hip_fp64, hip_fp32):
    """R > 127: pass 1 runs as the painter form, pass 2 still as the gather
    (whose lists live in global memory; the sweep form's LDS lists overflowed
    on such volumes)."""
    hip = hip_fp64 if prec == "fp64" else hip_fp32
    vol = _large_radius_volume(op, r)
    src = vol if prec == "fp64" else quantize_fp32(vol)
    got, _, _ = hip.apply(op, vol, r)
    ref = oracle.morph(src, op, r, threads=8)
    assert ref[0][-1] > 0
    bad, err = compare_csr(got.off, got.z, *ref)
    if prec == "fp64":
        assert len(bad) == 0 and err <= 1e-9, (op, r, bad[:5], err)
    else:
        assert narrow_gap_mismatches(got.off, got.z, ref[0], ref[1], bad) and err <= 1e-4, (op, r, bad[:5], err)


def test_repeated_dense_calls_keep_their_workspace_bounded(hip_fp64):
    """Dense windows at large radii make pass 1 retry with more room per
    column; the room grows only where a form overflowed, so a context that
    runs many such calls in a row keeps a bounded workspace (the sweep
    form's pass-2 arena once grew with every retry and ran the device out of
    memory after a few calls)."""
    from voroffset_amd import synth
    from .helpers import csr_valid
    last = None
    for lam in (6, 12, 20, 30):
        for r in (260.0, 300.0, 400.0):
            vol = synth.poisson_columns(48, 40, lam=lam, zmax=1500, seed=int(lam * 10 + r), z_margin=int(r) + 4)
            got, _, _ = hip_fp64.apply("dilation", vol, r)
            assert csr_valid(got.off, got.z), (lam, r)
            last = (vol, r, got)
    vol, r, got = last
    bad, err = compare_csr(got.off, got.z, *oracle.morph(vol, "dilation", r, threads=8))
    assert len(bad) == 0 and err <= 1e-9, (bad[:5], err)
