"""The comparison helpers the large-volume GPU tests rely on, checked against
per-column loops on small random volumes (CPU only)."""

import numpy as np

from voroffset_amd import synth

from .helpers import crop, csr_valid, inside, window_columns


def _inside_loop(inner, outer, tol=0.0):
    bad = 0
    for ci, co in zip(inner.columns(), outer.columns()):
        for k in range(0, len(ci), 2):
            a, b = ci[k] + tol, ci[k + 1] - tol
            if a > b:
                continue
            if not any(co[j] <= a and b <= co[j + 1] for j in range(0, len(co), 2)):
                bad += 1
    return bad


def test_crop_and_window_match_columns():
    vol = synth.poisson_columns(13, 9, lam=4, zmax=200, seed=3)
    c = crop(vol, 2, 3, 7, 5)
    for y in range(5):
        for x in range(7):
            assert np.array_equal(c.at(x, y), vol.at(2 + x, 3 + y))
    o, z = window_columns(vol.off, vol.z, vol.nx, 2, 3, 7, 5)
    assert np.array_equal(o, c.off) and np.array_equal(z, c.z)
    assert c.z_range() == vol.z_range()


def test_inside_agrees_with_loop():
    rng = np.random.default_rng(5)
    for seed in range(20):
        a = synth.poisson_columns(6, 5, lam=3, zmax=100, seed=seed)
        b = synth.poisson_columns(6, 5, lam=2, mean_len=60, zmax=100, seed=seed + 100)
        tol = float(rng.choice([0.0, 0.5]))
        assert inside(a.off, a.z, b.off, b.z, tol) == _inside_loop(a, b, tol)
        assert inside(a.off, a.z, a.off, a.z) == 0
    # boundaries: equal endpoints are inside, one ulp outside is not
    from voroffset_amd.volume import CompressedVolume
    outer = CompressedVolume.from_columns(2, 1, [[0.0, 1.0, 2.0, 3.0], [5.0, 6.0]])
    inner = CompressedVolume.from_columns(2, 1, [[0.0, 1.0, 2.5, 3.0], [5.0, np.nextafter(6.0, 7.0)]])
    assert inside(inner.off, inner.z, outer.off, outer.z) == 1
    across = CompressedVolume.from_columns(2, 1, [[0.5, 2.5], []])
    assert inside(across.off, across.z, outer.off, outer.z) == 1


def test_csr_valid():
    assert csr_valid(np.array([0, 4, 4]), np.array([0.0, 1.0, 2.0, 3.0]))
    assert not csr_valid(np.array([0, 4, 4]), np.array([0.0, 1.0, 1.0, 3.0]))   # touching, not fused
    assert not csr_valid(np.array([0, 3, 4]), np.array([0.0, 1.0, 2.0, 3.0]))   # odd column
    assert not csr_valid(np.array([0, 2, 4]), np.array([1.0, 0.0, 2.0, 3.0]))   # inverted
    assert csr_valid(np.array([0, 2, 4]), np.array([5.0, 6.0, 2.0, 3.0]))       # columns independent


def test_narrow_feature_mismatches():
    from .helpers import narrow_gap_mismatches
    off1, off2 = np.array([0, 2]), np.array([0, 4])
    one = np.array([0.0, 3.0])
    # a gap of 2e-6 in one column and none in the other: explained
    assert narrow_gap_mismatches(off1, one, off2, np.array([0.0, 1.0, 1.000002, 3.0]), [0])
    # a segment of width 0 or 1e-7 more in one column: explained
    assert narrow_gap_mismatches(np.array([0, 4]), np.array([0.0, 3.0, 5.0, 5.0]), off1, one, [0])
    assert narrow_gap_mismatches(np.array([0, 4]), np.array([0.0, 3.0, 5.0, 5.0000001]), off1, one, [0])
    # a gap or an extra segment of 1e-3: not explained
    assert not narrow_gap_mismatches(off1, one, off2, np.array([0.0, 1.0, 1.001, 3.0]), [0])
    assert not narrow_gap_mismatches(np.array([0, 4]), np.array([0.0, 3.0, 5.0, 5.001]), off1, one, [0])
