"""Host-side tests that need no GPU: the container, the generators and the
native library's exported interface."""

import ctypes
import os
import re

import numpy as np
import pytest

from voroffset_amd import synth
from voroffset_amd.volume import CompressedVolume

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_volume_roundtrip(tmp_path):
    vol = synth.poisson_columns(7, 5, lam=3, zmax=100, seed=2)
    p = tmp_path / "v.txt"
    vol.save(p)
    back = CompressedVolume.load(p)
    assert back.nx == 7 and back.ny == 5
    np.testing.assert_array_equal(back.off, vol.off)
    np.testing.assert_array_equal(back.z, vol.z)
    assert back.num_segments() == vol.num_segments()
    assert abs(back.get_volume() - vol.get_volume()) < 1e-9


def test_geometry_matches_reference_constructor():
    # CompressedVolume.cpp:11-23: origin shifted by padding, grid = ceil(extent/s) + 2p
    v = CompressedVolume.from_bbox((1.0, 2.0, 3.0), (10.0, 4.2, 7.0), 0.5, 3)
    assert (v.nx, v.ny) == (26, 15)
    assert v.origin == (-0.5, 0.5, 1.5)
    zlo, zhi = v.z_range()
    assert zlo == 3.0 and zhi == 3.0 + 6 + 14.0


def test_poisson_generator_statistics():
    vol = synth.poisson_columns(64, 64, lam=6, zmax=1024, seed=1)
    vol.validate()
    s = vol.num_segments() / vol.ncols
    assert 5.7 < s < 6.2
    lens = vol.z[1::2] - vol.z[0::2]
    assert np.all(lens >= 0) and 35 < lens.mean() < 45
    assert vol.z.min() >= 0 and vol.z.max() <= 1024


def test_validate_rejects_bad_columns():
    bad = CompressedVolume.from_columns(2, 1, [[3.0, 1.0], []])
    with pytest.raises(ValueError):
        bad.validate()
    odd = CompressedVolume(2, 1, np.array([0, 1, 1]), np.array([1.0]))
    with pytest.raises(ValueError):
        odd.validate()


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "voroffset_hip.h")).read()
    declared = set(re.findall(r"\b(vo_[a-z_0-9]+)\s*\(", header))
    assert declared, "no declarations found"
    from voroffset_amd import _native
    lib = ctypes.CDLL(_native.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(_native.EXPORTS)


def test_batch_specs_ranges():
    specs = synth.batch_specs(256)
    sides = np.array([s["side"] for s in specs])
    radii = np.array([s["radius"] for s in specs])
    assert sides.min() >= 64 and sides.max() <= 512
    assert radii.min() >= 2 and radii.max() <= 32
    assert {s["op"] for s in specs} == set(synth.OPS)
