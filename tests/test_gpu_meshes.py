"""End to end on meshes, through the C ABI: the GPU rasterizer feeds the GPU
morphology, checked against the oracle on the same dexel volume.

  * 32 icospheres (subdivision 3, centres U[0,1]^3, radii U[0.05, 0.2]),
    128 dexels along the longest axis, padding 8, dilation r = 4;
  * 2000 random 8-sided prisms in a 512-dexel box, voxel size 1, padding 16:
    dilation r = 10, erosion r = 3, opening and closing r = 5.
FP64 storage: counts identical, endpoints within 1e-9 dexel; FP32 storage:
the bar of the direct-volume tests (1e-4 dexel, gaps < 1e-5 counted).
"""

import numpy as np
import pytest

import oracle
from voroffset_amd import mesh as M

from .helpers import compare_csr, narrow_gap_mismatches, quantize_fp32

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spheres_volume():
    mesh = M.icosphere_union(32, level=3, seed=1)
    vol = M.create_dexels(mesh, 1.0, 8, 128)
    assert vol.nx == 144 or vol.ny == 144
    return vol


@pytest.fixture(scope="module")
def tubes_volume():
    mesh = M.prism_tubes(2000, box=512.0, seed=2)
    vol = M.create_dexels(mesh, 1.0, 16, -1)
    assert max(vol.nx, vol.ny) <= 544
    return vol


def _even(vol):
    assert np.all(np.diff(vol.off) % 2 == 0), "rasterized columns must pair up"


def test_spheres_dilation_fp64(hip_fp64, spheres_volume):
    vol = spheres_volume
    _even(vol)
    res, t1, t2 = hip_fp64.dilation(vol, 4.0)
    ref = oracle.morph(vol, "dilation", 4.0, threads=8)
    bad, err = compare_csr(res.off, res.z, *ref)
    assert len(bad) == 0 and err <= 1e-9, (bad[:8], err)


def test_spheres_dilation_fp32(hip_fp32, spheres_volume):
    vol = spheres_volume
    res, _, _ = hip_fp32.dilation(vol, 4.0)
    ref = oracle.morph(quantize_fp32(vol), "dilation", 4.0, threads=8)
    bad, err = compare_csr(res.off, res.z, *ref)
    assert narrow_gap_mismatches(res.off, res.z, ref[0], ref[1], bad) and err <= 1e-4, (bad[:8], err)


@pytest.mark.parametrize("op,r", [("dilation", 10.0), ("erosion", 3.0), ("opening", 5.0), ("closing", 5.0)])
def test_tubes_ops_fp64(hip_fp64, tubes_volume, op, r):
    vol = tubes_volume
    _even(vol)
    res, _, _ = hip_fp64.apply(op, vol, r)
    ref = oracle.morph(vol, op, r, threads=16)
    bad, err = compare_csr(res.off, res.z, *ref)
    assert len(bad) == 0 and err <= 1e-9, (op, bad[:8], err)
