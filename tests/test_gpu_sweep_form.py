"""The alternative kernel forms against the oracle, through the C ABI: the sweep
form of pass 2 (used whole with VO_PASS2=sweep, and for the tiles the gather
form hands back) and the culling and painter forms of pass 1 (VO_PASS1=cull,
VO_PASS1=paint); the default forms are the z-sweep pass 1 and the gather pass 2."""

import os

import pytest

import oracle
from voroffset_amd import synth

from .helpers import compare_csr

pytestmark = pytest.mark.gpu


def _context_with(**env):
    from voroffset_amd.morpho import VoronoiMorphoHip
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return VoronoiMorphoHip("fp64")
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


FORMS = {"sweep-pass2": {"VO_PASS2": "sweep"}, "cull-pass1": {"VO_PASS1": "cull"},
         "paint-pass1": {"VO_PASS1": "paint"}}


@pytest.fixture(scope="module", params=sorted(FORMS))
def hip_sweep(request):
    env = FORMS[request.param]
    m = _context_with(**env)
    yield m
    m.close()


@pytest.mark.parametrize("op", ["dilation", "erosion", "closing", "opening"])
@pytest.mark.parametrize("r", [1.5, 4.0, 9.0])
def test_sweep_form_random(hip_sweep, op, r):
    vol = synth.poisson_columns(36, 29, lam=5, zmax=400, seed=int(r * 7), z_margin=12)
    res, _, _ = hip_sweep.apply(op, vol, r)
    bad, err = compare_csr(res.off, res.z, *oracle.morph(vol, op, r))
    assert len(bad) == 0 and err < 1e-9, (bad[:5], err)


def test_gather_hands_back_overfull_tiles(hip_fp64):
    """A column with more output intervals than a gather list holds makes its
    tile fall back to the sweep form; the result must not change."""
    cols = [[] for _ in range(70 * 3)]
    cols[5 + 70] = [float(k) for k in range(0, 160, 2)]    # 40 separate segments
    cols[66 + 140] = [float(k) for k in range(0, 90, 3)]
    from voroffset_amd.volume import CompressedVolume
    vol = CompressedVolume.from_columns(70, 3, cols, origin=(0.0, 0.0, 0.0), extent=(70.0, 3.0, 200.0))
    for r in (0.4, 0.9, 2.0):
        res, _, _ = hip_fp64.dilation(vol, r)
        bad, err = compare_csr(res.off, res.z, *oracle.morph(vol, "dilation", r))
        assert len(bad) == 0 and err < 1e-9, (r, bad[:5], err)
